// The reference's C API front door (sources/c-interface/*.cpp) for the parts
// that sit on the hot path: operators and states in the reference's protobuf
// wire format, the simple least-eigenvalue solvers (operator.cpp:23-52) on the
// device-resident chain, state expectation values and overlaps (state.cpp:9-14),
// serialization objects (serialization.cpp) and the thread-local error string
// (errors.cpp:12-24).  Errors inside a call set the error string and return
// the reference's failure value (NULL, or the outputs left untouched).
#include "nc_chain.hpp"
#include "nc_formats.hpp"

#include "../../include/nutcracker_c_api.h"

#include <cstring>
#include <exception>
#include <memory>
#include <string>

struct NutcrackerOperator {
    nc::pb::OperatorBuffer buffer;  // the distinct sites and the sequence, as on the wire
};
struct NutcrackerState {
    nc::pb::StateBuffer buffer;  // first site MIDDLE, the rest RIGHT-normalised
};
struct NutcrackerSerialization {
    std::string bytes;
};

namespace {

thread_local std::unique_ptr<std::string> g_error;

void set_error(const char* m) { g_error.reset(new std::string(m)); }

template <typename R, typename F>
R guarded(R fail, F&& f) {
    try {
        return f();
    } catch (std::exception const& e) {
        set_error(e.what());
        return fail;
    }
}

using nc::pb::OperatorBuffer;
using nc::pb::StateBuffer;

// host checks of an operator message: sequence entries, block indices, bonds
void validate_operator(OperatorBuffer const& ob) {
    if (ob.sequence.empty()) throw nc::Error("operator: empty sequence", -93);
    for (auto const& s : ob.sites) {
        if (!s.physical_dimension || !s.left_dimension || !s.right_dimension || !s.number_of_matrices)
            throw nc::Error("operator site: bad dimensions", -2);
        for (uint32_t n = 0; n < s.number_of_matrices; ++n)
            if (s.index_data[2 * n] < 1 || s.index_data[2 * n] > s.left_dimension || s.index_data[2 * n + 1] < 1 ||
                s.index_data[2 * n + 1] > s.right_dimension)
                throw nc::Error("operator site: block index out of range", -3);
    }
    for (uint32_t id : ob.sequence)
        if (id >= ob.sites.size())
            throw nc::Error("operator: no such operator site number " + std::to_string(id + 1), -93);
    for (size_t i = 0; i + 1 < ob.sequence.size(); ++i)
        if (ob.sites[ob.sequence[i]].right_dimension != ob.sites[ob.sequence[i + 1]].left_dimension)
            throw nc::Error("operator bond mismatch between sites", -65);
    if (ob.sites[ob.sequence.front()].left_dimension != 1 || ob.sites[ob.sequence.back()].right_dimension != 1)
        throw nc::Error("the chain's outer operator bonds must be 1", -66);
}

// the chain's operator sites on the device, one per position of the sequence
std::vector<std::unique_ptr<nc::MpoSite>> expand_operator(OperatorBuffer const& ob) {
    validate_operator(ob);
    std::vector<std::unique_ptr<nc::MpoSite>> ops;
    for (uint32_t id : ob.sequence) {
        auto const& s = ob.sites[id];
        auto op = std::make_unique<nc::MpoSite>();
        op->build((int)s.physical_dimension, (int)s.left_dimension, (int)s.right_dimension,
                  (int)s.number_of_matrices, s.index_data.data(), reinterpret_cast<const nc::zd*>(s.matrix_data.data()));
        ops.push_back(std::move(op));
    }
    return ops;
}

// a state's sites uploaded to the device
struct DeviceState {
    std::vector<int> d, bonds;
    std::vector<nc::ZBuf> sites;
    std::vector<const double2*> ptrs;
    explicit DeviceState(StateBuffer const& sb) {
        if (sb.sites.empty()) throw nc::Error("state: no sites", -94);
        bonds.push_back((int)sb.sites[0].left_dimension);
        for (size_t i = 0; i < sb.sites.size(); ++i) {
            auto const& s = sb.sites[i];
            if ((int)s.left_dimension != bonds.back()) throw nc::Error("state: bond mismatch between sites", -94);
            d.push_back((int)s.physical_dimension);
            bonds.push_back((int)s.right_dimension);
            sites.emplace_back(s.data.size() / 2);
            nc::h2d(sites.back().ptr, s.data.data(), s.data.size() * sizeof(double));
            ptrs.push_back(sites.back().ptr);
        }
    }
};

// the chain's current state in the reference's canonical form (the chain is at
// its first site: that site MIDDLE, the others RIGHT-normalised)
StateBuffer capture_state(nc::Chain const& c) {
    StateBuffer sb;
    for (int i = 0; i < c.N; ++i) {
        nc::pb::StateSiteBuffer s;
        s.normalization = i == 0 ? nc::pb::MIDDLE : nc::pb::RIGHT;
        s.physical_dimension = (uint32_t)c.ops[i]->d;
        s.left_dimension = (uint32_t)c.bonds[i];
        s.right_dimension = (uint32_t)c.bonds[i + 1];
        s.data.resize(2 * (size_t)c.ops[i]->d * c.bonds[i] * c.bonds[i + 1]);
        nc::d2h(s.data.data(), c.sites[i].ptr, s.data.size() * sizeof(double));
        sb.sites.push_back(std::move(s));
    }
    return sb;
}

// the chain of operator.cpp:23-52: least values, initial bond dimension 2 grown
// by one until the energy settles, thresholds 1e-12 / 1e-9 / 1e-8 / 1e-8
std::unique_ptr<nc::Chain> simple_chain(OperatorBuffer const& ob) {
    auto ops = expand_operator(ob);
    std::vector<int> dims;
    for (auto const& o : ops) dims.push_back(o->d);
    const int chi = std::min(2, nc::maximum_bandwidth(dims));
    auto c = std::make_unique<nc::Chain>((int)ops.size(), std::move(ops), nc::bandwidth_sequence(chi, dims));
    auto& o = c->opts;
    o.which[0] = 'S';
    o.which[1] = 'R';
    o.sanity = 1e-12;
    o.site_tol = 1e-9;
    o.sweep_tol = 1e-8;
    o.chain_tol = 1e-8;
    o.grow = true;
    c->randomize(0);
    c->reset();
    return c;
}

}  // namespace

extern "C" {

void Nutcracker_clearError() { g_error.reset(); }
char const* Nutcracker_getError() { return g_error ? g_error->c_str() : nullptr; }

void Nutcracker_Serialization_free(NutcrackerSerialization* s) { delete s; }
uint32_t Nutcracker_Serialization_getSize(NutcrackerSerialization* s) { return (uint32_t)s->bytes.size(); }
void Nutcracker_Serialization_write(NutcrackerSerialization* s, void* buffer) {
    std::memcpy(buffer, s->bytes.data(), s->bytes.size());
}

NutcrackerOperator* Nutcracker_Operator_deserialize(uint32_t size, void const* data) {
    return guarded<NutcrackerOperator*>(nullptr, [&] {
        OperatorBuffer ob = nc::pb::decode_operator(data, size);
        validate_operator(ob);
        return new NutcrackerOperator{std::move(ob)};
    });
}
void Nutcracker_Operator_free(NutcrackerOperator* op) { delete op; }
NutcrackerSerialization* Nutcracker_Operator_serialize(NutcrackerOperator const* op) {
    return guarded<NutcrackerSerialization*>(nullptr, [&] {
        return new NutcrackerSerialization{nc::pb::encode_operator(op->buffer)};
    });
}

void Nutcracker_Operator_simpleSolveForLeastEigenvalues(NutcrackerOperator const* op, uint32_t number_of_levels,
                                                        float* eigenvalues) {
    guarded<int>(0, [&] {
        auto c = simple_chain(op->buffer);
        const std::vector<double> e = c->solve_for_multiple_levels((int)number_of_levels, 1);
        for (size_t k = 0; k < e.size(); ++k) eigenvalues[k] = (float)e[k];
        return 0;
    });
}

void Nutcracker_Operator_simpleSolveForLeastEigenvaluesWithEigenvectors(NutcrackerOperator const* op,
                                                                        uint32_t number_of_levels,
                                                                        float* eigenvalues,
                                                                        NutcrackerState** eigenvectors) {
    guarded<int>(0, [&] {
        auto c = simple_chain(op->buffer);
        std::vector<StateBuffer> states;
        const std::vector<double> e = c->solve_for_multiple_levels(
            (int)number_of_levels, 1, [&](int) { states.push_back(capture_state(*c)); });
        for (size_t k = 0; k < e.size(); ++k) {
            eigenvalues[k] = (float)e[k];
            eigenvectors[k] = new NutcrackerState{std::move(states[k])};
        }
        return 0;
    });
}

NutcrackerState* Nutcracker_State_deserialize(uint32_t size, void const* data) {
    return guarded<NutcrackerState*>(nullptr, [&] {
        StateBuffer sb = nc::pb::decode_state(data, size);
        if (sb.sites.empty()) throw nc::Error("state: no sites", -94);
        for (size_t i = 1; i < sb.sites.size(); ++i)
            if (sb.sites[i].left_dimension != sb.sites[i - 1].right_dimension)
                throw nc::Error("state: bond mismatch between sites", -94);
        return new NutcrackerState{std::move(sb)};
    });
}
void Nutcracker_State_free(NutcrackerState* state) { delete state; }
NutcrackerSerialization* Nutcracker_State_serialize(NutcrackerState const* state) {
    return guarded<NutcrackerSerialization*>(nullptr, [&] {
        return new NutcrackerSerialization{nc::pb::encode_state(state->buffer)};
    });
}

void Nutcracker_State_computeExpectation(NutcrackerState const* state, NutcrackerOperator const* op,
                                         nc_zcomplex* result) {
    guarded<int>(0, [&] {
        auto ops = expand_operator(op->buffer);
        DeviceState ds(state->buffer);
        if (ops.size() != ds.sites.size()) throw nc::Error("expectation: state and operator lengths differ", -95);
        std::vector<nc::MpoSite const*> po;
        for (size_t i = 0; i < ops.size(); ++i) {
            if (ops[i]->d != ds.d[i]) throw nc::Error("expectation: physical dimensions differ", -95);
            po.push_back(ops[i].get());
        }
        const nc::zd v = nc::mps_expectation(po, ds.bonds, ds.ptrs);
        result->re = v.real();
        result->im = v.imag();
        return 0;
    });
}

void Nutcracker_State_computeOverlap(NutcrackerState const* state1, NutcrackerState const* state2,
                                     nc_zcomplex* result) {
    guarded<int>(0, [&] {
        DeviceState a(state1->buffer), b(state2->buffer);
        if (a.sites.size() != b.sites.size()) throw nc::Error("overlap: states of different lengths", -73);
        if (a.d != b.d) throw nc::Error("overlap: physical dimensions differ", -74);
        const nc::zd v = nc::mps_overlap((int)a.sites.size(), a.d, a.bonds, a.ptrs, b.bonds, b.ptrs);
        result->re = v.real();
        result->im = v.imag();
        return 0;
    });
}

}  // extern "C"
