// The reference's protobuf messages for operators and states
// (sources/common/nutcracker.proto:10-35), encoded and decoded by hand in the
// protobuf wire format (base-128 varints; length-delimited sub-messages and
// packed repeated fields), so the library needs no protobuf runtime.  Field
// numbers and meanings follow the .proto file; the element orders follow
// protobuf.cpp:53-93 (matrix_data / data: the tensor's complex elements in
// storage order, real then imaginary part; index_data: the (from, to) pairs,
// 1-based) and io.cpp:26-50 (sequence: 0-based positions in the list of
// distinct sites).
#include "nc_formats.hpp"

#include <cstring>

namespace nc {
namespace pb {

namespace {

enum Wire { VARINT = 0, FIXED64 = 1, LEN = 2, FIXED32 = 5 };

struct Reader {
    const uint8_t* p;
    const uint8_t* end;
    bool done() const { return p >= end; }
    uint64_t varint() {
        uint64_t v = 0;
        for (int shift = 0; shift < 64; shift += 7) {
            if (p >= end) throw Error("protobuf: truncated varint", -90);
            const uint8_t b = *p++;
            v |= (uint64_t)(b & 0x7f) << shift;
            if (!(b & 0x80)) return v;
        }
        throw Error("protobuf: varint longer than 10 bytes", -90);
    }
    double fixed64() {
        if (end - p < 8) throw Error("protobuf: truncated 64-bit field", -90);
        double v;
        std::memcpy(&v, p, 8);  // little-endian on the host
        p += 8;
        return v;
    }
    Reader sub() {
        const uint64_t n = varint();
        if ((uint64_t)(end - p) < n) throw Error("protobuf: truncated length-delimited field", -90);
        Reader r{p, p + n};
        p += n;
        return r;
    }
    void skip(int wire) {
        switch (wire) {
            case VARINT: varint(); break;
            case FIXED64:
                if (end - p < 8) throw Error("protobuf: truncated field", -90);
                p += 8;
                break;
            case LEN: sub(); break;
            case FIXED32:
                if (end - p < 4) throw Error("protobuf: truncated field", -90);
                p += 4;
                break;
            default: throw Error("protobuf: unsupported wire type", -90);
        }
    }
};

// repeated double, packed (LEN) or one element per field (FIXED64)
void read_doubles(Reader& r, int wire, std::vector<double>& out) {
    if (wire == LEN) {
        Reader s = r.sub();
        if ((s.end - s.p) % 8) throw Error("protobuf: packed doubles of odd length", -90);
        while (!s.done()) out.push_back(s.fixed64());
    } else if (wire == FIXED64) {
        out.push_back(r.fixed64());
    } else {
        throw Error("protobuf: bad wire type for a double field", -90);
    }
}

// repeated uint32, packed (LEN) or one element per field (VARINT)
void read_u32s(Reader& r, int wire, std::vector<uint32_t>& out) {
    if (wire == LEN) {
        Reader s = r.sub();
        while (!s.done()) out.push_back((uint32_t)s.varint());
    } else if (wire == VARINT) {
        out.push_back((uint32_t)r.varint());
    } else {
        throw Error("protobuf: bad wire type for a uint32 field", -90);
    }
}

uint32_t read_u32(Reader& r, int wire) {
    if (wire != VARINT) throw Error("protobuf: bad wire type for a uint32 field", -90);
    return (uint32_t)r.varint();
}

struct Writer {
    std::string out;
    void varint(uint64_t v) {
        while (v >= 0x80) {
            out.push_back((char)(uint8_t)((v & 0x7f) | 0x80));
            v >>= 7;
        }
        out.push_back((char)(uint8_t)v);
    }
    void tag(int field, int wire) { varint(((uint64_t)field << 3) | (uint64_t)wire); }
    void u32(int field, uint32_t v) {
        tag(field, VARINT);
        varint(v);
    }
    void bytes(int field, std::string const& b) {
        tag(field, LEN);
        varint(b.size());
        out += b;
    }
    void packed_doubles(int field, std::vector<double> const& v) {
        if (v.empty()) return;
        tag(field, LEN);
        varint(v.size() * 8);
        const size_t at = out.size();
        out.resize(at + v.size() * 8);
        std::memcpy(&out[at], v.data(), v.size() * 8);
    }
    void packed_u32s(int field, std::vector<uint32_t> const& v) {
        if (v.empty()) return;
        Writer w;
        for (uint32_t x : v) w.varint(x);
        bytes(field, w.out);
    }
};

template <typename F>
void fields(Reader r, F&& f) {
    while (!r.done()) {
        const uint64_t key = r.varint();
        f((int)(key >> 3), (int)(key & 7), r);
    }
}

}  // namespace

OperatorBuffer decode_operator(const void* data, size_t size) {
    OperatorBuffer ob;
    Reader r{static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + size};
    fields(r, [&](int field, int wire, Reader& rr) {
        if (field == 1 && wire == LEN) {
            Reader s = rr.sub();
            OperatorSiteBuffer site;
            unsigned seen = 0;
            fields(s, [&](int f, int w, Reader& q) {
                switch (f) {
                    case 1: site.number_of_matrices = read_u32(q, w); seen |= 1; break;
                    case 2: site.physical_dimension = read_u32(q, w); seen |= 2; break;
                    case 3: site.left_dimension = read_u32(q, w); seen |= 4; break;
                    case 4: site.right_dimension = read_u32(q, w); seen |= 8; break;
                    case 5: read_doubles(q, w, site.matrix_data); break;
                    case 6: read_u32s(q, w, site.index_data); break;
                    default: q.skip(w);
                }
            });
            if (seen != 15) throw Error("protobuf: OperatorSiteBuffer without a required field", -91);
            const size_t d = site.physical_dimension, n = site.number_of_matrices;
            if (site.matrix_data.size() != 2 * n * d * d)
                throw Error("protobuf: OperatorSiteBuffer matrix_data has the wrong length", -91);
            if (site.index_data.size() != 2 * n)
                throw Error("protobuf: OperatorSiteBuffer index_data has the wrong length", -91);
            ob.sites.push_back(std::move(site));
        } else if (field == 2) {
            read_u32s(rr, wire, ob.sequence);
        } else {
            rr.skip(wire);
        }
    });
    return ob;
}

std::string encode_operator(OperatorBuffer const& ob) {
    Writer w;
    for (auto const& site : ob.sites) {
        Writer s;
        s.u32(1, site.number_of_matrices);
        s.u32(2, site.physical_dimension);
        s.u32(3, site.left_dimension);
        s.u32(4, site.right_dimension);
        s.packed_doubles(5, site.matrix_data);
        s.packed_u32s(6, site.index_data);
        w.bytes(1, s.out);
    }
    w.packed_u32s(2, ob.sequence);
    return w.out;
}

StateBuffer decode_state(const void* data, size_t size) {
    StateBuffer sb;
    Reader r{static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + size};
    fields(r, [&](int field, int wire, Reader& rr) {
        if (field == 1 && wire == LEN) {
            Reader s = rr.sub();
            StateSiteBuffer site;
            unsigned seen = 0;
            fields(s, [&](int f, int w, Reader& q) {
                switch (f) {
                    case 1: site.normalization = (int)read_u32(q, w); break;
                    case 2: site.physical_dimension = read_u32(q, w); seen |= 1; break;
                    case 3: site.left_dimension = read_u32(q, w); seen |= 2; break;
                    case 4: site.right_dimension = read_u32(q, w); seen |= 4; break;
                    case 5: read_doubles(q, w, site.data); break;
                    default: q.skip(w);
                }
            });
            if (seen != 7) throw Error("protobuf: StateSiteBuffer without a required field", -92);
            const size_t n = (size_t)site.physical_dimension * site.left_dimension * site.right_dimension;
            if (site.data.size() != 2 * n) throw Error("protobuf: StateSiteBuffer data has the wrong length", -92);
            sb.sites.push_back(std::move(site));
        } else {
            rr.skip(wire);
        }
    });
    return sb;
}

std::string encode_state(StateBuffer const& sb) {
    Writer w;
    for (auto const& site : sb.sites) {
        Writer s;
        s.u32(1, (uint32_t)site.normalization);
        s.u32(2, site.physical_dimension);
        s.u32(3, site.left_dimension);
        s.u32(4, site.right_dimension);
        s.packed_doubles(5, site.data);
        w.bytes(1, s.out);
    }
    return w.out;
}

}  // namespace pb
}  // namespace nc
