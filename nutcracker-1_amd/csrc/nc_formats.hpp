// Operator and state messages of the reference's protobuf schema
// (sources/common/nutcracker.proto:10-35), as plain structs with a hand-written
// wire codec (formats.cpp).
#pragma once

#include "nc_internal.hpp"

#include <string>
#include <vector>

namespace nc {
namespace pb {

// Normalization enum of the schema
enum { MIDDLE = 0, LEFT = 1, RIGHT = 2 };

struct OperatorSiteBuffer {
    uint32_t number_of_matrices = 0, physical_dimension = 0, left_dimension = 0, right_dimension = 0;
    std::vector<double> matrix_data;   // 2 * n * d * d: the blocks W_n(a, b) in storage order, re then im
    std::vector<uint32_t> index_data;  // 2 * n: (from, to) per block, 1-based
};
struct OperatorBuffer {
    std::vector<OperatorSiteBuffer> sites;  // the distinct sites
    std::vector<uint32_t> sequence;         // per chain position: 0-based index into sites
};
struct StateSiteBuffer {
    int normalization = MIDDLE;
    uint32_t physical_dimension = 0, left_dimension = 0, right_dimension = 0;
    std::vector<double> data;  // 2 * d * bl * br: the site (br, bl, d) in storage order, re then im
};
struct StateBuffer {
    std::vector<StateSiteBuffer> sites;
};

// decoders validate required fields and lengths (errors -90 .. -92)
OperatorBuffer decode_operator(const void* data, size_t size);
std::string encode_operator(OperatorBuffer const& ob);
StateBuffer decode_state(const void* data, size_t size);
std::string encode_state(StateBuffer const& sb);

}  // namespace pb
}  // namespace nc
