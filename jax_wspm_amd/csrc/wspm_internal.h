// Internal declarations shared by the host pattern builder and the HIP code.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wspm.h"

namespace wspm {

// Rows per assembly workgroup (one thread per row).
constexpr int ASM_BLOCK = 256;

struct HostPattern {
    std::vector<int32_t> row_ptr;    // n+1
    std::vector<int32_t> cols;       // nnz
    std::vector<int32_t> diag_pos;   // n
    std::vector<int32_t> inc_ptr;    // n+1
    std::vector<int32_t> inc_code;   // (element << 2) | local vertex
    std::vector<uint32_t> inc_slot;  // byte c = in-row offset of vertex c
    int max_row = 0;
    int max_block_nnz = 0;
};

int build_pattern(int dim, int n, int E, const int32_t* conn, HostPattern& P, std::string& err);

}  // namespace wspm
