// Internal declarations shared by the host pattern builder and the HIP code.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wspm.h"

namespace wspm {

struct HostPattern {
    std::vector<int32_t> row_ptr;    // n+1
    std::vector<int32_t> cols;       // nnz
    std::vector<int32_t> diag_pos;   // n
    std::vector<int32_t> inc_ptr;    // n+1
    std::vector<int32_t> inc_code;   // (element << 2) | local vertex
    std::vector<uint32_t> inc_slot;  // byte c = in-row offset of vertex c
    int max_row = 0;
};

int build_pattern(int dim, int n, int E, const int32_t* conn, HostPattern& P, std::string& err);
// Largest CSR-value and incidence counts spanned by R consecutive rows.
void block_maxima(const HostPattern& P, int R, int& max_nnz, int& max_inc);
// Reverse Cuthill-McKee ordering of the pattern's graph: perm[i] = new index
// of node i.  Each connected component starts from a pseudo-peripheral node
// of least degree; neighbours are visited by increasing degree.
std::vector<int32_t> rcm_order(const HostPattern& P);
// Lower and upper bandwidth of the pattern under perm (identity when empty).
void bandwidth(const HostPattern& P, const std::vector<int32_t>& perm, int& kl, int& ku);

}  // namespace wspm
