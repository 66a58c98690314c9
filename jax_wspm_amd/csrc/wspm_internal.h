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
    std::vector<int32_t> conn;       // the element vertices (2D meshes; the fan tables are built from them)
    int max_row = 0;
};

// Fan order of the rows of a 2D mesh (the ELL operator of the time-step
// solves and its assembly kernel): slot 0 is the node itself, slots 1..F
// its neighbours in their cyclic order around it, so that fan triangle u
// joins the node to the neighbours in slots u and u + 1.  A closed fan (an
// interior node, F triangles) also joins the last neighbour to the first:
// its triangle 0 takes the last neighbour from pred.  An open fan (a
// boundary node, F - 1 triangles) has no triangle 0.  Tables in ELL layout
// (entry k of row i at k * n + i): col [W][n] (padding: the node itself),
// inc [iw][n] = element << 2 | the node's local vertex of fan triangle u
// (padding 0), pred [n], info [n] = F | closed << 4 | (local vertex of the
// node in fan triangle u) << (5 + 2u).  Returns false (no fan
// order) unless every row is one fan of consistently oriented triangles
// whose nodes are exactly the row's columns, with at most W - 1 neighbours.
bool build_fan(const HostPattern& P, int W, std::vector<int32_t>& col, std::vector<int32_t>& inc,
               std::vector<int32_t>& pred, std::vector<uint32_t>& info, int& iw);

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
