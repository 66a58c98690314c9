// Host-side construction of the fixed CSR pattern and the row-owned
// incidence map used by the deterministic assembly kernel.
//
// The reference rebuilds its sparse structure on every Picard iteration
// (COO fill + BCOO.sum_duplicates, src/numerics/assembly_re_v2.py:119-152).
// Here it is built once per mesh:
//   * CSR pattern: union of the (i, j) pairs of every element, columns sorted;
//   * diag_pos[i]: CSR index of A(i, i);
//   * incidence lists: for each node i, the (element, local vertex) pairs
//     that touch it, in increasing element order, plus for each such pair the
//     in-row offsets of the element's vertices (one byte each, packed in a
//     uint32).  The assembly kernel walks a row's incidences in this fixed
//     order, so every CSR value is a sum in a fixed order: no atomics, and
//     bitwise-repeatable results.
#include "wspm_internal.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace wspm {

int build_pattern(int dim, int n, int E, const int32_t* conn, HostPattern& P, std::string& err) {
    const int K = dim + 1;
    if (dim != 2 && dim != 3) { err = "dim must be 2 or 3"; return WSPM_EINVAL; }
    if (n <= 0 || E <= 0) { err = "empty mesh"; return WSPM_EINVAL; }
    for (int64_t q = 0; q < (int64_t)E * K; ++q) {
        if (conn[q] < 0 || conn[q] >= n) {
            err = "element index out of bounds at entry " + std::to_string(q);
            return WSPM_EINVAL;
        }
    }
    if ((int64_t)E * 4 >= (int64_t)1 << 31) { err = "too many elements"; return WSPM_EINVAL; }

    // node -> incident elements (counting sort, element order preserved)
    P.inc_ptr.assign(n + 1, 0);
    for (int64_t q = 0; q < (int64_t)E * K; ++q) P.inc_ptr[conn[q] + 1]++;
    for (int i = 0; i < n; ++i) P.inc_ptr[i + 1] += P.inc_ptr[i];
    const int64_t ninc = P.inc_ptr[n];
    P.inc_code.resize(ninc);
    {
        std::vector<int32_t> fill(P.inc_ptr.begin(), P.inc_ptr.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int a = 0; a < K; ++a) {
                const int i = conn[(int64_t)e * K + a];
                P.inc_code[fill[i]++] = e * 4 + a;
            }
    }

    // CSR rows: sorted unique neighbours through incident elements
    P.row_ptr.assign(n + 1, 0);
    std::vector<int32_t> cols;
    cols.reserve((size_t)ninc * 2);
    std::vector<int32_t> tmp;
    int max_row = 0;
    for (int i = 0; i < n; ++i) {
        tmp.clear();
        for (int32_t q = P.inc_ptr[i]; q < P.inc_ptr[i + 1]; ++q) {
            const int e = P.inc_code[q] >> 2;
            for (int c = 0; c < K; ++c) tmp.push_back(conn[(int64_t)e * K + c]);
        }
        tmp.push_back(i);  // a node touched by no element still gets its diagonal
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        max_row = std::max<int>(max_row, (int)tmp.size());
        cols.insert(cols.end(), tmp.begin(), tmp.end());
        if ((int64_t)cols.size() >= ((int64_t)1 << 31) - 1) { err = "nnz exceeds int32"; return WSPM_EINVAL; }
        P.row_ptr[i + 1] = (int32_t)cols.size();
    }
    if (max_row > 255) { err = "row longer than 255 entries"; return WSPM_EINVAL; }
    P.cols.swap(cols);
    P.max_row = max_row;

    P.diag_pos.resize(n);
    for (int i = 0; i < n; ++i) {
        const int32_t* b = P.cols.data() + P.row_ptr[i];
        const int32_t* e = P.cols.data() + P.row_ptr[i + 1];
        P.diag_pos[i] = (int32_t)(std::lower_bound(b, e, i) - P.cols.data());
    }

    // in-row offsets of each incident element's vertices
    P.inc_slot.resize(ninc);
    for (int i = 0; i < n; ++i) {
        const int32_t* b = P.cols.data() + P.row_ptr[i];
        const int32_t* e = P.cols.data() + P.row_ptr[i + 1];
        for (int32_t q = P.inc_ptr[i]; q < P.inc_ptr[i + 1]; ++q) {
            const int el = P.inc_code[q] >> 2;
            uint32_t packed = 0;
            for (int c = 0; c < K; ++c) {
                const int j = conn[(int64_t)el * K + c];
                const uint32_t off = (uint32_t)(std::lower_bound(b, e, j) - b);
                packed |= off << (8 * c);
            }
            P.inc_slot[q] = packed;
        }
    }

    if (dim == 2) P.conn.assign(conn, conn + (size_t)E * K);
    return WSPM_OK;
}

bool build_fan(const HostPattern& P, int W, std::vector<int32_t>& col, std::vector<int32_t>& inc,
               std::vector<int32_t>& pred, std::vector<uint32_t>& info, int& iw) {
    const int n = (int)P.row_ptr.size() - 1;
    if (P.conn.empty() || W > 16) return false;
    // planes of the triangle table: triangle indices run to nq - 1 (closed)
    // or nq (open), and are checked against W - 2 per row below
    iw = 0;
    for (int i = 0; i < n; ++i) iw = std::max(iw, P.inc_ptr[i + 1] - P.inc_ptr[i] + 1);
    iw = std::min(iw, W - 1);
    col.assign((size_t)W * n, 0);
    inc.assign((size_t)std::max(iw, 1) * n, 0);
    pred.assign(n, 0);
    info.assign(n, 0);
    int la_[16], a_[16], b_[16], ord[16], nb[17];
    for (int i = 0; i < n; ++i)
        if (P.inc_ptr[i + 1] - P.inc_ptr[i] > 15) return false;
    for (int i = 0; i < n; ++i) {
        const int q0 = P.inc_ptr[i], nq = P.inc_ptr[i + 1] - q0;
        // the link of incidence u: the other two vertices in the element's own
        // cyclic order (node = vertex la, then la + 1, la + 2)
        for (int u = 0; u < nq; ++u) {
            const int code = P.inc_code[q0 + u], e = code >> 2, la = code & 3;
            la_[u] = code;
            a_[u] = P.conn[(size_t)e * 3 + (la + 1) % 3];
            b_[u] = P.conn[(size_t)e * 3 + (la + 2) % 3];
        }
        // every neighbour starts at most one link and ends at most one
        int start = -1, nstart = 0;
        for (int u = 0; u < nq; ++u) {
            int outs = 0, ins = 0;
            for (int v = 0; v < nq; ++v) {
                outs += a_[v] == a_[u];
                ins += b_[v] == a_[u];
                if (v != u && b_[v] == b_[u]) return false;
            }
            if (outs != 1 || ins > 1) return false;
            if (ins == 0) {
                start = u;
                ++nstart;
            }
        }
        if (nstart > 1) return false;
        const bool closed = nq > 0 && nstart == 0;
        if (closed) start = 0;
        // walk the links from the start
        int F = 0;
        if (nq > 0) {
            ord[0] = start;
            for (int k = 1; k < nq; ++k) {
                int nx = -1;
                for (int v = 0; v < nq; ++v)
                    if (a_[v] == b_[ord[k - 1]]) nx = v;
                if (nx < 0 || nx == start) return false;   // more than one fan
                ord[k] = nx;
            }
            if (closed && b_[ord[nq - 1]] != a_[start]) return false;
            for (int k = 0; k < nq; ++k) nb[1 + k] = a_[ord[k]];
            F = closed ? nq : nq + 1;
            if (!closed) nb[F] = b_[ord[nq - 1]];
        }
        if (F + 1 > W || F + 1 != P.row_ptr[i + 1] - P.row_ptr[i]) return false;
        if ((closed ? nq - 1 : nq) >= iw) return false;    // a triangle index beyond the table
        // the fan's nodes are exactly the row's columns
        int srt[17];
        srt[0] = i;
        for (int k = 1; k <= F; ++k) srt[k] = nb[k];
        std::sort(srt, srt + F + 1);
        for (int k = 0; k <= F; ++k)
            if (srt[k] != P.cols[P.row_ptr[i] + k]) return false;
        col[i] = i;
        for (int k = 1; k < W; ++k) col[(size_t)k * n + i] = k <= F ? nb[k] : i;
        // fan triangle u joins slots u and u + 1: closed, triangle 0 is the
        // last link (last neighbour -> first) and triangle u the u-th; open,
        // triangle u (1 <= u < F) is the u-th link
        uint32_t las = 0;
        for (int u = 0; u < nq; ++u) {
            const int slot = closed ? (u + 1) % nq : u + 1;   // triangle index of link ord[u]
            inc[(size_t)slot * n + i] = la_[ord[u]];
            las |= (uint32_t)(la_[ord[u]] & 3) << (5 + 2 * slot);
        }
        pred[i] = closed ? nb[F] : i;
        info[i] = (uint32_t)F | (closed ? 16u : 0u) | las;
    }
    return true;
}

void block_maxima(const HostPattern& P, int R, int& max_nnz, int& max_inc) {
    const int n = (int)P.row_ptr.size() - 1;
    max_nnz = 0;
    max_inc = 0;
    for (int r0 = 0; r0 < n; r0 += R) {
        const int r1 = std::min(n, r0 + R);
        max_nnz = std::max<int>(max_nnz, P.row_ptr[r1] - P.row_ptr[r0]);
        max_inc = std::max<int>(max_inc, P.inc_ptr[r1] - P.inc_ptr[r0]);
    }
}

namespace {

// Breadth-first order from s over nodes not yet done (neighbours by
// increasing degree); levels = number of levels, last = start of the last.
std::vector<int32_t> bfs(const HostPattern& P, const std::vector<int32_t>& deg, int s,
                         const std::vector<uint8_t>& done, int& levels, size_t& last) {
    std::vector<int32_t> order{s};
    std::vector<uint8_t> seen(done);
    seen[s] = 1;
    size_t lo = 0;
    levels = 0;
    last = 0;
    std::vector<int32_t> nb;
    while (lo < order.size()) {
        const size_t hi = order.size();
        last = lo;
        ++levels;
        for (size_t q = lo; q < hi; ++q) {
            const int i = order[q];
            nb.clear();
            for (int k = P.row_ptr[i]; k < P.row_ptr[i + 1]; ++k) {
                const int j = P.cols[k];
                if (!seen[j]) {
                    seen[j] = 1;
                    nb.push_back(j);
                }
            }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
            order.insert(order.end(), nb.begin(), nb.end());
        }
        lo = hi;
    }
    return order;
}

}  // namespace

std::vector<int32_t> rcm_order(const HostPattern& P) {
    const int n = (int)P.row_ptr.size() - 1;
    std::vector<int32_t> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = P.row_ptr[i + 1] - P.row_ptr[i];
    std::vector<uint8_t> done(n, 0);
    std::vector<int32_t> cm;
    cm.reserve(n);
    for (int s0 = 0; s0 < n; ++s0) {
        if (done[s0]) continue;
        int levels = 0;
        size_t last = 0;
        std::vector<int32_t> order = bfs(P, deg, s0, done, levels, last);
        int s = s0;                               // least degree in the component
        for (int v : order)
            if (deg[v] < deg[s]) s = v;
        order = bfs(P, deg, s, done, levels, last);
        // George-Liu: restart from a least-degree node of the last level while
        // that deepens the level structure
        for (int rep = 0; rep < 8; ++rep) {
            int c = order[last];
            for (size_t q = last; q < order.size(); ++q)
                if (deg[order[q]] < deg[c]) c = order[q];
            int lv2 = 0;
            size_t last2 = 0;
            std::vector<int32_t> o2 = bfs(P, deg, c, done, lv2, last2);
            if (lv2 <= levels) break;
            order.swap(o2);
            levels = lv2;
            last = last2;
        }
        for (int v : order) {
            done[v] = 1;
            cm.push_back(v);
        }
    }
    std::vector<int32_t> perm(n);
    for (int q = 0; q < n; ++q) perm[cm[n - 1 - q]] = q;   // reversed
    return perm;
}

void bandwidth(const HostPattern& P, const std::vector<int32_t>& perm, int& kl, int& ku) {
    kl = ku = 0;
    const int n = (int)P.row_ptr.size() - 1;
    for (int i = 0; i < n; ++i) {
        const int r = perm.empty() ? i : perm[i];
        for (int k = P.row_ptr[i]; k < P.row_ptr[i + 1]; ++k) {
            const int c = perm.empty() ? P.cols[k] : perm[P.cols[k]];
            kl = std::max(kl, r - c);
            ku = std::max(ku, c - r);
        }
    }
}

}  // namespace wspm
