// Fused P1 assembly of the mixed-form Richards system (gfx950).
//
// Reference: src/numerics/assembly_re_v2.py:13-166 (triangles) and
// src/numerics/assembly_re_3D.py:12-170 (tetrahedra).  Per element the
// reference loops over the quadrature points, inverts J at each, builds
//   stiff = sum_q w (N.K) G^T G detJ,  mass = sum_q w N N^T detJ (row-sum lumped),
//   src   = -sum_q w (N.K) G[dim-1] detJ,
//   A_e   = M_L diag(C)/dt + stiff,  b_e = src - M_L (theta_m - theta_n - C h)/dt,
// then scatters 9E (16E) COO entries and sums duplicates.
//
// Here a workgroup owns R consecutive rows (nodes).  Their incident
// (element, local vertex) pairs are contiguous in the precomputed incidence
// list, so:
//   1. every thread takes incidences of the workgroup (many independent
//      gathers in flight), recomputes the element geometry in closed form (J is
//      constant on a P1 element, one inverse instead of one per quadrature
//      point) and writes that element's row contribution (dim+1 stiffness
//      values, the gravity source and the lumped mass) to LDS;
//   2. one thread per CSR entry of the rows sums, in the fixed incidence
//      order, the contributions that land in its slot (identity rows for
//      Dirichlet nodes), writes it to HBM (coalesced) and to an LDS copy of
//      the rows' window; the diagonal's thread also sums the source term and
//      takes the Jacobi inverse diagonal;
//   3. optionally, LANES = 256/R threads per row fuse the first step of the
//      Krylov solve: r0 = b - A x0 and its dot products (the SpMV the solver
//      would otherwise start with, on values that are still in LDS).
// No atomics, no COO staging, fixed summation order: bitwise repeatable.
#include "assemble.h"
#include "wspm_device.h"
#include "wspm_internal.h"

namespace wspm {

// Row contribution of one incidence (element e seen from its vertex la):
//   out[c] (c <= dim)  stiffness K_e(la, c) = Kbar detJ G(:,la).G(:,c)
//   out[dim+1]         gravity source -Kbar detJ dN_la/d(last coordinate)
//   out[dim+2]         lumped mass detJ * ml[la]
// The mass term M_L C/dt on the diagonal and -M_L (theta_m - theta_n - C h)/dt
// in the source are linear in the row node's values: they are applied once per
// row with the summed lumped mass.
template <int D>
__device__ inline void element_compute(const AsmArgs& a, const double (&px)[D + 1], const double (&py)[D + 1],
                                       const double (&pz)[D + 1], const double (&kn)[D + 1], double ks, int la,
                                       double (&out)[D + 3]) {
    constexpr int K = D + 1;
    double G[D][K];
    double det;
    if constexpr (D == 2) {
        const double J00 = px[1] - px[0], J01 = py[1] - py[0];
        const double J10 = px[2] - px[0], J11 = py[2] - py[0];
        det = J00 * J11 - J01 * J10;
        const double id = 1.0 / det;
        G[0][1] = J11 * id;  G[1][1] = -J10 * id;
        G[0][2] = -J01 * id; G[1][2] = J00 * id;
    } else {
        double r[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r[c][0] = px[c + 1] - px[0];
            r[c][1] = py[c + 1] - py[0];
            r[c][2] = pz[c + 1] - pz[0];
        }
        // J rows r1, r2, r3; J^{-1} columns = (r2 x r3, r3 x r1, r1 x r2) / det
        double cx[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double* u = r[(c + 1) % 3];
            const double* v = r[(c + 2) % 3];
            cx[c][0] = u[1] * v[2] - u[2] * v[1];
            cx[c][1] = u[2] * v[0] - u[0] * v[2];
            cx[c][2] = u[0] * v[1] - u[1] * v[0];
        }
        det = r[0][0] * cx[0][0] + r[0][1] * cx[0][1] + r[0][2] * cx[0][2];
        const double id = 1.0 / det;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) G[d][c + 1] = cx[c][d] * id;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double sg = 0.0;
#pragma unroll
        for (int c = 1; c < K; ++c) sg += G[d][c];
        G[d][0] = -sg;
    }
    double kbar = 0.0;
#pragma unroll
    for (int c = 0; c < K; ++c) kbar += a.qc.ck[c] * kn[c];
    kbar *= ks;
    const double kd = kbar * det;
    double ga[D];
    double ml = 0.0;
#pragma unroll
    for (int c = 0; c < K; ++c)
        if (c == la) {
            ml = a.qc.ml[c];
#pragma unroll
            for (int d = 0; d < D; ++d) ga[d] = G[d][c];
        }
#pragma unroll
    for (int c = 0; c < K; ++c) {
        double g = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) g += ga[d] * G[d][c];
        out[c] = kd * g;
    }
    out[K] = -kd * ga[D - 1];
    out[K + 1] = det * ml;
}

// Row contribution of one incidence (element e seen from its vertex la):
//   out[c] (c <= dim)  stiffness K_e(la, c) = Kbar detJ G(:,la).G(:,c)
//   out[dim+1]         gravity source -Kbar detJ dN_la/d(last coordinate)
//   out[dim+2]         lumped mass detJ * ml[la]
// The mass term M_L C/dt on the diagonal and -M_L (theta_m - theta_n - C h)/dt
// in the source are linear in the row node's values: they are applied once per
// row with the summed lumped mass.  Returns the incidence's slot word.
template <int D>
__device__ inline uint32_t element_row(const AsmArgs& a, int q, double (&out)[D + 3]) {
    constexpr int K = D + 1;
    const int2 ic = a.inc[q];
    const int e = ic.x >> 2;
    const int4 cn = a.conn4[e];
    const int nd[4] = {cn.x, cn.y, cn.z, cn.w};
    double px[K], py[K], pz[K], kn[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const double4 r = a.nrec.template rec<D>(nd[c]);
        px[c] = r.x;
        py[c] = r.y;
        pz[c] = (D == 3) ? r.z : 0.0;
        kn[c] = (D == 3) ? r.w : r.z;
    }
    const double ks = a.Ks_elem != nullptr ? a.Ks_elem[e] : 1.0;
    element_compute<D>(a, px, py, pz, kn, ks, ic.x & 3, out);
    return (uint32_t)ic.y;
}

// Krylov start fused into the assembly epilogue (one row, lane 0 of its group):
// r = b - A x0 and the method's first vectors and dot products.
template <int MODE>
__device__ inline void krylov_start(const AsmArgs& a, const KrylovInit& ki, int i, double bi, double ax,
                                    double dinv, double (&acc)[3]) {
    const double r = bi - ax;
    if constexpr (MODE == ASM_INIT_BICG) {
        ki.x[i] = ki.x0[i];
        ki.r[i] = r;
        ki.rh[i] = r;
        acc[0] += r * r;
        acc[1] += bi * bi;
    } else if constexpr (MODE == ASM_INIT_CG) {
        ki.x[i] = ki.x0[i];
        ki.r[i] = r;
        const double z = ki.jacobi ? dinv * r : r;
        if (ki.jacobi) ki.z[i] = z;
        ki.p[i] = z;
        acc[0] += r * z;
        acc[1] += r * r;
        acc[2] += bi * bi;
    } else if constexpr (MODE == ASM_INIT_GMRES) {
        ki.x[i] = ki.x0[i];
        const double w = ki.jacobi ? dinv * r : r;
        ki.w[i] = w;
        const double mb = ki.jacobi ? dinv * bi : bi;
        acc[0] += bi * bi;
        acc[1] += mb * mb;
        acc[2] += w * w;
    } else if constexpr (MODE == ASM_INIT_RESID) {
        if (ki.x != nullptr) ki.x[i] = ki.x0[i];   // the refinement's iterate starts at x0
        ki.r[i] = r;
        acc[0] += r * r;
        acc[1] += bi * bi;
    }
}

template <int D, int R, int MODE>
__global__ __launch_bounds__(ASM_THREADS) void k_assemble(AsmArgs a, KrylovInit ki, double* partials, int prow) {
    constexpr int K = D + 1;
    constexpr int NC = K + 2;                 // values per incidence in LDS
    constexpr int LANES = ASM_THREADS / R;    // threads per row in the fused Krylov start
    static_assert(LANES >= 1 && LANES <= 64 && (LANES & (LANES - 1)) == 0, "bad assembly shape");
    extern __shared__ double lds[];
    __shared__ int rp_s[R + 1], ip_s[R + 1], dg_s[R];
    __shared__ uint8_t bc_s[R];
    __shared__ double bsh[R], dsh[R], cs_s[R], rs_s[R];
    // rows of logical block L: each XCD assembles one contiguous range of
    // rows (their neighbours' records come through one L2); partials in
    // column L
    const int L = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = L * R;
    const int r1 = min(a.n, r0 + R);
    const int nrows = r1 - r0;
    if (a.tflag != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *a.tflag = a.n;
    const int base = a.row_ptr[r0];
    const int wlen = a.row_ptr[r1] - base;
    const int q0 = a.inc_ptr[r0];
    const int nq = a.inc_ptr[r1] - q0;
    double* win = lds;                                    // [win_cap] finished CSR window
    double* xw = lds + a.win_cap;                         // [win_cap] x0 at the window's columns
    double* con = lds + 2 * a.win_cap;                    // [inc_cap][NC] incidence contributions
    uint32_t* slw = reinterpret_cast<uint32_t*>(con + (size_t)a.inc_cap * NC);  // [inc_cap] slot words

    // 0-1. all global reads: the row metadata, the x0 gathers of the window
    //      and the element contributions of every incidence of the workgroup.
    //      Each thread's first pass of the three (its row, its window entry,
    //      its first pair of incidences) is loaded at once with indices
    //      clamped into range, so that no load waits behind a branch or
    //      behind an independent one: after the block's row pointers, the
    //      incidences, metadata and columns, then the elements and x0, then
    //      the node records -- three round trips instead of six.  Longer
    //      blocks finish in the loops below.
    auto elem_pair = [&](int t0, const int2 ic0, const int2 ic1, const double4 (&r0c)[K], const double4 (&r1c)[K],
                         double ks0, double ks1) {
        const int t1 = t0 + ASM_THREADS;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int t = h == 0 ? t0 : t1;
            if (t >= nq) break;
            const double4* r = h == 0 ? r0c : r1c;
            double px[K], py[K], pz[K], kn[K];
#pragma unroll
            for (int c = 0; c < K; ++c) {
                px[c] = r[c].x;
                py[c] = r[c].y;
                pz[c] = (D == 3) ? r[c].z : 0.0;
                kn[c] = (D == 3) ? r[c].w : r[c].z;
            }
            double out[D + 3];
            const int2 ic = h == 0 ? ic0 : ic1;
            element_compute<D>(a, px, py, pz, kn, h == 0 ? ks0 : ks1, ic.x & 3, out);
            slw[t] = (uint32_t)ic.y;
#pragma unroll
            for (int c = 0; c < NC; ++c) con[t * NC + c] = out[c];
        }
    };
    {
        const int tid = threadIdx.x;
        // a block always has rows, window entries and (on a mesh) incidences
        const int qa = q0 + (nq > 0 ? min(tid, nq - 1) : 0);
        const int qb = q0 + (nq > 0 ? min(tid + ASM_THREADS, nq - 1) : 0);
        const int2 ic0 = a.inc[qa];
        const int2 ic1 = a.inc[qb];
        const int jm = min(tid, nrows), jr = min(tid, nrows - 1);
        const int rpv = a.row_ptr[r0 + jm];
        const int ipv = a.inc_ptr[r0 + jm];
        const int dgv = a.diag_pos[r0 + jr];
        const uint8_t* bf = a.bc_flag != nullptr ? a.bc_flag : reinterpret_cast<const uint8_t*>(a.row_ptr);
        const int fl = bf[r0 + jr];
        const int bcv = a.bc_flag != nullptr ? fl : 0;
        const double ci = a.C[r0 + jr];
        const double rsv = a.theta_m[r0 + jr] - a.theta_n[r0 + jr] - ci * a.h_m[r0 + jr];
        const int colk = MODE != ASM_INIT_NONE ? a.cols[base + min(tid, wlen - 1)] : 0;
        const int4 cn0 = a.conn4[ic0.x >> 2];
        const int4 cn1 = a.conn4[ic1.x >> 2];
        double x0k = 0.0;
        if constexpr (MODE != ASM_INIT_NONE) x0k = ki.x0[colk];
        const int nd0[4] = {cn0.x, cn0.y, cn0.z, cn0.w};
        const int nd1[4] = {cn1.x, cn1.y, cn1.z, cn1.w};
        double4 rc0[K], rc1[K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            rc0[c] = a.nrec.template rec<D>(nd0[c]);
            rc1[c] = a.nrec.template rec<D>(nd1[c]);
        }
        const double* ksp = a.Ks_elem != nullptr ? a.Ks_elem : nullptr;
        const double ks0 = ksp != nullptr ? ksp[ic0.x >> 2] : 1.0;
        const double ks1 = ksp != nullptr ? ksp[ic1.x >> 2] : 1.0;
        if (tid <= nrows) {
            rp_s[tid] = rpv - base;
            ip_s[tid] = ipv - q0;
        }
        if (tid < nrows) {
            dg_s[tid] = dgv - base;
            bc_s[tid] = (uint8_t)bcv;
            cs_s[tid] = ci;
            rs_s[tid] = rsv;
        }
        if constexpr (MODE != ASM_INIT_NONE)
            if (tid < wlen) xw[tid] = x0k;
        elem_pair(tid, ic0, ic1, rc0, rc1, ks0, ks1);
    }
    for (int j = threadIdx.x + ASM_THREADS; j <= nrows; j += ASM_THREADS) {
        rp_s[j] = a.row_ptr[r0 + j] - base;
        ip_s[j] = a.inc_ptr[r0 + j] - q0;
        if (j < nrows) {
            const int i = r0 + j;
            dg_s[j] = a.diag_pos[i] - base;
            bc_s[j] = a.bc_flag != nullptr ? a.bc_flag[i] : 0;
            const double ci = a.C[i];
            cs_s[j] = ci;
            rs_s[j] = a.theta_m[i] - a.theta_n[i] - ci * a.h_m[i];
        }
    }
    if constexpr (MODE != ASM_INIT_NONE)
        for (int k = threadIdx.x + ASM_THREADS; k < wlen; k += ASM_THREADS) xw[k] = ki.x0[a.cols[base + k]];
    for (int t0 = threadIdx.x + 2 * ASM_THREADS; t0 < nq; t0 += 2 * ASM_THREADS) {
        const int t1 = t0 + ASM_THREADS;
        const bool two = t1 < nq;
        const int2 ic0 = a.inc[q0 + t0];
        const int2 ic1 = two ? a.inc[q0 + t1] : ic0;
        const int4 cn0 = a.conn4[ic0.x >> 2];
        const int4 cn1 = a.conn4[ic1.x >> 2];
        const int nd0[4] = {cn0.x, cn0.y, cn0.z, cn0.w};
        const int nd1[4] = {cn1.x, cn1.y, cn1.z, cn1.w};
        double4 rc0[K], rc1[K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            rc0[c] = a.nrec.template rec<D>(nd0[c]);
            rc1[c] = a.nrec.template rec<D>(nd1[c]);
        }
        const double ks0 = a.Ks_elem != nullptr ? a.Ks_elem[ic0.x >> 2] : 1.0;
        const double ks1 = a.Ks_elem != nullptr ? a.Ks_elem[ic1.x >> 2] : 1.0;
        elem_pair(t0, ic0, ic1, rc0, rc1, ks0, ks1);
    }
    __syncthreads();

    // 2. one thread per CSR entry of the window: sum the contributions whose
    //    vertex lands in this entry's slot, in incidence order (the diagonal
    //    entry's thread also sums the row's source term).  Dirichlet rows are
    //    identity rows.  Written straight to HBM, coalesced.
    for (int k = threadIdx.x; k < wlen; k += ASM_THREADS) {
        int lo = 0, hi = nrows - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rp_s[mid] <= k) lo = mid; else hi = mid - 1;
        }
        const int j = lo;
        const uint32_t sl = (uint32_t)(k - rp_s[j]);
        const bool diag = (k == dg_s[j]);
        double v = 0.0, bi = 0.0;
        if (bc_s[j]) {
            v = diag ? 1.0 : 0.0;
            bi = a.bc_val[r0 + j];
        } else {
            double msum = 0.0;
            // an incidence's vertices are distinct, so at most one of its slot
            // bytes equals sl: one zero-byte test finds it (the lowest flagged
            // byte is the zero one; bytes beyond K are set to 0xff)
            // Four incidences per pass: their slot words, then the matching
            // contributions (+0.0 where none matches, which leaves the sum
            // bitwise as it is: the entries start at +0.0 and so never become
            // -0.0), each issued together, so a pass costs two LDS round
            // trips instead of two per incidence; the sum still runs in
            // incidence order.
            const uint32_t pat = sl * 0x01010101u;
            const uint32_t pad = K == 4 ? 0u : 0xff000000u;
            const int te = ip_s[j + 1];
            for (int t = ip_s[j]; t < te; t += 4) {
                uint32_t hz[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t m = ((t + u < te ? slw[t + u] : 0xffffffffu) | pad) ^ pat;
                    hz[u] = (m - 0x01010101u) & ~m & 0x80808080u;
                }
                double c[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    c[u] = hz[u] != 0u ? con[(t + u) * NC + ((__builtin_ctz(hz[u]) - 7) >> 3)] : 0.0;
#pragma unroll
                for (int u = 0; u < 4; ++u) v += c[u];
                if (diag) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (t + u < te) {
                            bi += con[(t + u) * NC + K];
                            msum += con[(t + u) * NC + K + 1];
                        }
                    }
                }
            }
            if (diag) {
                v += a.inv_dt * msum * cs_s[j];
                bi -= a.inv_dt * msum * rs_s[j];
            }
        }
        win[k] = v;
        // CSR order, or ELL (entry sl of row r0 + j at sl * ell_n + r0 + j)
        const size_t o = a.ell_n ? (size_t)sl * a.ell_n + (size_t)(r0 + j) : (size_t)base + k;
        a.A[o] = v;
        if (a.A32 != nullptr) a.A32[o] = (float)v;
        if (diag) {
            const int i = r0 + j;
            a.b[i] = bi;
            const double di = 1.0 / (v != 0.0 ? v : 1.0);
            if (a.dinv != nullptr) a.dinv[i] = di;
            if (a.dinv32 != nullptr) a.dinv32[i] = 1.0f / (v != 0.0 ? (float)v : 1.0f);
            bsh[j] = bi;
            dsh[j] = di;
        }
    }

    // 3. fused Krylov start on the finished window (LANES threads per row)
    if constexpr (MODE != ASM_INIT_NONE) {
        __syncthreads();
        double acc[3] = {0.0, 0.0, 0.0};
        const int g = threadIdx.x / LANES, l = threadIdx.x % LANES;
        double s = 0.0;
        if (g < nrows)
            for (int k = rp_s[g] + l; k < rp_s[g + 1]; k += LANES) s += win[k] * xw[k];
#pragma unroll
        for (int o = LANES / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LANES);
        if (g < nrows && l == 0) krylov_start<MODE>(a, ki, r0 + g, bsh[g], s, dsh[g], acc);
        store_partials<3, ASM_THREADS / 64>(acc, partials, prow, L);
    }
}

// Register variant: one thread per row, no LDS.  Used when every row has at
// most MAXR entries (8 in 2D, 16 in 3D covers the structured meshes); each
// contribution is added to its slot with a compile-time select, so the row
// never leaves registers and all 256 threads of a workgroup stay busy.  Same
// fixed per-slot summation order as k_assemble.
template <int D, int MAXR, int MODE>
__device__ inline void assemble_row(const AsmArgs& a, const KrylovInit& ki, int i, double (&acc)[3]) {
    constexpr int K = D + 1;
    const int rs = a.row_ptr[i];
    const int rl = a.row_ptr[i + 1] - rs;
    const int dg = a.diag_pos[i] - rs;
    double v[MAXR];
#pragma unroll
    for (int s = 0; s < MAXR; ++s) v[s] = 0.0;
    double bi;
    if (a.bc_flag != nullptr && a.bc_flag[i]) {
#pragma unroll
        for (int s = 0; s < MAXR; ++s) v[s] = (s == dg) ? 1.0 : 0.0;
        bi = a.bc_val[i];
    } else {
        bi = 0.0;
        double msum = 0.0;
        const int qe = a.inc_ptr[i + 1];
        // NB incidences per batch: their loads are issued together (three
        // dependent round trips per batch: incidence, element, node records)
        for (int qb = a.inc_ptr[i]; qb < qe; qb += ASM_REG_BATCH) {
            int2 ic[ASM_REG_BATCH];
            int4 cn[ASM_REG_BATCH];
            double4 nr[ASM_REG_BATCH][K];
            double ks[ASM_REG_BATCH];
#pragma unroll
            for (int u = 0; u < ASM_REG_BATCH; ++u)
                ic[u] = (qb + u < qe) ? a.inc[qb + u] : make_int2(0, 0);
#pragma unroll
            for (int u = 0; u < ASM_REG_BATCH; ++u)
                cn[u] = (qb + u < qe) ? a.conn4[ic[u].x >> 2] : make_int4(0, 0, 0, 0);
#pragma unroll
            for (int u = 0; u < ASM_REG_BATCH; ++u) {
                const int nd[4] = {cn[u].x, cn[u].y, cn[u].z, cn[u].w};
#pragma unroll
                for (int c = 0; c < K; ++c)
                    nr[u][c] = (qb + u < qe) ? a.nrec.template rec<D>(nd[c]) : make_double4(0.0, 0.0, 0.0, 0.0);
                ks[u] = (a.Ks_elem != nullptr && qb + u < qe) ? a.Ks_elem[ic[u].x >> 2] : 1.0;
            }
#pragma unroll
            for (int u = 0; u < ASM_REG_BATCH; ++u) {
                if (qb + u >= qe) break;
                double px[K], py[K], pz[K], kn[K];
#pragma unroll
                for (int c = 0; c < K; ++c) {
                    px[c] = nr[u][c].x;
                    py[c] = nr[u][c].y;
                    pz[c] = (D == 3) ? nr[u][c].z : 0.0;
                    kn[c] = (D == 3) ? nr[u][c].w : nr[u][c].z;
                }
                double out[D + 3];
                element_compute<D>(a, px, py, pz, kn, ks[u], ic[u].x & 3, out);
                add_slots<K, MAXR>(v, (uint32_t)ic[u].y, out);
                bi += out[K];
                msum += out[K + 1];
            }
        }
        const double ci = a.C[i];
        const double rhs = a.theta_m[i] - a.theta_n[i] - ci * a.h_m[i];
        const double madd = a.inv_dt * msum * ci;
#pragma unroll
        for (int s = 0; s < MAXR; ++s) v[s] += (s == dg) ? madd : 0.0;
        bi -= a.inv_dt * msum * rhs;
    }
    double dgv = 0.0;
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
        if (s < rl) {
            const size_t o = a.ell_n ? (size_t)s * a.ell_n + i : (size_t)rs + s;
            a.A[o] = v[s];
            if (a.A32 != nullptr) a.A32[o] = (float)v[s];
        }
        if (s == dg) dgv = v[s];
    }
    a.b[i] = bi;
    const double di = 1.0 / (dgv != 0.0 ? dgv : 1.0);
    if (a.dinv != nullptr) a.dinv[i] = di;
    if (a.dinv32 != nullptr) a.dinv32[i] = 1.0f / (dgv != 0.0 ? (float)dgv : 1.0f);
    if constexpr (MODE != ASM_INIT_NONE) {
        double ax = 0.0;
#pragma unroll
        for (int s = 0; s < MAXR; ++s)
            if (s < rl) ax += v[s] * ki.x0[a.cols[rs + s]];
        krylov_start<MODE>(a, ki, i, bi, ax, di, acc);
    }
}

// Neighbourhood variant (2D).  A row's CSR columns are exactly the nodes its
// incident elements touch, and every incidence's slot word already maps the
// element's vertices onto those columns.  So the row gathers the node records
// and x0 values of its columns once; each incidence then only adds its own
// word and Ks_e.  Three dependent memory round trips per row (row pointers;
// columns + incidences; records + x0 + Ks) instead of three per batch of
// incidences, and rl node-record loads instead of 3 per incidence.  The
// element arithmetic and the order of every sum are those of assemble_row,
// so the results are identical.  Needs at most MAXR - 1 incidences per row.
// LDS = 1 (default): the row's neighbour records are staged in a per-thread
// slice of LDS and each incidence reads its three vertices back by slot index
// (3 LDS reads per vertex).  LDS = 0: they are taken from registers by a
// one-hot blend over the slots (about 30 instructions per vertex).  Each
// thread touches only its own slice ([component][slot][thread], so the lanes
// of a wavefront hit distinct banks whatever their slots): no barrier.  Both
// give bitwise the same results; LDS = 1 measured 2 % (512^2) to 4 %
// (256-instance ensemble) faster per time step.  (Accumulating the row's
// entries in LDS as well was no faster and changed the FMA contraction of
// the sums.)
template <int MAXR, int MODE, int LDS = 1>
__device__ inline void assemble_row_nbr(const AsmArgs& a, const KrylovInit& ki, int i, double (&acc)[3]) {
    constexpr int MAXI = MAXR - 1;
    int rs, rl, dg, nq;
    int col[MAXR];
    int2 ic[MAXI];
    const bool dir = a.bc_flag != nullptr && a.bc_flag[i];
    if (a.ell_info != nullptr) {
        // all of the row's index data in one round trip (ELL layouts)
        const uint32_t inf = a.ell_info[i];
        rs = 0;
        rl = (int)(inf & 15u);
        dg = (int)((inf >> 4) & 15u);
        nq = (int)((inf >> 8) & 15u);
#pragma unroll
        for (int s = 0; s < MAXR; ++s) col[s] = s < a.ell_w ? a.ell_col[(size_t)s * a.n + i] : i;
#pragma unroll
        for (int u = 0; u < MAXI; ++u) ic[u] = u < a.ell_iw ? a.ell_inc[(size_t)u * a.n + i] : make_int2(0, 0);
    } else {
        rs = a.row_ptr[i];
        rl = a.row_ptr[i + 1] - rs;
        dg = a.diag_pos[i] - rs;
        const int q0 = a.inc_ptr[i];
        nq = a.inc_ptr[i + 1] - q0;
#pragma unroll
        for (int s = 0; s < MAXR; ++s) col[s] = s < rl ? a.cols[rs + s] : i;
#pragma unroll
        for (int u = 0; u < MAXI; ++u) ic[u] = (!dir && u < nq) ? a.inc[q0 + u] : make_int2(0, 0);
    }
    double4 nr[MAXR];
    double xv[MAXR];
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
        nr[s] = (!dir && s < rl) ? a.nrec.rec<2>(col[s]) : make_double4(0.0, 0.0, 0.0, 0.0);
        if constexpr (MODE != ASM_INIT_NONE) xv[s] = s < rl ? ki.x0[col[s]] : 0.0;
    }
    double ks[MAXI];
#pragma unroll
    for (int u = 0; u < MAXI; ++u) ks[u] = (a.Ks_elem != nullptr && !dir && u < nq) ? a.Ks_elem[ic[u].x >> 2] : 1.0;
    double v[MAXR];
#pragma unroll
    for (int s = 0; s < MAXR; ++s) v[s] = 0.0;
    double bi;
    if (dir) {
#pragma unroll
        for (int s = 0; s < MAXR; ++s) v[s] = (s == dg) ? 1.0 : 0.0;
        bi = a.bc_val[i];
    } else {
        bi = 0.0;
        double msum = 0.0;
        // per-thread LDS slice of the records (LDS == 1): [3][MAXR][BLOCK]
        __shared__ double lds[LDS == 1 ? 3 * MAXR * BLOCK : 1];
        double* const rec = lds + threadIdx.x;
        if constexpr (LDS == 1) {
#pragma unroll
            for (int s = 0; s < MAXR; ++s) {
                rec[(0 * MAXR + s) * BLOCK] = nr[s].x;
                rec[(1 * MAXR + s) * BLOCK] = nr[s].y;
                rec[(2 * MAXR + s) * BLOCK] = nr[s].z;
            }
        }
        // every iteration predicated (no early exit), so the loop unrolls fully
        // and ic / ks stay in registers
#pragma unroll
        for (int u = 0; u < MAXI; ++u) {
            if (u < nq) {
            const uint32_t w = (uint32_t)ic[u].y;
            double px[3], py[3], pz[3], kn[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const uint32_t sl = (w >> (8 * c)) & 0xffu;
                if constexpr (LDS == 1) {
                    px[c] = rec[(0 * MAXR + sl) * BLOCK];
                    py[c] = rec[(1 * MAXR + sl) * BLOCK];
                    kn[c] = rec[(2 * MAXR + sl) * BLOCK];
                } else {
                    // one-hot blend of the row's records (exact: one weight is 1,
                    // the others 0 against finite values).  A select chain here is
                    // turned by the compiler into an indexed load from a private
                    // array, i.e. scratch memory.
                    double rx = 0.0, ry = 0.0, rk = 0.0;
#pragma unroll
                    for (int s = 0; s < MAXR; ++s) {
                        const double m = (sl == (uint32_t)s) ? 1.0 : 0.0;
                        rx += m * nr[s].x;
                        ry += m * nr[s].y;
                        rk += m * nr[s].z;
                    }
                    px[c] = rx;
                    py[c] = ry;
                    kn[c] = rk;
                }
                pz[c] = 0.0;
            }
            double out[5];
            element_compute<2>(a, px, py, pz, kn, ks[u], ic[u].x & 3, out);
            add_slots<3, MAXR>(v, w, out);
            bi += out[3];
            msum += out[4];
            }
        }
        const double ci = a.C[i];
        const double rhs = a.theta_m[i] - a.theta_n[i] - ci * a.h_m[i];
        const double madd = a.inv_dt * msum * ci;
#pragma unroll
        for (int s = 0; s < MAXR; ++s) v[s] += (s == dg) ? madd : 0.0;
        bi -= a.inv_dt * msum * rhs;
    }
    double dgv = 0.0;
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
        if (s < rl) {
            const size_t o = a.ell_n ? (size_t)s * a.ell_n + i : (size_t)rs + s;
            a.A[o] = v[s];
            if (a.A32 != nullptr) a.A32[o] = (float)v[s];
        }
        if (s == dg) dgv = v[s];
    }
    a.b[i] = bi;
    const double di = 1.0 / (dgv != 0.0 ? dgv : 1.0);
    if (a.dinv != nullptr) a.dinv[i] = di;
    if (a.dinv32 != nullptr) a.dinv32[i] = 1.0f / (dgv != 0.0 ? (float)dgv : 1.0f);
    if constexpr (MODE != ASM_INIT_NONE) {
        double ax = 0.0;
#pragma unroll
        for (int s = 0; s < MAXR; ++s)
            if (s < rl) ax += v[s] * xv[s];
        krylov_start<MODE>(a, ki, i, bi, ax, di, acc);
    }
}

// 1 / d for an element determinant (finite, normal, non-zero): the hardware
// reciprocal refined by two Newton steps, within an ulp or so of the IEEE
// quotient in 5 instructions instead of the 10 of the full division (the
// fan kernels are bound by their instruction issue at ensemble scale).
__device__ inline double recip_nr(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// The fan kernels' quadrature weights by the node's local vertex la,
// (ck[la], ck[la + 1], ck[la + 2], ml[la]) at qw[4 * la], into LDS; read
// after a workgroup barrier.  (A lane-varying select of the kernel
// argument's weights compiles into loads from the argument segment.)
__device__ inline void fan_weights(const QuadConst& qc, double* qw) {
    if (threadIdx.x < 12) {
        const int la = threadIdx.x / 4, c = threadIdx.x % 4;
        qw[threadIdx.x] = c == 3 ? qc.ml[la] : qc.ck[(la + c) % 3];
    }
}

// Fan variant (2D, the time-step operator in fan order).  The row's slots
// are the node itself (slot 0) and its neighbours in their cyclic order
// around it, so fan triangle u is (node, slot u, slot u + 1), with the
// closed fan's triangle 0 taking the neighbour before slot 1 from pred.
// Every triangle therefore adds to slots known at compile time after the
// unroll: no slot selects, no staging of the records.  The triangle is
// taken in the rotated order (node, slot u, slot u + 1) = (v_la, v_la+1,
// v_la+2) of its own vertices, the same orientation, so det J, G and the
// stiffness row of the node are those of the element (to rounding), with
// the quadrature's conductivity and mass weights of vertices la, la + 1,
// la + 2.  Sums run over the triangles in fan order; a closed fan's
// triangle 0 contribution to the last neighbour is added after the others.
template <int MODE>
__device__ inline void assemble_row_fan(const AsmArgs& a, const KrylovInit& ki, int i, const double* qw,
                                        double (&acc)[3]) {
    constexpr int W = 8;          // the node + up to 7 neighbours
    constexpr int T = W - 1;      // fan triangles
    const uint32_t inf = a.fan_info[i];
    const int F = (int)(inf & 15u);
    const bool closed = (inf & 16u) != 0;
    const bool dir = a.bc_flag != nullptr && a.bc_flag[i];
    int col[W];
#pragma unroll
    for (int s = 0; s < W; ++s) col[s] = s < a.ell_w ? a.ell_col[(size_t)s * a.n + i] : i;
    const int pr = a.fan_pred[i];
    // every gather unconditional: padding columns, pred and padding triangle
    // codes all name valid nodes / elements (the row's own node, element 0),
    // and values of absent slots are never used.  (A select between a record
    // and a zero record compiles into a flat load through a selected pointer.)
    double4 nr[W];
#pragma unroll
    for (int s = 0; s < W; ++s) nr[s] = a.nrec.rec<2>(col[s]);
    const double4 np = a.nrec.rec<2>(pr);
    double xv[W];
    if constexpr (MODE != ASM_INIT_NONE) {
#pragma unroll
        for (int s = 0; s < W; ++s) xv[s] = ki.x0[col[s]];
    }
    // per-element Ks through the fan triangles' elements (the table is read
    // only then: the node's local vertex of each triangle is in the info word)
    double ks[T];
    if (a.Ks_elem != nullptr) {   // uniform
        // every code load issued (past fan_iw: plane 0's, unused), so that
        // none waits behind a branch
        int code[T];
#pragma unroll
        for (int u = 0; u < T; ++u) code[u] = a.fan_inc[(size_t)(u < a.fan_iw ? u : 0) * a.n + i];
#pragma unroll
        for (int u = 0; u < T; ++u) ks[u] = a.Ks_elem[(u < a.fan_iw ? code[u] : 0) >> 2];
    } else {
#pragma unroll
        for (int u = 0; u < T; ++u) ks[u] = 1.0;
    }
    double v[W];
#pragma unroll
    for (int s = 0; s < W; ++s) v[s] = 0.0;
    double bi;
    if (dir) {
        v[0] = 1.0;
        bi = a.bc_val[i];
    } else {
        bi = 0.0;
        double msum = 0.0, vp = 0.0;
        const double sx = nr[0].x, sy = nr[0].y, sk = nr[0].z;
#pragma unroll
        for (int u = 0; u < T; ++u) {
            const bool ex = closed ? (u < F) : (u >= 1 && u < F);
            if (ex) {
                const double4 P = (u == 0) ? np : nr[u];
                const double4 Q = nr[u + 1];
                const int la = (int)((inf >> (5 + 2 * u)) & 3u);
                const double4 w = *reinterpret_cast<const double4*>(qw + 4 * la);
                // edge vectors from the node: d1 = P - S, d2 = Q - S (the
                // rows of J), e = Q - P (the edge opposite the node).  With
                // G = J^-1 grad N, the node's stiffness row kbar det G0.Gc is
                // c (e.e, -e.d2, e.d1) with c = kbar / det (the three sum to
                // zero), and its gravity term kbar det dN0/dy = kbar (J10 - J00)
                // = kbar e.x needs no division
                const double d1x = P.x - sx, d1y = P.y - sy;
                const double d2x = Q.x - sx, d2y = Q.y - sy;
                const double det = d1x * d2y - d1y * d2x;
                const double ex = d2x - d1x, ey = d2y - d1y;
                const double ckS = w.x, ckP = w.y, ckQ = w.z, mlS = w.w;
                const double kbar = (ckS * sk + ckP * P.z + ckQ * Q.z) * ks[u];
                const double c = kbar * recip_nr(det);
                const double kss = c * (ex * ex + ey * ey);
                const double ksq = c * (ex * d1x + ey * d1y);
                const double ksp = -(kss + ksq);
                v[0] += kss;
                if (u == 0) vp += ksp;
                else v[u] += ksp;
                v[u + 1] += ksq;
                bi -= kbar * ex;
                msum += det * mlS;
            }
        }
        // the closed fan's triangle 0 joins the last neighbour (slot F)
#pragma unroll
        for (int s = 1; s < W; ++s) v[s] = (closed && s == F) ? v[s] + vp : v[s];
        const double ci = a.C[i];
        const double rhs = a.theta_m[i] - a.theta_n[i] - ci * a.h_m[i];
        v[0] += a.inv_dt * msum * ci;
        bi -= a.inv_dt * msum * rhs;
    }
#pragma unroll
    for (int s = 0; s < W; ++s) {
        if (s <= F && s < a.ell_w) {
            const size_t o = (size_t)s * a.ell_n + i;
            a.A[o] = v[s];
            if (a.A32 != nullptr) a.A32[o] = (float)v[s];
        }
    }
    a.b[i] = bi;
    const double dgv = v[0];
    const double di = 1.0 / (dgv != 0.0 ? dgv : 1.0);
    if (a.dinv != nullptr) a.dinv[i] = di;
    if (a.dinv32 != nullptr) a.dinv32[i] = 1.0f / (dgv != 0.0 ? (float)dgv : 1.0f);
    if constexpr (MODE != ASM_INIT_NONE) {
        double ax = 0.0;
#pragma unroll
        for (int s = 0; s < W; ++s)
            if (s <= F) ax += v[s] * xv[s];
        krylov_start<MODE>(a, ki, i, bi, ax, di, acc);
    }
}

// Windowed fan variant.  The workgroup stages its row block's node window
// (coordinates, K and the Krylov start's x0 of every node its rows touch, a
// few consecutive segments) in LDS with coalesced loads, then each row reads
// the records of its fan slots from LDS at the positions nwin_col gives.
// Every global load of a row block is issued before the window's barrier
// and waited for once: the row's own inputs (WinRow), the window, and, with
// per-element Ks, the fan triangles' element codes then their Ks -- no load
// sits behind a branch or behind another load it does not depend on.  The
// arithmetic and its order are those of assemble_row_fan, so the values are
// bitwise the same.
struct WinRow {
    uint32_t inf;      // F | closed << 4 | local vertices << 5
    int dir;
    int lc[8];         // window positions of the fan slots
    int lp;            // of the closed fan's predecessor
    double ks[7];      // per fan triangle (1 without per-element Ks)
    double C, thm, thn, hm;
};

// The row's own inputs.  bc_flag may be null (no Dirichlet rows): the flag
// byte is then read from fan_info and not used, so that no load sits behind
// a branch.  code receives the fan triangles' element codes (with per-element
// Ks; 0 past fan_iw), whose Ks load_win_ks reads once they have arrived.
__device__ inline void load_win_row(const AsmArgs& a, int i, WinRow& r, int (&code)[7]) {
    constexpr int W = 8, T = W - 1;
    r.inf = a.fan_info[i];
    const uint8_t* bf = a.bc_flag != nullptr ? a.bc_flag : reinterpret_cast<const uint8_t*>(a.fan_info);
    const int fl = bf[i];
    r.dir = a.bc_flag != nullptr ? fl : 0;
#pragma unroll
    for (int s = 0; s < W; ++s) r.lc[s] = (int)a.nwin_col[(size_t)(s < a.ell_w ? s : 0) * a.n + i];
    r.lp = (int)a.nwin_pred[i];
    r.C = a.C[i];
    r.thm = a.theta_m[i];
    r.thn = a.theta_n[i];
    r.hm = a.h_m[i];
    if (a.Ks_elem != nullptr) {   // uniform
#pragma unroll
        for (int u = 0; u < T; ++u) code[u] = a.fan_inc[(size_t)(u < a.fan_iw ? u : 0) * a.n + i];
    }
}

__device__ inline void load_win_ks(const AsmArgs& a, const int (&code)[7], WinRow& r) {
    constexpr int T = 7;
    if (a.Ks_elem != nullptr) {   // uniform
#pragma unroll
        for (int u = 0; u < T; ++u) r.ks[u] = a.Ks_elem[(u < a.fan_iw ? code[u] : 0) >> 2];
    } else {
#pragma unroll
        for (int u = 0; u < T; ++u) r.ks[u] = 1.0;
    }
}

// qw: the quadrature weights by local vertex (fan_weights).
template <int MODE>
__device__ inline void assemble_row_win(const AsmArgs& a, const KrylovInit& ki, int i, const WinRow& r,
                                        const double* lx, const double* ly, const double* lk, const double* lx0,
                                        const double* qw, double (&acc)[3]) {
    constexpr int W = 8;          // the node + up to 7 neighbours
    constexpr int T = W - 1;      // fan triangles
    const uint32_t inf = r.inf;
    const int F = (int)(inf & 15u);
    const bool closed = (inf & 16u) != 0;
    double v[W];
#pragma unroll
    for (int s = 0; s < W; ++s) v[s] = 0.0;
    double bi;
    if (r.dir) {
        v[0] = 1.0;
        bi = a.bc_val[i];   // the few Dirichlet rows read theirs here
    } else {
        bi = 0.0;
        double msum = 0.0, vp = 0.0;
        const double sx = lx[r.lc[0]], sy = ly[r.lc[0]], sk = lk[r.lc[0]];
#pragma unroll
        for (int u = 0; u < T; ++u) {
            const bool ex = closed ? (u < F) : (u >= 1 && u < F);
            if (ex) {
                const int pp = (u == 0) ? r.lp : r.lc[u];
                const int qq = r.lc[u + 1];
                const int la = (int)((inf >> (5 + 2 * u)) & 3u);
                const double Px = lx[pp], Py = ly[pp], Pk = lk[pp];
                const double Qx = lx[qq], Qy = ly[qq], Qk = lk[qq];
                const double d1x = Px - sx, d1y = Py - sy;
                const double d2x = Qx - sx, d2y = Qy - sy;
                const double det = d1x * d2y - d1y * d2x;
                const double ex_ = d2x - d1x, ey = d2y - d1y;
                const double4 w = *reinterpret_cast<const double4*>(qw + 4 * la);
                const double ckS = w.x, ckP = w.y, ckQ = w.z, mlS = w.w;
                const double kbar = (ckS * sk + ckP * Pk + ckQ * Qk) * r.ks[u];
                const double c = kbar * recip_nr(det);
                const double kss = c * (ex_ * ex_ + ey * ey);
                const double ksq = c * (ex_ * d1x + ey * d1y);
                const double ksp = -(kss + ksq);
                v[0] += kss;
                if (u == 0) vp += ksp;
                else v[u] += ksp;
                v[u + 1] += ksq;
                bi -= kbar * ex_;
                msum += det * mlS;
            }
        }
        // the closed fan's triangle 0 joins the last neighbour (slot F)
#pragma unroll
        for (int s = 1; s < W; ++s) v[s] = (closed && s == F) ? v[s] + vp : v[s];
        const double ci = r.C;
        const double rhs = r.thm - r.thn - ci * r.hm;
        v[0] += a.inv_dt * msum * ci;
        bi -= a.inv_dt * msum * rhs;
    }
#pragma unroll
    for (int s = 0; s < W; ++s) {
        if (s <= F && s < a.ell_w) {
            const size_t o = (size_t)s * a.ell_n + i;
            a.A[o] = v[s];
            if (a.A32 != nullptr) a.A32[o] = (float)v[s];
        }
    }
    a.b[i] = bi;
    const double dgv = v[0];
    const double di = 1.0 / (dgv != 0.0 ? dgv : 1.0);
    if (a.dinv != nullptr) a.dinv[i] = di;
    if (a.dinv32 != nullptr) a.dinv32[i] = 1.0f / (dgv != 0.0 ? (float)dgv : 1.0f);
    if constexpr (MODE != ASM_INIT_NONE) {
        double ax = 0.0;
#pragma unroll
        for (int s = 0; s < W; ++s)
            if (s <= F) ax += v[s] * lx0[r.lc[s]];
        krylov_start<MODE>(a, ki, i, bi, ax, di, acc);
    }
}

// The node window of a row block: up to WIN_SEG segments (lo, len) of
// consecutive nodes, staged at their running offsets in the x, y, K, x0
// planes of wl, WIN_CAP doubles apart.  Each thread loads its (at most
// WIN_CAP / BLOCK rounded up) window positions at once; positions past the
// window load the first node again and are not stored.  The caller's barrier
// publishes the planes.
constexpr int WIN_PER = (WIN_CAP + BLOCK - 1) / BLOCK;
struct WinStage {
    double2 p[WIN_PER];
    double kk[WIN_PER], xx[WIN_PER];
    int tot;
};

__device__ inline void load_win_words(const AsmArgs& a, int blk, int2 (&w)[WIN_SEG]) {
    const int2* src = reinterpret_cast<const int2*>(a.nwin + (size_t)WIN_WORDS * blk);
#pragma unroll
    for (int k = 0; k < WIN_SEG; ++k) w[k] = src[k];
}

template <int MODE>
__device__ inline void stage_load(const AsmArgs& a, const KrylovInit& ki, const int2 (&w)[WIN_SEG], WinStage& st) {
    static_assert(WIN_SEG == 3, "three segments");
    const double2* xy = reinterpret_cast<const double2*>(a.nrec.coord);
    const int l0 = w[0].y, l1 = w[1].y;
    st.tot = l0 + l1 + w[2].y;
#pragma unroll
    for (int m = 0; m < WIN_PER; ++m) {
        const int q = (int)threadIdx.x + m * BLOCK;
        int g = q < l0 ? w[0].x + q : (q < l0 + l1 ? w[1].x + (q - l0) : w[2].x + (q - l0 - l1));
        g = q < st.tot ? g : w[0].x;
        st.p[m] = xy[g];
        st.kk[m] = a.nrec.K[g];
        if constexpr (MODE != ASM_INIT_NONE) st.xx[m] = ki.x0[g];
    }
}

template <int MODE>
__device__ inline void stage_store(const WinStage& st, double* wl) {
#pragma unroll
    for (int m = 0; m < WIN_PER; ++m) {
        const int q = (int)threadIdx.x + m * BLOCK;
        if (q < st.tot) {
            wl[q] = st.p[m].x;
            wl[WIN_CAP + q] = st.p[m].y;
            wl[2 * WIN_CAP + q] = st.kk[m];
            if constexpr (MODE != ASM_INIT_NONE) wl[3 * WIN_CAP + q] = st.xx[m];
        }
    }
}

// Everything a row block reads from global memory past its window words,
// issued in dependency order (the row's inputs, then the window and the Ks),
// the window stored to LDS, then the barrier.
template <int MODE>
__device__ inline void load_block(const AsmArgs& a, const KrylovInit& ki, const int2 (&w)[WIN_SEG], int i, WinRow& r,
                                  double* wl) {
    int code[7];
    load_win_row(a, i, r, code);
    WinStage st;
    stage_load<MODE>(a, ki, w, st);
    load_win_ks(a, code, r);
    stage_store<MODE>(st, wl);
    __syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_assemble_win(AsmArgs a, KrylovInit ki, double* partials, int prow) {
    // the window's planes at fixed offsets (one address per node, immediate
    // offsets for the planes)
    __shared__ double wl[4 * WIN_CAP];
    const double* lx = wl;
    const double* ly = wl + WIN_CAP;
    const double* lk = wl + 2 * WIN_CAP;
    const double* lx0 = wl + 3 * WIN_CAP;
    // quadrature weights by local vertex (read after the window's barrier)
    __shared__ __align__(16) double qw[12];
    fan_weights(a.qc, qw);
    if (a.tab == nullptr) {
        if (a.gate != nullptr && *a.gate == 0) return;   // uniform: a Picard iteration enqueued ahead, not needed
        if (a.tflag != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *a.tflag = a.n;
        // rows of logical block L, as in k_assemble_reg; its window is block L's
        const int L = xcd_block(blockIdx.x, gridDim.x);
        const int i = L * BLOCK + threadIdx.x;
        int2 w[WIN_SEG];
        load_win_words(a, L, w);
        WinRow r;
        load_block<MODE>(a, ki, w, i < a.n ? i : a.n - 1, r, wl);
        double acc[3] = {0.0, 0.0, 0.0};
        if (i < a.n) assemble_row_win<MODE>(a, ki, i, r, lx, ly, lk, lx0, qw, acc);
        if constexpr (MODE != ASM_INIT_NONE) store_partials<3>(acc, partials, prow, L);
        return;
    }
    // ensemble: the walk of k_assemble_reg, the window of each entry
    int v0, v1, vstep;
    xcd_walk(a.act ? a.act[0] : a.n_tab, v0, v1, vstep);
    for (int v = v0; v < v1; v += vstep) {   // uniform over the workgroup
        const int t = __builtin_amdgcn_readfirstlane(a.act ? a.act[1 + v] : v);
        int2 w[WIN_SEG];
        load_win_words(a, t, w);
        const int4 tb = a.tab[t];
        const int r0 = __builtin_amdgcn_readfirstlane(tb.x), r1 = __builtin_amdgcn_readfirstlane(tb.y);
        // the instance's Picard flag is read with the entry's data (listed
        // entries are not converged; the flag only guards)
        const int pdv = a.ens_st[__builtin_amdgcn_readfirstlane(tb.z)].pdone;
        const int i = r0 + threadIdx.x;
        WinRow r;
        load_block<MODE>(a, ki, w, i < r1 ? i : r1 - 1, r, wl);
        if (!__builtin_amdgcn_readfirstlane(pdv)) {   // uniform
            if (a.tflag != nullptr && threadIdx.x == 0) atomicAdd(a.tflag, r1 - r0);
            double acc[3] = {0.0, 0.0, 0.0};
            if (i < r1) assemble_row_win<MODE>(a, ki, i, r, lx, ly, lk, lx0, qw, acc);
            if constexpr (MODE != ASM_INIT_NONE) store_partials<3>(acc, partials, prow, t);
        }
        __syncthreads();   // the next entry's window replaces this one
    }
}

template <int MODE>
static void launch_win_mode(dim3 g, hipStream_t s, const AsmArgs& a, const KrylovInit& ki, double* partials,
                            int prow) {
    hipLaunchKernelGGL((k_assemble_win<MODE>), g, dim3(BLOCK), 0, s, a, ki, partials, prow);
}

static hipError_t launch_win(int mode, dim3 g, hipStream_t s, const AsmArgs& a, const KrylovInit& ki,
                             double* partials, int prow) {
    switch (mode) {
        case ASM_INIT_BICG: launch_win_mode<ASM_INIT_BICG>(g, s, a, ki, partials, prow); break;
        case ASM_INIT_CG: launch_win_mode<ASM_INIT_CG>(g, s, a, ki, partials, prow); break;
        case ASM_INIT_GMRES: launch_win_mode<ASM_INIT_GMRES>(g, s, a, ki, partials, prow); break;
        case ASM_INIT_RESID: launch_win_mode<ASM_INIT_RESID>(g, s, a, ki, partials, prow); break;
        default: launch_win_mode<ASM_INIT_NONE>(g, s, a, ki, partials, prow); break;
    }
    return hipGetLastError();
}

// NB: 0 incidence gathers; 1, 2 neighbourhood gathers with LDS = NB - 1; 3 fan order
// (qw: the fan kernel's quadrature weights by local vertex, fan_weights)
template <int D, int MAXR, int MODE, int NB>
__device__ inline void assemble_row_any(const AsmArgs& a, const KrylovInit& ki, int i, const double* qw,
                                        double (&acc)[3]) {
    if constexpr (NB == 3 && D == 2) assemble_row_fan<MODE>(a, ki, i, qw, acc);
    else if constexpr (NB > 0 && D == 2) assemble_row_nbr<MAXR, MODE, NB - 1>(a, ki, i, acc);
    else assemble_row<D, MAXR, MODE>(a, ki, i, acc);
}

template <int D, int MAXR, int MODE, int NBR>
__global__ __launch_bounds__(BLOCK) void k_assemble_reg(AsmArgs a, KrylovInit ki, double* partials, int prow) {
    __shared__ __align__(16) double qw[NBR == 3 ? 12 : 1];
    if constexpr (NBR == 3) {
        fan_weights(a.qc, qw);
        __syncthreads();
    }
    if (a.tab == nullptr) {
        if (a.gate != nullptr && *a.gate == 0) return;   // uniform: a Picard iteration enqueued ahead, not needed
        if (a.tflag != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *a.tflag = a.n;
        // rows of logical block L (each XCD assembles one contiguous range of
        // rows, whose neighbours' records then come through one L2); its
        // partial sums in column L, as without the map
        const int L = xcd_block(blockIdx.x, gridDim.x);
        const int i = L * BLOCK + threadIdx.x;
        double acc[3] = {0.0, 0.0, 0.0};
        if (i < a.n) assemble_row_any<D, MAXR, MODE, NBR>(a, ki, i, qw, acc);
        if constexpr (MODE != ASM_INIT_NONE) store_partials<3>(acc, partials, prow, L);
        return;
    }
    // ensemble: fixed grid over the listed table entries (instances still
    // iterating Picard), each XCD on one block of the list, partial sums in
    // the entry's column
    int v0, v1, vstep;
    xcd_walk(a.act ? a.act[0] : a.n_tab, v0, v1, vstep);
    for (int v = v0; v < v1; v += vstep) {   // uniform over the workgroup
        const int t = a.act ? a.act[1 + v] : v;
        const int4 tb = a.tab[t];
        if (a.ens_st[tb.z].pdone) continue;
        if (a.tflag != nullptr && threadIdx.x == 0) atomicAdd(a.tflag, tb.y - tb.x);
        const int i = tb.x + threadIdx.x;
        double acc[3] = {0.0, 0.0, 0.0};
        if (i < tb.y) assemble_row_any<D, MAXR, MODE, NBR>(a, ki, i, qw, acc);
        if constexpr (MODE != ASM_INIT_NONE) store_partials<3>(acc, partials, prow, t);
    }
}

void compute_quad_const(int dim, QuadConst& qc) {
    // Literal rules of the reference.
    const double tq[3][2] = {{0.166666666666667, 0.166666666666667},
                             {0.166666666666667, 0.666666666666667},
                             {0.666666666666667, 0.166666666666667}};
    const double tw = 0.166666666666667;
    const double A3 = 0.585410196624969, B3 = 0.138196601125011;
    const double qq[4][3] = {{A3, B3, B3}, {B3, A3, B3}, {B3, B3, A3}, {B3, B3, B3}};
    for (int c = 0; c < 4; ++c) qc.ck[c] = qc.ml[c] = 0.0;
    if (dim == 2) {
        for (int q = 0; q < 3; ++q) {
            const double N[3] = {1.0 - tq[q][0] - tq[q][1], tq[q][0], tq[q][1]};
            const double sN = N[0] + N[1] + N[2];
            for (int c = 0; c < 3; ++c) {
                qc.ck[c] += tw * N[c];
                qc.ml[c] += tw * N[c] * sN;
            }
        }
    } else {
        for (int q = 0; q < 4; ++q) {
            const double N[4] = {1.0 - qq[q][0] - qq[q][1] - qq[q][2], qq[q][0], qq[q][1], qq[q][2]};
            const double sN = N[0] + N[1] + N[2] + N[3];
            for (int c = 0; c < 4; ++c) {
                qc.ck[c] += 0.25 * N[c];
                qc.ml[c] += 0.25 * N[c] * sN;
            }
        }
    }
}

int asm_rows_per_block(int dim) { return dim == 2 ? ASM_ROWS_2D : ASM_ROWS_3D; }

size_t asm_lds_bytes(int dim, int max_block_nnz, int max_block_inc) {
    return (2 * (size_t)max_block_nnz + (size_t)max_block_inc * (dim + 3)) * sizeof(double) +
           (size_t)max_block_inc * sizeof(uint32_t);
}

template <int D, int R>
static hipError_t launch_mode(int mode, dim3 g, size_t lds, hipStream_t s, const AsmArgs& a,
                              const KrylovInit& ki, double* partials, int prow) {
    switch (mode) {
        case ASM_INIT_BICG:
            hipLaunchKernelGGL((k_assemble<D, R, ASM_INIT_BICG>), g, dim3(ASM_THREADS), lds, s, a, ki, partials, prow); break;
        case ASM_INIT_CG:
            hipLaunchKernelGGL((k_assemble<D, R, ASM_INIT_CG>), g, dim3(ASM_THREADS), lds, s, a, ki, partials, prow); break;
        case ASM_INIT_GMRES:
            hipLaunchKernelGGL((k_assemble<D, R, ASM_INIT_GMRES>), g, dim3(ASM_THREADS), lds, s, a, ki, partials, prow); break;
        case ASM_INIT_RESID:
            hipLaunchKernelGGL((k_assemble<D, R, ASM_INIT_RESID>), g, dim3(ASM_THREADS), lds, s, a, ki, partials, prow); break;
        default:
            hipLaunchKernelGGL((k_assemble<D, R, ASM_INIT_NONE>), g, dim3(ASM_THREADS), lds, s, a, ki, partials, prow); break;
    }
    return hipGetLastError();
}

template <int D, int MAXR, int NBR>
static hipError_t launch_reg(int mode, dim3 g, hipStream_t s, const AsmArgs& a, const KrylovInit& ki,
                             double* partials, int prow) {
    switch (mode) {
        case ASM_INIT_BICG:
            hipLaunchKernelGGL((k_assemble_reg<D, MAXR, ASM_INIT_BICG, NBR>), g, dim3(BLOCK), 0, s, a, ki, partials, prow); break;
        case ASM_INIT_CG:
            hipLaunchKernelGGL((k_assemble_reg<D, MAXR, ASM_INIT_CG, NBR>), g, dim3(BLOCK), 0, s, a, ki, partials, prow); break;
        case ASM_INIT_GMRES:
            hipLaunchKernelGGL((k_assemble_reg<D, MAXR, ASM_INIT_GMRES, NBR>), g, dim3(BLOCK), 0, s, a, ki, partials, prow); break;
        case ASM_INIT_RESID:
            hipLaunchKernelGGL((k_assemble_reg<D, MAXR, ASM_INIT_RESID, NBR>), g, dim3(BLOCK), 0, s, a, ki, partials, prow); break;
        default:
            hipLaunchKernelGGL((k_assemble_reg<D, MAXR, ASM_INIT_NONE, NBR>), g, dim3(BLOCK), 0, s, a, ki, partials, prow); break;
    }
    return hipGetLastError();
}

int asm_reg_max_row(int dim) { return dim == 2 ? ASM_REG_ROW_2D : ASM_REG_ROW_3D; }

hipError_t launch_assemble(int dim, const AsmArgs& a, int mode, const KrylovInit& ki, double* partials,
                           int prow, size_t lds, int* grid_out, hipStream_t s, int kernel) {
    if (kernel != ASM_K_LDS) {
        const int grid = a.tab ? a.ens_grid : (a.n + BLOCK - 1) / BLOCK;
        if (grid_out) *grid_out = grid;
        if (kernel == ASM_K_WIN) return launch_win(mode, dim3(grid), s, a, ki, partials, prow);
        if (kernel == ASM_K_FAN) return launch_reg<2, ASM_REG_ROW_2D, 3>(mode, dim3(grid), s, a, ki, partials, prow);
        if (kernel == ASM_K_NBR) return launch_reg<2, ASM_REG_ROW_2D, 2>(mode, dim3(grid), s, a, ki, partials, prow);
        if (kernel == ASM_K_NBR_BLEND) return launch_reg<2, ASM_REG_ROW_2D, 1>(mode, dim3(grid), s, a, ki, partials, prow);
        return launch_reg<2, ASM_REG_ROW_2D, 0>(mode, dim3(grid), s, a, ki, partials, prow);
    }
    const int R = asm_rows_per_block(dim);
    const int grid = (a.n + R - 1) / R;
    if (grid_out) *grid_out = grid;
    if (dim == 2) return launch_mode<2, ASM_ROWS_2D>(mode, dim3(grid), lds, s, a, ki, partials, prow);
    return launch_mode<3, ASM_ROWS_3D>(mode, dim3(grid), lds, s, a, ki, partials, prow);
}

}  // namespace wspm
