// Krylov building blocks for gfx950: CSR SpMV with fused epilogues and fused
// vector updates with deterministic dot products.
//
// Each Krylov step of jax.scipy.sparse.linalg.{cg,bicgstab,gmres} (called by
// the reference at src/solvers/jax_linear_solver.py:282,305,329) is mapped to
// a handful of kernels; every dot product the recurrence needs is computed in
// the kernel that produces its operands, and the recurrence scalars are
// updated on the device by the workgroup that finishes the reduction.  All
// kernels return immediately once the solve is marked done, so the host can
// enqueue several iterations ahead without reading anything back.
#pragma once

#include <type_traits>

#include "wspm_device.h"

namespace wspm {

// A functor may decide in load() that its launch does no row work (a folded
// Krylov kernel whose solve has ended, or whose step is cut short): active()
// false.  Functors without active() always work.
template <class Op>
__device__ inline auto active_impl(const Op& op, int) -> decltype(op.active()) { return op.active(); }
template <class Op>
__device__ inline bool active_impl(const Op&, long) { return true; }
template <class Op>
__device__ inline bool active_of(const Op& op) { return active_impl(op, 0); }

template <class T>
struct KVec {
    int n;
    const T* b;
    T* x;
    const T* x0;
    T *r, *rh, *p, *ph, *q, *sh, *t, *z, *w;
    T* V;         // GMRES basis, column-major, leading dimension ldv
    long ldv;
    const T* dinv;  // Jacobi inverse diagonal or nullptr
    int gpc;        // a general preconditioner is applied between the kernels
                    // (dinv is then null; ph, sh, z are separate vectors)
};

// Kernels with ND > 0 leave one partial row per workgroup; the host follows
// them with k_finish<ND>, which runs the functor's fin() on the totals.
// ---------------------------------------------------------------------------
// CSR SpMV: LANES consecutive lanes own one row (LANES | 64), the row sum is
// finished with a width-LANES butterfly, then lane 0 runs the epilogue.
// ---------------------------------------------------------------------------
// tflag (nullable): the benchmark's launch-timing slot; receives the number of
// rows the launch worked on (skipped instances count none), so timed averages
// count working launches only and weigh them by their rows.
// ew: ensemble walk (ew.tab null: one system; workgroup b owns the contiguous
// rows of logical block xcd_block(b), with the state red.st[0]).
//
// SPMV_UNROLL rows per lane group and pass: their row pointers, then their
// first LANES entries (values + columns), then the x gathers are each issued
// together, so a pass costs three dependent memory round trips for
// SPMV_UNROLL rows instead of for one.
constexpr int SPMV_UNROLL = 4;

// Rows r0, r0 + stride, ... below rend (this lane group's rows).
template <int LANES, class T, class Epi>
__device__ inline void spmv_rows(int r0, int rend, int stride, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci, const T* __restrict__ av, const Epi& epi,
                                 double* acc) {
    constexpr int U = SPMV_UNROLL;
    const int l = threadIdx.x % LANES;
    for (; r0 < rend; r0 += stride * U) {
        int kb[U], ke[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int row = r0 + u * stride;
            kb[u] = row < rend ? rp[row] : 0;
            ke[u] = row < rend ? rp[row + 1] : 0;
        }
        int col[U];
        double val[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = kb[u] + l;
            col[u] = k < ke[u] ? ci[k] : -1;
            val[u] = k < ke[u] ? (double)av[k] : 0.0;
        }
        double s[U];
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = col[u] >= 0 ? val[u] * epi.xv(col[u]) : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            for (int k = kb[u] + l + LANES; k < ke[u]; k += LANES) s[u] += (double)av[k] * epi.xv(ci[k]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int o = LANES / 2; o > 0; o >>= 1) s[u] += __shfl_xor(s[u], o, LANES);
            const int row = r0 + u * stride;
            if (l == 0 && row < rend) epi.row(row, s[u], acc);
        }
    }
}

template <int LANES, class T, class Epi>
__global__ __launch_bounds__(BLOCK) void k_spmv(int n, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci,
                                                 const T* __restrict__ av, Epi epi, Red red, int* tflag,
                                                 EnsWalk ew) {
    constexpr int GPB = BLOCK / LANES;
    constexpr int ND = Epi::ND;
    constexpr int NA = ND > 0 ? ND : 1;
    const int g = threadIdx.x / LANES;
    if (ew.tab == nullptr) {
        if (epi.skip(red.st)) return;
        double acc[NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[j] = 0.0;
        epi.load(red.st);
        if (active_of(epi)) {
            if (tflag != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *tflag = n;
            const int chunk = (n + gridDim.x - 1) / gridDim.x;
            const int rb = xcd_block(blockIdx.x, gridDim.x) * chunk;
            spmv_rows<LANES, T>(rb + g, min(n, rb + chunk), GPB, rp, ci, av, epi, acc);
        }
        if constexpr (ND > 0) store_partials<ND>(acc, red.partials, red.prow);
        return;
    }
    int v0, v1, vstep;
    xcd_walk(ew.count(), v0, v1, vstep);
    for (int v = v0; v < v1; v += vstep) {   // uniform over the workgroup
        const int t = ew.entry(v);
        const RowTab tb = ew.tab[t];
        KState* st = red.st + tb.z;
        // the state's skip flags and the epilogue's scalars in one round trip
        Epi e = epi;
        const bool sk = epi.skip(st);
        e.load(st);
        if (sk) continue;
        if (tflag != nullptr && threadIdx.x == 0 && tb.w > 0) atomicAdd(tflag, tb.w);
        double acc[NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[j] = 0.0;
        spmv_rows<LANES, T>(tb.x + g, tb.y, GPB, rp, ci, av, e, acc);
        if constexpr (ND > 0) store_partials<ND>(acc, red.partials, red.prow, t);
    }
}

// ---------------------------------------------------------------------------
// ELL SpMV: one lane per row.  Entry k of row i sits at k * n + i
// (column-major over the rows, width W = the longest row; shorter rows are
// padded with value 0 and their own column).  A wavefront's loads of entry k
// cover 64 consecutive rows, so every value and column load is coalesced,
// and all W of them are issued together before the W gathers of x: two
// dependent memory round trips per row, no row pointers and no cross-lane
// reduction.  The row sum runs over k in order.
// ---------------------------------------------------------------------------
// CI int32_t: the column table holds columns; int16_t: it holds each
// column's offset from the row (2 bytes an entry instead of 4).
template <int W, class T, class CI, class Epi>
__device__ inline void ell_rows(int rb, int re, int n, const CI* __restrict__ ci, const T* __restrict__ av,
                                const Epi& epi, double* acc) {
    for (int i = rb + (int)threadIdx.x; i < re; i += BLOCK) {
        int c[W];
        double v[W];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if constexpr (std::is_same<CI, int16_t>::value) c[k] = i + (int)ci[(size_t)k * n + i];
            else c[k] = ci[(size_t)k * n + i];
            v[k] = (double)av[(size_t)k * n + i];
        }
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < W; ++k) s += v[k] * epi.xv(c[k]);
        epi.row(i, s, acc);
    }
}

template <int W, class T, class Epi, class CI = int32_t>
__global__ __launch_bounds__(BLOCK) void k_spmv_ell(int n, const CI* __restrict__ ci, const T* __restrict__ av,
                                                     Epi epi, Red red, int* tflag, EnsWalk ew) {
    constexpr int ND = Epi::ND;
    constexpr int NA = ND > 0 ? ND : 1;
    if (ew.tab == nullptr) {
        if (epi.skip(red.st)) return;
        double acc[NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[j] = 0.0;
        epi.load(red.st);
        if (active_of(epi)) {
            if (tflag != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *tflag = n;
            // contiguous chunk per workgroup, a multiple of BLOCK rows
            const int chunk = ((n + gridDim.x - 1) / gridDim.x + BLOCK - 1) / BLOCK * BLOCK;
            const int rb = xcd_block(blockIdx.x, gridDim.x) * chunk;
            ell_rows<W, T, CI>(rb, min(n, rb + chunk), n, ci, av, epi, acc);
        }
        if constexpr (ND > 0) store_partials<ND>(acc, red.partials, red.prow);
        return;
    }
    int v0, v1, vstep;
    xcd_walk(ew.count(), v0, v1, vstep);
    for (int v = v0; v < v1; v += vstep) {   // uniform over the workgroup
        const int t = ew.entry(v);
        const RowTab tb = ew.tab[t];
        KState* st = red.st + tb.z;
        // the state's skip flags and the epilogue's scalars in one round trip
        Epi e = epi;
        const bool sk = epi.skip(st);
        e.load(st);
        if (sk) continue;
        if (tflag != nullptr && threadIdx.x == 0 && tb.w > 0) atomicAdd(tflag, tb.w);
        double acc[NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[j] = 0.0;
        ell_rows<W, T, CI>(tb.x, tb.y, n, ci, av, e, acc);
        if constexpr (ND > 0) store_partials<ND>(acc, red.partials, red.prow, t);
    }
}

// Elementwise kernel with fused dot products.
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_vec(int n, Op op, Red red, EnsWalk ew) {
    constexpr int ND = Op::ND;
    constexpr int NA = ND > 0 ? ND : 1;
    if (ew.tab == nullptr) {
        if (op.skip(red.st)) return;
        double acc[NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[j] = 0.0;
        op.load(red.st);
        if (active_of(op)) {
            // contiguous chunk per workgroup, a multiple of BLOCK rows
            const int chunk = ((n + gridDim.x - 1) / gridDim.x + BLOCK - 1) / BLOCK * BLOCK;
            const int rb = xcd_block(blockIdx.x, gridDim.x) * chunk, re = min(n, rb + chunk);
            for (int i = rb + threadIdx.x; i < re; i += BLOCK) op.row(i, acc);
        }
        if constexpr (ND > 0) store_partials<ND>(acc, red.partials, red.prow);
        return;
    }
    int v0, v1, vstep;
    xcd_walk(ew.count(), v0, v1, vstep);
    for (int v = v0; v < v1; v += vstep) {   // uniform over the workgroup
        const int t = ew.entry(v);
        const RowTab tb = ew.tab[t];
        KState* st = red.st + tb.z;
        // the state's skip flags and the functor's scalars in one round trip
        Op o = op;
        const bool sk = op.skip(st);
        o.load(st);
        if (sk) continue;
        double acc[NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[j] = 0.0;
        for (int i = tb.x + threadIdx.x; i < tb.y; i += BLOCK) o.row(i, acc);
        if constexpr (ND > 0) store_partials<ND>(acc, red.partials, red.prow, t);
    }
}

template <class T>
__device__ inline double mval(const T* dinv, int i, double v) {
    return dinv ? (double)dinv[i] * v : v;
}

// ------------------------------- plain SpMV --------------------------------
template <class T>
struct EpiPlain {
    static constexpr int ND = 0;
    const T* x;
    T* y;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)x[j]; }
    __device__ void row(int i, double ax, double*) const { y[i] = (T)ax; }
    __device__ void fin(const double*, KState*) const {}
};

// ------------------------------- BiCGSTAB ----------------------------------
template <class T>
struct BicgInit {  // r = b - A x0; x = x0; r_hat = p = q = r
    static constexpr int ND = 2;
    KVec<T> v;
    double tol, atol;
    int maxiter;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)v.x0[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        const double bi = (double)v.b[i];
        const T rT = (T)(bi - ax);
        v.x[i] = v.x0[i];
        v.r[i] = rT; v.rh[i] = rT;
        const double rr = (double)rT;
        acc[0] += rr * rr;
        acc[1] += bi * bi;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->tol = tol; s->atol = atol; s->maxiter = maxiter;
        s->bs = t[1];
        s->atol2 = fmax(tol * tol * t[1], atol * atol);
        s->rs = t[0];
        s->rho_next = t[0];
        s->rho = s->alpha = s->omega = 1.0;
        s->k = 0; s->steps = 0; s->code = 0; s->exit_early = 0; s->nan_flag = 0;
        s->done = !(s->rs > s->atol2 && 0 < maxiter);
    }
};

template <class T>
struct BicgP {  // p = r + beta (p - omega q); p_hat = M p
    static constexpr int ND = 0;
    KVec<T> v;
    double beta, omega;
    int first;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) {
        beta = s->rho_next / s->rho * s->alpha / s->omega;
        omega = s->omega;
        first = s->k == 0;
    }
    __device__ void row(int i, double*) const {
        // first iteration: p = q = r0 and omega = 1, so r + beta (p - omega q) = r
        // exactly; the start kernels therefore do not write p and q.
        const double pi = first ? (double)v.r[i]
                                : (double)v.r[i] + beta * ((double)v.p[i] - omega * (double)v.q[i]);
        v.p[i] = (T)pi;
        if (v.dinv) v.ph[i] = (T)((double)v.dinv[i] * (double)(T)pi);
    }
    __device__ void fin(const double*, KState*) const {}
};

template <class T>
struct BicgQ {  // q = A p_hat; alpha = rho_ / (r_hat . q)
    static constexpr int ND = 1;
    KVec<T> v;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)v.ph[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        // the row-local operand is read before the store (q may alias it for
        // all the compiler knows), so its load issues with the SpMV's loads
        const double rh = (double)v.rh[i];
        const T qT = (T)ax;
        v.q[i] = qT;
        acc[0] += rh * (double)qT;
    }
    __device__ void fin(const double* t, KState* s) const { s->alpha = s->rho_next / t[0]; }
};

template <class T>
struct BicgS {  // s = r - alpha q (in place of r); s_hat = M s; |s|^2
    static constexpr int ND = 1;
    KVec<T> v;
    double alpha;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) { alpha = s->alpha; }
    __device__ void row(int i, double* acc) const {
        const T sT = (T)((double)v.r[i] - alpha * (double)v.q[i]);
        v.r[i] = sT;
        const double sd = (double)sT;
        if (v.dinv) v.sh[i] = (T)((double)v.dinv[i] * sd);
        acc[0] += sd * sd;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->ss = t[0];
        s->exit_early = t[0] < s->atol2;
    }
};

template <class T>
struct BicgT {  // t = A s_hat; omega = (t.s)/(t.t)
    static constexpr int ND = 2;
    KVec<T> v;
    __device__ bool skip(const KState* s) const { return (s->done != 0) | (s->exit_early != 0); }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)v.sh[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        const double ri = (double)v.r[i];   // before the store, as in BicgQ
        const T tT = (T)ax;
        v.t[i] = tT;
        const double td = (double)tT;
        acc[0] += td * ri;
        acc[1] += td * td;
    }
    __device__ void fin(const double* t, KState* s) const { s->omega = t[0] / t[1]; }
};

template <class T>
struct BicgXR {  // x, r update; |r|^2, r_hat.r; loop control
    static constexpr int ND = 2;
    KVec<T> v;
    double alpha, omega;
    int early;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) {
        alpha = s->alpha; omega = s->omega; early = s->exit_early;
    }
    __device__ void row(int i, double* acc) const {
        const double phi = (double)v.ph[i];
        double ri;
        if (early) {
            v.x[i] = (T)((double)v.x[i] + alpha * phi);
            ri = (double)v.r[i];
        } else {
            const double shi = (double)v.sh[i];
            v.x[i] = (T)((double)v.x[i] + (alpha * phi + omega * shi));
            const T rT = (T)((double)v.r[i] - omega * (double)v.t[i]);
            v.r[i] = rT;
            ri = (double)rT;
        }
        acc[0] += ri * ri;
        acc[1] += (double)v.rh[i] * ri;
    }
    __device__ void fin(const double* t, KState* s) const {
        int kk = s->k + 1;
        if ((!s->exit_early && s->omega == 0.0) || s->alpha == 0.0) kk = -11;
        if (s->rho_next == 0.0) kk = -10;
        s->k = kk;
        s->code = kk;
        s->steps += 1;
        s->rho = s->rho_next;
        s->rho_next = t[1];
        s->rs = t[0];
        s->done = !(s->rs > s->atol2 && kk < s->maxiter && kk >= 0);
    }
};

// ------------------- CG / BiCGSTAB without finish kernels -------------------
// One system, Jacobi or no preconditioner.  Instead of a one-workgroup
// finish kernel after each reduction, every workgroup of the next kernel sums
// the producer's partial rows itself (reduce_all: the same rows in the same
// order as k_finish, so the scalars are bitwise those of the finish), and the
// scalars of iteration k live in an iteration history H[k] (written by
// workgroup 0 of the kernel that derives them, read by later kernels only).
// A BiCGSTAB iteration is then five kernels instead of nine, CG three instead
// of five.  The solver state (KState) the host reads is written by workgroup
// 0 of the kernel that decides whether the solve goes on.
struct KHist {
    double rho, alpha, omega, rs, gamma;
    int32_t done, early, code, pad;
};

// The producing kernel's partial rows: dot j of workgroup b at p[j * prow + b].
struct FoldIn {
    const double* p;
    int prow, nblk;
};

template <int ND>
__device__ inline void reduce_all(const FoldIn& in, double (&tot)[ND]) {
    double acc[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) acc[j] = 0.0;
    strided_sum<ND>(in.p, in.prow, 0, in.nblk, acc);
    __shared__ double sh[ND * NWAVE];
    __shared__ double bc[ND];
    block_sum<ND>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) bc[j] = acc[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ND; ++j) tot[j] = bc[j];
}

__device__ inline bool lead() { return blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0; }

// Iteration k's first kernel: decide whether the solve goes on (k = 0: from
// the start's state; else from the previous iteration's r-update sums), then
// p = r + beta (p - omega q), p_hat = M p.  With n = 0 (one workgroup) it is
// the end-of-chunk finish that leaves the state for the host.
// start.p != nullptr (k = 0): the solve was started by the assembly epilogue,
// whose partial rows (r0.r0, b.b) this kernel reduces in place of that start's
// finish kernel, setting the state as BicgInit::fin does.
template <class T>
struct FBicgP {
    static constexpr int ND = 0;
    KVec<T> v;
    KHist* H;
    KState* stw;
    FoldIn in;          // BicgXR of iteration k - 1 (or the assembly's start)
    int k;
    double beta, omega;
    int first, go;
    FoldIn start;
    double tol, atol;
    int maxiter;
    const int* gate = nullptr;   // k == 0: *gate == 0 ends the solve before it starts (state untouched)
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState* s) {
        first = k == 0;
        if (k == 0 && gate != nullptr && *gate == 0) {
            go = 0;
            beta = 0.0;
            omega = 1.0;
            if (lead()) H[0].done = 1;
            return;
        }
        if (k == 0 && start.p != nullptr) {
            double t[2];
            reduce_all<2>(start, t);
            const double atol2 = fmax(tol * tol * t[1], atol * atol);
            const int done = !(t[0] > atol2 && 0 < maxiter);
            go = !done;
            beta = 0.0;
            omega = 1.0;
            if (lead()) {
                stw->tol = tol; stw->atol = atol; stw->maxiter = maxiter;
                stw->bs = t[1];
                stw->atol2 = atol2;
                stw->rs = t[0];
                stw->rho_next = t[0];
                stw->rho = stw->alpha = stw->omega = 1.0;
                stw->k = 0; stw->steps = 0; stw->code = 0; stw->exit_early = 0; stw->nan_flag = 0;
                stw->done = done;
                H[0].rho = t[0];
                H[0].rs = t[0];
                H[0].alpha = 1.0;
                H[0].omega = 1.0;
                H[0].done = done;
            }
            return;
        }
        if (k == 0) {
            go = !s->done;
            if (lead()) {
                H[0].rho = s->rho_next;
                H[0].rs = s->rs;
                H[0].alpha = s->alpha;
                H[0].omega = s->omega;
                H[0].done = s->done;
            }
            beta = 0.0;
            omega = s->omega;
            return;
        }
        if (H[k - 1].done) {
            go = 0;
            if (lead()) H[k].done = 1;
            return;
        }
        double t[2];
        reduce_all<2>(in, t);
        const KHist& hp = H[k - 1];
        const int kk = hp.code;
        const double rs = t[0], rho = t[1];
        const int done = !(rs > s->atol2 && kk < s->maxiter && kk >= 0);
        go = !done;
        beta = rho / hp.rho * hp.alpha / hp.omega;
        omega = hp.omega;
        if (lead()) {
            H[k].rho = rho;
            H[k].rs = rs;
            H[k].done = done;
            stw->k = kk;
            stw->code = kk;
            stw->steps = k;
            stw->rho = hp.rho;
            stw->rho_next = rho;
            stw->rs = rs;
            stw->alpha = hp.alpha;
            stw->omega = hp.omega;
            stw->done = done;
        }
    }
    __device__ bool active() const { return go; }
    __device__ void row(int i, double*) const {
        const double pi = first ? (double)v.r[i]
                                : (double)v.r[i] + beta * ((double)v.p[i] - omega * (double)v.q[i]);
        v.p[i] = (T)pi;
        if (v.dinv) v.ph[i] = (T)((double)v.dinv[i] * (double)(T)pi);
    }
    __device__ void fin(const double*, KState*) const {}
};

template <class T>
struct FBicgQ : BicgQ<T> {   // q = A p_hat; r_hat.q
    const KHist* H;
    int k;
    __device__ bool skip(const KState*) const { return H[k].done; }
};

template <class T>
struct FBicgS : BicgS<T> {   // alpha from q's sums; s = r - alpha q; s_hat; |s|^2
    KHist* H;
    FoldIn in;
    int k;
    __device__ bool skip(const KState*) const { return H[k].done; }
    __device__ void load(const KState*) {
        double t[1];
        reduce_all<1>(in, t);
        this->alpha = H[k].rho / t[0];
        if (lead()) H[k].alpha = this->alpha;
    }
};

template <class T>
struct FBicgT : BicgT<T> {   // exit test from |s|^2; t = A s_hat; t.s, t.t
    KHist* H;
    KState* stw;
    FoldIn in;
    int k;
    int go;
    __device__ bool skip(const KState*) const { return H[k].done; }
    __device__ void load(const KState* s) {
        double t[1];
        reduce_all<1>(in, t);
        const int early = t[0] < s->atol2;
        go = !early;
        if (lead()) {
            H[k].early = early;
            stw->ss = t[0];
            stw->exit_early = early;
        }
    }
    __device__ bool active() const { return go; }
};

template <class T>
struct FBicgXR : BicgXR<T> {  // omega from t's sums; x, r update; |r|^2, r_hat.r
    KHist* H;
    FoldIn in;
    int k;
    __device__ bool skip(const KState*) const { return H[k].done; }
    __device__ void load(const KState*) {
        const KHist& h = H[k];
        this->alpha = h.alpha;
        this->early = h.early;
        // early exit: omega keeps its value (t was not formed)
        double om = k > 0 ? H[k - 1].omega : 1.0;
        if (!h.early) {
            double t[2];
            reduce_all<2>(in, t);
            om = t[0] / t[1];
        }
        this->omega = om;
        if (lead()) {
            int kk = k + 1;
            if ((!h.early && om == 0.0) || h.alpha == 0.0) kk = -11;
            if (h.rho == 0.0) kk = -10;
            H[k].omega = om;
            H[k].code = kk;
        }
    }
};

// ----------------------------------- CG ------------------------------------
template <class T>
struct CgInit {  // r = b - A x0; z = M r; p = z
    static constexpr int ND = 3;
    KVec<T> v;
    double tol, atol;
    int maxiter;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)v.x0[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        const double bi = (double)v.b[i];
        const T rT = (T)(bi - ax);
        const double rd = (double)rT;
        v.x[i] = v.x0[i];
        v.r[i] = rT;
        const T zT = v.dinv ? (T)((double)v.dinv[i] * rd) : rT;
        if (v.dinv) v.z[i] = zT;
        v.p[i] = zT;
        acc[0] += rd * (double)zT;
        acc[1] += rd * rd;
        acc[2] += bi * bi;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->tol = tol; s->atol = atol; s->maxiter = maxiter;
        s->bs = t[2];
        s->atol2 = fmax(tol * tol * t[2], atol * atol);
        s->gamma = t[0];
        s->rs = (v.dinv ? t[1] : t[0]);
        s->k = 0; s->steps = 0; s->code = 0; s->nan_flag = 0;
        s->done = !(s->rs > s->atol2 && 0 < maxiter);
    }
};

template <class T>
struct CgAp {  // q = A p; alpha = gamma / (p . q)
    static constexpr int ND = 1;
    KVec<T> v;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)v.p[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        const double pi = (double)v.p[i];   // before the store, as in BicgQ
        const T qT = (T)ax;
        v.q[i] = qT;
        acc[0] += pi * (double)qT;
    }
    __device__ void fin(const double* t, KState* s) const { s->alpha = s->gamma / t[0]; }
};

template <class T>
struct CgXR {  // x += alpha p; r -= alpha q; z = M r; gamma', |r|^2
    static constexpr int ND = 2;
    KVec<T> v;
    double alpha;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) { alpha = s->alpha; }
    __device__ void row(int i, double* acc) const {
        v.x[i] = (T)((double)v.x[i] + alpha * (double)v.p[i]);
        const T rT = (T)((double)v.r[i] - alpha * (double)v.q[i]);
        v.r[i] = rT;
        const double rd = (double)rT;
        const T zT = v.dinv ? (T)((double)v.dinv[i] * rd) : rT;
        if (v.dinv) v.z[i] = zT;
        acc[0] += rd * (double)zT;
        acc[1] += rd * rd;
    }
    __device__ void fin(const double* t, KState* s) const {
        if (!v.gpc) {   // with a general M, gamma = r.z follows in CgPcRZ
            s->beta = t[0] / s->gamma;
            s->gamma = t[0];
        }
        s->rs = (v.dinv || v.gpc) ? t[1] : t[0];
        s->k += 1;
        s->steps = s->k;
        s->done = !(s->rs > s->atol2 && s->k < s->maxiter);
    }
};

template <class T>
struct CgP {  // p = z + beta p (the new p of JAX's body, computed after the test)
    static constexpr int ND = 0;
    KVec<T> v;
    double beta;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) { beta = s->beta; }
    __device__ void row(int i, double*) const {
        const double zi = (v.dinv || v.gpc) ? (double)v.z[i] : (double)v.r[i];
        v.p[i] = (T)(zi + beta * (double)v.p[i]);
    }
    __device__ void fin(const double*, KState*) const {}
};

// CG: the first two kernels of iteration k go on while H[k] (k = 0: the
// start's state) says so; CgP's load decides for iteration k + 1.
__device__ inline bool cg_done(const KHist* H, int k, const KState* s) { return k == 0 ? s->done : H[k].done; }

template <class T>
struct FCgAp : CgAp<T> {      // q = A p; p.q
    const KHist* H;
    int k;
    __device__ bool skip(const KState* s) const { return cg_done(H, k, s); }
};

template <class T>
struct FCgXR : CgXR<T> {      // alpha from p.q; x += alpha p; r -= alpha q; z = M r; r.z, |r|^2
    KHist* H;
    FoldIn in;
    int k;
    __device__ bool skip(const KState* s) const { return cg_done(H, k, s); }
    __device__ void load(const KState* s) {
        double t[1];
        reduce_all<1>(in, t);
        const double gamma = k == 0 ? s->gamma : H[k].gamma;
        this->alpha = gamma / t[0];
        if (lead()) {
            H[k].alpha = this->alpha;
            if (k == 0) H[0].gamma = gamma;
        }
    }
};

template <class T>
struct FCgP {                 // beta, |r|^2 and the loop test from CgXR's sums; p = z + beta p
    static constexpr int ND = 0;
    KVec<T> v;
    KHist* H;
    KState* stw;
    FoldIn in;
    int k;
    double beta;
    int go;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState* s) {
        if (cg_done(H, k, s)) {   // ended earlier: carry the flag to iteration k + 1
            go = 0;
            if (lead()) H[k + 1].done = 1;
            return;
        }
        double t[2];
        reduce_all<2>(in, t);
        const double gamma = H[k].gamma;
        const double g1 = t[0];
        const double rs = v.dinv ? t[1] : t[0];
        beta = g1 / gamma;
        const int kk = k + 1;
        const int done = !(rs > s->atol2 && kk < s->maxiter);
        go = !done;
        if (lead()) {
            H[k + 1].gamma = g1;
            H[k + 1].rs = rs;
            H[k + 1].done = done;
            stw->beta = beta;
            stw->gamma = g1;
            stw->alpha = H[k].alpha;
            stw->rs = rs;
            stw->k = kk;
            stw->steps = kk;
            stw->done = done;
        }
    }
    __device__ bool active() const { return go; }
    __device__ void row(int i, double*) const {
        const double zi = v.dinv ? (double)v.z[i] : (double)v.r[i];
        v.p[i] = (T)(zi + beta * (double)v.p[i]);
    }
    __device__ void fin(const double*, KState*) const {}
};

// ---------------------------------- GMRES ----------------------------------
__device__ inline double safe_norm(double sq, double thresh) {
    const double nrm = sqrt(sq);
    return nrm > thresh ? nrm : 0.0;
}

template <class T>
struct GmInit {  // w = M (b - A x0); x = x0; |b|, |M b|, |w|
    static constexpr int ND = 3;
    KVec<T> v;
    double tol, atol;
    int maxiter, restart;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)v.x0[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        const double bi = (double)v.b[i];
        const T rT = (T)(bi - ax);
        const T wT = v.dinv ? (T)((double)v.dinv[i] * (double)rT) : rT;
        v.x[i] = v.x0[i];
        v.w[i] = wT;
        const double mb = (double)(v.dinv ? (T)((double)v.dinv[i] * bi) : (T)bi);
        acc[0] += bi * bi;
        acc[1] += mb * mb;
        acc[2] += (double)wT * (double)wT;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->tol = tol; s->atol = atol; s->maxiter = maxiter; s->restart = restart;
        s->bnorm = sqrt(t[0]);
        const double atol_eff = fmax(tol * s->bnorm, atol);
        s->mbnorm = sqrt(t[1]);
        s->ptol = s->mbnorm * fmin(1.0, atol_eff / s->bnorm);
        s->rnorm = safe_norm(t[2], 2.220446049250313e-16);
        s->k = 0; s->cycles = 0; s->arnoldi = 0; s->code = 0; s->nan_flag = 0;
        s->done = !(0 < maxiter && s->rnorm > s->ptol);
    }
};

template <class T>
struct GmResid {  // after a cycle: w = M (b - A x); |w|; loop control
    static constexpr int ND = 1;
    KVec<T> v;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return (double)v.x[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        const T rT = (T)((double)v.b[i] - ax);
        const T wT = v.dinv ? (T)((double)v.dinv[i] * (double)rT) : rT;
        v.w[i] = wT;
        acc[0] += (double)wT * (double)wT;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->k += 1;
        s->cycles = s->k;
        if (v.gpc) return;   // |M r| and the loop test follow in GmPcResid
        s->rnorm = safe_norm(t[0], 2.220446049250313e-16);
        s->done = !(s->k < s->maxiter && s->rnorm > s->ptol);
    }
};

template <class T>
struct GmScale {  // V[:, 0] = w / rnorm (0 when rnorm == 0)
    static constexpr int ND = 0;
    KVec<T> v;
    double rn;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) { rn = s->rnorm; }
    __device__ void row(int i, double*) const {
        v.V[i] = rn > 0.0 ? (T)((double)v.w[i] / rn) : (T)0;
    }
    __device__ void fin(const double*, KState*) const {}
};

template <class T>
struct GmArnoldi {  // w = M (A V[:, j]); v0 = |w|
    static constexpr int ND = 1;
    KVec<T> v;
    int j;
    __device__ bool skip(const KState* s) const { return (s->done != 0) | (s->active == 0); }
    __device__ void load(const KState*) {}
    __device__ double xv(int c) const { return (double)v.V[(long)j * v.ldv + c]; }
    __device__ void row(int i, double ax, double* acc) const {
        const T wT = v.dinv ? (T)((double)v.dinv[i] * (double)(T)ax) : (T)ax;
        v.w[i] = wT;
        acc[0] += (double)wT * (double)wT;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->v0 = safe_norm(t[0], 2.220446049250313e-16);
    }
};

template <class T>
struct GmUpdate {  // w -= V[:, :j+1] h; v1 = |w| (threshold eps * v0); H row j
    static constexpr int ND = 1;
    KVec<T> v;
    int j;
    double* H;   // [restart][restart+1], row-major
    __device__ bool skip(const KState* s) const { return (s->done != 0) | (s->active == 0); }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double* acc) const {
        double wi = (double)v.w[i];
        double qh = 0.0;
        for (int l = 0; l <= j; ++l) qh += (double)v.V[(long)l * v.ldv + i] * hb[l];
        const T wT = (T)(wi - qh);
        v.w[i] = wT;
        acc[0] += (double)wT * (double)wT;
    }
    const double* hb;
    __device__ void fin(const double* t, KState* s) const {
        const int R = s->restart;
        const double v1 = safe_norm(t[0], 2.220446049250313e-16 * s->v0);
        s->v1 = v1;
        double* Hj = H + (long)j * (R + 1);
        for (int l = 0; l <= R; ++l) Hj[l] = (l <= j) ? s->hbuf[l] : 0.0;
        Hj[j + 1] = v1;
        s->arnoldi += 1;
        s->jdone = j;
        if (v1 == 0.0) s->active = 0;
    }
};

template <class T>
struct GmNormalize {  // V[:, j+1] = w / v1 (0 when v1 == 0)
    static constexpr int ND = 0;
    KVec<T> v;
    int j;
    double v1;
    __device__ bool skip(const KState* s) const { return (s->done != 0) | (s->jdone != j); }
    __device__ void load(const KState* s) { v1 = s->v1; }
    __device__ void row(int i, double*) const {
        v.V[(long)(j + 1) * v.ldv + i] = v1 > 0.0 ? (T)((double)v.w[i] / v1) : (T)0;
    }
    __device__ void fin(const double*, KState*) const {}
};

template <class T>
struct GmUpdX {  // x += V[:, :R] y
    static constexpr int ND = 0;
    KVec<T> v;
    int R;
    const double* y;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double*) const {
        double d = 0.0;
        for (int l = 0; l < R; ++l) d += (double)v.V[(long)l * v.ldv + i] * y[l];
        v.x[i] = (T)((double)v.x[i] + d);
    }
    __device__ void fin(const double*, KState*) const {}
};

// ------------------- general preconditioner steps (gpc) ---------------------
// With a general M (block Jacobi, SSOR, ILU) the host applies M between the
// kernels; these finish the dot products that involve M's output.
template <class T>
struct CgPcStart {  // after z = M r0: p = z; gamma = r.z
    static constexpr int ND = 1;
    KVec<T> v;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double* acc) const {
        const T zi = v.z[i];
        v.p[i] = zi;
        acc[0] += (double)v.r[i] * (double)zi;
    }
    __device__ void fin(const double* t, KState* s) const { s->gamma = t[0]; }
};

template <class T>
struct CgPcRZ {  // after z = M r: beta = r.z / gamma; gamma = r.z
    static constexpr int ND = 1;
    KVec<T> v;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double* acc) const { acc[0] += (double)v.r[i] * (double)v.z[i]; }
    __device__ void fin(const double* t, KState* s) const {
        s->beta = t[0] / s->gamma;
        s->gamma = t[0];
    }
};

template <class T>
struct GmPcStart {  // w = M r0 (in t), M b (in t2): |M b|, |M r0|, loop test
    static constexpr int ND = 2;
    KVec<T> v;
    const T* mr;
    const T* mb;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double* acc) const {
        const T wi = mr[i];
        v.w[i] = wi;
        acc[0] += (double)mb[i] * (double)mb[i];
        acc[1] += (double)wi * (double)wi;
    }
    __device__ void fin(const double* t, KState* s) const {
        const double atol_eff = fmax(s->tol * s->bnorm, s->atol);
        s->mbnorm = sqrt(t[0]);
        s->ptol = s->mbnorm * fmin(1.0, atol_eff / s->bnorm);
        s->rnorm = safe_norm(t[1], 2.220446049250313e-16);
        s->done = !(0 < s->maxiter && s->rnorm > s->ptol);
    }
};

template <class T>
struct GmPcNorm {  // Arnoldi: w = M (A v) (in t); v0 = |w|
    static constexpr int ND = 1;
    KVec<T> v;
    const T* mw;
    __device__ bool skip(const KState* s) const { return (s->done != 0) | (s->active == 0); }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double* acc) const {
        const T wi = mw[i];
        v.w[i] = wi;
        acc[0] += (double)wi * (double)wi;
    }
    __device__ void fin(const double* t, KState* s) const { s->v0 = safe_norm(t[0], 2.220446049250313e-16); }
};

template <class T>
struct GmPcResid {  // after a cycle: w = M (b - A x) (in t); |w|; loop test
    static constexpr int ND = 1;
    KVec<T> v;
    const T* mw;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double* acc) const {
        const T wi = mw[i];
        v.w[i] = wi;
        acc[0] += (double)wi * (double)wi;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->rnorm = safe_norm(t[0], 2.220446049250313e-16);
        s->done = !(s->k < s->maxiter && s->rnorm > s->ptol);
    }
};

}  // namespace wspm
