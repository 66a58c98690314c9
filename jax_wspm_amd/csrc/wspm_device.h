// Device-side building blocks: deterministic reductions, soil models.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wspm {

constexpr int BLOCK = 256;          // threads per workgroup (4 wavefronts of 64)
constexpr int NWAVE = BLOCK / 64;
constexpr int MAX_BLOCKS = 2048;    // cap on the grids of SpMV / vector reductions
constexpr int ELL_W3D = 16;         // widest 3D ELL row (a tetrahedral grid's interior rows have 15)
constexpr int MAX_DOTS = 8;         // dot products per reduction kernel
constexpr int MAX_RESTART = 64;

// Soil model and parameters (phi layout of the C ABI).
struct SoilParams {
    int model;        // 0 gardner, 1 van Genuchten
    double p[6];
};

// Scalars of one Krylov solve, resident in device memory.  Written only by
// the single finishing workgroup of a reduction kernel, read by all later
// kernels of the stream.
struct KState {
    double rho, alpha, omega, gamma, beta;   // CG / BiCGSTAB recurrences
    double rho_next;                         // r_hat . r of the newest r
    double bs, atol2, rs, ss;                // |b|^2, stop threshold, |r|^2, |s|^2
    double bnorm, mbnorm, ptol, rnorm;       // GMRES
    double v0, v1, qnorm;                    // GMRES Arnoldi step
    double scale;                            // mixed precision residual scale
    double picard_err, picard_d2, picard_h2;
    double tol, atol;
    int32_t maxiter, k, steps, done;
    int32_t exit_early, pass2, active, code;
    int32_t cycles, arnoldi, restart, nan_flag;
    int32_t jdone, refine;
    int32_t rdone;                           // ensemble mixed precision: refinement converged
    int32_t rstall;                          // mixed precision: a Jacobi sweep failed to halve |r|
    // Ensemble mode (one KState per instance): Picard control of the instance
    // within the current time step, and its soil parameters.
    int32_t pdone;                           // Picard converged or capped this step
    int32_t pev;                             // Picard error of this iteration taken
    int32_t pit, kry;                        // Picard / Krylov iterations this step
    double picard_tol;
    int32_t picard_maxit, picard_rel;
    SoilParams sp;
    double hbuf[MAX_RESTART + 1];            // CGS projection of the current pass
    double y[MAX_RESTART];                   // GMRES least-squares solution
};

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin by index,
// and each XCD has its own L2.  xcd_block maps a workgroup's index to a
// logical index such that every XCD receives one contiguous range of logical
// blocks: neighbouring rows (which gather the same node and element records)
// then share one L2 instead of being fetched by several.  A bijection of
// [0, nb) for any nb.  Used by the one-system SpMV, vector and assembly
// kernels (it brings the SpMV's fetched bytes down to its algorithmic
// bytes); the ensemble walks use xcd_walk below.
constexpr int NXCD = 8;
__device__ inline int xcd_block(int b, int nb) {
    const int q = nb / NXCD, r = nb % NXCD;
    const int x = b % NXCD, k = b / NXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// ---------------------------------------------------------------------------
// Deterministic reductions.  Wavefront butterfly (fixed xor order), then the
// NWAVE wave sums added in wave order by thread 0.  Grid level: every
// workgroup stores its partial row with plain stores; a one-workgroup finish
// kernel, launched right after on the same stream, adds the rows in workgroup
// order and updates the solver scalars.  The kernel boundary is the only
// device-wide synchronisation: no atomics, no in-kernel fences (an agent-scope
// release on gfx950 writes back the XCD's L2, which for a streaming kernel is
// most of its output, once per workgroup).  Results depend only on the launch
// shape, never on scheduling.
// ---------------------------------------------------------------------------
template <int ND, int NW = NWAVE>
__device__ inline void block_sum(double (&v)[ND], double* sh) {
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        double s = v[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        v[j] = s;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) sh[j * NW + wid] = v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double s = 0.0;
            for (int w = 0; w < NW; ++w) s += sh[j * NW + w];
            v[j] = s;
        }
    }
}

// Node data of the element kernels: static coordinates (2D: (x, y) as
// double2, 16 bytes a node; 3D: (x, y, z, -) as double4) and the nodal
// conductivity K in an array of its own, rewritten by the soil-model kernels
// every Picard iteration.  (K inside a 32-byte coordinate record made those
// stores partial-sector writes: a Picard update over 7.2M nodes took 145
// against 102 us with K contiguous, on MI355X.)  rec<D>(j) is (x, y, K, -)
// in 2D and (x, y, z, K) in 3D; xy(j) the first two coordinates.
struct NodeRecs {
    const double* coord;
    const double* K;
    int dim;
    template <int D>
    __device__ double4 rec(int j) const {
        if constexpr (D == 2) {
            const double2 p = reinterpret_cast<const double2*>(coord)[j];
            return make_double4(p.x, p.y, K[j], 0.0);
        } else {
            double4 r = reinterpret_cast<const double4*>(coord)[j];
            r.w = K[j];
            return r;
        }
    }
    __device__ double2 xy(int j) const {
        return *reinterpret_cast<const double2*>(coord + (size_t)j * (dim == 2 ? 2 : 4));
    }
};

struct Red {
    double* partials;     // column-major [MAX_DOTS][prow]: dot j of workgroup b at j * prow + b
    int prow;
    KState* st;           // solver state (ensemble: array indexed by instance)
};

// Ensemble row table: entry t covers rows [x, y) of instance z; w is the
// instance's row count in its first entry, 0 in the others.  The entries of
// one instance are consecutive; iblk[i] is the first of instance i
// (iblk[n_inst] = number of entries).  Partial sums of entry t go to column t.
using RowTab = int4;

// How an ensemble launch walks the table: a fixed grid strides over the
// entries listed in act (act[0] = count, act[1..count] = entry indices,
// rebuilt on the device as instances converge), or over all n_tab entries
// when act is null.  tab == nullptr: not an ensemble launch.
struct EnsWalk {
    const RowTab* tab;
    const int* act;
    int n_tab;
    __device__ int count() const { return act ? act[0] : n_tab; }
    __device__ int entry(int v) const { return act ? act[1 + v] : v; }
};

// The listed positions [v0, v1) this workgroup walks, with stride step, in
// an ensemble launch of cnt listed entries.  Every XCD takes one contiguous
// block of ceil(cnt / 8) positions, walked by its own workgroups (index b
// runs on XCD b % 8), so the neighbouring entries of an instance -- whose
// rows gather each other's node records and x values -- share one L2, and
// the XCDs get equal shares of the list whatever its length.  Each entry
// still writes its own partial-row column, so results do not depend on
// the walk.  Grids below 8 workgroups walk the list plainly.
__device__ inline void xcd_walk(int cnt, int& v0, int& v1, int& step) {
    const int nb = gridDim.x;
    if (nb < NXCD) {
        v0 = blockIdx.x;
        v1 = cnt;
        step = nb;
        return;
    }
    const int x = blockIdx.x % NXCD, k = blockIdx.x / NXCD;
    const int chunk = (cnt + NXCD - 1) / NXCD;
    v0 = x * chunk + k;
    v1 = min(cnt, (x + 1) * chunk);
    step = nb / NXCD + (x < nb % NXCD ? 1 : 0);
}

// Store this workgroup's ND block sums in column bid of the partials
// (default: the workgroup's index; NW = wavefronts per workgroup).
template <int ND, int NW = NWAVE>
__device__ inline void store_partials(double (&v)[ND], double* partials, int prow, int bid = -1) {
    __shared__ double sh[ND * NW];
    block_sum<ND, NW>(v, sh);
    if (threadIdx.x == 0) {
        const unsigned b = bid >= 0 ? (unsigned)bid : blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
        for (int j = 0; j < ND; ++j) partials[(size_t)j * prow + b] = v[j];
    }
}

// acc[j] += p[j * prow + b] for this thread's columns b = b0 + threadIdx.x,
// + BLOCK, ... below b1, in that order.  The loads of SUM_U columns are
// issued before their adds, so a thread waits for one memory round trip per
// SUM_U columns instead of per column; each acc[j] still adds its columns in
// column order, so the sums are bitwise those of the one-column loop.
constexpr int SUM_U = 8;
template <int ND>
__device__ inline void strided_sum(const double* __restrict__ p, int prow, int b0, int b1, double (&acc)[ND]) {
    int b = b0 + (int)threadIdx.x;
    for (; b + (SUM_U - 1) * BLOCK < b1; b += SUM_U * BLOCK) {
        double v[SUM_U][ND];
#pragma unroll
        for (int u = 0; u < SUM_U; ++u)
#pragma unroll
            for (int j = 0; j < ND; ++j) v[u][j] = p[(size_t)j * prow + b + u * BLOCK];
#pragma unroll
        for (int u = 0; u < SUM_U; ++u)
#pragma unroll
            for (int j = 0; j < ND; ++j) acc[j] += v[u][j];
    }
    for (; b < b1; b += BLOCK) {
#pragma unroll
        for (int j = 0; j < ND; ++j) acc[j] += p[(size_t)j * prow + b];
    }
}

// Sum nblk partials per dot in workgroup order and run op.fin(totals, state).
// One workgroup; the reads are coalesced.
template <int ND, class Op>
__global__ __launch_bounds__(BLOCK) void k_finish(Op op, const double* __restrict__ partials, int prow, int nblk,
                                                   KState* st) {
    if (op.skip(st)) return;
    double acc[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) acc[j] = 0.0;
    strided_sum<ND>(partials, prow, 0, nblk, acc);
    __shared__ double sh[ND * NWAVE];
    block_sum<ND>(acc, sh);
    if (threadIdx.x == 0) op.fin(acc, st);
}

// Sum of the partial columns [b0, b1) per dot in the order every ensemble
// finish uses (thread-strided columns, then block_sum); the total is in
// thread 0's acc.
template <int ND>
__device__ inline void ens_reduce(const double* __restrict__ partials, int prow, int b0, int b1, double (&acc)[ND]) {
#pragma unroll
    for (int j = 0; j < ND; ++j) acc[j] = 0.0;
    strided_sum<ND>(partials, prow, b0, b1, acc);
    __shared__ double sh[ND * NWAVE];
    block_sum<ND>(acc, sh);
}

// Ensemble finish: workgroup i adds the partials of instance i's workgroups
// [iblk[i], iblk[i+1]) in order and runs op.fin on state st[i].  Instances
// whose Picard iteration has converged are left alone.  RESET_PEV: this
// reduction starts a Picard iteration (assembly + solver start).
template <int ND, class Op, bool RESET_PEV = false>
__global__ __launch_bounds__(BLOCK) void k_finish_ens(Op op, const double* __restrict__ partials, int prow,
                                                       const int* __restrict__ iblk, KState* st0) {
    KState* st = st0 + blockIdx.x;
    const bool pd = st->pdone != 0, sk = op.skip(st);   // loaded together
    if (pd | sk) return;
    double acc[ND];
    ens_reduce<ND>(partials, prow, iblk[blockIdx.x], iblk[blockIdx.x + 1], acc);
    if (threadIdx.x == 0) {
        op.fin(acc, st);
        if (RESET_PEV) st->pev = 0;
    }
}

// The K vertex contributions of one incidence added to the row's entries:
// vertex c lands in slot (w >> 8c) & 0xff.  An element's vertices are
// distinct nodes, so a slot receives at most one of them: each slot selects
// its value (or +0.0) and adds once.  That is the sum K one-hot adds per slot
// give, bit for bit (adding +0.0 is exact, and the entries, which start at
// +0.0, never become -0.0), with a third of the FP64 adds (measured: the
// one-hot adds and selects were most of this kernel's VALU instructions).
template <int K, int MAXR>
__device__ inline void add_slots(double (&v)[MAXR], uint32_t w, const double* out) {
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
        double a = 0.0;
#pragma unroll
        for (int c = 0; c < K; ++c) a = (((w >> (8 * c)) & 0xffu) == (uint32_t)s) ? out[c] : a;
        v[s] += a;
    }
}

// ---------------------------------------------------------------------------
// Soil models (src/models/van_genuchten.py:9-47, src/models/exponential.py:10-42)
// ---------------------------------------------------------------------------

__device__ inline void soil_eval(const SoilParams& sp, double h, double& C, double& K, double& th) {
    if (sp.model == 1) {
        const double alpha = sp.p[0], ts = sp.p[1], tr = sp.p[2], n = sp.p[3], m = sp.p[4], ks = sp.p[5];
        if (h < 0.0) {
            // With s = (a|h|)^n = exp(n log(a|h|)) (no pow: its
            // extended-precision logarithm made this kernel compute-bound),
            // L = log(1+s) and Q = log(1+1/s) = L - log s, both from ONE
            // log1p without cancellation: s >= 1: Q = log1p(1/s), L = n
            // log(a|h|) + Q (two non-negative terms); s < 1: L = log1p(s),
            // Q = L - n log(a|h|) (likewise).  Then
            //   Se = (1+s)^(-m) = exp(-m L),  theta = (ts-tr) Se + tr,
            //   1 - (1 - Se^(1/m))^m = 1 - (s/(1+s))^m = -expm1(-m Q),
            //   K = Ks sqrt(Se) (1 - (1 - Se^(1/m))^m)^2,
            //   C = (ts-tr) a m n s/(a|h|) Se/(1+s) (one division).
            // The reference's Mualem form cancels when Se is small (dry soil)
            // and takes Se back from theta - theta_r; this one does neither:
            // C, K and theta are within 1e-14 relative of 40-digit values over
            // a in [0.01, 0.2], n in [1.1, 3], |h| in [1e-3, 1e4] (measured;
            // the reference's K form is within 1e-11 there), and it costs six
            // transcendental calls and two divisions per node instead of
            // seven and four (a Picard update writing contiguous arrays over
            // 7.2M nodes: 136 -> 102 us on MI355X).
            const double ah = alpha * fabs(h);
            const double nl = n * log(ah);
            const double ahn = exp(nl);
            const bool big = ahn >= 1.0;
            const double lp = log1p(big ? 1.0 / ahn : ahn);
            const double L = big ? nl + lp : lp;
            const double Q = big ? lp : lp - nl;
            const double se = exp(-m * L);
            th = (ts - tr) * se + tr;
            const double inner = -expm1(-m * Q);
            K = ks * (sqrt(se) * (inner * inner));
            // a|h| == 0 (underflow): s = 0, Se = 1, s/(a|h|) -> (a|h|)^(n-1)
            const double g = ah > 0.0 ? (ahn * se) / (ah * (1.0 + ahn)) : pow(ah, n - 1.0);
            C = (ts - tr) * alpha * m * n * g;
        } else {
            th = ts;
            K = ks;
            C = 0.0;
        }
    } else {
        const double a0 = sp.p[0], ts = sp.p[1], tr = sp.p[2], ks = sp.p[3];
        if (h <= 0.0) {
            const double e = exp(a0 * h);
            th = (ts - tr) * e + tr;
            K = ks * e;
            C = a0 * (ts - tr) * e;
        } else {
            th = ts;
            K = ks;
            C = 0.0;
        }
    }
}

}  // namespace wspm
