// Device-side building blocks: deterministic reductions, soil models.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wspm {

constexpr int BLOCK = 256;          // threads per workgroup (4 wavefronts of 64)
constexpr int NWAVE = BLOCK / 64;
constexpr int MAX_BLOCKS = 2048;    // cap on reduction grids (partials buffer rows)
constexpr int MAX_DOTS = 72;        // partials buffer columns
constexpr int MAX_RESTART = 64;

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
    int32_t jdone, refine, pad0, pad1;
    double hbuf[MAX_RESTART + 1];            // CGS projection of the current pass
    double y[MAX_RESTART];                   // GMRES least-squares solution
};

// ---------------------------------------------------------------------------
// Deterministic reductions.  Wavefront butterfly (fixed xor order), then the
// NWAVE wave sums added in wave order by thread 0.  The grid-level reduction
// writes one partial row per workgroup; the last workgroup to arrive (integer
// ticket) adds the partials in workgroup order and runs the finishing step.
// The result depends only on the launch shape, never on scheduling.
// ---------------------------------------------------------------------------
template <int ND>
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
        for (int j = 0; j < ND; ++j) sh[j * NWAVE + wid] = v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double s = 0.0;
            for (int w = 0; w < NWAVE; ++w) s += sh[j * NWAVE + w];
            v[j] = s;
        }
    }
}

struct Red {
    double* partials;     // [MAX_BLOCKS][MAX_DOTS]
    unsigned* counter;    // ticket, 0 between kernels
    KState* st;
};

// Sum ND per-thread values over the whole grid; the finishing workgroup's
// thread 0 calls fin(tot) with the grid totals.  Every workgroup of the grid
// must call this exactly once.
template <int ND, class Fin>
__device__ inline void grid_sum(double (&v)[ND], const Red& red, Fin&& fin) {
    __shared__ double sh[(ND > 0 ? ND : 1) * NWAVE];
    __shared__ int is_last;
    block_sum<ND>(v, sh);
    const unsigned nblk = gridDim.x * gridDim.y;
    const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) red.partials[(size_t)bid * MAX_DOTS + j] = v[j];
        __threadfence();
        const unsigned t = atomicAdd(red.counter, 1u);
        is_last = (t == nblk - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double acc[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) acc[j] = 0.0;
    const volatile double* P = red.partials;
    for (unsigned b = threadIdx.x; b < nblk; b += BLOCK) {
#pragma unroll
        for (int j = 0; j < ND; ++j) acc[j] += P[(size_t)b * MAX_DOTS + j];
    }
    block_sum<ND>(acc, sh);
    if (threadIdx.x == 0) {
        fin(acc);
        *red.counter = 0u;
        __threadfence();
    }
}

// ---------------------------------------------------------------------------
// Soil models (src/models/van_genuchten.py:9-47, src/models/exponential.py:10-42)
// ---------------------------------------------------------------------------
struct SoilParams {
    int model;        // 0 gardner, 1 van Genuchten
    double p[6];
};

__device__ inline void soil_eval(const SoilParams& sp, double h, double& C, double& K, double& th) {
    if (sp.model == 1) {
        const double alpha = sp.p[0], ts = sp.p[1], tr = sp.p[2], n = sp.p[3], m = sp.p[4], ks = sp.p[5];
        if (h < 0.0) {
            const double ah = alpha * fabs(h);
            const double ahn = pow(ah, n);
            const double base = 1.0 + ahn;
            th = (ts - tr) * pow(base, -m) + tr;
            const double se = (th - tr) / (ts - tr);
            const double inner = 1.0 - pow(1.0 - pow(se, 1.0 / m), m);
            K = ks * (pow(se, 0.5) * (inner * inner));
            C = (ts - tr) * alpha * m * n * pow(ah, n - 1.0) * pow(base, -m - 1.0);
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
