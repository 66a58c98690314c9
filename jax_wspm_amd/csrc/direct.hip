// Banded LU with partial pivoting for gfx950 (see direct.h).
//
// The factorisation is a sequence of n dependent pivot steps; each step's
// work (pivot search over <= kl + 1 rows, a row exchange and a rank-1 update
// of kl x (kl + ku) entries) is spread over one 1024-thread workgroup, with a
// workgroup barrier between the phases.  No workgroup ever waits on another,
// and every barrier is reached by all threads of the workgroup.  Row-major
// band storage keeps the update's columns contiguous across threads.
#include "direct.h"

namespace wspm {

namespace {

__device__ inline double& AB(const Band& b, int i, int j) {
    return b.ab[(size_t)i * b.W + (j - i + b.kl)];
}

__global__ __launch_bounds__(256) void k_band_fill(Band b, const int32_t* __restrict__ row_ptr,
                                                    const int32_t* __restrict__ cols, const double* __restrict__ A) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= b.n) return;
    const int r = b.perm[i];
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) AB(b, r, b.perm[cols[k]]) = A[k];
}

// Step k (gbtrf without blocking): p = first row of max |a_ik|, i in
// [k, k + kl]; exchange rows k and p over columns [k, k + kl + ku]; scale the
// column below the pivot; subtract the rank-1 product from the trailing block.
__global__ __launch_bounds__(BAND_THREADS) void k_band_factor(Band b) {
    constexpr int NW = BAND_THREADS / 64;
    __shared__ double sv[NW];
    __shared__ int si[NW];
    __shared__ int piv;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int n = b.n, KU = b.kl + b.ku;
    int info = 0;
    for (int k = 0; k < n; ++k) {
        const int m = min(b.kl, n - 1 - k);
        double best = -1.0;
        int bi = k;
        for (int r = t; r <= m; r += BAND_THREADS) {
            const double v = fabs(AB(b, k + r, k));
            if (v > best) {
                best = v;
                bi = k + r;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            sv[wid] = best;
            si[wid] = bi;
        }
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < NW; ++w)
                if (sv[w] > best || (sv[w] == best && si[w] < bi)) {
                    best = sv[w];
                    bi = si[w];
                }
            piv = bi;
            b.ipiv[k] = bi;
            if (!(best > 0.0) && info == 0) info = k + 1;
        }
        __syncthreads();
        const int p = piv;
        const int jend = min(k + KU, n - 1);
        if (p != k) {
            for (int j = k + t; j <= jend; j += BAND_THREADS) {
                const double x = AB(b, k, j);
                AB(b, k, j) = AB(b, p, j);
                AB(b, p, j) = x;
            }
        }
        __syncthreads();
        const double d = AB(b, k, k);
        if (d != 0.0) {
            for (int r = 1 + t; r <= m; r += BAND_THREADS) AB(b, k + r, k) /= d;
            __syncthreads();
            const int w = jend - k;
            for (int q = t; q < m * w; q += BAND_THREADS) {
                const int r = 1 + q / w, j = k + 1 + q % w;
                AB(b, k + r, j) -= AB(b, k + r, k) * AB(b, k, j);
            }
        }
        __syncthreads();
    }
    if (t == 0) *b.info = info;
}

// gbtrs: y = P rhs; forward with the row exchanges and the unit lower
// factor; back substitution with U (bandwidth kl + ku); x = P^T y.
__global__ __launch_bounds__(256) void k_band_solve(Band b, const double* __restrict__ rhs, double* __restrict__ x) {
    const int t = threadIdx.x;
    const int n = b.n, KU = b.kl + b.ku;
    double* y = b.y;
    for (int i = t; i < n; i += 256) y[b.perm[i]] = rhs[i];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (t == 0) {
            const int p = b.ipiv[k];
            if (p != k) {
                const double s = y[k];
                y[k] = y[p];
                y[p] = s;
            }
        }
        __syncthreads();
        const double yk = y[k];
        const int m = min(b.kl, n - 1 - k);
        for (int r = 1 + t; r <= m; r += 256) y[k + r] -= AB(b, k + r, k) * yk;
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        if (t == 0) y[k] /= AB(b, k, k);
        __syncthreads();
        const double yk = y[k];
        const int m = min(KU, k);
        for (int r = 1 + t; r <= m; r += 256) y[k - r] -= AB(b, k - r, k) * yk;
        __syncthreads();
    }
    for (int i = t; i < n; i += 256) x[i] = y[b.perm[i]];
}

}  // namespace

hipError_t launch_band_fill(const Band& b, const int32_t* row_ptr, const int32_t* cols, const double* A,
                            hipStream_t s) {
    hipError_t e = hipMemsetAsync(b.ab, 0, (size_t)b.n * b.W * sizeof(double), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_band_fill, dim3((b.n + 255) / 256), dim3(256), 0, s, b, row_ptr, cols, A);
    return hipGetLastError();
}

hipError_t launch_band_factor(const Band& b, hipStream_t s) {
    hipLaunchKernelGGL(k_band_factor, dim3(1), dim3(BAND_THREADS), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_band_solve(const Band& b, const double* rhs, double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_band_solve, dim3(1), dim3(256), 0, s, b, rhs, x);
    return hipGetLastError();
}

}  // namespace wspm
