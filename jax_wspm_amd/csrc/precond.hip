// General preconditioners for gfx950 (see precond.h).
//
// The triangular sweeps are inherently ordered.  Rows are grouped into level
// sets on the host (a row's level is one more than the highest level among
// the rows it depends on), and one 1024-thread workgroup walks the levels,
// each level's rows in parallel, with a workgroup barrier between levels:
// no inter-workgroup waiting, and every row sees exactly the values a
// sequential sweep would give it.  Each row's sum runs over its CSR entries in
// column order.
#include "precond.h"

#include "wspm_internal.h"

namespace wspm {

__device__ inline bool pc_skip(const KState* st, int mode) {
    if (!st || mode == 0) return false;
    return mode == 1 ? st->done != 0 : (st->done != 0 || !st->active);
}

// value of A(i, j) in row i (0 if not in the pattern)
__device__ inline double entry(const PcArgs& a, int i, int j) {
    for (int k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k)
        if (a.cols[k] == j) return a.A[k];
    return 0.0;
}

// Block Jacobi (jax_linear_solver.py:39-87): B = A[2k:2k+2, 2k:2k+2] + 1e-12 I,
// inverted; a trailing 1x1 block for odd n.
__global__ void k_pc_block_setup(PcArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = 2 * k;
    if (s >= a.n) return;
    if (s + 1 < a.n) {
        const double p = entry(a, s, s) + 1e-12, q = entry(a, s, s + 1);
        const double r = entry(a, s + 1, s), t = entry(a, s + 1, s + 1) + 1e-12;
        const double det = p * t - q * r;
        a.binv[4 * k + 0] = t / det;
        a.binv[4 * k + 1] = -q / det;
        a.binv[4 * k + 2] = -r / det;
        a.binv[4 * k + 3] = p / det;
    } else {
        a.binv[4 * k] = 1.0 / (entry(a, s, s) + 1e-12);
    }
}

// Diagonal; ilu = 1 replaces exact zeros by 1 (jax_linear_solver.py:196, 248).
__global__ void k_pc_diag(PcArgs a, int ilu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double d = a.A[a.diag_pos[i]];
    a.dg[i] = (ilu && d == 0.0) ? 1.0 : d;
}

__global__ void k_pc_block_apply(PcArgs a, const double* src, double* dst, const KState* st, int skipmode) {
    if (pc_skip(st, skipmode)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = 2 * k;
    if (s >= a.n) return;
    if (s + 1 < a.n) {
        const double x0 = src[s], x1 = src[s + 1];
        dst[s] = a.binv[4 * k + 0] * x0 + a.binv[4 * k + 1] * x1;
        dst[s + 1] = a.binv[4 * k + 2] * x0 + a.binv[4 * k + 3] * x1;
    } else {
        dst[s] = a.binv[4 * k] * src[s];
    }
}

// FORM 0: SSOR forward  y_i = (x_i - sum_{j<i} A_ij y_j) / d_i, and
//         z_i = (d_i y_i) / d_i (the reference's backward sweep, see the oracle)
// FORM 1: ILU forward   y_i = x_i - sum_{j<i} (A_ij / dg_j) y_j
// FORM 2: ILU backward  y_i = (x_i - sum_{j>i} A_ij y_j) / dg_i
template <int FORM>
__global__ __launch_bounds__(TRSV_THREADS) void k_trsv(PcArgs a, Levels lv, const double* x, double* y, double* z,
                                                       const KState* st, int skipmode) {
    if (pc_skip(st, skipmode)) return;
    for (int L = 0; L < lv.n_lev; ++L) {
        const int k1 = lv.lvl_ptr[L + 1];
        for (int k = lv.lvl_ptr[L] + threadIdx.x; k < k1; k += TRSV_THREADS) {
            const int i = lv.lvl_rows[k];
            double acc = 0.0;
            for (int e = a.row_ptr[i]; e < a.row_ptr[i + 1]; ++e) {
                const int j = a.cols[e];
                if (FORM == 2 ? j > i : j < i) {
                    const double c = FORM == 1 ? a.A[e] / a.dg[j] : a.A[e];
                    acc += c * y[j];
                }
            }
            if (FORM == 0) {
                const double d = a.dg[i];
                const double yi = (x[i] - acc) / d;
                y[i] = yi;
                z[i] = (d * yi) / d;
            } else if (FORM == 1) {
                y[i] = x[i] - acc;
            } else {
                y[i] = (x[i] - acc) / a.dg[i];
            }
        }
        __syncthreads();
    }
}

// Jacobi inverse diagonal, 0 -> 1 (jax_linear_solver.py:12-36).
__global__ void k_pc_jacobi(PcArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double d = a.A[a.diag_pos[i]];
    a.dg[i] = 1.0 / (d != 0.0 ? d : 1.0);
}

hipError_t launch_pc_setup(int kind, const PcArgs& a, hipStream_t s) {
    const int b = 256;
    if (kind == WSPM_PRECOND_JACOBI) {
        hipLaunchKernelGGL(k_pc_jacobi, dim3((a.n + b - 1) / b), dim3(b), 0, s, a);
    } else if (kind == WSPM_PRECOND_BLOCK_JACOBI) {
        const int pairs = (a.n + 1) / 2;
        hipLaunchKernelGGL(k_pc_block_setup, dim3((pairs + b - 1) / b), dim3(b), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_pc_diag, dim3((a.n + b - 1) / b), dim3(b), 0, s, a, kind == WSPM_PRECOND_ILU ? 1 : 0);
    }
    return hipGetLastError();
}

hipError_t launch_pc_apply(int kind, const PcArgs& a, const Levels& lo, const Levels& up, const double* src,
                           double* dst, double* tmp, const KState* st, int skipmode, hipStream_t s) {
    if (kind == WSPM_PRECOND_BLOCK_JACOBI) {
        const int pairs = (a.n + 1) / 2, b = 256;
        hipLaunchKernelGGL(k_pc_block_apply, dim3((pairs + b - 1) / b), dim3(b), 0, s, a, src, dst, st, skipmode);
    } else if (kind == WSPM_PRECOND_SSOR) {
        hipLaunchKernelGGL(k_trsv<0>, dim3(1), dim3(TRSV_THREADS), 0, s, a, lo, src, tmp, dst, st, skipmode);
    } else {
        hipLaunchKernelGGL(k_trsv<1>, dim3(1), dim3(TRSV_THREADS), 0, s, a, lo, src, tmp, nullptr, st, skipmode);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_trsv<2>, dim3(1), dim3(TRSV_THREADS), 0, s, a, up, tmp, dst, nullptr, st, skipmode);
    }
    return hipGetLastError();
}

}  // namespace wspm
