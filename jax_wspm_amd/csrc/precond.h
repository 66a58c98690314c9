// General preconditioners of src/solvers/jax_linear_solver.py (precond.hip):
// 2x2 block Jacobi, SSOR (omega = 1) and the reference's "ilu0"
// (I + L D^-1)(D + U).  Jacobi itself stays fused into the Krylov kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wspm_device.h"

namespace wspm {

constexpr int TRSV_THREADS = 1024;   // one workgroup walks the level sets

// Level sets of the lower / upper triangular dependencies of the pattern
// (row i depends on its columns j < i, resp. j > i); rows of one level are
// independent.  lvl_ptr[L]..lvl_ptr[L+1] index lvl_rows.
struct Levels {
    int n_lev = 0;
    int32_t* lvl_ptr = nullptr;
    int32_t* lvl_rows = nullptr;
};

struct PcArgs {
    int n;
    const int32_t* row_ptr;
    const int32_t* cols;
    const int32_t* diag_pos;
    const double* A;
    double* binv;          // block Jacobi: 4 per pair (row-major 2x2)
    double* dg;            // diagonal (ILU: 0 replaced by 1)
};

// skipmode: 0 never; 1 when st->done; 2 when st->done or !st->active
hipError_t launch_pc_setup(int kind, const PcArgs& a, hipStream_t s);
// dst = M src (dst != src); tmp: [n] scratch for the two-sweep forms.
hipError_t launch_pc_apply(int kind, const PcArgs& a, const Levels& lo, const Levels& up, const double* src,
                           double* dst, double* tmp, const KState* st, int skipmode, hipStream_t s);

}  // namespace wspm
