// Direct solve of a system on the context's CSR pattern (direct.hip): banded
// LU with partial pivoting (the LAPACK gbtrf / gbtrs scheme) in a row-major
// band, after a bandwidth-reducing symmetric permutation chosen on the host.
//
// The reference's `direct` solver is a dense LU with partial pivoting
// (jnp.linalg.solve on the densified matrix, src/solvers/solver_types.py:
// 60-63).  Pivoting within the band is the same elimination restricted to
// the rows that can hold nonzeros, so the factor is that of the dense LU of
// the permuted matrix up to the order of its row exchanges.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wspm {

constexpr int BAND_THREADS = 1024;   // one workgroup walks the pivots

struct Band {
    int n, kl, ku;           // ku: upper bandwidth of the pattern; U gets kl + ku after pivoting
    int W;                   // row stride 2 kl + ku + 1; entry (i, j) at i * W + (j - i + kl)
    double* ab;              // [n * W]
    int32_t* ipiv;           // [n] row exchanged with k at step k
    const int32_t* perm;     // [n] new index of original row / column i
    double* y;               // [n] work vector (new numbering)
    int32_t* info;           // device int: 0, or k + 1 for the first zero pivot
};

// ab = P A P^T in band form (A: CSR values on the pattern).
hipError_t launch_band_fill(const Band& b, const int32_t* row_ptr, const int32_t* cols, const double* A,
                            hipStream_t s);
// In-place banded LU with partial pivoting.
hipError_t launch_band_factor(const Band& b, hipStream_t s);
// x = A^{-1} rhs with the factor (rhs and x in the original numbering).
hipError_t launch_band_solve(const Band& b, const double* rhs, double* x, hipStream_t s);

}  // namespace wspm
