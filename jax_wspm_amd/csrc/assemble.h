// Fused Richards assembly: argument block and launcher (assemble.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wspm {

// Quadrature-derived constants, computed on the host from the reference's
// literal points/weights (src/numerics/gauss.py:31-37, gauss_3D.py:29-39):
//   ck[c] = sum_q w_q N_c(q)               conductivity weight of vertex c
//   ml[c] = sum_q w_q N_c(q) sum_d N_d(q)   lumped-mass factor of vertex c
struct QuadConst {
    double ck[4];
    double ml[4];
};

struct AsmArgs {
    int n;
    const int32_t* row_ptr;
    const int32_t* diag_pos;
    const int32_t* inc_ptr;
    const int32_t* inc_code;
    const uint32_t* inc_slot;
    const int32_t* conn;
    const double* X;
    const double* Y;
    const double* Z;
    const double* Kn;        // nodal K (or Kr when Ks_elem is set)
    const double* Ks_elem;   // nullable
    const double* C;
    const double* theta_m;
    const double* theta_n;
    const double* h_m;
    double inv_dt;
    const uint8_t* bc_flag;  // nullable
    const double* bc_val;
    double* A;
    float* A32;              // nullable: also write an FP32 copy of A
    double* b;
    double* dinv;            // nullable
    float* dinv32;           // nullable
    QuadConst qc;
};

void compute_quad_const(int dim, QuadConst& qc);
hipError_t launch_assemble(int dim, const AsmArgs& a, int max_block_nnz, hipStream_t s);

}  // namespace wspm
