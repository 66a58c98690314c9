// Discrete adjoint of the implicit Richards step (adjoint.hip).
//
// The converged step k solves F_k(h_k, h_{k-1}; p) = 0 with, per element e
// and vertex a (mixed form; the C(h) terms of the Picard matrix cancel at the
// fixed point of src/solvers/richards_solver_v2.py:141-168):
//   F_{e,a} = detJ Kbar_e (sum_b G_a.G_b h_b + dN_a/dz) + ml_{e,a} (theta(h_a) - theta(h_prev_a)) / dt
//   Kbar_e  = Ks_e sum_c ck_c Kr(h_c)
// and F_i = h_i - h_D on Dirichlet rows.  Parameters p: log Ks_e per element,
// van Genuchten alpha and n (m = 1 - 1/n).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "assemble.h"
#include "wspm_device.h"

namespace wspm {

// Node quantities of one adjoint step (all [n]).
struct AdjNodal {
    double *kr, *dkr_dh, *dkr_da, *dkr_dn;          // at h_k
    double *dth_dh, *dth_da, *dth_dn;               // at h_k
    double *dthp_da, *dthp_dn;                      // at h_{k-1}
};

struct AdjArgs {
    int n, E;
    const int32_t* row_ptr;
    const int32_t* diag_pos;
    const int32_t* inc_ptr;
    const int2* inc;
    const int4* conn4;
    const double4* nrec;
    const uint8_t* bc_flag;     // nullable
    const double* ks_e;         // [E]
    const double* h;            // h_k
    const double* hp;           // h_{k-1}
    SoilParams sp;              // van Genuchten, Ks = 1 (Kr), m tied to n
    double inv_dt;
    QuadConst qc;
    AdjNodal nd;
    const double* ml;           // [n] global lumped mass
};

void adj_tie_m(SoilParams& sp);   // m = 1 - 1/n
hipError_t launch_adj_nodal(const AdjArgs& a, hipStream_t s);
hipError_t launch_adj_mass(const AdjArgs& a, double* ml, hipStream_t s);
// J = dF_k/dh_k on the CSR pattern (max_row <= 8 or <= 16), its inverse
// diagonal, then JT[k] = J[tpos[k]] (tpos: position of the transposed entry).
hipError_t launch_adj_jacobian(const AdjArgs& a, int max_row, double* J, double* JT, double* dinv,
                               const int32_t* tpos, int64_t nnz, hipStream_t s);
// rhs = -dL/dh_k + ml dtheta/dh(h_k) lam_next / dt on non-Dirichlet rows.
hipError_t launch_adj_rhs(const AdjArgs& a, const double* dldh, const double* lam_next, double* rhs, hipStream_t s);
// grad_logks[e] += Kbar_e t_e; per-element / per-node alpha and n terms into
// ge[2E] / gn[2n] (summed by the caller's reduction).
hipError_t launch_adj_params(const AdjArgs& a, const double* lam, double* grad_logks, double* ge, double* gn,
                             hipStream_t s);
// Observations: dldh[obs[j]] += dldy[j] dtheta/dh; gan[0..1] += sum dldy dtheta/d(alpha, n).
hipError_t launch_adj_obs(const AdjArgs& a, int n_obs, const int32_t* obs, const double* dldy, double* dldh,
                          double* gan, hipStream_t s);

}  // namespace wspm
