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
#include "transport.h"
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
    NodeRecs nrec;
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

// JT[k] = J[tpos[k]] for nnz CSR slots.
hipError_t launch_transpose_values(const double* J, const int32_t* tpos, double* JT, int64_t nnz, hipStream_t s);

// ---------------------------------------------------------------------------
// Adjoint of the nitrogen species steps (NH4 -> NO2 -> NO3) that follow the
// Richards step k (src/solvers/nitrogen_solver.py:148-238 at fixed dt, lumped
// mass, Cauchy inlet).  Species s solves R_s = A_s(f_k) c_s^k - b_s = 0 with
//   A_s = M_s / dt + K_s(f_k) (+ Cauchy edge terms, constant),
//   b_s = M_s c_s^{k-1} / dt - mu_{s-1} int theta c_{s-1}^k N (s > 0),
// where f_k = (theta, qx, qz, Dxx, Dxz, Dzz) are the node fields of h_k
// (k_transport_fields) and M_s the row-sum lumped -int (theta + rho Kd_s) N N^T.
// The adjoints mu_s^k solve A_s^T mu_s^k = dL/dc_s^k + M_s^{k+1} mu_s^{k+1} / dt
// - (dR_{s+1}^k / dc_s^k)^T mu_{s+1}^k, NO3 first; the fields then receive
// g_f = -sum_s mu_s^T dR_s/df, which is carried to h_k (dL/dh_k, handed to
// the Richards adjoint of the same step), to alpha, n and to log Ks_e.
// ---------------------------------------------------------------------------
struct TrAdjArgs {
    int n, E;
    const int32_t* inc_ptr;
    const int2* inc;
    const int4* conn4;
    NodeRecs nrec;
    TrQuad tq;
    const double* ks_e;         // [E]
    const double* h;            // h_k
    SoilParams sp;              // van Genuchten, Ks = 1, m tied to n
    double DL, DT, Dm;
    double inv_dt;
    double mu[3];               // first-order rates of the species
    double rho_kd;              // sorption of species 0
    const double *th, *qx, *qz; // fields of h_k
    const double* c;            // [3n] species at step k
    const double* cp;           // [3n] species at step k-1
    const double* lam;          // [3n] species adjoints at step k
};

// Lumped species mass of step k+1 from its theta: ml_s = mth + rho_kd_s * mT.
hipError_t launch_tr_adj_mass(const TrAdjArgs& a, const double* th_next, double* mth, double* mT, hipStream_t s);
// rhs of species s: M_s^{k+1} lam_next / dt (lam_next nullable) plus the
// coupling term of species s + 1 (lam_up = its adjoint, nullable).
hipError_t launch_tr_adj_rhs(const TrAdjArgs& a, int species, const double* mth, const double* mT,
                             const double* lam_next, const double* lam_up, double* rhs, hipStream_t s);
// rhs[obs[j]] += dldy[j]
hipError_t launch_tr_adj_obs(int n_obs, const int32_t* obs, const double* dldy, double* rhs, hipStream_t s);
// Field gradients gf[6n] = (theta, qx, qz, Dxx, Dxz, Dzz) of -sum_s mu_s^T R_s.
hipError_t launch_tr_adj_fields(const TrAdjArgs& a, double* gf, hipStream_t s);
// Fields -> h_k, alpha, n, log Ks: dldh[n] (written), gn[2n] (alpha / n
// terms per node, summed by the caller), grad_logks[E] (added to); scratch
// w[3n].
hipError_t launch_tr_adj_chain(const TrAdjArgs& a, const double* gf, double* w, double* dldh, double* gn,
                               double* grad_logks, hipStream_t s);

}  // namespace wspm
