// Solute / nitrogen transport on the Richards mesh context (transport.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wspm_device.h"

namespace wspm {

// Quadrature tensors of the reference's 3-point triangle rule, computed on
// the host from its literal points and weights (src/numerics/gauss.py:31-37):
//   ck[c]       = sum_q w_q N_c(q)
//   T[a][c]     = sum_q w_q N_a(q) N_c(q)
//   U[a][c][v]  = sum_q w_q N_a(q) N_c(q) N_v(q)
// Each element term is such a sum times det J (signed, as the reference's
// jnp.linalg.det); the integrands are polynomials in N, so these reproduce
// the reference's quadrature sums.
struct TrQuad {
    double ck[3];
    double T[3][3];
    double U[3][3][3];
};
void compute_tr_quad(TrQuad& q);

// Node values of src/models/transport_utilities.py:119-192.
struct FieldArgs {
    int n;
    const int32_t* inc_ptr;
    const int2* inc;
    const int4* conn4;
    NodeRecs nrec;
    const double* h;
    const double* ks_e;       // nullable: per-element Ks; K = soil K(h) * mean of the incident Ks_e
    SoilParams sp;
    double DL, DT, Dm;
    double *qx, *qz, *aq, *dxx, *dxz, *dzz, *th;
};

// One species system (assembly_solute.py / assembly_nitrogen.py), one thread
// per row over the node -> (element, vertex) incidences.
struct TrArgs {
    int n;
    const int32_t* row_ptr;
    const int32_t* cols;
    const int32_t* diag_pos;
    const int32_t* inc_ptr;
    const int2* inc;
    const int4* conn4;
    NodeRecs nrec;
    const double *th, *qx, *qz, *dxx, *dxz, *dzz;
    const double* c_n;        // concentration at the start of the step
    const double* c_src;      // coupling species (nullable)
    double mu, rho_kd, mu_src, inv_dt;
    int lumped;
    TrQuad tq;
    double *A, *b, *dinv;     // dinv nullable
};

// Edge terms on a boundary node list (src/models/boundary.py:108-421): list
// entry k couples nodes[k-1] -> nodes[k] -> nodes[k+1]; thread k adds both
// edges' contributions to row nodes[k] (edge k-1 first).
enum { EDGE_NEUMANN = 0, EDGE_FREEDRAIN_WATER = 1, EDGE_CAUCHY = 2, EDGE_FREEDRAIN_N = 3 };
struct EdgeArgs {
    int cnt;
    const int32_t* nodes;
    const int32_t* slot_prev;   // CSR position of (nodes[k], nodes[k-1]), -1 for k = 0
    const int32_t* slot_next;   // CSR position of (nodes[k], nodes[k+1]), -1 for the last
    const int32_t* diag_pos;
    NodeRecs nrec;        // node coordinates (xy)
    int kind;
    double flux, c_in;
    const double* K;            // K[i * kstride]
    int kstride;
    const double* th;
    double* A;
    double* b;
    double* dinv;               // nullable: Jacobi inverse diagonal of the rows touched
};

hipError_t launch_transport_fields(const FieldArgs& a, hipStream_t s);
hipError_t launch_assemble_transport(const TrArgs& a, int max_row, hipStream_t s);
hipError_t launch_edges(const EdgeArgs& a, hipStream_t s);

}  // namespace wspm
