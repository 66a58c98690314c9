// Fused Richards assembly: argument blocks and launcher (assemble.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "wspm_device.h"

namespace wspm {

struct KState;

// Rows per assembly workgroup.  2D rows have ~6 incident triangles and ~7
// entries, 3D rows ~24 incident tets and ~15 entries; R sets the LDS staging
// (incidences x (dim+2) doubles + the rows' CSR window) and the write-out
// lanes per row (256 / R).
#ifndef WSPM_ASM_ROWS_2D
#define WSPM_ASM_ROWS_2D 64
#endif
#ifndef WSPM_ASM_ROWS_3D
#define WSPM_ASM_ROWS_3D 16
#endif
#ifndef WSPM_ASM_THREADS
#define WSPM_ASM_THREADS 256
#endif
constexpr int ASM_ROWS_2D = WSPM_ASM_ROWS_2D;
constexpr int ASM_ROWS_3D = WSPM_ASM_ROWS_3D;
constexpr int ASM_THREADS = WSPM_ASM_THREADS;   // threads per assembly workgroup
// Longest row the one-thread-per-row register kernel handles (else LDS kernel).
constexpr int ASM_REG_ROW_2D = 8;
constexpr int ASM_REG_ROW_3D = 0;      // 3D rows (~24 incidences) use the LDS kernel
constexpr int ASM_REG_BATCH = 3;       // incidences whose loads are issued together

// What the assembly epilogue also computes (the first step of the solver).
enum { ASM_INIT_NONE = 0, ASM_INIT_BICG = 1, ASM_INIT_CG = 2, ASM_INIT_GMRES = 3, ASM_INIT_RESID = 4 };

// Quadrature-derived constants, computed on the host from the reference's
// literal points/weights (src/numerics/gauss.py:31-37, gauss_3D.py:29-39):
//   ck[c] = sum_q w_q N_c(q)               conductivity weight of vertex c
//   ml[c] = sum_q w_q N_c(q) sum_d N_d(q)   lumped-mass factor of vertex c
struct QuadConst {
    double ck[4];
    double ml[4];
};

// Packed device layout read by the assembly (built once per mesh):
//   inc[q]   = (element << 2 | local vertex, in-row offsets of the element's
//              vertices, one byte each) for the q-th incidence;
//   conn4[e] = element vertices, padded to 4;
//   nrec     = node coordinates and the nodal conductivity K (or Kr with a
//              per-element Ks) written by the soil-model kernels (NodeRecs).
// One incidence then costs 1 + 1 + 2 (dim+1) (+1 for Ks_e) vector loads.
struct AsmArgs {
    int n;
    int win_cap;             // LDS doubles reserved for the CSR window
    int inc_cap;             // LDS incidence records reserved
    const int32_t* row_ptr;
    const int32_t* cols;
    const int32_t* diag_pos;
    const int32_t* inc_ptr;
    const int2* inc;
    const int4* conn4;
    NodeRecs nrec;
    const double* Ks_elem;   // nullable
    const double* C;         // row-node values (read once per row)
    const double* theta_m;
    const double* theta_n;
    const double* h_m;
    double inv_dt;
    const uint8_t* bc_flag;  // nullable
    const double* bc_val;
    double* A;
    float* A32;              // nullable: also write an FP32 copy of A
    int ell_n;               // 0: A (and A32) in CSR order; else ELL, entry s of row i at s * ell_n + i
    const int* gate;         // non-null (one system): the register kernels return at once while *gate == 0
                             // (register kernels only)
    double* b;
    double* dinv;            // nullable
    float* dinv32;           // nullable
    // Row-major-free row access for the neighbourhood kernel (nullable): the
    // ELL column table (entry s of row i at s * n + i, width ell_w), the
    // incidences in the same layout (width ell_iw, padding (0, 0)) and per
    // row rl | dg << 4 | nq << 8 (entries, diagonal slot, incidences).  Every
    // load that depends only on the row index is then issued at once: two
    // dependent memory round trips per row instead of three.
    const int32_t* ell_col;
    const int2* ell_inc;
    const uint32_t* ell_info;
    int ell_w, ell_iw;
    // Fan kernel (ASM_K_FAN): ell_col in fan order (slot 0 the node, slots
    // 1..F its neighbours around it), the element << 2 | local vertex of fan
    // triangle u at u * n + i (fan_iw planes), the neighbour before slot 1
    // (closed fans) and F | closed << 4 per row.
    const int32_t* fan_inc;
    const int32_t* fan_pred;
    const uint32_t* fan_info;
    int fan_iw;
    // Windowed fan kernel (ASM_K_WIN): per row block (BLOCK rows from 0, or
    // an ensemble entry) up to WIN_SEG segments of consecutive nodes, at
    // nwin[WIN_WORDS * block] = (lo, len) pairs, that cover every column of
    // its rows; staged in LDS (coordinates, K and x0) at the segments' running
    // offsets, nwin_cap <= WIN_CAP nodes at most.  nwin_col[s * n + i]: the LDS position
    // of fan slot s of row i (nwin_pred: of its predecessor).
    const int32_t* nwin;
    const uint16_t* nwin_col;
    const uint16_t* nwin_pred;
    int nwin_cap;
    QuadConst qc;
    // Ensemble mode (register kernel only): a grid of ens_grid workgroups
    // walks the row-table entries listed in act (act[0] = count; null: all
    // n_tab entries); entry t = rows [tab[t].x, tab[t].y) of instance
    // tab[t].z.  Instances whose Picard iteration has converged
    // (ens_st[z].pdone) are skipped.
    const int4* tab;         // nullable
    const int* act;
    const KState* ens_st;
    int n_tab;
    int ens_grid;
    int* tflag;              // nullable: the benchmark's timing slot, receives the rows this launch assembled
};

// Vectors written by the fused Krylov start (FP64).
struct KrylovInit {
    const double* x0;
    double* x;
    double *r, *rh, *p, *z, *w;
    int jacobi;
};

void compute_quad_const(int dim, QuadConst& qc);
int asm_rows_per_block(int dim);
size_t asm_lds_bytes(int dim, int max_block_nnz, int max_block_inc);
// Launch the assembly; grid_out receives the number of workgroups (partial
// rows written when mode != ASM_INIT_NONE).
int asm_reg_max_row(int dim);
// Assembly kernels: LDS-staged (any rows), register with incidence gathers
// (2D rows <= 8 entries), register with the row's neighbourhood gathered once
// (2D rows <= 8 entries and <= 7 incidences).
// ASM_K_NBR stages the neighbourhood in LDS; ASM_K_NBR_BLEND selects from
// registers (WSPM_ASM_KERNEL=blend).
// ASM_K_FAN: the time-step assembly into an ELL operator in fan order
// (every row a fan of consistently oriented triangles; chosen per launch).
// ASM_K_WIN: the same with each row block's node records staged in LDS.
enum { ASM_K_LDS = 0, ASM_K_REG = 1, ASM_K_NBR = 2, ASM_K_NBR_BLEND = 3, ASM_K_FAN = 4, ASM_K_WIN = 5 };
// Node windows of the windowed fan kernel: segments per block, ints per
// block in AsmArgs::nwin, and the most nodes a window holds (4 planes of
// doubles: 30 KiB, so that 5 workgroups share a CU's 160 KiB of LDS, the
// occupancy its registers allow).
constexpr int WIN_SEG = 3;
constexpr int WIN_WORDS = 2 * WIN_SEG;
constexpr int WIN_CAP = 960;
inline bool asm_is_nbr(int k) { return k == ASM_K_NBR || k == ASM_K_NBR_BLEND; }
hipError_t launch_assemble(int dim, const AsmArgs& a, int mode, const KrylovInit& ki, double* partials,
                           int prow, size_t lds, int* grid_out, hipStream_t s, int kernel);

}  // namespace wspm
