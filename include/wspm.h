/*
 * wspm.h -- C ABI of the MI355X (gfx950) implicit P1-FEM Richards time step.
 *
 * Plain C linkage, plain pointers and sizes.  "dev" pointers are HIP device
 * pointers owned by the caller (e.g. torch tensor data_ptr()); "host" pointers
 * are ordinary host memory.  Every call returns 0 on success or a negative
 * WSPM_E* code; wspm_last_error() gives the message.  Numerical
 * non-convergence is not an error: it is reported in the stats structs, as the
 * reference reports nothing and accepts the iterate.
 *
 * One context = one mesh + one HIP stream; a context is not thread-safe,
 * distinct contexts are independent.  All work is enqueued on the context's
 * stream; entry points that return host-visible results synchronise it.
 *
 * Reference interfaces replaced (paths relative to the JAX-WSPM checkout):
 *   wspm_ctx_create        src/mesh/loader.py:11-108 (mesh arrays) + the BCOO
 *                          index construction of
 *                          src/numerics/assembly_re_v2.py:114-152
 *   wspm_bc_set            src/models/boundary.py:38-53, 603-686 (ordered
 *                          Dirichlet lists, later list wins)
 *   wspm_soil              src/models/van_genuchten.py:9-47,
 *                          src/models/exponential.py:10-42 (vmapped)
 *   wspm_assemble_richards src/numerics/assembly_re_v2.py:62-166 and
 *                          src/numerics/assembly_re_3D.py:72-170, fused with
 *                          the Dirichlet rows and the Jacobi diagonal of
 *                          src/solvers/jax_linear_solver.py:12-36
 *   wspm_spmv              the BCOO matvec inside jax.scipy.sparse.linalg
 *   wspm_krylov            src/solvers/solver_types.py:7-68 ->
 *                          jax_linear_solver.py:262-330 (cg/bicgstab/gmres)
 *   wspm_richards_*        src/solvers/richards_solver_v2.py:107-168 and
 *                          src/solvers/richards_solver_3D_v2.py:177-222
 *                          (Picard loop of one implicit time step)
 *   wspm_edge_set,         src/models/boundary.py:108-173 (Neumann, free
 *   wspm_richards_edges    drainage of water), :217-277 (Cauchy), :336-421
 *                          (free drainage of nitrogen): edge terms on
 *                          consecutive boundary-list pairs
 *   wspm_transport_fields  src/models/transport_utilities.py:119-192
 *   wspm_assemble_transport src/numerics/assembly_solute.py:10-186,
 *                          src/numerics/assembly_nitrogen.py:33-607 with the
 *                          lumping / mass products / M/dt + K of
 *                          src/solvers/nitrogen_solver.py:36-66,148-304 and
 *                          src/solvers/coupled_solver_v2.py:154-200
 *   wspm_adj_step          no counterpart (the reference has no gradient code):
 *                          one backward step of the discrete adjoint of
 *                          wspm_richards_step
 *   wspm_adj_transport_step  no counterpart: one backward step of the adjoint
 *                          of the nitrogen species steps
 *   wspm_domain_integral,  richards_mass_balance-checkpoint.py:25-65,
 *   wspm_element_size,     src/models/transport_utilities.py:196-242
 *   wspm_peclet_courant    (diagnostics)
 *   wspm_ens_*             no single counterpart: the reference runs one
 *                          instance per process (src/cli.py:87-150); an
 *                          ensemble context time-steps many independent
 *                          instances of wspm_richards_* at once.
 */
#ifndef WSPM_H
#define WSPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wspm_ctx wspm_ctx;

enum {
    WSPM_OK = 0,
    WSPM_EINVAL = -1,   /* bad argument (shape, index, enum) */
    WSPM_EHIP = -2,     /* HIP runtime error */
    WSPM_ENOMEM = -3,   /* allocation failed */
    WSPM_ESTATE = -4    /* call out of order (e.g. step before init) */
};

enum { WSPM_SOIL_GARDNER = 0, WSPM_SOIL_VAN_GENUCHTEN = 1 };
enum { WSPM_CG = 0, WSPM_BICGSTAB = 1, WSPM_GMRES = 2 };
/* Picard steps only (wspm_richards_params.method): each linear system solved
 * by wspm_direct's banded LU instead of a Krylov method. */
enum { WSPM_DIRECT = 3 };
/* Jacobi is fused into the Krylov kernels; the others (src/solvers/
 * jax_linear_solver.py:39-252) are applied between them: 2x2 block Jacobi,
 * SSOR (omega = 1) and the reference's "ilu0" (I + L D^-1)(D + U). */
enum { WSPM_PRECOND_NONE = 0, WSPM_PRECOND_JACOBI = 1, WSPM_PRECOND_BLOCK_JACOBI = 2,
       WSPM_PRECOND_SSOR = 3, WSPM_PRECOND_ILU = 4 };
enum { WSPM_FP64 = 0, WSPM_MIXED = 1 };
/* Mixed precision: the FP32 inner solve of each refinement sweep.  KRYLOV:
 * the method's own CG / BiCGSTAB / GMRES to inner_tol.  JACOBI: inner_iters
 * Jacobi-preconditioned Richardson steps d <- d + M (r - A d) from d = 0 (no
 * dot products; needs the Jacobi preconditioner), switched to KRYLOV for the
 * rest of the context's life once a sweep fails to halve the FP64 residual. */
enum { WSPM_INNER_KRYLOV = 0, WSPM_INNER_JACOBI = 1 };
enum { WSPM_PICARD_REL = 0, WSPM_PICARD_ABS = 1 };

typedef struct {
    int32_t iters;        /* CG/BiCGSTAB iterations; GMRES restart cycles    */
    int32_t arnoldi;      /* GMRES Arnoldi steps (0 for CG/BiCGSTAB)          */
    int32_t code;         /* BiCGSTAB k (-10/-11 breakdown); GMRES info (-1 NaN) */
    int32_t refinements;  /* mixed precision: FP64 refinement sweeps          */
    double res_norm;      /* final |r| (|M r| for GMRES)                      */
    double b_norm;        /* |b|                                              */
} wspm_krylov_stats;

typedef struct {
    int32_t picard_iters;
    int32_t krylov_iters; /* summed over the Picard iterations of the step   */
    int32_t converged;    /* 1 if err < picard_tol, 0 if the cap was hit      */
    int32_t reserved;
    double err;           /* last Picard error (relative or absolute)        */
} wspm_step_stats;

typedef struct {
    int32_t soil_model;   /* WSPM_SOIL_*                                      */
    int32_t method;       /* WSPM_CG / WSPM_BICGSTAB / WSPM_GMRES / WSPM_DIRECT */
    int32_t precond;      /* WSPM_PRECOND_* (general ones: FP64, one mesh)    */
    int32_t precision;    /* WSPM_FP64 or WSPM_MIXED                          */
    double phi[6];        /* gardner: alpha0, theta_s, theta_r, Ks
                             van Genuchten: alpha, theta_s, theta_r, n, m, Ks
                             (config/settings.py:19-22, 47-48)               */
    double tol;           /* Krylov rtol (reference 1e-10)                    */
    double atol;          /* Krylov atol (reference 0)                        */
    int32_t maxiter;      /* reference 100                                    */
    int32_t restart;      /* GMRES restart (reference 30, <= 64)              */
    double picard_tol;
    int32_t picard_maxit;
    int32_t picard_mode;  /* WSPM_PICARD_REL | WSPM_PICARD_ABS               */
    double inner_tol;     /* mixed precision: FP32 inner relative tolerance   */
    int32_t inner_method; /* mixed precision: WSPM_INNER_KRYLOV | _JACOBI     */
    int32_t inner_iters;  /* WSPM_INNER_JACOBI: steps per sweep (0: 3; <= 16) */
} wspm_richards_params;

/* Library ------------------------------------------------------------------ */
int32_t wspm_version(void);
/* Last error of ctx, or of the last failed wspm_ctx_create when ctx is NULL. */
const char* wspm_last_error(const wspm_ctx* ctx);

/* Mesh context ------------------------------------------------------------ */
/* conn: host int32 [n_elems * (dim+1)], 0-based; coords: host double
 * [n_nodes * dim] row-major (x0,y0[,z0],x1,...).  stream: hipStream_t or NULL. */
int32_t wspm_ctx_create(int32_t dim, int32_t n_nodes, int32_t n_elems,
                        const int32_t* conn, const double* coords, void* stream,
                        wspm_ctx** out);
void wspm_ctx_destroy(wspm_ctx* ctx);
/* Reverse Cuthill-McKee renumbering of a mesh (host only, no GPU call; no
 * counterpart in the reference, whose unstructured solute meshes,
 * data/mesh/solute read by src/mesh/loader.py:33-36, run in their file order).
 * conn as wspm_ctx_create.  Out (host int32): node_new[n_nodes], the new index
 * of every node (each connected component from a pseudo-peripheral node of
 * least degree, neighbours by increasing degree, the order reversed);
 * elem_old[n_elems], the original element at each new position (elements by
 * their smallest new vertex index, ties in file order); band (int32[4] or
 * NULL): lower and upper bandwidth of the node graph before and after.
 * Errors: wspm_last_error(NULL). */
int32_t wspm_mesh_rcm(int32_t dim, int32_t n_nodes, int32_t n_elems, const int32_t* conn, int32_t* node_new,
                      int32_t* elem_old, int32_t* band);
int32_t wspm_set_stream(wspm_ctx* ctx, void* stream);
/* info (host int64[8]): n_nodes, n_elems, nnz, dim, max_row_len, n_bc, the
 * CSR SpMV's lanes per row, and 1 if the time-step solves of the last
 * wspm_richards_init / wspm_ens_richards_init read the operator in ELL order
 * with each row's columns sorted, 2 in ELL order with each row's columns in
 * fan order (the node, then its neighbours around it), 3 in fan order with
 * the assembly staging each row block's node window in LDS, 0 in CSR order. */
int32_t wspm_ctx_info(const wspm_ctx* ctx, int64_t* info);
/* Copy the CSR pattern to host: row_ptr int32[n+1], cols int32[nnz]. */
int32_t wspm_get_pattern(const wspm_ctx* ctx, int32_t* row_ptr, int32_t* cols);

/* Dirichlet rows: host arrays, applied in list order (a node listed twice takes
 * the later value, as top-then-bottom in the reference). n_bc = 0 clears. */
int32_t wspm_bc_set(wspm_ctx* ctx, int32_t n_bc, const int32_t* nodes, const double* values);

/* Pointwise soil model over n values (any n; dev pointers).  params as phi above. */
int32_t wspm_soil(wspm_ctx* ctx, int32_t model, const double* params, int32_t n,
                  const double* h, double* C, double* K, double* theta);

/* Fused P1 assembly of the mixed-form Richards system into the fixed CSR
 * pattern, with Dirichlet rows applied.  All dev pointers.  Ks_elem: per-element
 * saturated conductivity (then K holds the relative conductivity) or NULL.
 * dinv (Jacobi inverse diagonal) may be NULL. */
int32_t wspm_assemble_richards(wspm_ctx* ctx, const double* Ks_elem, const double* h_m,
                               const double* C, const double* K, const double* theta_m,
                               const double* theta_n, double dt,
                               double* A_vals, double* b, double* dinv);

/* y = A x on the context's pattern (dev pointers). */
int32_t wspm_spmv(wspm_ctx* ctx, const double* A_vals, const double* x, double* y);

/* Krylov solve that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The CG or
 * BiCGSTAB of wspm_krylov (Jacobi with dinv, or none), on a CSR matrix the
 * caller passes (dev: row_ptr [n+1], cols [nnz], A_vals [nnz]; nnz host),
 * enqueued on `stream` (hipStream_t or NULL) without a context.  Its
 * workspace is a file-scope device array of the library (40 MiB per GPU,
 * held until the process ends): n <= 524288 rows and maxiter <= 10000, else
 * WSPM_EINVAL.  No hipMalloc, hipFree or allocating call is made; the host
 * reads the solver state back by a copy to its stack and a stream
 * synchronisation.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (other threads, other streams) race.  Errors:
 * wspm_last_error(NULL). */
int32_t wspm_krylov_noalloc(int32_t method, int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* cols,
                            const double* A_vals, const double* dinv, const double* b, const double* x0, double* x,
                            double tol, double atol, int32_t maxiter, void* stream, wspm_krylov_stats* stats);

/* Preconditioned Krylov solve A x = b from x0 (dev pointers; x0 may equal x).
 * dinv NULL = no preconditioner, else Jacobi with that inverse diagonal. */
int32_t wspm_krylov(wspm_ctx* ctx, int32_t method, const double* A_vals, const double* dinv,
                    const double* b, const double* x0, double* x, double tol, double atol,
                    int32_t maxiter, int32_t restart, wspm_krylov_stats* stats);

/* Krylov solve with any WSPM_PRECOND_* built from A_vals (dev pointers). */
int32_t wspm_krylov_pc(wspm_ctx* ctx, int32_t method, int32_t precond, const double* A_vals,
                       const double* b, const double* x0, double* x, double tol, double atol,
                       int32_t maxiter, int32_t restart, wspm_krylov_stats* stats);

/* Mixed precision: FP32 Krylov inner solves inside FP64 iterative refinement
 * on the true residual; stops when |b - A x| <= max(tol |b|, atol) (FP64). */
int32_t wspm_krylov_mixed(wspm_ctx* ctx, int32_t method, const double* A_vals, const double* dinv,
                          const double* b, const double* x0, double* x, double tol, double atol,
                          int32_t maxiter, int32_t restart, double inner_tol, int32_t max_refine,
                          wspm_krylov_stats* stats);

/* Device-resident Richards time stepping ----------------------------------- */
/* h0 dev [n].  Ks_elem dev [n_elems] or NULL.  Copies both. */
int32_t wspm_richards_init(wspm_ctx* ctx, const wspm_richards_params* prm,
                           const double* Ks_elem, const double* h0);
/* One implicit step of size dt: Picard to convergence (or the cap). */
int32_t wspm_richards_step(wspm_ctx* ctx, double dt, wspm_step_stats* stats);
/* Copy state (dev pointers, each may be NULL): h, theta(h), theta at the start
 * of the last step. */
int32_t wspm_richards_state(wspm_ctx* ctx, double* h, double* theta, double* theta_n);

/* Timing hooks for the benchmark (no counterpart in the reference): when
 * enabled, the SpMV launches of the Krylov solvers are bracketed by HIP events
 * on the context's stream, and each launch records on the device whether it
 * did its work or returned early (solve already converged).  on: 0 off,
 * 1 the FP64 SpMV launches, 2 the FP32 ones (mixed precision inner solves),
 * plus 4: the assembly launches as well (wspm_timing_read_asm). */
int32_t wspm_timing_enable(wspm_ctx* ctx, int32_t on);
/* Number of SpMV launches that did their work since enable/last read, their
 * summed event-to-event duration (ms) and the rows they worked on (rows may
 * be NULL; an ensemble launch skips the rows of converged instances). */
int32_t wspm_timing_read(wspm_ctx* ctx, int64_t* launches, double* total_ms, int64_t* rows);
/* The same figures over the SpMV launches that worked on at least min_rows
 * rows, without resetting them (call before wspm_timing_read): what the
 * kernel reaches when a launch has enough rows to stream, apart from the
 * fixed cost of small launches. */
int32_t wspm_timing_read_over(wspm_ctx* ctx, int64_t min_rows, int64_t* launches, double* total_ms,
                              int64_t* rows);
/* With on & 4 in wspm_timing_enable, the assembly launches (a Picard
 * iteration's fused assembly + solver start) are bracketed by HIP events
 * too; this reads them as wspm_timing_read does (launches that assembled no
 * row, e.g. a Picard iteration enqueued ahead and not needed, are left out). */
int32_t wspm_timing_read_asm(wspm_ctx* ctx, int64_t* launches, double* total_ms, int64_t* rows);
/* Kernel launches (device-to-device copies and HIP graph replays counted as
 * one each) the context has enqueued on its stream since its creation, and
 * the Picard rounds its time steps have run: an ensemble step's rounds over
 * the instances still iterating, or one system's Picard iterations.  No
 * counterpart in the reference (the benchmark reports launches per round). */
int32_t wspm_launch_count(const wspm_ctx* ctx, int64_t* launches, int64_t* picard_rounds);

/* Transport (2D) -------------------------------------------------------------
 * Boundary node lists for edge terms, in the reference's list order: an edge
 * joins consecutive entries, its length is x_J - x_I.  Consecutive entries must
 * be joined by a mesh edge (their matrix entries live in the CSR pattern). */
enum { WSPM_EDGE_INLET = 0, WSPM_EDGE_OUTLET = 1 };
int32_t wspm_edge_set(wspm_ctx* ctx, int32_t list, int32_t n, const int32_t* nodes);
/* Richards steps of this context add -flux * int N on the inlet list
 * (Neumann) and/or -int K N on the outlet list (free drainage) to b. */
int32_t wspm_richards_edges(wspm_ctx* ctx, int32_t inlet_neumann, double neumann_flux,
                            int32_t outlet_freedrain);
/* Node fields for transport from h (dev arrays [n]): water fluxes qx, qz, |q|,
 * dispersion Dxx, Dxz, Dzz and theta; van Genuchten phi (host double[6]).
 * Ks_elem (dev [E]) or NULL: with a per-element Ks field (an extension; the
 * reference's Ks is phi[5]) a node's K is the soil model's K times the mean
 * Ks_elem of its incident elements. */
int32_t wspm_transport_fields(wspm_ctx* ctx, const double* phi, const double* h, const double* Ks_elem,
                              double DL, double DT, double Dm, double* qx, double* qz, double* absq,
                              double* dxx, double* dxz, double* dzz, double* theta);
typedef struct {
    int32_t lumped;           /* 1: row-sum lumped mass (nitrogen), 0: consistent (solute) */
    int32_t inlet_cauchy;     /* add Cauchy terms on the inlet list                      */
    int32_t outlet_freedrain; /* add -K/theta free-drainage terms on the outlet list     */
    int32_t reserved;
    double dt;
    double mu;                /* + mu int theta N N^T in the stiffness (reaction)         */
    double rho_kd;            /* - rho Kd int N N^T in the mass (NH4 sorption)            */
    double mu_src;            /* source -mu_src int theta c_src N (coupled species)       */
    double cauchy_flux, c_inlet;
} wspm_transport_params;
/* One species system A = M/dt + K (+ edge terms), b = M c_n / dt + source (+
 * Cauchy), on the context's pattern.  Fields from wspm_transport_fields; K
 * (dev [n]) only for free drainage; c_src (dev [n]) or NULL; dinv may be NULL. */
int32_t wspm_assemble_transport(wspm_ctx* ctx, const wspm_transport_params* tp,
                                const double* theta, const double* qx, const double* qz,
                                const double* dxx, const double* dxz, const double* dzz,
                                const double* K, const double* c_n, const double* c_src,
                                double* A_vals, double* b, double* dinv);

/* Discrete adjoint (2D, van Genuchten, per-element Ks) -------------------------
 * Step k solved F_k(h_k, h_{k-1}; p) = 0 (the converged Picard fixed point of
 * wspm_richards_step; m = 1 - 1/n).  One call does, for a loss L:
 *   J_k^T lam_k = -dL/dh_k + ML dtheta/dh(h_k) lam_{k+1} / dt   (Jacobi-BiCGSTAB)
 *   grad_logks[e] += lam_k . dF_k/dlogKs_e,  grad_an[0..1] += lam_k . dF_k/d(alpha, n)
 * plus the observation terms of L = l(theta(h_k)[obs]): dL/dh_k and the direct
 * dtheta/d(alpha, n) parts, from dldy = dl/dy (y = theta(h_k)[obs]).
 * Device pointers: Ks_elem [E], h = h_k, h_prev = h_{k-1}, obs [n_obs] int32,
 * dldy [n_obs] (both NULL when n_obs = 0), lam_next [n] (NULL for the last
 * step), dldh_in [n] (NULL, or further dL/dh_k terms, e.g. from
 * wspm_adj_transport_step), lam [n] out, grad_logks [E] and grad_an [2]
 * accumulated. */
typedef struct {
    double dt;
    double phi[6];        /* alpha, theta_s, theta_r, n, (m = 1 - 1/n), (Ks: Ks_elem) */
    double tol;           /* relative tolerance of the transpose solve           */
    int32_t maxiter;
    int32_t n_obs;
} wspm_adj_params;
int32_t wspm_adj_step(wspm_ctx* ctx, const wspm_adj_params* prm, const double* Ks_elem,
                      const double* h, const double* h_prev, const int32_t* obs,
                      const double* dldy, const double* lam_next, const double* dldh_in,
                      double* lam, double* grad_logks, double* grad_an, wspm_krylov_stats* stats);

/* Adjoint of the nitrogen species steps NH4 -> NO2 -> NO3 that follow the
 * Richards step k (no counterpart: the reference's nitrogen_solver.py:148-238
 * is never differentiated).  Forward model at fixed dt, lumped mass, Cauchy
 * inlet (the inlet list of wspm_edge_set), fields of wspm_transport_fields
 * with Ks_elem:
 *   A_s(f_k) c_s^k = M_s(f_k) c_s^{k-1} / dt - mu_{s-1} int theta c_{s-1}^k N.
 * One call solves the three transposed systems (NO3 first, Jacobi-BiCGSTAB):
 *   A_s^T mu_s^k = dL/dc_s^k + M_s(f_{k+1}) mu_s^{k+1} / dt
 *                  - (dR_{s+1}/dc_s)^T mu_{s+1}^k
 * and carries -sum_s mu_s^T dR_s/df_k through the fields to dldh (out: the
 * dL/dh_k to pass to wspm_adj_step's dldh_in), grad_logks and grad_an.
 * Device pointers: Ks_elem [E]; h = h_k; h_next = h_{k+1} (NULL at the last
 * step); c, c_prev [3n] species at steps k and k-1 (NH4, NO2, NO3 blocks);
 * obs [n_obs], dldy [n_obs] = dL/dNO3^k[obs]; lam_next [3n] (NULL at the last
 * step); lam [3n] out; dldh [n] out; grad_logks [E], grad_an [2] accumulated.
 * stats: the three solves' iterations summed, the largest residual. */
typedef struct {
    double dt;
    double phi[6];        /* van Genuchten as wspm_adj_params (Ks: Ks_elem)     */
    double DL, DT, Dm;    /* dispersivities and molecular diffusion              */
    double mu[3];         /* first-order rates of NH4, NO2, NO3                  */
    double rho_kd;        /* rho * Kd of NH4 (sorption)                          */
    double cauchy_flux;   /* Cauchy flux on the inlet list (0: no inlet term)    */
    double tol;           /* relative tolerance of the transpose solves          */
    int32_t maxiter;
    int32_t n_obs;
} wspm_adj_tr_params;
int32_t wspm_adj_transport_step(wspm_ctx* ctx, const wspm_adj_tr_params* prm, const double* Ks_elem,
                                const double* h, const double* h_next, const double* c,
                                const double* c_prev, const int32_t* obs, const double* dldy,
                                const double* lam_next, double* lam, double* dldh,
                                double* grad_logks, double* grad_an, wspm_krylov_stats* stats);

/* Ensembles ------------------------------------------------------------------
 * n_inst independent 2D instances (own mesh, soil parameters, Dirichlet rows)
 * in one context.  Nodes and elements are numbered instance after instance:
 * instance i owns global nodes [sum(node_counts[:i]), +node_counts[i]) and
 * elements likewise.  conn (host int32 [sum(elem_counts) * 3]) holds each
 * instance's connectivity in its own LOCAL numbering; coords (host double
 * [sum(node_counts) * 2]) are concatenated.  Every instance converges on its
 * own: its Krylov solve and Picard loop stop independently, and its
 * reductions cover its own rows only.  On an ensemble context use
 * wspm_bc_set / wspm_spmv / wspm_soil with global numbering, and
 * wspm_richards_step / wspm_richards_state (concatenated arrays);
 * wspm_assemble_richards, wspm_krylov, wspm_krylov_mixed and
 * wspm_richards_init return WSPM_EINVAL.  The SpMV layout follows from the
 * instances' row lengths (ELL when they allow it); an autotuner times the
 * launch grid at creation (wspm_ens_shape). */
int32_t wspm_ens_create(int32_t dim, int32_t n_inst, const int32_t* node_counts,
                        const int32_t* elem_counts, const int32_t* conn,
                        const double* coords, void* stream, wspm_ctx** out);
/* prm: shared settings (CG or BiCGSTAB; no preconditioner or Jacobi; FP64,
 * or WSPM_MIXED: per-instance iterative refinement, FP32 corrections of the
 * FP64 residual until it meets prm->tol, with prm->inner_tol for the FP32
 * solves); phi: host double [n_inst * 6]
 * per-instance soil parameters (layout of wspm_richards_params.phi), or NULL
 * for prm->phi everywhere.  Ks_elem dev [total elements] or NULL; h0 dev
 * [total nodes]. */
int32_t wspm_ens_richards_init(wspm_ctx* ctx, const wspm_richards_params* prm,
                               const double* phi, const double* Ks_elem, const double* h0);
/* Per-instance results of the last wspm_richards_step (host arrays [n_inst],
 * each may be NULL).  The step's own wspm_step_stats aggregates them: max
 * Picard iterations, summed Krylov iterations, all converged, max error. */
int32_t wspm_ens_step_stats(wspm_ctx* ctx, int32_t* picard_iters, int32_t* krylov_iters,
                            int32_t* converged, double* err);

/* Launch shape of an ensemble context (host int32[8]): [0] 1 if the
 * time-step operator is ELL (0: CSR), [1] its width (ELL) or lanes per row
 * (CSR), [2] workgroups walking the entry lists, [3] rows per entry, [4]
 * entries, [5] grids the autotuner timed (0: fixed by WSPM_ENS_GRID), [6]
 * the chosen grid's SpMV time over the whole operator in nanoseconds, [7] 1
 * if mixed-precision Jacobi inner steps have stalled and Krylov inner solves
 * took over (WSPM_INNER_JACOBI), else 0.
 * The format follows from the meshes alone, so results are repeatable; the
 * tuned grid changes no bit of them. */
int32_t wspm_ens_shape(const wspm_ctx* ctx, int32_t* out);

/* ---- Direct solve -------------------------------------------------------------
 * x = A^{-1} b for CSR values A on the context's pattern (all dev): the
 * `direct` solver of src/solvers/solver_types.py:60-63 (jnp.linalg.solve on
 * the densified matrix, LU with partial pivoting).  Here: banded LU with
 * partial pivoting in the band of the pattern after a reverse Cuthill-McKee
 * permutation (kept when it narrows the band), factor and solve on the GPU.
 * The band (n x (2 kl + ku + 1) doubles, kl/ku the permuted bandwidths) is
 * allocated on first use; WSPM_ENOMEM beyond 32 GiB.  Returns WSPM_EINVAL
 * with "singular matrix" when a pivot column is zero (info: the column + 1,
 * may be NULL). */
int32_t wspm_direct(wspm_ctx* ctx, const double* A_vals, const double* b, double* x, int32_t* info);
/* Bandwidths of the direct solver's ordering (host out: kl, ku, 1 if the
 * RCM permutation is used). */
int32_t wspm_direct_info(wspm_ctx* ctx, int32_t* kl, int32_t* ku, int32_t* reordered);

/* ---- Diagnostics (2D, single-mesh contexts) ---- */

/* Domain integral of the nodal field f (dev [n]) with the 3-point rule of
 * integrate_2d_domain (src/solvers/.ipynb_checkpoints/
 * richards_mass_balance-checkpoint.py:25-65): sum over triangles of
 * |det J| * sum_q (1/3) f(x_q), i.e. twice the integral, as the reference.
 * out: host double.  The mass-balance driver of that file (:210-335)
 * evaluates it on theta and on the exact theta every step. */
int32_t wspm_domain_integral(wspm_ctx* ctx, const double* f, double* out);
/* calculate_element_size (src/models/transport_utilities.py:196-204): the
 * smallest edge length of every triangle; size dev [E]. */
int32_t wspm_element_size(wspm_ctx* ctx, double* size);
/* calculate_peclet_courant (transport_utilities.py:206-242) on the node
 * fields of wspm_transport_fields (all dev [n]).  Per element, with h its
 * smallest edge and v = q / theta at its nodes: Pe = max over the nodes of
 * |v_x| h / (2 Dxx) and |v_z| h / (2 Dzz); Cr = max of |v| dt / h.
 * out (host [2]): all_elements = 0 reproduces the reference's result
 * (max(Pe_0, Cr_0), max(Pe_1, Cr_1)), which its tuple iteration at :237-240
 * yields; all_elements = 1 gives (max_e Pe_e, max_e Cr_e).  NaN propagates
 * as in jnp.max.  pe_elem, cr_elem: dev [E] per-element values, or NULL. */
int32_t wspm_peclet_courant(wspm_ctx* ctx, const double* qx, const double* qz, const double* theta,
                            const double* dxx, const double* dzz, double dt, int32_t all_elements,
                            double* pe_elem, double* cr_elem, double* out);

#ifdef __cplusplus
}
#endif
#endif /* WSPM_H */
