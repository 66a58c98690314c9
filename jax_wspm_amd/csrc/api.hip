// C ABI, per-mesh context and host drivers of the implicit time step.
//
// The host side only enqueues work on the context's stream.  It reads the
// device back at three places: after a Krylov solve (iteration chunk done?),
// after each Picard iteration (the Picard error, which decides the loop), and
// after each mixed-precision refinement sweep.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "assemble.h"
#include "adjoint.h"
#include "diagnostics.h"
#include "direct.h"
#include "krylov_kernels.h"
#include "precond.h"
#include "transport.h"
#include "wspm_device.h"
#include "wspm_internal.h"

using namespace wspm;

namespace {

thread_local std::string g_create_error;

struct Err {};
// An error with its own code and message (thrown inside API_BEGIN / API_END).
struct ApiFail {
    int code;
    std::string msg;
};

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
            throw Err{};                                                                \
        }                                                                               \
    } while (0)

// A kernel launch on the context's stream, counted (wspm_launch_count).
#define LAUNCH(...)                      \
    do {                                 \
        ++ctx->n_launch;                 \
        hipLaunchKernelGGL(__VA_ARGS__); \
    } while (0)

bool env_is(const char* name, const char* value) {
    const char* e = std::getenv(name);
    return e && std::string(e) == value;
}

// ---------------------------------------------------------------------------
// Soil / Picard kernels
// ---------------------------------------------------------------------------
// K goes to K[i] and/or to the assembly's nodal K (krec[i]).
// ens: ensemble mode, the instance's soil parameters come from its state.
struct SoilOp {
    static constexpr int ND = 0;
    SoilParams sp;
    const double* h;
    double *C, *K, *th;
    double* krec;
    int ens;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState* s) {
        if (ens) sp = s->sp;
    }
    __device__ void row(int i, double*) const {
        double c, k, t;
        soil_eval(sp, h[i], c, k, t);
        C[i] = c; th[i] = t;
        if (K) K[i] = k;
        if (krec) krec[i] = k;
    }
    __device__ void fin(const double*, KState*) const {}
};

// |h - h_m|, |h| and the soil model at the new iterate (used by the next
// Picard iteration, and as theta(h) once the step has converged).
// Ensemble mode (ens): an instance is processed once per Picard iteration,
// after its Krylov solve is done, with its own soil parameters; h is copied
// into hm (instances finish their iterations at different times, so the two
// buffers are not swapped); the finish counts its Picard/Krylov iterations and
// marks it converged (or capped), as the loop condition of
// src/solvers/richards_solver_v2.py:141-168.
struct PicardOp {
    static constexpr int ND = 2;
    SoilParams sp;
    const double* h;
    double* hm;                 // ensemble: receives h (the new iterate becomes current)
    double *C, *K, *th;
    double* krec;
    int rel;
    int ens;
    const int* gate = nullptr;   // one system: *gate == 0 skips this update (an iteration enqueued ahead)
    // one system, mixed precision: the refinement's state, whose convergence
    // flag, stall flag and Jacobi inner steps this update's state takes over
    // (rdone, rstall, kry), so that one read-back tells the host all three
    const KState* ref = nullptr;
    __device__ bool skip(const KState* s) const {
        if (!ens) return gate != nullptr && *gate == 0;   // uniform
        // the three flags loaded together (no load behind another's test)
        const int pd = s->pdone, dn = s->done, pv = s->pev;
        return (pd != 0) | (dn == 0) | (pv != 0);
    }
    __device__ void load(const KState* s) {
        if (ens) sp = s->sp;
    }
    __device__ void row(int i, double* acc) const {
        const double hi = h[i];
        const double d = hi - hm[i];
        if (ens) hm[i] = hi;
        acc[0] += d * d;
        acc[1] += hi * hi;
        double c, k, t;
        soil_eval(sp, hi, c, k, t);
        C[i] = c; th[i] = t;
        if (K) K[i] = k;
        if (krec) krec[i] = k;
    }
    __device__ void fin(const double* t, KState* s) const {
        s->picard_d2 = t[0];
        s->picard_h2 = t[1];
        s->picard_err = rel ? sqrt(t[0]) / sqrt(t[1]) : sqrt(t[0]);
        if (ref != nullptr) {
            s->rdone = ref->rdone;
            s->rstall = ref->rstall;
            s->kry = ref->steps;
        }
        if (ens) {
            s->pev = 1;
            s->pit += 1;
            s->kry += s->steps;
            s->pdone = !(s->picard_err >= s->picard_tol && s->pit < s->picard_maxit);
        }
    }
};

// Ensemble bookkeeping --------------------------------------------------------
// Start of a time step: every instance iterates again.
__global__ void k_ens_step_begin(KState* st, int n_inst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) {
        st[i].pdone = 0;
        st[i].pev = 0;
        st[i].pit = 0;
        st[i].kry = 0;
    }
}

// Host read-back without a stream synchronisation: copy nwords 32-bit words
// of device state into pinned host memory, then raise the sequence word
// there.  The sequence value comes from a device counter (dseq, incremented
// once per publication), so a captured graph that publishes writes a new value
// on every replay; the host counts publications the same way and spins on the
// word.  Every thread makes its stores visible at system scope before the
// workgroup barrier, and the sequence word is stored after it, so a host that
// sees the new value reads the new state.  On MI355X this read-back takes
// about 6 us against 14 us for a copy + hipStreamSynchronize.
__device__ inline void publish_seq(int* dseq, int* hseq) {
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        const int s = *dseq + 1;
        *dseq = s;
        *reinterpret_cast<volatile int*>(hseq) = s;
        __threadfence_system();
    }
}

__global__ __launch_bounds__(BLOCK) void k_publish(const uint32_t* __restrict__ src, int nwords, uint32_t* dst,
                                                   int* dseq, int* hseq) {
    for (int k = threadIdx.x; k < nwords; k += BLOCK) dst[k] = src[k];
    publish_seq(dseq, hseq);
}

// k_finish (one workgroup) that then publishes the whole state as k_publish
// does: the Picard error's reduction and the host read-back in one kernel.
// gate_in non-null and *gate_in == 0: this iteration was enqueued ahead and
// is not needed; nothing is reduced or published.  gate_out non-null: opens
// the next Picard iteration, enqueued ahead (*gate_out = 1), when this one
// has not converged (err >= gtol, not NaN), its Krylov solve is done and
// gallow (another iteration is allowed); the host evaluates the same test on
// the published state.
template <int ND, class Op>
__global__ __launch_bounds__(BLOCK) void k_finish_pub(Op op, const double* __restrict__ partials, int prow,
                                                      int nblk, KState* st, uint32_t* hdst, int* dseq, int* hseq,
                                                      const int* gate_in, int* gate_out, double gtol, int gallow) {
    if (gate_in != nullptr && *gate_in == 0) return;   // uniform
    if (!op.skip(st)) {   // uniform over the workgroup
        double acc[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) acc[j] = 0.0;
        strided_sum<ND>(partials, prow, 0, nblk, acc);
        __shared__ double sh[ND * NWAVE];
        block_sum<ND>(acc, sh);
        if (threadIdx.x == 0) {
            op.fin(acc, st);
            if (gate_out != nullptr) *gate_out = st->picard_err >= gtol && st->done && gallow;
        }
    }
    // thread 0's state stores, seen by the whole workgroup (one CU)
    __threadfence_block();
    __syncthreads();
    const uint32_t* src = reinterpret_cast<const uint32_t*>(st);
    for (int k = threadIdx.x; k < (int)(sizeof(KState) / 4); k += BLOCK) hdst[k] = src[k];
    publish_seq(dseq, hseq);
}

// Counts of running instances (one workgroup): out[0] instances whose Krylov
// solve is still running, out[1] instances still iterating Picard, out[2] the
// most Krylov steps of a solve, out[3] instances still iterating Picard whose
// Jacobi refinement sweep has stalled.  hseq non-null: out is pinned host memory and
// the counts are published as by k_publish.
__global__ __launch_bounds__(BLOCK) void k_ens_status(const KState* __restrict__ st, int n_inst, int* out,
                                                      int* dseq, int* hseq) {
    // every flag of an instance loaded at once (no load behind another's
    // test), counts combined by wave butterflies and the 4 wave totals
    int kact = 0, pact = 0, ms = 0, sta = 0;
    for (int i = threadIdx.x; i < n_inst; i += BLOCK) {
        const int pd = st[i].pdone, dn = st[i].done, rs = st[i].rstall, sp = st[i].steps;
        pact += pd == 0;
        kact += (pd == 0) & (dn == 0);
        sta += (pd == 0) & (rs != 0);
        ms = max(ms, sp);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        kact += __shfl_xor(kact, o, 64);
        pact += __shfl_xor(pact, o, 64);
        sta += __shfl_xor(sta, o, 64);
        ms = max(ms, __shfl_xor(ms, o, 64));
    }
    __shared__ int wv[4][NWAVE];
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        wv[0][w] = kact;
        wv[1][w] = pact;
        wv[2][w] = ms;
        wv[3][w] = sta;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) {
            c0 += wv[0][w];
            c1 += wv[1][w];
            c2 = max(c2, wv[2][w]);
            c3 += wv[3][w];
        }
        out[0] = c0;
        out[1] = c1;
        out[2] = c2;
        out[3] = c3;
    }
    if (hseq != nullptr) publish_seq(dseq, hseq);
}

// Active-entry list of the row table: act[0] = number of entries, act[1..] =
// the entries of the instances still active, in instance order.  mode 0:
// Krylov solve running (!pdone && !done); mode 1: Picard iterating (!pdone).
// Workgroup i places instance i: its offset is the entry count of the active
// instances before it (one pass over their flags), its entries are written
// in parallel.
// mode 0: Krylov solve running; 1: Picard iterating; 2: mixed-precision
// refinement still running (rdone clear).
__device__ inline bool ens_on(const KState& s, int mode) {
    const int pd = s.pdone, rd = s.rdone, dn = s.done;   // loaded together
    return (pd == 0) & ((mode == 1) | (mode == 2 ? rd == 0 : dn == 0));
}

__global__ __launch_bounds__(BLOCK) void k_ens_compact(const KState* __restrict__ st, int n_inst,
                                                       const int* __restrict__ iblk, int* act, int mode,
                                                       int* act_zero) {
    const int i = blockIdx.x;
    int c = 0;
    for (int j = threadIdx.x; j < i; j += BLOCK) {
        const int nb = iblk[j + 1] - iblk[j];
        c += ens_on(st[j], mode) ? nb : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    __shared__ int ws[NWAVE];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    int pos = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) pos += ws[w];
    const bool on = ens_on(st[i], mode);
    const int b0 = iblk[i], nb = iblk[i + 1] - b0;
    if (on)
        for (int b = threadIdx.x; b < nb; b += BLOCK) act[1 + pos + b] = b0 + b;
    if (i == n_inst - 1 && threadIdx.x == 0) act[0] = pos + (on ? nb : 0);
    // the other Krylov list's count: zero for the next list built by
    // k_finish_ens_k (nothing reads that list here)
    if (act_zero != nullptr && i == 0 && threadIdx.x == 0) *act_zero = 0;
}

// k_finish_ens that also rebuilds the Krylov active list (mode as ens_on) from
// the states it has just finished, in place of the k_ens_compact launch that
// would follow it.  Workgroup i appends instance i's entries at a position
// taken from the list's counter (act_out[0], zero on entry), so the entry
// order of the list varies from run to run; the results do not, since every
// entry is computed on its own and stores its partial rows in its own
// column.  act_zero: the count of the list being replaced, set to zero for
// the next rebuild (nothing reads that list in this kernel).
template <int ND, class Op>
__global__ __launch_bounds__(BLOCK) void k_finish_ens_k(Op op, const double* __restrict__ partials, int prow,
                                                        const int* __restrict__ iblk, KState* st0, int* act_out,
                                                        int* act_zero, int mode) {
    KState* st = st0 + blockIdx.x;
    const bool pd = st->pdone != 0, sk = op.skip(st);   // loaded together
    if (!(pd | sk)) {   // uniform over the workgroup
        double acc[ND];
        ens_reduce<ND>(partials, prow, iblk[blockIdx.x], iblk[blockIdx.x + 1], acc);
        if (threadIdx.x == 0) op.fin(acc, st);
    }
    __shared__ int pos;
    const int b0 = iblk[blockIdx.x], nb = iblk[blockIdx.x + 1] - b0;
    if (threadIdx.x == 0) {   // the state as this thread has just left it
        pos = ens_on(*st, mode) ? atomicAdd(act_out, nb) : -1;
        if (blockIdx.x == 0) *act_zero = 0;
    }
    __syncthreads();
    if (pos >= 0)
        for (int b = threadIdx.x; b < nb; b += BLOCK) act_out[1 + pos + b] = b0 + b;
}

// Per-instance statistics of the last time step.
__global__ void k_ens_collect(const KState* __restrict__ st, int n_inst, int* ip, int* ik, int* ic, double* er) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) {
        ip[i] = st[i].pit;
        ik[i] = st[i].kry;
        ic[i] = st[i].picard_err < st[i].picard_tol;
        er[i] = st[i].picard_err;
    }
}

// Iterative-refinement sweeps per solve (single system and ensembles).
constexpr int MAX_REFINE = 50;
// Jacobi inner steps per refinement sweep: at most, and by default.
constexpr int MAX_JAC_STEPS = 16;
constexpr int JAC_STEPS = 3;
// A Jacobi sweep stalls unless it brings |r| below this fraction of |r| before it.
constexpr double JAC_STALL = 0.5;

// Ensemble mixed precision, per instance: FP64 residual r = b - A x of the
// instances still refining; the finish records |r| and |b| and marks the
// refinement converged (rdone; done too, which ends the instance's solve for
// the loops that test done) once |r| <= max(tol |b|, atol) or after
// MAX_REFINE sweeps.  first: the residual of the Picard iteration's start.
struct RefResidE {
    static constexpr int ND = 2;
    const double* b;
    const double* x;
    double* r;
    double tol, atol;
    int first;
    // stall > 0: the residual after a sweep of Jacobi inner steps (steps of
    // them); rstall is raised when |r| has not fallen below stall times the
    // previous |r| and the refinement goes on.  The steps add to the
    // iteration count.
    double stall = 0.0;
    int steps = 0;
    __device__ bool skip(const KState* s) const { return !first && s->rdone; }
    __device__ void load(const KState*) {}
    __device__ double xv(int j) const { return x[j]; }
    __device__ void row(int i, double ax, double* acc) const {
        const double bi = b[i];
        const double ri = bi - ax;
        r[i] = ri;
        acc[0] += ri * ri;
        acc[1] += bi * bi;
    }
    __device__ void fin(const double* t, KState* s) const {
        const double prev = s->rnorm;
        s->rnorm = sqrt(t[0]);
        s->bnorm = sqrt(t[1]);
        s->refine = first ? 0 : s->refine + 1;
        s->rdone = !(s->rnorm > fmax(tol * s->bnorm, atol)) || s->refine >= MAX_REFINE;
        s->done = s->rdone;
        if (first) {
            s->rstall = 0;
            s->steps = 0;
        } else if (stall > 0.0) {
            s->steps += steps;
            if (!s->rdone && !(s->rnorm <= stall * prev)) s->rstall = 1;
        }
    }
};

// Mixed precision with Jacobi inner steps.  A refinement sweep solves the
// correction equation A d = r~ (r~ = r / |r| in FP32, A and M = diag(A)^-1 in
// FP32) by k steps of d <- d + M (r~ - A d) from d = 0, with no dot products,
// then adds |r| d to the FP64 iterate.  The first step's d = M r~ is formed at
// the gathered columns of the second step, and the last step adds its result
// to x, so a sweep is k - 1 SpMV launches.  On the time-step operators of
// small dt (diagonally dominant: |I - M A| about 1e-3 to 1e-2) three steps
// reach the FP32 floor of the correction; the FP64 residual after the sweep
// decides whether another sweep follows.  gated: the state says whether the
// refinement has converged (rdone), and then the launch does nothing.
struct JacFirst {     // d = M r~ + M (r~ - A M r~); x += |r| d (x set) or d, r~ stored
    static constexpr int ND = 0;
    const double* r;
    const float* dinv;
    float* r32;
    float* dout;
    double* x;
    int gated;
    double inv, scale;
    __device__ bool skip(const KState* s) const { return gated && s->rdone; }
    __device__ void load(const KState* s) {
        scale = s->rnorm;
        inv = scale > 0.0 ? 1.0 / scale : 0.0;
    }
    __device__ float rt(int j) const { return (float)(r[j] * inv); }
    __device__ double xv(int j) const { return (double)(dinv[j] * rt(j)); }
    __device__ void row(int i, double ax, double*) const {
        const float ri = rt(i);
        const float mi = dinv[i];
        const double d = (double)(mi * ri) + (double)mi * ((double)ri - ax);
        if (x != nullptr) {
            x[i] += scale * d;
        } else {
            r32[i] = ri;
            dout[i] = (float)d;
        }
    }
    __device__ void fin(const double*, KState*) const {}
};

struct JacOnly {      // one step: x += |r| M r~
    static constexpr int ND = 0;
    const double* r;
    const float* dinv;
    double* x;
    int gated;
    double inv, scale;
    __device__ bool skip(const KState* s) const { return gated && s->rdone; }
    __device__ void load(const KState* s) {
        scale = s->rnorm;
        inv = scale > 0.0 ? 1.0 / scale : 0.0;
    }
    __device__ void row(int i, double*) const { x[i] += scale * (double)(dinv[i] * (float)(r[i] * inv)); }
    __device__ void fin(const double*, KState*) const {}
};

struct JacStep {      // d' = d + M (r~ - A d); x += |r| d' (x set) or d' stored
    static constexpr int ND = 0;
    const float* din;
    const float* r32;
    const float* dinv;
    float* dout;
    double* x;
    int gated;
    double scale;
    __device__ bool skip(const KState* s) const { return gated && s->rdone; }
    __device__ void load(const KState* s) { scale = s->rnorm; }
    __device__ double xv(int j) const { return (double)din[j]; }
    __device__ void row(int i, double ax, double*) const {
        const double d = (double)din[i] + (double)dinv[i] * ((double)r32[i] - ax);
        if (x != nullptr) x[i] += scale * d;
        else dout[i] = (float)d;
    }
    __device__ void fin(const double*, KState*) const {}
};

struct RefScaleE {  // r32 = r / |r|, d32 = 0 on the instances whose inner solve starts
    static constexpr int ND = 0;
    const double* r;
    float* r32;
    float* d32;
    double inv;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) { inv = s->rnorm > 0.0 ? 1.0 / s->rnorm : 0.0; }
    __device__ void row(int i, double*) const {
        r32[i] = (float)(r[i] * inv);
        d32[i] = 0.0f;
    }
    __device__ void fin(const double*, KState*) const {}
};

// Start of an inner FP32 solve from a zero guess, fused with the scaling of
// its right-hand side: r32 = r / |r|, d32 = 0, then the method's start row
// (BicgInit / CgInit) with A d32 = 0, so no SpMV.  Instances whose solve does
// not start (done set) keep their state.
template <class Init>
struct RefStartE {
    static constexpr int ND = Init::ND;
    Init init;            // init.v: b = r32, x0 = x = d32
    const double* r;
    float* r32;
    float* d32;
    double inv;
    __device__ bool skip(const KState* s) const { return s->done; }
    __device__ void load(const KState* s) { inv = s->rnorm > 0.0 ? 1.0 / s->rnorm : 0.0; }
    __device__ void row(int i, double* acc) const {
        r32[i] = (float)(r[i] * inv);
        d32[i] = 0.0f;
        init.row(i, 0.0, acc);
    }
    __device__ void fin(const double* t, KState* s) const { init.fin(t, s); }
};

struct RefCorrectE {  // x += |r| d32 on the instances refined in this sweep
    static constexpr int ND = 0;
    double* x;
    const float* d32;
    double scale;
    __device__ bool skip(const KState* s) const { return s->rdone; }
    __device__ void load(const KState* s) { scale = s->rnorm; }
    __device__ void row(int i, double*) const { x[i] += scale * (double)d32[i]; }
    __device__ void fin(const double*, KState*) const {}
};

// A solver-start functor restricted to the instances whose solve starts
// (done clear): the others keep their state.
template <class Op>
struct ActiveOnly : Op {
    __device__ bool skip(const KState* s) const { return s->done; }
};

struct RefScale {  // r32 = r / |r|; d32 = 0
    static constexpr int ND = 0;
    const double* r;
    float* r32;
    float* d32;
    double inv;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState* s) { inv = s->rnorm > 0.0 ? 1.0 / s->rnorm : 0.0; }
    __device__ void row(int i, double*) const {
        r32[i] = (float)(r[i] * inv);
        d32[i] = 0.0f;
    }
    __device__ void fin(const double*, KState*) const {}
};

// RefStartE for one system: the scale |r| comes from the refinement state,
// the solver state written by the start is the inner one.
template <class Init>
struct RefStart {
    static constexpr int ND = Init::ND;
    Init init;
    const double* r;
    float* r32;
    float* d32;
    const double* rnorm;  // device address of the refinement state's |r|
    double inv;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) { inv = *rnorm > 0.0 ? 1.0 / *rnorm : 0.0; }
    __device__ void row(int i, double* acc) const {
        r32[i] = (float)(r[i] * inv);
        d32[i] = 0.0f;
        init.row(i, 0.0, acc);
    }
    __device__ void fin(const double* t, KState* s) const { init.fin(t, s); }
};

struct RefCorrect {  // x += |r| d32
    static constexpr int ND = 0;
    double* x;
    const float* d32;
    double scale;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState* s) { scale = s->rnorm; }
    __device__ void row(int i, double*) const { x[i] += scale * (double)d32[i]; }
    __device__ void fin(const double*, KState*) const {}
};

struct SetK {  // the assembly's nodal K from a caller array
    static constexpr int ND = 0;
    const double* K;
    double* krec;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double*) const { krec[i] = K[i]; }
    __device__ void fin(const double*, KState*) const {}
};

struct ToF32 {
    static constexpr int ND = 0;
    const double* a;
    float* o;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double*) const { o[i] = (float)a[i]; }
    __device__ void fin(const double*, KState*) const {}
};

// GMRES pieces that are not elementwise --------------------------------------
template <class T>
__global__ __launch_bounds__(BLOCK) void k_gm_dots(KVec<T> v, int j, Red red) {
    KState* s = red.st;
    if (s->done || !s->active) return;
    const int c0 = blockIdx.y * 8;
    const int nb = gridDim.x;
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0.0;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < v.n; i += nb * BLOCK) {
        const double wi = (double)v.w[i];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c0 + u <= j) acc[u] += (double)v.V[(long)(c0 + u) * v.ldv + i] * wi;
    }
    store_partials<8>(acc, red.partials, red.prow);
}

// h[l] = sum over the nb row-blocks of column l's partials (l <= j), in order.
__global__ __launch_bounds__(BLOCK) void k_gm_dots_finish(KState* s, const double* __restrict__ partials, int prow,
                                                           int j, int nb) {
    if (s->done || !s->active) return;
    for (int l = threadIdx.x; l <= j; l += BLOCK) {
        const int by = l / 8, u = l % 8;
        double sum = 0.0;
        for (int bx = 0; bx < nb; ++bx) sum += partials[(size_t)u * prow + by * nb + bx];
        s->hbuf[l] = sum;
    }
}

__global__ void k_gm_cycle_begin(KState* s, double* H, int R) {
    if (s->done) return;
    for (int idx = threadIdx.x; idx < R * (R + 1); idx += blockDim.x) {
        const int i = idx / (R + 1), l = idx % (R + 1);
        H[idx] = (i == l) ? 1.0 : 0.0;
    }
    if (threadIdx.x == 0) {
        s->active = 1;
        s->jdone = -1;
    }
}

// y = (H H^T)^{-1} (H beta e1) by Cholesky (JAX _lstsq: normal equations +
// solve(assume_a='pos')).  One workgroup, R <= 64.
__global__ __launch_bounds__(64) void k_gm_lstsq(KState* s, const double* H, int R) {
    if (s->done) return;
    __shared__ double a2[MAX_RESTART * MAX_RESTART];
    __shared__ double b2[MAX_RESTART];
    const double beta = s->rnorm;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < R * R; idx += blockDim.x) {
        const int i = idx / R, k = idx % R;
        double acc = 0.0;
        for (int l = 0; l <= R; ++l) acc += H[i * (R + 1) + l] * H[k * (R + 1) + l];
        a2[idx] = acc;
    }
    for (int i = tid; i < R; i += blockDim.x) b2[i] = H[i * (R + 1)] * beta;
    __syncthreads();
    // in-place lower Cholesky
    for (int c = 0; c < R; ++c) {
        if (tid == 0) {
            double d = a2[c * R + c];
            for (int p = 0; p < c; ++p) d -= a2[c * R + p] * a2[c * R + p];
            a2[c * R + c] = sqrt(d);
        }
        __syncthreads();
        for (int i = c + 1 + tid; i < R; i += blockDim.x) {
            double v = a2[i * R + c];
            for (int p = 0; p < c; ++p) v -= a2[i * R + p] * a2[c * R + p];
            a2[i * R + c] = v / a2[c * R + c];
        }
        __syncthreads();
    }
    if (tid == 0) {
        double z[MAX_RESTART];
        for (int i = 0; i < R; ++i) {
            double v = b2[i];
            for (int p = 0; p < i; ++p) v -= a2[i * R + p] * z[p];
            z[i] = v / a2[i * R + i];
        }
        for (int i = R - 1; i >= 0; --i) {
            double v = z[i];
            for (int p = i + 1; p < R; ++p) v -= a2[p * R + i] * s->y[p];
            s->y[i] = v / a2[i * R + i];
        }
    }
}

__global__ void k_nan_check(KState* s, const double* x, int n) {
    // GMRES info: -1 if x has a NaN (jnp.isnan(norm(x))); run on one block.
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    int local = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) local |= (x[i] != x[i]);
    if (local) atomicOr(&bad, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        s->nan_flag = bad;
        if (bad) s->code = -1;
    }
}

// Adjoint parameter reduction: out[0..1] += (sum a[0:len], sum a[len:2 len]).
struct AdjSum {
    static constexpr int ND = 2;
    const double* a;
    int len;
    double* out;
    __device__ bool skip(const KState*) const { return false; }
    __device__ void load(const KState*) {}
    __device__ void row(int i, double* acc) const {
        acc[0] += a[i];
        acc[1] += a[len + i];
    }
    __device__ void fin(const double* t, KState*) const {
        out[0] += t[0];
        out[1] += t[1];
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
template <class T>
struct KWork {
    int n = 0;
    T *r = nullptr, *rh = nullptr, *p = nullptr, *ph = nullptr, *q = nullptr, *sh = nullptr,
      *t = nullptr, *z = nullptr, *w = nullptr, *V = nullptr;
    int vcols = 0;
};

struct wspm_ctx {
    std::string err;
    hipStream_t stream = nullptr;
    int dim = 0, n = 0, E = 0;
    int64_t nnz = 0;
    int max_row = 0;
    int asm_win = 0, asm_inc = 0, asm_grid = 0;   // assembly LDS window / incidence capacity
    size_t asm_lds = 0;
    bool asm_reg = false;                         // rows fit the register assembly kernel
    // ELL copy of the operator for the time-step solves (rows of 7 or 8
    // entries): column table, width, whether this run's steps use it, and
    // whether the SpMV launches read ELL right now
    int32_t* ell_col = nullptr;
    int16_t* ell_col16 = nullptr;   // the same table as offsets from the row, when they all fit in 16 bits
    int ell_w = 0;
    int2* ell_inc = nullptr;        // neighbourhood assembly: incidences in ELL layout
    uint32_t* ell_info = nullptr;   // and rl | dg << 4 | nq << 8 per row
    int ell_iw = 0;
    // the ELL operator in fan order and the fan kernel's tables (build_fan)
    bool fan = false;
    int32_t *fan_inc = nullptr, *fan_pred = nullptr;
    uint32_t* fan_info = nullptr;
    // node windows of the windowed fan assembly (build_node_windows)
    int32_t* nwin = nullptr;
    uint16_t *nwin_col = nullptr, *nwin_pred = nullptr;
    int nwin_cap = 0;
    bool nwin_ok = false;
    std::vector<RowTab> htab;       // ensembles: the row table on the host
    int fan_iw = 0;
    // finish-free CG / BiCGSTAB: second partial-row region and the iteration
    // histories of the main (0) and FP32 inner (1) solves
    double* partials2 = nullptr;
    KHist* khist[2] = {nullptr, nullptr};
    int khist_cap[2] = {0, 0};
    // a second Picard iteration enqueued ahead of the first one's read-back:
    // its gate (opened by the first iteration's finish), its Krylov history
    // and its pinned state mirror
    int* gate = nullptr;
    KHist* khist_ahead = nullptr;
    int khist_ahead_cap = 0;
    KState* hstate_ahead = nullptr;
    bool ell = false;
    bool spmv_ell = false;
    int asm_kernel = ASM_K_LDS;                   // ASM_K_*
    int prow = MAX_BLOCKS;                        // partials rows
    int lanes = 8;
    int spmv_grid = 1, vec_grid = 1;
    HostPattern hp;
    // device mesh/pattern
    int32_t *row_ptr = nullptr, *cols = nullptr, *diag_pos = nullptr, *inc_ptr = nullptr;
    int2* inc = nullptr;            // (element << 2 | vertex, slot word) per incidence
    int4* conn4 = nullptr;          // padded connectivity
    double* ncoord = nullptr;       // node coordinates: (x, y) / (x, y, z, -) per node (NodeRecs)
    double* Kn = nullptr;           // nodal conductivity read by the assembly (NodeRecs)
    bool nrec_dirty = false;        // Kn changed by a direct assembly call
    uint8_t* bc_flag = nullptr;
    double* bc_val = nullptr;
    int n_bc = 0;
    QuadConst qc;
    // reductions
    double* partials = nullptr;
    KState* states = nullptr;     // [3]: main, inner, refinement
    KState* hstate = nullptr;     // pinned mirror
    double* H = nullptr;          // GMRES Hessenberg [MAX_RESTART][MAX_RESTART+1]
    double* H32 = nullptr;
    KWork<double> w64;
    KWork<float> w32;
    // mixed precision
    double* ref_r = nullptr;
    float *ref_r32 = nullptr, *ref_d32 = nullptr, *A32 = nullptr, *dinv32 = nullptr;
    float* ref_d32b = nullptr;      // second iterate of the Jacobi inner steps (more than 3)
    bool jac_inner_off = false;     // a Jacobi sweep stalled: Krylov inner solves from then on
    double jac_stall = 0.5;         // a Jacobi sweep stalls unless |r| falls below this times |r| before it
    // Richards state
    bool rich_init = false;
    bool has_ks = false;
    wspm_richards_params prm{};
    SoilParams sp{};
    double *hA = nullptr, *hB = nullptr, *Cn = nullptr, *thm = nullptr,
           *thn = nullptr, *Ks = nullptr, *Av = nullptr, *bv = nullptr, *dinv = nullptr;
    int last_iters[3] = {1, 1, 1};
    // timing
    bool timing = false;
    bool timing_fp32 = false;       // time the FP32 SpMV launches instead of the FP64 ones
    bool timing_asm = false;        // also time the assembly launches (separate events and slots)
    int64_t n_launch = 0;           // kernel launches and device copies enqueued (wspm_launch_count)
    int64_t n_picard = 0;           // Picard rounds of the time steps
    int* asm_tflags = nullptr;      // [TIMING_SLOTS] rows each timed assembly launch worked on
    std::vector<hipEvent_t> asm_ev;
    size_t asm_ev_used = 0;
    int* tflags = nullptr;          // [TIMING_SLOTS] rows each timed launch worked on (0: skipped)
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    // ensemble (tab != nullptr): instances concatenated, one state per instance
    int n_inst = 1;
    RowTab* tab = nullptr;          // [n_tab] rows and instance of each workgroup
    int n_tab = 0;
    int* iblk = nullptr;            // [n_inst + 1] first workgroup of each instance
    int* ens_status = nullptr;      // device int[4] (k_ens_status)
    int* hstatus = nullptr;         // pinned mirror
    int* ens_out = nullptr;         // k_ens_collect output: 3 int + 1 double per instance
    int* kact = nullptr;            // [1 + n_tab] entries of instances whose Krylov solve runs
    int* kact_alt = nullptr;        // the list k_finish_ens_k builds next (count zero between builds)
    bool ens_list_kernel = false;   // WSPM_ENS_LIST=kernel: Krylov lists rebuilt by k_ens_compact only
    int* pact = nullptr;            // [1 + n_tab] entries of instances still iterating Picard
    int ens_grid = 0;               // workgroups of an ensemble launch
    int ens_grid_cands = 0;         // grids the autotuner timed (0: fixed by WSPM_ENS_GRID)
    float ens_tune_us = 0.0f;       // the chosen grid's SpMV time over the whole operator
    int ens_compact_every = 1;      // Krylov iterations between rebuilds of the active list
    // boundary node lists for edge terms: [0] inlet (top), [1] outlet (bottom)
    struct EdgeList {
        int cnt = 0;
        int32_t *nodes = nullptr, *slot_prev = nullptr, *slot_next = nullptr;
        std::string unjoined;   // the first consecutive pair not joined by a mesh edge ("" if none)
    } edges[2];
    // discrete adjoint workspace (allocated on first use)
    struct {
        double *nodal = nullptr, *ml = nullptr, *J = nullptr, *JT = nullptr, *dinv = nullptr, *rhs = nullptr,
               *dldh = nullptr, *ge = nullptr, *gn = nullptr;
        int32_t* tpos = nullptr;
    } adj;
    // species adjoint workspace (allocated on first use)
    struct {
        double *f = nullptr, *fnext = nullptr, *mth = nullptr, *mT = nullptr, *A = nullptr, *AT = nullptr,
               *dinv = nullptr, *b = nullptr, *rhs = nullptr, *gf = nullptr, *w = nullptr, *gn = nullptr;
    } adjtr;
    double* diag_out = nullptr;     // device scalars of the diagnostics reductions
    // direct solver (allocated on first use)
    struct {
        Band b{};
        bool ready = false, reordered = false;
        std::vector<int32_t> perm;
    } band;
    // general preconditioners (allocated on first use)
    struct {
        Levels lo, up;
        bool levels = false;
        double *binv = nullptr, *dg = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
    } pc;
    bool rich_neumann = false, rich_freedrain = false;
    double rich_flux = 0.0;
    TrQuad tq{};
    // host read-back (k_publish): device publication counter, pinned
    // sequence word, the host's count of publications; fast_poll false: copy
    // + stream synchronisation instead (WSPM_POLL=sync)
    int* dseq = nullptr;
    int* hseq = nullptr;
    int seq = 0;
    bool fast_poll = true;
    int spin_us = 1000;         // spin on the word this long before blocking on the stream (WSPM_SPIN_US)
    // HIP graphs of whole Picard iterations, replayed while the step's
    // parameters repeat (WSPM_GRAPHS=1; by default every kernel is launched
    // on the stream)
    struct StepGraph {
        int kind;               // GRAPH_* below
        double dt;
        const double* hA;       // the Picard iterate's buffer (they alternate)
        int K;                  // Krylov iterations in the graph
        int epoch;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> graphs;
    int graph_epoch = 0;        // bumped by every call that changes what a graph captured
    bool use_graphs = false;
    hipStream_t cap_stream = nullptr;   // capture stream (the context's may be the null stream)
    std::vector<void*> allocs;
    std::vector<void*> host_allocs;
};

namespace {

template <class P>
P* dalloc(wspm_ctx* ctx, size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(P));
    if (e != hipSuccess) {
        ctx->err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
        throw Err{};
    }
    ctx->allocs.push_back(p);
    return static_cast<P*>(p);
}

template <class P>
P* dupload(wspm_ctx* ctx, const P* src, size_t count) {
    P* d = dalloc<P>(ctx, count);
    HIPCHK(hipMemcpy(d, src, count * sizeof(P), hipMemcpyHostToDevice));
    return d;
}

// The nodal K the assembly reads, and the node data of the element kernels.
double* krec_of(wspm_ctx* ctx) { return ctx->Kn; }
NodeRecs nrec_of(const wspm_ctx* ctx) { return NodeRecs{ctx->ncoord, ctx->Kn, ctx->dim}; }

// Address of a field of a device-resident KState (no host dereference).
template <class F>
F* state_field(KState* ds, size_t off) {
    return reinterpret_cast<F*>(reinterpret_cast<char*>(ds) + off);
}

// State array `which` (0 main, 1 inner, 2 refinement): one KState per instance.
KState* states_of(wspm_ctx* ctx, int which) { return ctx->states + (size_t)which * ctx->n_inst; }

Red red_of(wspm_ctx* ctx, int which) {
    return Red{ctx->partials, ctx->prow, states_of(ctx, which)};
}

// Which table entries an ensemble launch walks.
enum { LIST_ALL = 0, LIST_K = 1, LIST_P = 2 };

EnsWalk walk_of(wspm_ctx* ctx, int list) {
    return EnsWalk{ctx->tab, list == LIST_K ? ctx->kact : list == LIST_P ? ctx->pact : nullptr, ctx->n_tab};
}

// Rebuild an active-entry list (LIST_K or LIST_P) from the instance states;
// mode as ens_on (default: 0 for LIST_K, 1 for LIST_P).
void ens_compact(wspm_ctx* ctx, int list, int mode = -1) {
    if (mode < 0) mode = list == LIST_K ? 0 : 1;
    LAUNCH(k_ens_compact, dim3(ctx->n_inst), dim3(BLOCK), 0, ctx->stream, (const KState*)ctx->states,
                       ctx->n_inst, (const int*)ctx->iblk, list == LIST_K ? ctx->kact : ctx->pact, mode,
                       list == LIST_K ? ctx->kact_alt : nullptr);
    HIPCHK(hipGetLastError());
}

// The Krylov list is rebuilt by the finish kernel of the iteration's last
// reduction (k_finish_ens_k) instead of a k_ens_compact launch after it,
// except under graph replay (the captured kernels keep the list pointers of
// the capture, and this swaps them).
// WSPM_ENS_LIST=kernel keeps the separate launch.
bool fused_compact(const wspm_ctx* ctx) { return ctx->tab && !ctx->use_graphs && !ctx->ens_list_kernel; }

// Finish of a reduction kernel: one workgroup, or one per instance (ensemble).
template <class Op, bool RESET_PEV = false>
void launch_finish(wspm_ctx* ctx, const Op& op, int nblk, int which, bool compact_k = false, int cmode = 0) {
    if constexpr (Op::ND > 0) {
        if (ctx->tab && compact_k) {
            LAUNCH((k_finish_ens_k<Op::ND, Op>), dim3(ctx->n_inst), dim3(BLOCK), 0, ctx->stream, op,
                               ctx->partials, ctx->prow, ctx->iblk, states_of(ctx, which), ctx->kact_alt,
                               ctx->kact, cmode);
            std::swap(ctx->kact, ctx->kact_alt);
        } else if (ctx->tab) {
            LAUNCH((k_finish_ens<Op::ND, Op, RESET_PEV>), dim3(ctx->n_inst), dim3(BLOCK), 0,
                               ctx->stream, op, ctx->partials, ctx->prow, ctx->iblk, states_of(ctx, which));
        } else {
            LAUNCH((k_finish<Op::ND, Op>), dim3(1), dim3(BLOCK), 0, ctx->stream, op, ctx->partials,
                               ctx->prow, nblk, states_of(ctx, which));
        }
        HIPCHK(hipGetLastError());
    }
}

// compact_k: the finish also rebuilds the Krylov list (ensembles; cmode as ens_on)
template <class Op>
void launch_vec(wspm_ctx* ctx, const Op& op, int which, int list = LIST_ALL, bool compact_k = false,
                int cmode = 0) {
    const int grid = ctx->tab ? ctx->ens_grid : ctx->vec_grid;
    LAUNCH(k_vec<Op>, dim3(grid), dim3(BLOCK), 0, ctx->stream, ctx->n, op, red_of(ctx, which),
                       walk_of(ctx, list));
    HIPCHK(hipGetLastError());
    launch_finish(ctx, op, grid, which, compact_k, cmode);
}

hipEvent_t next_event(wspm_ctx* ctx) {
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

constexpr size_t TIMING_SLOTS = (size_t)1 << 18;

// SpMV launch; a timed launch (benchmark timing on) records its start and
// stop events as part of the kernel's own dispatch (hipExtLaunchKernelGGL),
// so the elapsed time is the kernel's execution, as a kernel trace reports it,
// without the dispatch gaps of separately recorded events.
// part: partial rows to write (default ctx->partials); finish false: no
// finish kernel (a folded Krylov kernel's consumer reduces the rows itself).
template <class T, class Epi>
void launch_spmv(wspm_ctx* ctx, const T* av, const Epi& epi, int which, bool timed = false, int list = LIST_ALL,
                 double* part = nullptr, bool finish = true, bool compact_k = false) {
    const bool t = timed && ctx->timing && ctx->timing_fp32 == std::is_same<T, float>::value &&
                   ctx->ev_used / 2 < TIMING_SLOTS;
    int* tflag = t ? ctx->tflags + ctx->ev_used / 2 : nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (t) {
        e0 = next_event(ctx);
        e1 = next_event(ctx);
    }
    Red red = red_of(ctx, which);
    if (part) red.partials = part;
    const EnsWalk tab = walk_of(ctx, list);
    auto go = [&](auto kernel, dim3 g, auto... args) {
        ++ctx->n_launch;
        if (t) hipExtLaunchKernelGGL(kernel, g, dim3(BLOCK), 0, ctx->stream, e0, e1, 0, args...);
        else hipLaunchKernelGGL(kernel, g, dim3(BLOCK), 0, ctx->stream, args...);
        HIPCHK(hipGetLastError());
    };
    if (ctx->spmv_ell) {
        const int eg = ctx->tab ? ctx->ens_grid : ctx->vec_grid;
        if (ctx->ell_col16) {
            const int16_t* c16 = ctx->ell_col16;
            if (ctx->ell_w == 7) go(k_spmv_ell<7, T, Epi, int16_t>, dim3(eg), ctx->n, c16, av, epi, red, tflag, tab);
            else if (ctx->ell_w == 8) go(k_spmv_ell<8, T, Epi, int16_t>, dim3(eg), ctx->n, c16, av, epi, red, tflag, tab);
            else if (ctx->ell_w == ELL_W3D - 1)
                go(k_spmv_ell<ELL_W3D - 1, T, Epi, int16_t>, dim3(eg), ctx->n, c16, av, epi, red, tflag, tab);
            else go(k_spmv_ell<ELL_W3D, T, Epi, int16_t>, dim3(eg), ctx->n, c16, av, epi, red, tflag, tab);
        } else if (ctx->ell_w == 7) {
            go(k_spmv_ell<7, T, Epi>, dim3(eg), ctx->n, (const int32_t*)ctx->ell_col, av, epi, red, tflag, tab);
        } else {
            go(k_spmv_ell<8, T, Epi>, dim3(eg), ctx->n, (const int32_t*)ctx->ell_col, av, epi, red, tflag, tab);
        }
        if (finish) launch_finish(ctx, epi, eg, which, compact_k);
        return;
    }
    const int grid = ctx->tab ? ctx->ens_grid : ctx->spmv_grid;
    const dim3 g(grid);
    const int32_t* rp = ctx->row_ptr;
    const int32_t* ci = ctx->cols;
    switch (ctx->lanes) {
        case 2: go(k_spmv<2, T, Epi>, g, ctx->n, rp, ci, av, epi, red, tflag, tab); break;
        case 4: go(k_spmv<4, T, Epi>, g, ctx->n, rp, ci, av, epi, red, tflag, tab); break;
        case 8: go(k_spmv<8, T, Epi>, g, ctx->n, rp, ci, av, epi, red, tflag, tab); break;
        case 16: go(k_spmv<16, T, Epi>, g, ctx->n, rp, ci, av, epi, red, tflag, tab); break;
        case 32: go(k_spmv<32, T, Epi>, g, ctx->n, rp, ci, av, epi, red, tflag, tab); break;
        default: go(k_spmv<64, T, Epi>, g, ctx->n, rp, ci, av, epi, red, tflag, tab); break;
    }
    if (finish) launch_finish(ctx, epi, grid, which, compact_k);
}

// Wait for the next publication (k_publish / k_ens_status with a sequence
// word): spin on the pinned word; if the stream drains or ctx->spin_us
// passes, block on the stream instead, after which the word must be there.
// A word past the one awaited is accepted: a publication enqueued behind the
// awaited one (a Picard iteration enqueued ahead) may already have run when
// the host falls back to synchronising the stream, and it writes its state
// into a mirror of its own.
void wait_published(wspm_ctx* ctx) {
    const int want = ++ctx->seq;
    volatile int* f = ctx->hseq;
    const auto t0 = std::chrono::steady_clock::now();
    const auto limit = std::chrono::microseconds(ctx->spin_us);
    for (unsigned spin = 1; (int)((unsigned)*f - (unsigned)want) < 0; ++spin) {
        if (ctx->spin_us > 0 && (spin & 1023u) != 0) continue;
        const hipError_t q = ctx->spin_us > 0 ? hipStreamQuery(ctx->stream) : hipSuccess;
        if (q != hipSuccess && q != hipErrorNotReady) HIPCHK(q);
        if (q == hipSuccess || std::chrono::steady_clock::now() - t0 > limit) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            if ((int)((unsigned)*f - (unsigned)want) < 0) {
                ctx->err = "status read-back: sequence word " + std::to_string(*f) + ", expected " +
                           std::to_string(want);
                ctx->seq = *f;
                throw Err{};
            }
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
}

// Enqueue the publication of state `which` into ctx->hstate[which].
void publish_state(wspm_ctx* ctx, int which) {
    LAUNCH(k_publish, dim3(1), dim3(BLOCK), 0, ctx->stream,
                       reinterpret_cast<const uint32_t*>(states_of(ctx, which)), (int)(sizeof(KState) / 4),
                       reinterpret_cast<uint32_t*>(ctx->hstate + which), ctx->dseq, ctx->hseq);
    HIPCHK(hipGetLastError());
}

void poll(wspm_ctx* ctx, int which) {
    if (ctx->fast_poll) {
        publish_state(ctx, which);
        wait_published(ctx);
        return;
    }
    HIPCHK(hipMemcpyAsync(ctx->hstate + which, states_of(ctx, which), sizeof(KState),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

// Ensemble: counts of running instances (k_ens_status), published straight
// into ctx->hstatus (fast_poll) or copied there.
void ens_status_publish(wspm_ctx* ctx, int which) {
    LAUNCH(k_ens_status, dim3(1), dim3(BLOCK), 0, ctx->stream, states_of(ctx, which), ctx->n_inst,
                       ctx->fast_poll ? ctx->hstatus : ctx->ens_status, ctx->dseq,
                       ctx->fast_poll ? ctx->hseq : nullptr);
    HIPCHK(hipGetLastError());
}

void poll_ens(wspm_ctx* ctx, int which) {
    ens_status_publish(ctx, which);
    if (ctx->fast_poll) {
        wait_published(ctx);
        return;
    }
    HIPCHK(hipMemcpyAsync(ctx->hstatus, ctx->ens_status, 4 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

// The Picard error and the soil model at the new iterate (one system), with
// the state published for the host when fast_poll: the error's finish kernel
// does the publication (one kernel less per Picard iteration).  picard_poll
// also waits for it.
// hdst: the pinned mirror to publish into (default ctx->hstate); gate_out /
// gtol / gallow and pop.gate as k_finish_pub's gate_out and gate_in.
void picard_enqueue(wspm_ctx* ctx, const PicardOp& pop, KState* hdst = nullptr, int* gate_out = nullptr,
                    double gtol = 0.0, int gallow = 0) {
    if (!ctx->fast_poll) {
        launch_vec(ctx, pop, 0);
        return;
    }
    const int grid = ctx->vec_grid;
    LAUNCH(k_vec<PicardOp>, dim3(grid), dim3(BLOCK), 0, ctx->stream, ctx->n, pop, red_of(ctx, 0),
                       walk_of(ctx, LIST_ALL));
    HIPCHK(hipGetLastError());
    LAUNCH((k_finish_pub<PicardOp::ND, PicardOp>), dim3(1), dim3(BLOCK), 0, ctx->stream, pop,
                       ctx->partials, ctx->prow, grid, states_of(ctx, 0),
                       reinterpret_cast<uint32_t*>(hdst ? hdst : ctx->hstate), ctx->dseq, ctx->hseq, pop.gate,
                       gate_out, gtol, gallow);
    HIPCHK(hipGetLastError());
}

void picard_poll(wspm_ctx* ctx, const PicardOp& pop) {
    picard_enqueue(ctx, pop);
    if (ctx->fast_poll) wait_published(ctx);
    else poll(ctx, 0);
}

enum { GRAPH_PICARD = 0, GRAPH_ENS_PICARD = 1, GRAPH_ENS_CHUNK = 2 };

// The graph of `record` (which enqueues on ctx->stream) for this key, captured
// on first use and replayed afterwards.  Replays issue each kernel in about
// 1.8 us on MI355X against 2.9 us for a stream launch (measured).  Capture
// records kernels only: no memset, copy or allocation is enqueued by the
// recorded functions.
template <class F>
hipGraphExec_t step_graph(wspm_ctx* ctx, int kind, double dt, int K, F&& record) {
    for (const auto& g : ctx->graphs)
        if (g.kind == kind && g.dt == dt && g.hA == ctx->hA && g.K == K && g.epoch == ctx->graph_epoch)
            return g.exec;
    if (!ctx->cap_stream) HIPCHK(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
    struct Restore {
        wspm_ctx* c;
        hipStream_t s;
        ~Restore() { c->stream = s; }
    } restore{ctx, ctx->stream};
    ctx->stream = ctx->cap_stream;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    try {
        record();
    } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(ctx->stream, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    if (ctx->graphs.size() >= 16) {
        HIPCHK(hipStreamSynchronize(restore.s));
        (void)hipGraphExecDestroy(ctx->graphs.front().exec);
        ctx->graphs.erase(ctx->graphs.begin());
    }
    ctx->graphs.push_back({kind, dt, ctx->hA, K, ctx->graph_epoch, ex});
    return ex;
}

template <class T>
KWork<T>& work_of(wspm_ctx* ctx);
template <>
KWork<double>& work_of<double>(wspm_ctx* ctx) { return ctx->w64; }
template <>
KWork<float>& work_of<float>(wspm_ctx* ctx) { return ctx->w32; }

template <class T>
KWork<T>& ensure_work(wspm_ctx* ctx, int vcols) {
    KWork<T>& w = work_of<T>(ctx);
    const int n = ctx->n;
    if (w.r == nullptr) {
        w.n = n;
        w.r = dalloc<T>(ctx, n); w.rh = dalloc<T>(ctx, n); w.p = dalloc<T>(ctx, n);
        w.ph = dalloc<T>(ctx, n); w.q = dalloc<T>(ctx, n); w.sh = dalloc<T>(ctx, n);
        w.t = dalloc<T>(ctx, n); w.z = dalloc<T>(ctx, n); w.w = dalloc<T>(ctx, n);
    }
    if (vcols > w.vcols) {
        w.V = dalloc<T>(ctx, (size_t)n * vcols);
        HIPCHK(hipMemsetAsync(w.V, 0, (size_t)n * vcols * sizeof(T), ctx->stream));
        w.vcols = vcols;
    }
    return w;
}

template <class T>
KVec<T> make_kvec(wspm_ctx* ctx, KWork<T>& w, const T* dinv, const T* b, const T* x0, T* x, bool gpc = false) {
    KVec<T> v{};
    v.n = ctx->n;
    v.b = b; v.x = x; v.x0 = x0;
    v.r = w.r; v.rh = w.rh; v.q = w.q; v.t = w.t; v.w = w.w;
    v.p = w.p;
    const bool sep = dinv != nullptr || gpc;
    v.ph = sep ? w.ph : w.p;      // M = identity: p_hat is p
    v.sh = sep ? w.sh : w.r;      // s lives in r
    v.z = sep ? w.z : w.r;
    v.gpc = gpc ? 1 : 0;
    v.V = w.V; v.ldv = ctx->n;
    v.dinv = dinv;
    return v;
}

// General preconditioners: workspace, level sets (once per mesh), setup from A.
void pc_ensure(wspm_ctx* ctx, bool levels) {
    auto& P = ctx->pc;
    if (!P.dg) {
        const int n = ctx->n;
        P.dg = dalloc<double>(ctx, n);
        P.binv = dalloc<double>(ctx, 4 * (size_t)((n + 1) / 2));
        P.t1 = dalloc<double>(ctx, n);
        P.t2 = dalloc<double>(ctx, n);
        P.t3 = dalloc<double>(ctx, n);
    }
    if (levels && !P.levels) {
        const HostPattern& H = ctx->hp;
        const int n = ctx->n;
        auto build = [&](bool lower, Levels& L) {
            std::vector<int> lev(n, 0);
            int nl = 0;
            for (int t = 0; t < n; ++t) {
                const int i = lower ? t : n - 1 - t;
                int m = -1;
                for (int k = H.row_ptr[i]; k < H.row_ptr[i + 1]; ++k) {
                    const int j = H.cols[k];
                    if (lower ? j < i : j > i) m = std::max(m, lev[j]);
                }
                lev[i] = m + 1;
                nl = std::max(nl, lev[i] + 1);
            }
            std::vector<int32_t> ptr(nl + 1, 0), rows(n);
            for (int i = 0; i < n; ++i) ++ptr[lev[i] + 1];
            for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
            std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
            for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
            L.n_lev = nl;
            L.lvl_ptr = dupload(ctx, ptr.data(), ptr.size());
            L.lvl_rows = dupload(ctx, rows.data(), rows.size());
        };
        build(true, P.lo);
        build(false, P.up);
        P.levels = true;
    }
}

PcArgs pc_args(wspm_ctx* ctx, const double* A) {
    return PcArgs{ctx->n, ctx->row_ptr, ctx->cols, ctx->diag_pos, A, ctx->pc.binv, ctx->pc.dg};
}

// One Krylov solve, split so callers can enqueue other work between the
// iterations and the read-back.  Kernels early-exit once the device marks the
// solve done, so iterations are enqueued ahead in chunks.
template <class T>
struct KrylovRun {
    wspm_ctx* ctx;
    int method;
    const T* A;
    const T* dinv;
    KVec<T> v;
    int which;
    int maxiter;
    int R = 0;          // GMRES restart actually used
    double* H = nullptr;
    KState* ds = nullptr;
    int nbx = 0;
    int enq = 0;        // iterations (CG/BiCGSTAB) or cycles (GMRES) enqueued
    int pc = 0;         // general preconditioner kind (WSPM_PRECOND_BLOCK_JACOBI..ILU) or 0
    // CG / BiCGSTAB of one system with Jacobi or none: no finish kernels
    // (FBicg* / FCg*); WSPM_KRYLOV=unfolded keeps a finish after each reduction
    bool fold = false;
    KHist* hist = nullptr;
    // folded BiCGSTAB started by the assembly with its finish deferred: the
    // start's partial rows, reduced by the first iteration's p kernel
    FoldIn start_in{nullptr, 0, 0};
    // folded BiCGSTAB enqueued ahead: its first p kernel ends the solve while *gate == 0
    const int* gate = nullptr;
    // ensembles, >= 0: the last iteration of each enqueue() call rebuilds the
    // Krylov list with this ens_on mode in its last finish (where that finish
    // rebuilds lists)
    int last_mode = -1;
    double tol0 = 0.0, atol0 = 0.0;

    // a folded kernel: no finish, partial rows into part
    template <class Op>
    void vec_nf(const Op& op, double* part, int rows, int grid) {
        Red red = red_of(ctx, which);
        red.partials = part ? part : ctx->partials;
        LAUNCH(k_vec<Op>, dim3(grid), dim3(BLOCK), 0, ctx->stream, rows, op, red,
                           EnsWalk{nullptr, nullptr, 0});
        HIPCHK(hipGetLastError());
    }

    // dst = M src for a general preconditioner (FP64 only)
    void apply(const T* src, T* dst, int skipmode) {
        if constexpr (std::is_same<T, double>::value) {
            HIPCHK(launch_pc_apply(pc, pc_args(ctx, A), ctx->pc.lo, ctx->pc.up, src, dst, ctx->pc.t3, ds, skipmode,
                                   ctx->stream));
        }
    }

    KrylovRun(wspm_ctx* c, int m, const T* A_, const T* dinv_, const T* b, const T* x0, T* x, double tol,
              double atol, int maxit, int restart, int wh, bool started, int pc_kind = 0)
        : ctx(c), method(m), A(A_), dinv(dinv_), which(wh), maxiter(maxit),
          pc(pc_kind >= WSPM_PRECOND_BLOCK_JACOBI ? pc_kind : 0), tol0(tol), atol0(atol) {
        const int n = ctx->n;
        ds = states_of(ctx, which);
        const bool gpc = pc != 0;
        fold = !gpc && !ctx->tab && method != WSPM_GMRES && which < 2 && !env_is("WSPM_KRYLOV", "unfolded");
        if (fold) {
            if (!ctx->partials2) ctx->partials2 = dalloc<double>(ctx, (size_t)ctx->prow * MAX_DOTS);
            const int need = std::max(maxiter, 0) + 2;
            if (ctx->khist_cap[which] < need) {
                ctx->khist[which] = dalloc<KHist>(ctx, need);
                ctx->khist_cap[which] = need;
            }
            hist = ctx->khist[which];
        }
        if (gpc) {
            if constexpr (std::is_same<T, double>::value) {
                pc_ensure(ctx, pc == WSPM_PRECOND_SSOR || pc == WSPM_PRECOND_ILU);
                HIPCHK(launch_pc_setup(pc, pc_args(ctx, A), ctx->stream));
            }
        }
        if (method == WSPM_GMRES) {
            R = std::min(std::min(restart, n), MAX_RESTART);
            KWork<T>& w = ensure_work<T>(ctx, R + 1);
            v = make_kvec<T>(ctx, w, dinv, b, x0, x, gpc);
            H = (which == 1) ? ctx->H32 : ctx->H;
            nbx = std::min((n + BLOCK - 1) / BLOCK, 256);
            if (!started) launch_spmv<T>(ctx, A, GmInit<T>{v, tol, atol, maxiter, R}, which, true);
            if (gpc) {
                T* t1 = (T*)ctx->pc.t1;
                T* t2 = (T*)ctx->pc.t2;
                apply(v.w, t1, 0);
                apply(b, t2, 0);
                launch_vec(ctx, GmPcStart<T>{v, t1, t2}, which);
            }
            launch_vec(ctx, GmScale<T>{v, 0.0}, which);
        } else {
            KWork<T>& w = ensure_work<T>(ctx, 0);
            v = make_kvec<T>(ctx, w, dinv, b, x0, x, gpc);
            if (started) {
            } else if (ctx->tab) {
                // ensemble: start the instances of the active list only
                if (method == WSPM_BICGSTAB)
                    launch_spmv<T>(ctx, A, ActiveOnly<BicgInit<T>>{{v, tol, atol, maxiter}}, which, true, LIST_K);
                else
                    launch_spmv<T>(ctx, A, ActiveOnly<CgInit<T>>{{v, tol, atol, maxiter}}, which, true, LIST_K);
            } else if (method == WSPM_BICGSTAB) {
                launch_spmv<T>(ctx, A, BicgInit<T>{v, tol, atol, maxiter}, which, true);
            } else {
                launch_spmv<T>(ctx, A, CgInit<T>{v, tol, atol, maxiter}, which, true);
                if (gpc) {
                    apply(v.r, v.z, 1);
                    launch_vec(ctx, CgPcStart<T>{v}, which);
                }
            }
        }
    }

    // Folded iterations: partial rows alternate between two regions so that
    // a kernel never overwrites the rows its consumer has not read yet
    // (q, t -> A; s, x/r -> B; p reads B).
    void enqueue_folded(int count) {
        double* PA = ctx->partials;
        double* PB = ctx->partials2;
        const int sb = ctx->spmv_ell ? ctx->vec_grid : ctx->spmv_grid;   // SpMV workgroups
        const int vb = ctx->vec_grid;
        const int pr = ctx->prow;
        KState* stw = states_of(ctx, which);
        int c = 0;
        for (; c < count && enq < maxiter; ++c, ++enq) {
            const int k = enq;
            if (method == WSPM_BICGSTAB) {
                FBicgP<T> P{v, hist, stw, FoldIn{PB, pr, vb}, k, 0.0, 0.0, 0, 0, FoldIn{nullptr, 0, 0}, tol0, atol0,
                            maxiter};
                if (k == 0) {
                    P.start = start_in;
                    P.gate = gate;
                }
                vec_nf(P, nullptr, ctx->n, vb);
                launch_spmv<T>(ctx, A, FBicgQ<T>{{v}, hist, k}, which, true, LIST_ALL, PA, false);
                vec_nf(FBicgS<T>{{v, 0.0}, hist, FoldIn{PA, pr, sb}, k}, PB, ctx->n, vb);
                launch_spmv<T>(ctx, A, FBicgT<T>{{v}, hist, stw, FoldIn{PB, pr, vb}, k, 0}, which, true, LIST_ALL,
                               PA, false);
                vec_nf(FBicgXR<T>{{v, 0.0, 0.0, 0}, hist, FoldIn{PA, pr, sb}, k}, PB, ctx->n, vb);
            } else {
                launch_spmv<T>(ctx, A, FCgAp<T>{{v}, hist, k}, which, true, LIST_ALL, PA, false);
                vec_nf(FCgXR<T>{{v, 0.0}, hist, FoldIn{PA, pr, sb}, k}, PB, ctx->n, vb);
                vec_nf(FCgP<T>{v, hist, stw, FoldIn{PB, pr, vb}, k, 0.0, 0}, nullptr, ctx->n, vb);
            }
        }
        // BiCGSTAB: the loop test of the next iteration, run alone, leaves the
        // state for the host (CG's last kernel already did)
        if (method == WSPM_BICGSTAB && c > 0)
            vec_nf(FBicgP<T>{v, hist, stw, FoldIn{PB, pr, vb}, enq, 0.0, 0.0, 0, 0, FoldIn{nullptr, 0, 0}, tol0, atol0,
                             maxiter},
                   nullptr, 0, 1);
    }

    void enqueue(int count) {
        if (fold) {
            enqueue_folded(count);
            return;
        }
        for (int c = 0; c < count && enq < maxiter; ++c, ++enq) {
            // drop the instances that converged (every ens_compact_every-th
            // iteration; in between, their entries are walked and skipped),
            // by the finish of the iteration's last reduction where it can
            const bool last = last_mode >= 0 && (c == count - 1 || enq + 1 == maxiter);
            const bool compact = ctx->tab && ((enq + 1) % ctx->ens_compact_every == 0 || last);
            const bool fc = compact && fused_compact(ctx) && (method == WSPM_BICGSTAB || !pc);
            const int cm = last ? last_mode : 0;
            if (method == WSPM_BICGSTAB) {
                launch_vec(ctx, BicgP<T>{v, 0.0, 0.0, 0}, which, LIST_K);
                if (pc) apply(v.p, v.ph, 1);
                launch_spmv<T>(ctx, A, BicgQ<T>{v}, which, true, LIST_K);
                launch_vec(ctx, BicgS<T>{v, 0.0}, which, LIST_K);
                if (pc) apply(v.r, v.sh, 1);      // s lives in r
                launch_spmv<T>(ctx, A, BicgT<T>{v}, which, true, LIST_K);
                launch_vec(ctx, BicgXR<T>{v, 0.0, 0.0, 0}, which, LIST_K, fc, cm);
            } else if (method == WSPM_CG) {
                launch_spmv<T>(ctx, A, CgAp<T>{v}, which, true, LIST_K);
                launch_vec(ctx, CgXR<T>{v, 0.0}, which, LIST_K, fc, cm);   // decides done; CgP skips those
                if (pc) {
                    apply(v.r, v.z, 1);
                    launch_vec(ctx, CgPcRZ<T>{v}, which, LIST_K);
                }
                launch_vec(ctx, CgP<T>{v, 0.0}, which, LIST_K);
            } else {
                gmres_cycle();
            }
            if (compact && !fc) ens_compact(ctx, LIST_K, cm);
        }
    }

    void gmres_cycle() {
        LAUNCH(k_gm_cycle_begin, dim3(1), dim3(256), 0, ctx->stream, ds, H, R);
        HIPCHK(hipGetLastError());
        for (int j = 0; j < R; ++j) {
            launch_spmv<T>(ctx, A, GmArnoldi<T>{v, j}, which, true);
            if (pc) {
                apply(v.w, (T*)ctx->pc.t1, 2);
                launch_vec(ctx, GmPcNorm<T>{v, (T*)ctx->pc.t1}, which);
            }
            LAUNCH(k_gm_dots<T>, dim3(nbx, (j + 1 + 7) / 8), dim3(BLOCK), 0, ctx->stream, v, j,
                               red_of(ctx, which));
            HIPCHK(hipGetLastError());
            LAUNCH(k_gm_dots_finish, dim3(1), dim3(BLOCK), 0, ctx->stream, ds, ctx->partials,
                               ctx->prow, j, nbx);
            HIPCHK(hipGetLastError());
            GmUpdate<T> up{v, j, H};
            up.hb = state_field<double>(ds, offsetof(KState, hbuf));
            launch_vec(ctx, up, which);
            launch_vec(ctx, GmNormalize<T>{v, j, 0.0}, which);
        }
        LAUNCH(k_gm_lstsq, dim3(1), dim3(64), 0, ctx->stream, ds, H, R);
        HIPCHK(hipGetLastError());
        launch_vec(ctx, GmUpdX<T>{v, R, state_field<double>(ds, offsetof(KState, y))}, which);
        launch_spmv<T>(ctx, A, GmResid<T>{v}, which, true);
        if (pc) {
            apply(v.w, (T*)ctx->pc.t1, 1);
            launch_vec(ctx, GmPcResid<T>{v, (T*)ctx->pc.t1}, which);
        }
        launch_vec(ctx, GmScale<T>{v, 0.0}, which);
    }

    // First chunk: as many iterations as the previous solve of this method took.
    int first_chunk() const { return method == WSPM_GMRES ? 1 : std::max(1, ctx->last_iters[method]); }

    // After a poll: can more iterations change anything?
    bool more() const { return !ctx->hstate[which].done && enq < maxiter; }

    void stats(wspm_krylov_stats* st) {
        const KState& hs = ctx->hstate[which];
        if (method != WSPM_GMRES) ctx->last_iters[method] = std::max(1, hs.steps);
        if (!st) return;
        if (method == WSPM_GMRES) {
            st->iters = hs.cycles;
            st->arnoldi = hs.arnoldi;
            st->code = 0;
            st->res_norm = hs.rnorm;
            st->b_norm = hs.bnorm;
        } else {
            st->iters = hs.steps;
            st->arnoldi = 0;
            st->code = hs.code;
            st->res_norm = std::sqrt(hs.rs);
            st->b_norm = std::sqrt(hs.bs);
        }
    }
};

// started: the first step (r0 = b - A x0, its dots and the solver scalars) was
// already done by the assembly epilogue (k_assemble with a Krylov start mode).
template <class T>
void krylov_solve(wspm_ctx* ctx, int method, const T* A, const T* dinv, const T* b, const T* x0, T* x,
                  double tol, double atol, int maxiter, int restart, int which, wspm_krylov_stats* st,
                  bool started = false, int pc = 0) {
    KrylovRun<T> run(ctx, method, A, dinv, b, x0, x, tol, atol, maxiter, restart, which, started, pc);
    run.enqueue(run.first_chunk());
    poll(ctx, which);
    while (run.more()) {
        run.enqueue(1);
        poll(ctx, which);
    }
    run.stats(st);
}

}  // namespace

// ---------------------------------------------------------------------------
// Krylov solve that allocates nothing (wspm_krylov_noalloc)
// ---------------------------------------------------------------------------
// Its workspace -- the Krylov vectors, two regions of partial rows, the
// solver states and the iteration history -- lives in this file-scope
// array, one copy per GPU, held until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the current device
// (the array's name is a host-side stand-in); every call on one GPU shares
// it, so calls that overlap on one GPU race.
static constexpr int NOALLOC_MAX_N = 1 << 19;          // rows
static constexpr int NOALLOC_MAX_ITER = 10000;         // maxiter
static constexpr size_t NOALLOC_VEC = 9;               // r, r_hat, p, p_hat, q, s_hat, t, z, w
static constexpr size_t NOALLOC_PROW = MAX_BLOCKS;     // partial rows (vec_grid <= 2048 at NOALLOC_MAX_N)
static constexpr size_t NOALLOC_DOUBLES =
    NOALLOC_VEC * NOALLOC_MAX_N + 2 * NOALLOC_PROW * MAX_DOTS +
    (3 * sizeof(KState) + (NOALLOC_MAX_ITER + 2) * sizeof(KHist) + 255) / sizeof(double) + 64;
static_assert((NOALLOC_MAX_N + BLOCK - 1) / BLOCK <= (int)NOALLOC_PROW, "partial rows");
static_assert(sizeof(KState) % 8 == 0 && sizeof(KHist) % 8 == 0, "scratch layout");
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {

// Per-device address of scratch (set once per device, with that device
// current) and the iterations of its last solve per method (the first chunk
// of the next one).
struct NoallocDev {
    double* ptr = nullptr;
    int last_iters[2] = {1, 1};
};
std::mutex g_noalloc_mu;
NoallocDev g_noalloc[64];

__global__ void k_fill_zero(double* p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
#define API_BEGIN try {
#define API_END                                     \
    }                                               \
    catch (const Err&) {                            \
        return ctx->err.find("hipMalloc") != std::string::npos ? WSPM_ENOMEM : WSPM_EHIP; \
    }                                               \
    catch (const ApiFail& f_) {                     \
        ctx->err = f_.msg;                          \
        return f_.code;                             \
    }                                               \
    catch (const std::exception& ex) {              \
        ctx->err = ex.what();                       \
        return WSPM_EHIP;                           \
    }

static int fail(wspm_ctx* ctx, int code, const std::string& msg) {
    ctx->err = msg;
    return code;
}

extern "C" {

int32_t wspm_version(void) { return 10000; }

const char* wspm_last_error(const wspm_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

static void free_ctx(wspm_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    else (void)hipDeviceSynchronize();
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->asm_ev) (void)hipEventDestroy(e);
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
    if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
    for (void* p : ctx->allocs) (void)hipFree(p);
    for (void* p : ctx->host_allocs) (void)hipHostFree(p);
    delete ctx;
}

int32_t wspm_mesh_rcm(int32_t dim, int32_t n_nodes, int32_t n_elems, const int32_t* conn, int32_t* node_new,
                      int32_t* elem_old, int32_t* band) {
    if (!conn || !node_new || !elem_old) {
        g_create_error = "null argument";
        return WSPM_EINVAL;
    }
    try {
        HostPattern P;
        std::string err;
        const int rc = build_pattern(dim, n_nodes, n_elems, conn, P, err);
        if (rc != WSPM_OK) {
            g_create_error = err;
            return rc;
        }
        const std::vector<int32_t> perm = rcm_order(P);
        std::copy(perm.begin(), perm.end(), node_new);
        // elements by their first vertex in the new order (ties: original order)
        const int K = dim + 1;
        std::vector<int32_t> key(n_elems);
        for (int e = 0; e < n_elems; ++e) {
            int m = INT32_MAX;
            for (int c = 0; c < K; ++c) m = std::min(m, perm[conn[(size_t)e * K + c]]);
            key[e] = m;
        }
        std::vector<int32_t> ord(n_elems);
        for (int e = 0; e < n_elems; ++e) ord[e] = e;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
        std::copy(ord.begin(), ord.end(), elem_old);
        if (band) {
            bandwidth(P, {}, band[0], band[1]);
            bandwidth(P, perm, band[2], band[3]);
        }
    } catch (const std::exception& ex) {
        g_create_error = ex.what();
        return WSPM_ENOMEM;
    }
    return WSPM_OK;
}

// Workgroups of a one-system CSR SpMV: each takes a chunk of rows that fills
// one pass of its lane groups (BLOCK / lanes groups, SPMV_UNROLL rows each),
// at most MAX_BLOCKS workgroups.  (Chunks of one row per group left three
// quarters of each pass idle and multiplied the partial rows the next
// kernel reduces: the 66k-row species SpMVs of the nitrogen runs.)
static int csr_spmv_grid(long n, int lanes) {
    const long rows = (long)(BLOCK / lanes) * SPMV_UNROLL;
    return (int)std::min<long>(std::max<long>((n + rows - 1) / rows, 1), MAX_BLOCKS);
}

int32_t wspm_ctx_create(int32_t dim, int32_t n_nodes, int32_t n_elems, const int32_t* conn,
                        const double* coords, void* stream, wspm_ctx** out) {
    if (!out || !conn || !coords) {
        g_create_error = "null argument";
        return WSPM_EINVAL;
    }
    *out = nullptr;
    wspm_ctx* ctx = new wspm_ctx();
    ctx->stream = (hipStream_t)stream;
    ctx->dim = dim;
    ctx->n = n_nodes;
    ctx->E = n_elems;
    std::string perr;
    int rc = build_pattern(dim, n_nodes, n_elems, conn, ctx->hp, perr);
    if (rc != WSPM_OK) {
        g_create_error = perr;
        delete ctx;
        return rc;
    }
    try {
        const HostPattern& P = ctx->hp;
        ctx->nnz = (int64_t)P.cols.size();
        ctx->max_row = P.max_row;
        {
            const int R = asm_rows_per_block(dim);
            block_maxima(P, R, ctx->asm_win, ctx->asm_inc);
            ctx->asm_lds = asm_lds_bytes(dim, ctx->asm_win, ctx->asm_inc);
            ctx->asm_grid = (n_nodes + R - 1) / R;
            // WSPM_ASM_KERNEL=lds selects the LDS-staged kernel even for short rows
            const char* ak = std::getenv("WSPM_ASM_KERNEL");
            ctx->asm_reg = P.max_row <= asm_reg_max_row(dim) && !(ak && std::string(ak) == "lds");
            if (ctx->asm_reg) ctx->asm_grid = (n_nodes + BLOCK - 1) / BLOCK;
            int max_row_inc = 0;
            for (int i = 0; i < n_nodes; ++i) max_row_inc = std::max(max_row_inc, P.inc_ptr[i + 1] - P.inc_ptr[i]);
            // WSPM_ASM_KERNEL=inc keeps the incidence-gather register kernel
            const bool nbr = dim == 2 && max_row_inc <= asm_reg_max_row(dim) - 1 && !(ak && std::string(ak) == "inc");
            // WSPM_ASM_KERNEL=blend: the neighbourhood kernel selecting its
            // records from registers instead of LDS
            const int nbk = (ak && std::string(ak) == "blend") ? ASM_K_NBR_BLEND : ASM_K_NBR;
            ctx->asm_kernel = !ctx->asm_reg ? ASM_K_LDS : nbr ? nbk : ASM_K_REG;
            ctx->prow = std::max(MAX_BLOCKS, std::max(ctx->asm_grid, (n_nodes + R - 1) / R));
        }
        if (!ctx->asm_reg && ctx->asm_lds > 64 * 1024) {
            g_create_error = "assembly LDS staging exceeds 64 KiB (rows too long or too many incident elements)";
            delete ctx;
            return WSPM_EINVAL;
        }
        const double avg = (double)ctx->nnz / n_nodes;
        int lanes = 2;
        while (lanes < 64 && lanes < avg) lanes *= 2;
        ctx->lanes = lanes;
        ctx->spmv_grid = csr_spmv_grid(n_nodes, lanes);
        ctx->vec_grid = std::min((n_nodes + BLOCK - 1) / BLOCK, 4096);
        ctx->prow = std::max(ctx->prow, ctx->vec_grid);
        ctx->row_ptr = dupload(ctx, P.row_ptr.data(), P.row_ptr.size());
        ctx->cols = dupload(ctx, P.cols.data(), P.cols.size());
        ctx->diag_pos = dupload(ctx, P.diag_pos.data(), P.diag_pos.size());
        ctx->inc_ptr = dupload(ctx, P.inc_ptr.data(), P.inc_ptr.size());
        {
            std::vector<int2> inc(P.inc_code.size());
            for (size_t q = 0; q < inc.size(); ++q) inc[q] = make_int2(P.inc_code[q], (int)P.inc_slot[q]);
            ctx->inc = dupload(ctx, inc.data(), inc.size());
            std::vector<int4> c4(n_elems);
            for (int e = 0; e < n_elems; ++e) {
                const int32_t* c = conn + (size_t)e * (dim + 1);
                c4[e] = make_int4(c[0], c[1], c[2], dim == 3 ? c[3] : c[0]);
            }
            ctx->conn4 = dupload(ctx, c4.data(), c4.size());
            const int cw = dim == 2 ? 2 : 4;
            std::vector<double> nc((size_t)n_nodes * cw, 0.0);
            for (int i = 0; i < n_nodes; ++i)
                for (int d = 0; d < dim; ++d) nc[(size_t)i * cw + d] = coords[(size_t)i * dim + d];
            ctx->ncoord = dupload(ctx, nc.data(), nc.size());
            ctx->Kn = dalloc<double>(ctx, n_nodes);
            HIPCHK(hipMemset(ctx->Kn, 0, (size_t)n_nodes * sizeof(double)));
        }
        ctx->bc_flag = dalloc<uint8_t>(ctx, n_nodes);
        ctx->bc_val = dalloc<double>(ctx, n_nodes);
        HIPCHK(hipMemset(ctx->bc_flag, 0, n_nodes));
        HIPCHK(hipMemset(ctx->bc_val, 0, n_nodes * sizeof(double)));
        ctx->partials = dalloc<double>(ctx, (size_t)ctx->prow * MAX_DOTS);
        ctx->states = dalloc<KState>(ctx, 3);
        HIPCHK(hipMemset(ctx->states, 0, 3 * sizeof(KState)));
        void* hp = nullptr;
        HIPCHK(hipHostMalloc(&hp, 3 * sizeof(KState), hipHostMallocDefault));
        ctx->host_allocs.push_back(hp);
        ctx->hstate = static_cast<KState*>(hp);
        std::memset(ctx->hstate, 0, 3 * sizeof(KState));
        ctx->dseq = dalloc<int>(ctx, 1);
        HIPCHK(hipMemset(ctx->dseq, 0, sizeof(int)));
        hp = nullptr;
        HIPCHK(hipHostMalloc(&hp, 64, hipHostMallocDefault));
        ctx->host_allocs.push_back(hp);
        ctx->hseq = static_cast<int*>(hp);
        std::memset(ctx->hseq, 0, 64);
        {
            const char* pe = std::getenv("WSPM_POLL");
            ctx->fast_poll = !(pe && std::string(pe) == "sync");
            const char* su = std::getenv("WSPM_SPIN_US");
            if (su) ctx->spin_us = std::max(0, std::atoi(su));
        }
        ctx->H = dalloc<double>(ctx, (size_t)MAX_RESTART * (MAX_RESTART + 1));
        ctx->H32 = dalloc<double>(ctx, (size_t)MAX_RESTART * (MAX_RESTART + 1));
        compute_quad_const(dim, ctx->qc);
        compute_tr_quad(ctx->tq);
        HIPCHK(hipDeviceSynchronize());
    } catch (const Err&) {
        g_create_error = ctx->err;
        free_ctx(ctx);
        return WSPM_EHIP;
    }
    *out = ctx;
    return WSPM_OK;
}

void wspm_ctx_destroy(wspm_ctx* ctx) { free_ctx(ctx); }

int32_t wspm_set_stream(wspm_ctx* ctx, void* stream) {
    if (!ctx) return WSPM_EINVAL;
    ctx->stream = (hipStream_t)stream;
    ++ctx->graph_epoch;
    return WSPM_OK;
}

int32_t wspm_ctx_info(const wspm_ctx* ctx, int64_t* info) {
    if (!ctx || !info) return WSPM_EINVAL;
    info[0] = ctx->n; info[1] = ctx->E; info[2] = ctx->nnz; info[3] = ctx->dim;
    info[4] = ctx->max_row; info[5] = ctx->n_bc; info[6] = ctx->lanes; info[7] = ctx->ell ? (ctx->fan ? (ctx->nwin_ok ? 3 : 2) : 1) : 0;
    return WSPM_OK;
}

int32_t wspm_get_pattern(const wspm_ctx* ctx, int32_t* row_ptr, int32_t* cols) {
    if (!ctx || !row_ptr || !cols) return WSPM_EINVAL;
    std::memcpy(row_ptr, ctx->hp.row_ptr.data(), ctx->hp.row_ptr.size() * sizeof(int32_t));
    std::memcpy(cols, ctx->hp.cols.data(), ctx->hp.cols.size() * sizeof(int32_t));
    return WSPM_OK;
}

int32_t wspm_bc_set(wspm_ctx* ctx, int32_t n_bc, const int32_t* nodes, const double* values) {
    if (!ctx) return WSPM_EINVAL;
    if (n_bc < 0 || (n_bc > 0 && (!nodes || !values))) return fail(ctx, WSPM_EINVAL, "bad BC arrays");
    std::vector<uint8_t> flag(ctx->n, 0);
    std::vector<double> val(ctx->n, 0.0);
    for (int q = 0; q < n_bc; ++q) {
        const int i = nodes[q];
        if (i < 0 || i >= ctx->n) return fail(ctx, WSPM_EINVAL, "BC node out of range");
        flag[i] = 1;
        val[i] = values[q];  // later entries win
    }
    API_BEGIN
    HIPCHK(hipMemcpyAsync(ctx->bc_flag, flag.data(), ctx->n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->bc_val, val.data(), ctx->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->n_bc = n_bc;
    ++ctx->graph_epoch;
    return WSPM_OK;
    API_END
}

static SoilParams make_soil(int model, const double* p) {
    SoilParams sp{};
    sp.model = model;
    for (int k = 0; k < 6; ++k) sp.p[k] = p[k];
    return sp;
}

int32_t wspm_soil(wspm_ctx* ctx, int32_t model, const double* params, int32_t n, const double* h,
                  double* C, double* K, double* theta) {
    if (!ctx) return WSPM_EINVAL;
    if (model != WSPM_SOIL_GARDNER && model != WSPM_SOIL_VAN_GENUCHTEN) return fail(ctx, WSPM_EINVAL, "unknown soil model");
    if (!params || !h || !C || !K || !theta || n < 0) return fail(ctx, WSPM_EINVAL, "bad soil arguments");
    API_BEGIN
    double p[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < (model == WSPM_SOIL_GARDNER ? 4 : 6); ++k) p[k] = params[k];
    if (n > 0) {
        const SoilOp op{make_soil(model, p), h, C, K, theta, nullptr, 0};
        const int grid = std::min((n + BLOCK - 1) / BLOCK, 1024);
        LAUNCH(k_vec<SoilOp>, dim3(grid), dim3(BLOCK), 0, ctx->stream, n, op, red_of(ctx, 0),
                           EnsWalk{nullptr, nullptr, 0});
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

// Edge terms of one boundary list on (A, b) (see transport.h).
static void launch_edge_list(wspm_ctx* ctx, int list, int kind, double flux, double c_in, const double* K,
                             int kstride, const double* th, double* A, double* b, double* dinv) {
    const auto& L = ctx->edges[list];
    if (L.cnt < 2) return;
    if ((kind == EDGE_CAUCHY || kind == EDGE_FREEDRAIN_N) && !L.unjoined.empty())
        throw ApiFail{WSPM_EINVAL, "consecutive edge nodes " + L.unjoined +
                                       " are not joined by a mesh edge (matrix edge terms need one)"};
    EdgeArgs e{};
    e.cnt = L.cnt;
    e.nodes = L.nodes;
    e.slot_prev = L.slot_prev;
    e.slot_next = L.slot_next;
    e.diag_pos = ctx->diag_pos;
    e.nrec = nrec_of(ctx);
    e.kind = kind;
    e.flux = flux;
    e.c_in = c_in;
    e.K = K;
    e.kstride = kstride;
    e.th = th;
    e.A = A;
    e.b = b;
    e.dinv = dinv;
    HIPCHK(launch_edges(e, ctx->stream));
}

// Assembly (+ Dirichlet rows + Jacobi diagonal); with mode != ASM_INIT_NONE
// the epilogue also starts the Krylov solve (BiCGSTAB / CG / GMRES state for
// x0 -> x, or the FP64 refinement residual) and its finish kernel runs here.
static void do_assemble(wspm_ctx* ctx, const double* Ks_elem, const double* h_m, const double* C,
                        const double* K, const double* theta_m, const double* theta_n, double dt,
                        double* A, double* b, double* dinv, float* A32, float* dinv32,
                        int mode = ASM_INIT_NONE, double* x = nullptr, double tol = 0.0, double atol = 0.0,
                        int maxiter = 0, int restart = 0, bool ell = false, int* start_grid = nullptr,
                        const int* gate = nullptr) {
    AsmArgs a{};
    a.gate = gate;
    a.n = ctx->n;
    a.win_cap = ctx->asm_win;
    a.inc_cap = ctx->asm_inc;
    a.row_ptr = ctx->row_ptr; a.cols = ctx->cols; a.diag_pos = ctx->diag_pos; a.inc_ptr = ctx->inc_ptr;
    a.inc = ctx->inc; a.conn4 = ctx->conn4; a.nrec = nrec_of(ctx);
    if (K) launch_vec(ctx, SetK{K, krec_of(ctx)}, 0);   // caller's K into the assembly's nodal K
    a.Ks_elem = Ks_elem; a.C = C; a.theta_m = theta_m; a.theta_n = theta_n; a.h_m = h_m;
    a.inv_dt = 1.0 / dt;
    a.bc_flag = ctx->n_bc > 0 ? ctx->bc_flag : nullptr;
    a.bc_val = ctx->bc_val;
    a.A = A; a.A32 = A32; a.b = b; a.dinv = dinv; a.dinv32 = dinv32;
    a.ell_n = ell ? ctx->n : 0;    // the register kernels, or the 3D LDS kernel
    int kernel = ctx->asm_kernel;
    if (ell && ctx->fan) {
        a.ell_col = ctx->ell_col;       // 32-bit columns: the 16-bit offsets measured slower here
        a.ell_w = ctx->ell_w;
        a.fan_inc = ctx->fan_inc;
        a.fan_pred = ctx->fan_pred;
        a.fan_info = ctx->fan_info;
        a.fan_iw = ctx->fan_iw;
        kernel = ASM_K_FAN;
        if (ctx->nwin_ok) {
            a.nwin = ctx->nwin;
            a.nwin_col = ctx->nwin_col;
            a.nwin_pred = ctx->nwin_pred;
            a.nwin_cap = ctx->nwin_cap;
            kernel = ASM_K_WIN;
        }
    } else if (ell && ctx->ell_info && asm_is_nbr(ctx->asm_kernel)) {
        a.ell_col = ctx->ell_col;
        a.ell_inc = ctx->ell_inc;
        a.ell_info = ctx->ell_info;
        a.ell_w = ctx->ell_w;
        a.ell_iw = ctx->ell_iw;
    }
    a.qc = ctx->qc;
    a.tab = ctx->tab;
    a.act = ctx->pact;
    a.ens_st = states_of(ctx, 0);
    a.n_tab = ctx->n_tab;
    a.ens_grid = ctx->ens_grid;
    KrylovInit ki{};
    if (mode != ASM_INIT_NONE) {
        KWork<double>& w = ensure_work<double>(ctx, 0);
        ki.x0 = h_m;
        ki.x = x;
        ki.r = mode == ASM_INIT_RESID ? ctx->ref_r : w.r;
        ki.rh = w.rh; ki.p = w.p; ki.z = w.z; ki.w = w.w;
        ki.jacobi = dinv != nullptr;
    }
    int grid = 0;
    // benchmark timing of the assembly: events recorded around the launch on
    // the context's stream, and the rows the launch worked on
    const bool timed = ctx->timing && ctx->timing_asm && ctx->asm_ev_used / 2 < TIMING_SLOTS;
    hipEvent_t ae0 = nullptr, ae1 = nullptr;
    if (timed) {
        while (ctx->asm_ev.size() < ctx->asm_ev_used + 2) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ctx->asm_ev.push_back(e);
        }
        ae0 = ctx->asm_ev[ctx->asm_ev_used];
        ae1 = ctx->asm_ev[ctx->asm_ev_used + 1];
        a.tflag = ctx->asm_tflags + ctx->asm_ev_used / 2;
        ctx->asm_ev_used += 2;
        HIPCHK(hipEventRecord(ae0, ctx->stream));
    }
    ++ctx->n_launch;
    HIPCHK(launch_assemble(ctx->dim, a, mode, ki, ctx->partials, ctx->prow, ctx->asm_lds, &grid, ctx->stream,
                           kernel));
    if (timed) HIPCHK(hipEventRecord(ae1, ctx->stream));
    if (mode == ASM_INIT_NONE) return;
    KWork<double>& w = ensure_work<double>(ctx, 0);
    KVec<double> v = make_kvec<double>(ctx, w, dinv, b, h_m, x);
    // the assembly leaves 3 partial columns per workgroup; each finish reads
    // the ones its functor uses (column sums are independent)
    switch (mode) {
        case ASM_INIT_BICG:
            // start_grid: the finish-free BiCGSTAB reduces the start's rows in
            // its first kernel
            if (start_grid && !ctx->tab) *start_grid = grid;
            else launch_finish<BicgInit<double>, true>(ctx, BicgInit<double>{v, tol, atol, maxiter}, grid, 0);
            break;
        case ASM_INIT_CG:
            launch_finish<CgInit<double>, true>(ctx, CgInit<double>{v, tol, atol, maxiter}, grid, 0);
            break;
        case ASM_INIT_GMRES: {
            const int R = std::min(std::min(restart, ctx->n), MAX_RESTART);
            launch_finish<GmInit<double>, true>(ctx, GmInit<double>{v, tol, atol, maxiter, R}, grid, 0);
            break;
        }
        default:
            if (ctx->tab)
                launch_finish<RefResidE, true>(ctx, RefResidE{b, x, ctx->ref_r, tol, atol, 1}, grid, 0);
            else
                launch_finish<RefResidE, true>(ctx, RefResidE{b, x, ctx->ref_r, tol, atol, 1}, grid, 2);
            break;
    }
    if (ctx->tab) ens_compact(ctx, LIST_K);
}

int32_t wspm_assemble_richards(wspm_ctx* ctx, const double* Ks_elem, const double* h_m, const double* C,
                               const double* K, const double* theta_m, const double* theta_n, double dt,
                               double* A_vals, double* b, double* dinv) {
    if (!ctx) return WSPM_EINVAL;
    if (!h_m || !C || !K || !theta_m || !theta_n || !A_vals || !b) return fail(ctx, WSPM_EINVAL, "null array");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    if (!(dt > 0.0)) return fail(ctx, WSPM_EINVAL, "dt must be positive");
    API_BEGIN
    do_assemble(ctx, Ks_elem, h_m, C, K, theta_m, theta_n, dt, A_vals, b, dinv, nullptr, nullptr);
    ctx->nrec_dirty = true;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

int32_t wspm_spmv(wspm_ctx* ctx, const double* A_vals, const double* x, double* y) {
    if (!ctx) return WSPM_EINVAL;
    if (!A_vals || !x || !y) return fail(ctx, WSPM_EINVAL, "null array");
    API_BEGIN
    launch_spmv<double>(ctx, A_vals, EpiPlain<double>{x, y}, 0, true);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

static int check_krylov_args(wspm_ctx* ctx, int method, int maxiter, int restart) {
    if (method != WSPM_CG && method != WSPM_BICGSTAB && method != WSPM_GMRES)
        return fail(ctx, WSPM_EINVAL, "Unsupported solver type");
    if (maxiter < 0) return fail(ctx, WSPM_EINVAL, "maxiter must be >= 0");
    if (method == WSPM_GMRES && (restart < 1 || restart > MAX_RESTART))
        return fail(ctx, WSPM_EINVAL, "restart must be in [1, 64]");
    return WSPM_OK;
}

// Band workspace of the direct solver: RCM when it narrows the band.
static void ensure_band(wspm_ctx* ctx) {
    auto& B = ctx->band;
    if (B.ready) return;
    const HostPattern& P = ctx->hp;
    int kl0, ku0, kl1, ku1;
    bandwidth(P, {}, kl0, ku0);
    std::vector<int32_t> rc = rcm_order(P);
    bandwidth(P, rc, kl1, ku1);
    B.reordered = (2 * kl1 + ku1) < (2 * kl0 + ku0);
    if (B.reordered) {
        B.perm = rc;
    } else {
        B.perm.resize(ctx->n);
        for (int i = 0; i < ctx->n; ++i) B.perm[i] = i;
    }
    const int kl = B.reordered ? kl1 : kl0, ku = B.reordered ? ku1 : ku0;
    const size_t W = 2 * (size_t)kl + ku + 1;
    const size_t bytes = (size_t)ctx->n * W * sizeof(double);
    if (bytes > (size_t(32) << 30))
        throw ApiFail{WSPM_ENOMEM, "direct solve: band of " + std::to_string(bytes >> 20) + " MiB exceeds 32 GiB"};
    Band& b = B.b;
    b.n = ctx->n;
    b.kl = kl;
    b.ku = ku;
    b.W = (int)W;
    b.ab = dalloc<double>(ctx, (size_t)ctx->n * W);
    b.ipiv = dalloc<int32_t>(ctx, ctx->n);
    b.perm = dupload(ctx, B.perm.data(), B.perm.size());
    b.y = dalloc<double>(ctx, ctx->n);
    b.info = dalloc<int32_t>(ctx, 1);
    B.ready = true;
}

// Enqueue fill + factor + solve (no synchronisation).
static void direct_solve(wspm_ctx* ctx, const double* A, const double* rhs, double* x) {
    ensure_band(ctx);
    const Band& b = ctx->band.b;
    HIPCHK(launch_band_fill(b, ctx->row_ptr, ctx->cols, A, ctx->stream));
    HIPCHK(launch_band_factor(b, ctx->stream));
    HIPCHK(launch_band_solve(b, rhs, x, ctx->stream));
}

static int32_t direct_info(wspm_ctx* ctx) {
    int32_t info = 0;
    HIPCHK(hipMemcpyAsync(&info, ctx->band.b.info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return info;
}

int32_t wspm_direct(wspm_ctx* ctx, const double* A_vals, const double* b, double* x, int32_t* info) {
    if (!ctx) return WSPM_EINVAL;
    if (!A_vals || !b || !x) return fail(ctx, WSPM_EINVAL, "null array");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    API_BEGIN
    direct_solve(ctx, A_vals, b, x);
    const int32_t inf = direct_info(ctx);
    if (info) *info = inf;
    if (inf) return fail(ctx, WSPM_EINVAL, "singular matrix (zero pivot column " + std::to_string(inf - 1) + ")");
    return WSPM_OK;
    API_END
}

int32_t wspm_direct_info(wspm_ctx* ctx, int32_t* kl, int32_t* ku, int32_t* reordered) {
    if (!ctx) return WSPM_EINVAL;
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    API_BEGIN
    ensure_band(ctx);
    if (kl) *kl = ctx->band.b.kl;
    if (ku) *ku = ctx->band.b.ku;
    if (reordered) *reordered = ctx->band.reordered ? 1 : 0;
    return WSPM_OK;
    API_END
}

int32_t wspm_krylov(wspm_ctx* ctx, int32_t method, const double* A_vals, const double* dinv, const double* b,
                    const double* x0, double* x, double tol, double atol, int32_t maxiter, int32_t restart,
                    wspm_krylov_stats* stats) {
    if (!ctx) return WSPM_EINVAL;
    if (!A_vals || !b || !x0 || !x) return fail(ctx, WSPM_EINVAL, "null array");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    int rc = check_krylov_args(ctx, method, maxiter, restart);
    if (rc) return rc;
    API_BEGIN
    wspm_krylov_stats st{};
    krylov_solve<double>(ctx, method, A_vals, dinv, b, x0, x, tol, atol, maxiter, restart, 0, &st);
    if (method == WSPM_GMRES) {
        LAUNCH(k_nan_check, dim3(1), dim3(256), 0, ctx->stream, ctx->states, x, ctx->n);
        HIPCHK(hipGetLastError());
        poll(ctx, 0);
        st.code = ctx->hstate[0].nan_flag ? -1 : 0;
    }
    if (stats) *stats = st;
    return WSPM_OK;
    API_END
}

int32_t wspm_krylov_noalloc(int32_t method, int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* cols,
                            const double* A_vals, const double* dinv, const double* b, const double* x0, double* x,
                            double tol, double atol, int32_t maxiter, void* stream, wspm_krylov_stats* stats) {
    if (method != WSPM_CG && method != WSPM_BICGSTAB) {
        g_create_error = "wspm_krylov_noalloc: method must be WSPM_CG or WSPM_BICGSTAB";
        return WSPM_EINVAL;
    }
    if (n < 1 || n > NOALLOC_MAX_N) {
        g_create_error = "wspm_krylov_noalloc: n = " + std::to_string(n) + " outside [1, " +
                         std::to_string(NOALLOC_MAX_N) + "] (the static workspace's rows)";
        return WSPM_EINVAL;
    }
    if (maxiter < 0 || maxiter > NOALLOC_MAX_ITER) {
        g_create_error = "wspm_krylov_noalloc: maxiter outside [0, " + std::to_string(NOALLOC_MAX_ITER) + "]";
        return WSPM_EINVAL;
    }
    if (nnz < n || !row_ptr || !cols || !A_vals || !b || !x0 || !x) {
        g_create_error = "wspm_krylov_noalloc: null array or nnz < n";
        return WSPM_EINVAL;
    }
    int dev = 0;
    hipError_t he = hipGetDevice(&dev);
    if (he != hipSuccess || dev < 0 || dev >= 64) {
        g_create_error = std::string("wspm_krylov_noalloc: hipGetDevice: ") + hipGetErrorString(he);
        return WSPM_EHIP;
    }
    NoallocDev* nd = &g_noalloc[dev];
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!nd->ptr) {
            void* p = nullptr;
            he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (he != hipSuccess || !p) {
                g_create_error = std::string("wspm_krylov_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(he);
                return WSPM_EHIP;
            }
            nd->ptr = static_cast<double*>(p);
        }
    }
    // a context on the host stack over the caller's matrix and the scratch
    // array: no device or pinned allocation anywhere below
    wspm_ctx c;
    wspm_ctx* ctx = &c;
    KState hst[3];
    std::memset(hst, 0, sizeof(hst));
    c.stream = static_cast<hipStream_t>(stream);
    c.dim = 2;
    c.n = n;
    c.nnz = nnz;
    const double avg = (double)nnz / n;
    int lanes = 2;
    while (lanes < 64 && lanes < avg) lanes *= 2;
    c.lanes = lanes;
    c.spmv_grid = csr_spmv_grid(n, lanes);
    c.vec_grid = std::min((n + BLOCK - 1) / BLOCK, 4096);
    c.prow = (int)NOALLOC_PROW;
    c.row_ptr = const_cast<int32_t*>(row_ptr);
    c.cols = const_cast<int32_t*>(cols);
    double* q = nd->ptr;
    KWork<double>& w = c.w64;
    w.n = n;
    double** vecs[NOALLOC_VEC] = {&w.r, &w.rh, &w.p, &w.ph, &w.q, &w.sh, &w.t, &w.z, &w.w};
    for (size_t k = 0; k < NOALLOC_VEC; ++k) *vecs[k] = q + k * (size_t)NOALLOC_MAX_N;
    q += NOALLOC_VEC * (size_t)NOALLOC_MAX_N;
    c.partials = q;
    c.partials2 = q + NOALLOC_PROW * MAX_DOTS;
    q += 2 * NOALLOC_PROW * MAX_DOTS;
    double* zero_from = q;
    c.states = reinterpret_cast<KState*>(q);
    q += 3 * sizeof(KState) / sizeof(double);
    c.khist[0] = reinterpret_cast<KHist*>(q);
    c.khist_cap[0] = maxiter + 2;
    q += (size_t)(maxiter + 2) * sizeof(KHist) / sizeof(double);
    c.hstate = hst;
    c.fast_poll = false;
    c.last_iters[method] = nd->last_iters[method];
    try {
        // the states and the iteration history start at zero on every call
        // (a fill kernel, not a memset node, so callers may capture the call)
        hipLaunchKernelGGL(k_fill_zero, dim3(64), dim3(BLOCK), 0, c.stream, zero_from, (size_t)(q - zero_from));
        HIPCHK(hipGetLastError());
        wspm_krylov_stats st{};
        krylov_solve<double>(ctx, method, A_vals, dinv, b, x0, x, tol, atol, maxiter, 30, 0, &st);
        nd->last_iters[method] = c.last_iters[method];
        if (stats) *stats = st;
    } catch (const Err&) {
        g_create_error = c.err;
        return WSPM_EHIP;
    } catch (const ApiFail& f) {
        g_create_error = f.msg;
        return f.code;
    }
    return WSPM_OK;
}

int32_t wspm_krylov_pc(wspm_ctx* ctx, int32_t method, int32_t precond, const double* A_vals, const double* b,
                       const double* x0, double* x, double tol, double atol, int32_t maxiter, int32_t restart,
                       wspm_krylov_stats* stats) {
    if (!ctx) return WSPM_EINVAL;
    if (!A_vals || !b || !x0 || !x) return fail(ctx, WSPM_EINVAL, "null array");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    if (precond < WSPM_PRECOND_NONE || precond > WSPM_PRECOND_ILU) return fail(ctx, WSPM_EINVAL, "unknown preconditioner");
    int rc = check_krylov_args(ctx, method, maxiter, restart);
    if (rc) return rc;
    API_BEGIN
    const double* dinv = nullptr;
    if (precond == WSPM_PRECOND_JACOBI) {
        pc_ensure(ctx, false);
        HIPCHK(launch_pc_setup(WSPM_PRECOND_JACOBI, pc_args(ctx, A_vals), ctx->stream));
        dinv = ctx->pc.dg;
    }
    wspm_krylov_stats st{};
    krylov_solve<double>(ctx, method, A_vals, dinv, b, x0, x, tol, atol, maxiter, restart, 0, &st, false, precond);
    if (method == WSPM_GMRES) {
        LAUNCH(k_nan_check, dim3(1), dim3(256), 0, ctx->stream, ctx->states, x, ctx->n);
        HIPCHK(hipGetLastError());
        poll(ctx, 0);
        st.code = ctx->hstate[0].nan_flag ? -1 : 0;
    }
    if (stats) *stats = st;
    return WSPM_OK;
    API_END
}

static size_t op_cap(const wspm_ctx* ctx);

static void ensure_mixed(wspm_ctx* ctx) {
    if (!ctx->ref_r) {
        ctx->ref_r = dalloc<double>(ctx, ctx->n);
        ctx->ref_r32 = dalloc<float>(ctx, ctx->n);
        ctx->ref_d32 = dalloc<float>(ctx, ctx->n);
        ctx->ref_d32b = dalloc<float>(ctx, ctx->n);
    }
}

static int env_int(const char* name, int dflt) {
    const char* e = std::getenv(name);
    return e ? std::max(1, std::atoi(e)) : dflt;
}

// Jacobi inner steps for the mixed-precision solves of a time step: asked for,
// Jacobi-preconditioned, and no sweep of this context has stalled yet.
static bool jacobi_on(const wspm_ctx* ctx, const float* dinv32) {
    return ctx->prm.inner_method == WSPM_INNER_JACOBI && dinv32 != nullptr && !ctx->jac_inner_off;
}

// Refinement sweeps enqueued per read-back with Jacobi inner steps (each is a
// few launches; the ones after convergence return at once).  Measured, 2 to
// 5 sweeps of 3 steps, same box: the 256-instance ensemble is fastest at 4
// (+3 % over 3); one 512^2 system at 2 (5,690-5,780 steps/s against
// 4,960-5,650 at 3 and about 4,150 at 4), one 64^3 system alike at 2 and 3.
// WSPM_JAC_SWEEPS overrides both.
static int jac_sweeps(bool ens) { return env_int("WSPM_JAC_SWEEPS", ens ? 4 : 2); }

// One refinement sweep with Jacobi inner steps: ctx->prm.inner_iters steps on
// the correction, x += |r| d, then the FP64 residual r = b - A x with its
// finish into state `which` (ensembles: per instance, compact_k rebuilds the
// list of instances still refining).  Every launch skips a refinement that
// has converged.
static void jacobi_sweep(wspm_ctx* ctx, const double* A, const float* A32, const float* dinv32, const double* b,
                         double* x, double tol, double atol, int which, bool compact_k) {
    const int k = ctx->prm.inner_iters > 0 ? ctx->prm.inner_iters : JAC_STEPS;
    const int list = ctx->tab ? LIST_K : LIST_ALL;
    if (k == 1) {
        launch_vec(ctx, JacOnly{ctx->ref_r, dinv32, x, 1, 0.0, 0.0}, which, list);
    } else {
        float* da = ctx->ref_d32;
        float* db = ctx->ref_d32b;
        launch_spmv<float>(ctx, A32, JacFirst{ctx->ref_r, dinv32, ctx->ref_r32, da, k == 2 ? x : nullptr, 1, 0.0, 0.0},
                           which, true, list);
        for (int j = 3; j <= k; ++j) {
            launch_spmv<float>(ctx, A32, JacStep{da, ctx->ref_r32, dinv32, db, j == k ? x : nullptr, 1, 0.0}, which,
                               true, list);
            std::swap(da, db);
        }
    }
    launch_spmv<double>(ctx, A, RefResidE{b, x, ctx->ref_r, tol, atol, 0, ctx->jac_stall, k}, which, true, list,
                        nullptr, true, compact_k);
}

// Iterative refinement of one system: FP32 corrections of the FP64 residual
// until |r| <= max(tol |b|, atol) or max_refine sweeps.  started: the first
// residual r = b - A x and its norms are already in ctx->ref_r / state 2
// (assembly epilogue, or sweeps enqueued by the caller behind it).  The inner
// solves are the method's Krylov solve to inner_tol, or Jacobi steps
// (jacobi_on, time steps only: jac), several sweeps per read-back.
static void mixed_solve(wspm_ctx* ctx, int method, const double* A, const double* dinv, const float* A32,
                        const float* dinv32, const double* b, double* x, double tol, double atol,
                        int maxiter, int restart, double inner_tol, int max_refine, wspm_krylov_stats* st,
                        bool started = false, bool jac = false) {
    ensure_mixed(ctx);
    int inner_total = 0, arn = 0, sweeps = 0;
    KState& hs = ctx->hstate[2];
    if (!started) launch_spmv<double>(ctx, A, RefResidE{b, x, ctx->ref_r, tol, atol, 1}, 2, true);
    while (true) {
        poll(ctx, 2);
        const double thresh = std::max(tol * hs.bnorm, atol);
        if (!(hs.rnorm > thresh) || sweeps >= max_refine) break;
        if (hs.rstall) ctx->jac_inner_off = true;
        if (jac && jacobi_on(ctx, dinv32)) {
            const int ns = std::min(jac_sweeps(false), max_refine - sweeps);
            for (int k = 0; k < ns; ++k) jacobi_sweep(ctx, A, A32, dinv32, b, x, tol, atol, 2, false);
            sweeps += ns;
            continue;
        }
        wspm_krylov_stats ist{};
        if (method == WSPM_GMRES) {
            launch_vec(ctx, RefScale{ctx->ref_r, ctx->ref_r32, ctx->ref_d32, 0.0}, 2);
            krylov_solve<float>(ctx, method, A32, dinv32, ctx->ref_r32, ctx->ref_d32, ctx->ref_d32, inner_tol, 0.0,
                                maxiter, restart, 1, &ist);
        } else {
            // zero initial guess: scaling and solver start in one pass, no SpMV
            KrylovRun<float> run(ctx, method, A32, dinv32, ctx->ref_r32, ctx->ref_d32, ctx->ref_d32, inner_tol, 0.0,
                                 maxiter, restart, 1, true);
            const double* rn = state_field<double>(states_of(ctx, 2), offsetof(KState, rnorm));
            if (method == WSPM_BICGSTAB)
                launch_vec(ctx, RefStart<BicgInit<float>>{{run.v, inner_tol, 0.0, maxiter}, ctx->ref_r, ctx->ref_r32,
                                                          ctx->ref_d32, rn, 0.0}, 1);
            else
                launch_vec(ctx, RefStart<CgInit<float>>{{run.v, inner_tol, 0.0, maxiter}, ctx->ref_r, ctx->ref_r32,
                                                        ctx->ref_d32, rn, 0.0}, 1);
            run.enqueue(run.first_chunk());
            poll(ctx, 1);
            while (run.more()) {
                run.enqueue(1);
                poll(ctx, 1);
            }
            run.stats(&ist);
        }
        inner_total += ist.iters;
        arn += ist.arnoldi;
        launch_vec(ctx, RefCorrect{x, ctx->ref_d32, 0.0}, 2);
        launch_spmv<double>(ctx, A, RefResidE{b, x, ctx->ref_r, tol, atol, 0}, 2, true);
        ++sweeps;
    }
    if (st) {
        st->iters = inner_total + hs.steps;   // state 2 counts the Jacobi steps since the start
        st->arnoldi = arn;
        st->code = 0;
        st->refinements = sweeps;
        st->res_norm = hs.rnorm;
        st->b_norm = hs.bnorm;
    }
}

int32_t wspm_krylov_mixed(wspm_ctx* ctx, int32_t method, const double* A_vals, const double* dinv,
                          const double* b, const double* x0, double* x, double tol, double atol, int32_t maxiter,
                          int32_t restart, double inner_tol, int32_t max_refine, wspm_krylov_stats* stats) {
    if (!ctx) return WSPM_EINVAL;
    if (!A_vals || !b || !x0 || !x) return fail(ctx, WSPM_EINVAL, "null array");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    int rc = check_krylov_args(ctx, method, maxiter, restart);
    if (rc) return rc;
    if (max_refine < 0) return fail(ctx, WSPM_EINVAL, "max_refine must be >= 0");
    API_BEGIN
    if (!ctx->A32) {
        ctx->A32 = dalloc<float>(ctx, op_cap(ctx));
        ctx->dinv32 = dalloc<float>(ctx, ctx->n);
    }
    // FP32 copies of the operator (the Richards driver gets them from assembly)
    {
        ToF32 op{A_vals, ctx->A32};
        const int g = (int)std::min<int64_t>((ctx->nnz + BLOCK - 1) / BLOCK, 4096);
        LAUNCH(k_vec<ToF32>, dim3(g), dim3(BLOCK), 0, ctx->stream, (int)ctx->nnz, op, red_of(ctx, 2),
                           EnsWalk{nullptr, nullptr, 0});  // ND = 0
        HIPCHK(hipGetLastError());
        if (dinv) launch_vec(ctx, ToF32{dinv, ctx->dinv32}, 2);
    }
    if (x != x0) HIPCHK(hipMemcpyAsync(x, x0, ctx->n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    wspm_krylov_stats st{};
    mixed_solve(ctx, method, A_vals, dinv, ctx->A32, dinv ? ctx->dinv32 : nullptr, b, x, tol, atol, maxiter, restart,
                inner_tol, max_refine, &st);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (stats) *stats = st;
    return WSPM_OK;
    API_END
}

static int check_richards_params(wspm_ctx* ctx, const wspm_richards_params* prm) {
    if (prm->soil_model != WSPM_SOIL_GARDNER && prm->soil_model != WSPM_SOIL_VAN_GENUCHTEN)
        return fail(ctx, WSPM_EINVAL, "unknown soil model");
    const bool direct = prm->method == WSPM_DIRECT;
    int rc = direct ? WSPM_OK : check_krylov_args(ctx, prm->method, prm->maxiter, prm->restart);
    if (rc) return rc;
    if (direct && (ctx->tab || prm->precision != WSPM_FP64))
        return fail(ctx, WSPM_EINVAL, "the direct solver is FP64 on a single-mesh context");
    if (prm->precond < WSPM_PRECOND_NONE || prm->precond > WSPM_PRECOND_ILU)
        return fail(ctx, WSPM_EINVAL, "unknown preconditioner");
    if (prm->precision != WSPM_FP64 && prm->precision != WSPM_MIXED)
        return fail(ctx, WSPM_EINVAL, "unknown precision");
    if (prm->precision == WSPM_MIXED && prm->precond > WSPM_PRECOND_JACOBI)
        return fail(ctx, WSPM_EINVAL, "mixed precision takes no preconditioner or Jacobi");
    if (prm->inner_method != WSPM_INNER_KRYLOV && prm->inner_method != WSPM_INNER_JACOBI)
        return fail(ctx, WSPM_EINVAL, "unknown inner method");
    if (prm->precision == WSPM_MIXED && prm->inner_method == WSPM_INNER_JACOBI &&
        prm->precond != WSPM_PRECOND_JACOBI)
        return fail(ctx, WSPM_EINVAL, "Jacobi inner iterations need the Jacobi preconditioner");
    if (prm->inner_iters < 0 || prm->inner_iters > MAX_JAC_STEPS)
        return fail(ctx, WSPM_EINVAL, "inner_iters must be in [0, 16]");
    if (prm->picard_maxit < 1) return fail(ctx, WSPM_EINVAL, "picard_maxit must be >= 1");
    return WSPM_OK;
}

// Values held for the time-step operator: CSR, or ELL of width max_row.
static size_t op_cap(const wspm_ctx* ctx) {
    const size_t ell = ((ctx->dim == 2 && ctx->max_row <= 8) || (ctx->dim == 3 && ctx->max_row <= ELL_W3D))
                           ? (size_t)ctx->max_row * ctx->n : 0;
    return std::max((size_t)ctx->nnz, ell);
}

// ELL for the time steps: 2D register assembly with rows of 7 or 8 entries,
// or 3D (the LDS assembly) with rows of 15 or 16 (a tetrahedral grid's
// interior rows have 15), at most 25 % padding, and a Krylov solve with no
// preconditioner or Jacobi (the general preconditioners and the direct
// solver read CSR).  3D also needs every column within 16-bit offset range
// of its row (checked when the table is built).  WSPM_SPMV=csr keeps CSR.
static bool ell_shape_ok(const wspm_ctx* ctx) {
    const char* env = std::getenv("WSPM_SPMV");
    if (env && std::string(env) == "csr") return false;
    if (ctx->dim == 2) {
        if (!ctx->asm_reg || ctx->asm_kernel == ASM_K_LDS) return false;
        if (ctx->max_row != 7 && ctx->max_row != 8) return false;
    } else {
        if (ctx->asm_kernel != ASM_K_LDS || ctx->max_row < ELL_W3D - 1 || ctx->max_row > ELL_W3D) return false;
        if (env_is("WSPM_ELL_COLS", "32")) return false;
    }
    return (double)ctx->max_row * ctx->n <= 1.25 * (double)ctx->nnz;
}

static bool ell_ok(const wspm_ctx* ctx, const wspm_richards_params* prm) {
    if (!ell_shape_ok(ctx)) return false;
    return prm->method != WSPM_DIRECT && prm->precond <= WSPM_PRECOND_JACOBI;
}

// The ELL column table as 16-bit offsets from the row, when every column is
// within 32767 rows of its row (WSPM_ELL_COLS=32 keeps 32-bit columns): the
// SpMV then reads 10 bytes an FP64 entry instead of 12 (6 instead of 8 in
// FP32); measured 9.95 -> 9.10 us per launch at 512^2.  The fan assembly
// keeps the 32-bit table (with the offsets it measured 24.3 -> 26.4 us).
static void ell_offsets(wspm_ctx* ctx, const std::vector<int32_t>& c) {
    if (env_is("WSPM_ELL_COLS", "32")) return;
    const size_t n = ctx->n;
    std::vector<int16_t> o(c.size());
    for (size_t q = 0; q < c.size(); ++q) {
        const long d = (long)c[q] - (long)(q % n);
        if (d < -32767 || d > 32767) return;
        o[q] = (int16_t)d;
    }
    ctx->ell_col16 = dupload(ctx, o.data(), o.size());
}

// Node windows of the windowed fan assembly.  Row blocks are BLOCK rows
// from 0 on one mesh and the row-table entries on an ensemble.  A block's
// window is the sorted set of nodes its rows touch (fan slots and
// predecessors), cut at its WIN_SEG - 1 widest gaps into segments of
// consecutive nodes; every position of a segment is staged, so a structured
// mesh's block takes its own rows and the grid rows either side (one
// segment when they touch, three on wide grids).  nwin_ok stays false if a
// window would exceed WIN_CAP nodes (the fan kernel then gathers from
// global memory); WSPM_ASM_WIN=0 keeps that kernel.
static void build_node_windows(wspm_ctx* ctx, const std::vector<int32_t>& fc, const std::vector<int32_t>& fp) {
    ctx->nwin_ok = false;
    if (env_is("WSPM_ASM_WIN", "0")) return;
    const int n = ctx->n, W = ctx->max_row;
    std::vector<std::pair<int, int>> blocks;
    if (ctx->tab) {
        for (const RowTab& t : ctx->htab) blocks.emplace_back(t.x, t.y);
    } else {
        for (int r = 0; r < n; r += BLOCK) blocks.emplace_back(r, std::min(n, r + BLOCK));
    }
    std::vector<int32_t> win(WIN_WORDS * blocks.size(), 0);
    std::vector<uint16_t> wcol((size_t)W * n), wpred(n);
    std::vector<int> cols;
    int cap = 0;
    for (size_t b = 0; b < blocks.size(); ++b) {
        const int r0 = blocks[b].first, r1 = blocks[b].second;
        cols.clear();
        for (int i = r0; i < r1; ++i) {
            for (int s = 0; s < W; ++s) cols.push_back(fc[(size_t)s * n + i]);
            cols.push_back(fp[i]);
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        // cut at the widest gaps (earliest first among equals)
        std::vector<int> cut;   // index k: a segment ends after cols[k]
        for (int s = 0; s < WIN_SEG - 1; ++s) {
            int best = -1, gap = 1;
            for (size_t k = 0; k + 1 < cols.size(); ++k) {
                const int g = cols[k + 1] - cols[k];
                if (g > gap && std::find(cut.begin(), cut.end(), (int)k) == cut.end()) {
                    gap = g;
                    best = (int)k;
                }
            }
            if (best >= 0) cut.push_back(best);
        }
        std::sort(cut.begin(), cut.end());
        int lo = cols.front(), off = 0, seg = 0;
        int segs[WIN_SEG][3];   // lo, len, offset
        for (size_t q = 0; q <= cut.size(); ++q) {
            const int hi = q < cut.size() ? cols[cut[q]] : cols.back();
            segs[seg][0] = lo;
            segs[seg][1] = hi - lo + 1;
            segs[seg][2] = off;
            off += hi - lo + 1;
            ++seg;
            if (q < cut.size()) lo = cols[cut[q] + 1];
        }
        if (off > WIN_CAP) return;
        cap = std::max(cap, off);
        for (int k = 0; k < seg; ++k) {
            win[WIN_WORDS * b + 2 * k] = segs[k][0];
            win[WIN_WORDS * b + 2 * k + 1] = segs[k][1];
        }
        auto pos = [&](int c) {
            for (int k = 0; k < seg; ++k)
                if (c >= segs[k][0] && c < segs[k][0] + segs[k][1]) return segs[k][2] + (c - segs[k][0]);
            return -1;
        };
        for (int i = r0; i < r1; ++i) {
            for (int s = 0; s < W; ++s) wcol[(size_t)s * n + i] = (uint16_t)pos(fc[(size_t)s * n + i]);
            wpred[i] = (uint16_t)pos(fp[i]);
        }
    }
    ctx->nwin = dupload(ctx, win.data(), win.size());
    ctx->nwin_col = dupload(ctx, wcol.data(), wcol.size());
    ctx->nwin_pred = dupload(ctx, wpred.data(), wpred.size());
    ctx->nwin_cap = cap;
    ctx->nwin_ok = true;
}

// Column table of the ELL layout: entry k of row i at k * n + i; padding
// points at the row itself (its value stays 0).  With the neighbourhood
// assembly, the rows are put in fan order when every row of the mesh is one
// fan of consistently oriented triangles (build_fan; WSPM_ASM_FAN=0 keeps
// the sorted order): the fan kernel then assembles into slots known at
// compile time.  The CSR pattern of the public entry points stays sorted.
static void ensure_ell(wspm_ctx* ctx) {
    if (ctx->ell_col) return;
    const HostPattern& P = ctx->hp;
    const int n = ctx->n, W = ctx->max_row;
    if (asm_is_nbr(ctx->asm_kernel) && W <= ASM_REG_ROW_2D && !env_is("WSPM_ASM_FAN", "0")) {
        std::vector<int32_t> fc, fi, fp;
        std::vector<uint32_t> finfo;
        int iw = 0;
        if (build_fan(P, W, fc, fi, fp, finfo, iw)) {
            ell_offsets(ctx, fc);
            ctx->ell_col = dupload(ctx, fc.data(), fc.size());
            ctx->ell_w = W;
            ctx->fan_inc = dupload(ctx, fi.data(), fi.size());
            ctx->fan_pred = dupload(ctx, fp.data(), fp.size());
            ctx->fan_info = dupload(ctx, finfo.data(), finfo.size());
            ctx->fan_iw = iw;
            ctx->fan = true;
            build_node_windows(ctx, fc, fp);
            return;
        }
    }
    std::vector<int32_t> c((size_t)W * n);
    for (int i = 0; i < n; ++i) {
        const int rs = P.row_ptr[i], rl = P.row_ptr[i + 1] - rs;
        for (int k = 0; k < W; ++k) c[(size_t)k * n + i] = k < rl ? P.cols[rs + k] : i;
    }
    ell_offsets(ctx, c);
    ctx->ell_col = dupload(ctx, c.data(), c.size());
    ctx->ell_w = W;
    // the neighbourhood assembly's incidences and row shape in the same layout
    // (WSPM_ASM_ROWS=csr: it reads them through the CSR row pointers instead)
    const char* ar = std::getenv("WSPM_ASM_ROWS");
    if (asm_is_nbr(ctx->asm_kernel) && !(ar && std::string(ar) == "csr")) {
        int iw = 0;
        for (int i = 0; i < n; ++i) iw = std::max(iw, P.inc_ptr[i + 1] - P.inc_ptr[i]);
        std::vector<int2> ic((size_t)iw * n, make_int2(0, 0));
        std::vector<uint32_t> info(n);
        for (int i = 0; i < n; ++i) {
            const int q0 = P.inc_ptr[i], nq = P.inc_ptr[i + 1] - q0;
            for (int u = 0; u < nq; ++u) ic[(size_t)u * n + i] = make_int2(P.inc_code[q0 + u], (int)P.inc_slot[q0 + u]);
            const int rl = P.row_ptr[i + 1] - P.row_ptr[i], dg = P.diag_pos[i] - P.row_ptr[i];
            info[i] = (uint32_t)rl | (uint32_t)dg << 4 | (uint32_t)nq << 8;
        }
        ctx->ell_inc = dupload(ctx, ic.data(), ic.size());
        ctx->ell_info = dupload(ctx, info.data(), info.size());
        ctx->ell_iw = iw;
    }
}

// The SpMV launches of a time step read the ELL operator while this lives.
struct EllScope {
    wspm_ctx* ctx;
    EllScope(wspm_ctx* c, bool on) : ctx(c) { ctx->spmv_ell = on; }
    ~EllScope() { ctx->spmv_ell = false; }
};

// Time-stepping arrays, Ks_e and h0 (both device, copied).
static void richards_alloc(wspm_ctx* ctx, const wspm_richards_params* prm, const double* Ks_elem,
                           const double* h0) {
    const int n = ctx->n;
    ++ctx->graph_epoch;
    {
        // off by default: replaying whole Picard iterations measured 2 % slower
        // on the 512^2 step and within noise on the 256-instance ensemble
        // (the host enqueues far enough ahead of these kernels)
        const char* ge = std::getenv("WSPM_GRAPHS");
        ctx->use_graphs = ge && std::string(ge) == "1";
    }
    if (!ctx->hA) {
        ctx->hA = dalloc<double>(ctx, n); ctx->hB = dalloc<double>(ctx, n);
        ctx->Cn = dalloc<double>(ctx, n);
        ctx->thm = dalloc<double>(ctx, n); ctx->thn = dalloc<double>(ctx, n);
        ctx->Av = dalloc<double>(ctx, op_cap(ctx)); ctx->bv = dalloc<double>(ctx, n);
        ctx->dinv = dalloc<double>(ctx, n);
    }
    if (Ks_elem) {
        if (!ctx->Ks) ctx->Ks = dalloc<double>(ctx, ctx->E);
        HIPCHK(hipMemcpyAsync(ctx->Ks, Ks_elem, ctx->E * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (prm->precision == WSPM_MIXED && !ctx->A32) {
        ctx->A32 = dalloc<float>(ctx, op_cap(ctx));
        ctx->dinv32 = dalloc<float>(ctx, n);
    }
    ctx->prm = *prm;
    ctx->jac_inner_off = false;
    {
        // WSPM_JAC_STALL overrides the stall ratio (a tiny one forces the
        // switch to Krylov inner solves at the first sweep that does not converge)
        const char* js = std::getenv("WSPM_JAC_STALL");
        ctx->jac_stall = js ? std::atof(js) : JAC_STALL;
    }
    // the time-step solves read the operator in ELL order when the rows allow
    ctx->ell = ell_ok(ctx, prm);
    if (ctx->ell) ensure_ell(ctx);
    if (ctx->ell && ctx->dim == 3 && !ctx->ell_col16) ctx->ell = false;   // columns beyond 16-bit offsets
    if (ctx->ell) {
        // padding slots stay zero: the assembly writes only the real entries
        HIPCHK(hipMemsetAsync(ctx->Av, 0, op_cap(ctx) * sizeof(double), ctx->stream));
        if (ctx->A32) HIPCHK(hipMemsetAsync(ctx->A32, 0, op_cap(ctx) * sizeof(float), ctx->stream));
    }
    HIPCHK(hipMemcpyAsync(ctx->hA, h0, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
}

int32_t wspm_richards_init(wspm_ctx* ctx, const wspm_richards_params* prm, const double* Ks_elem,
                           const double* h0) {
    if (!ctx) return WSPM_EINVAL;
    if (!prm || !h0) return fail(ctx, WSPM_EINVAL, "null argument");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "ensemble context: use wspm_ens_richards_init");
    int rc = check_richards_params(ctx, prm);
    if (rc) return rc;
    API_BEGIN
    richards_alloc(ctx, prm, Ks_elem, h0);
    double p[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 6; ++k) p[k] = prm->phi[k];
    if (Ks_elem) p[prm->soil_model == WSPM_SOIL_GARDNER ? 3 : 5] = 1.0;  // nodal model gives Kr
    ctx->sp = make_soil(prm->soil_model, p);
    launch_vec(ctx, SoilOp{ctx->sp, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx), 0}, 0);
    ctx->nrec_dirty = false;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->rich_init = true;
    ctx->has_ks = Ks_elem != nullptr;
    return WSPM_OK;
    API_END
}

static void ens_step(wspm_ctx* ctx, double dt, wspm_step_stats* stats);

// Gate, pinned mirror and Krylov history of a Picard iteration enqueued ahead.
static void ensure_ahead(wspm_ctx* ctx) {
    if (!ctx->gate) {
        ctx->gate = dalloc<int>(ctx, 1);
        HIPCHK(hipMemset(ctx->gate, 0, sizeof(int)));
    }
    if (!ctx->hstate_ahead) {
        void* hp = nullptr;
        HIPCHK(hipHostMalloc(&hp, sizeof(KState), hipHostMallocDefault));
        ctx->host_allocs.push_back(hp);
        ctx->hstate_ahead = static_cast<KState*>(hp);
        std::memset(ctx->hstate_ahead, 0, sizeof(KState));
    }
}

static KHist* khist_ahead(wspm_ctx* ctx, int maxiter) {
    const int need = std::max(maxiter, 0) + 2;
    if (ctx->khist_ahead_cap < need) {
        ctx->khist_ahead = dalloc<KHist>(ctx, need);
        ctx->khist_ahead_cap = need;
    }
    return ctx->khist_ahead;
}

int32_t wspm_richards_step(wspm_ctx* ctx, double dt, wspm_step_stats* stats) {
    if (!ctx) return WSPM_EINVAL;
    if (!ctx->rich_init) return fail(ctx, WSPM_ESTATE, "wspm_richards_init not called");
    if (!(dt > 0.0)) return fail(ctx, WSPM_EINVAL, "dt must be positive");
    API_BEGIN
    if (ctx->tab) {
        ens_step(ctx, dt, stats);
        return WSPM_OK;
    }
    const int n = ctx->n;
    const wspm_richards_params& P = ctx->prm;
    const bool mixed = P.precision == WSPM_MIXED;
    const bool jac = P.precond == WSPM_PRECOND_JACOBI;
    const double* Ks = ctx->has_ks ? ctx->Ks : nullptr;
    if (ctx->nrec_dirty) {  // a direct assembly call replaced the nodal K
        launch_vec(ctx, SoilOp{ctx->sp, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx), 0}, 0);
        ctx->nrec_dirty = false;
    }
    // theta_n = theta(h_n): the soil state at the start of the step is current
    ++ctx->n_launch;
    HIPCHK(hipMemcpyAsync(ctx->thn, ctx->thm, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    int it = 0, kry = 0;
    double err = INFINITY;
    // edge terms change b after the assembly, and a general preconditioner
    // needs M r0: in both cases the solver starts on its own
    // the direct solver ignores the preconditioner, as the reference's does
    const bool direct = P.method == WSPM_DIRECT;
    const bool gpc = !direct && P.precond > WSPM_PRECOND_JACOBI;
    const bool own_start = ctx->rich_neumann || ctx->rich_freedrain || gpc || direct;
    // whole Picard iterations replayed from HIP graphs: CG / BiCGSTAB started
    // by the assembly, FP64, no timing events
    const bool graphs = ctx->use_graphs && ctx->fast_poll && !own_start && !mixed && !ctx->timing &&
                        (P.method == WSPM_BICGSTAB || P.method == WSPM_CG);
    const EllScope ell_scope(ctx, ctx->ell);
    // the second Picard iteration enqueued ahead of the first one's read-back:
    // folded BiCGSTAB started by the register assembly, one system, spin-wait
    // read-back (WSPM_AHEAD=0: off)
    const bool ahead = !graphs && ctx->fast_poll && !own_start && !mixed && !ctx->tab && ctx->asm_reg &&
                       P.method == WSPM_BICGSTAB && P.picard_maxit >= 2 && !env_is("WSPM_KRYLOV", "unfolded") &&
                       !env_is("WSPM_AHEAD", "0");
    if (ahead) ensure_ahead(ctx);
    while (err >= P.picard_tol && it < P.picard_maxit) {
        const int mode = own_start ? ASM_INIT_NONE
                         : mixed ? ASM_INIT_RESID
                         : P.method == WSPM_BICGSTAB ? ASM_INIT_BICG
                         : P.method == WSPM_CG ? ASM_INIT_CG : ASM_INIT_GMRES;
        if (graphs) {
            // one graph: assembly + solver start, the expected number of
            // Krylov iterations, the Picard error, the state publication
            wspm_krylov_stats st{};
            KrylovRun<double> run(ctx, P.method, ctx->Av, jac ? ctx->dinv : nullptr, ctx->bv, ctx->hA, ctx->hB,
                                  P.tol, P.atol, P.maxiter, P.restart, 0, true, P.precond);
            const PicardOp pop{ctx->sp, ctx->hB, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx),
                               P.picard_mode == WSPM_PICARD_REL};
            const int K = std::min(run.first_chunk(), P.maxiter);
            const bool defer = run.fold && P.method == WSPM_BICGSTAB;
            const hipGraphExec_t ex = step_graph(ctx, GRAPH_PICARD, dt, K, [&] {
                int sg = 0;
                do_assemble(ctx, Ks, ctx->hA, ctx->Cn, nullptr, ctx->thm, ctx->thn, dt, ctx->Av, ctx->bv,
                            jac ? ctx->dinv : nullptr, nullptr, nullptr, mode, ctx->hB, P.tol, P.atol, P.maxiter,
                            P.restart, ctx->ell, defer ? &sg : nullptr);
                if (defer) run.start_in = FoldIn{ctx->partials, ctx->prow, sg};
                run.enqueue(K);
                picard_enqueue(ctx, pop);
            });
            run.enq = K;
            ++ctx->n_launch;
            HIPCHK(hipGraphLaunch(ex, ctx->stream));
            wait_published(ctx);
            if (run.more()) {
                while (run.more()) {
                    run.enqueue(1);
                    poll(ctx, 0);
                }
                picard_poll(ctx, pop);
            }
            run.stats(&st);
            kry += st.iters;
            err = ctx->hstate[0].picard_err;
            std::swap(ctx->hA, ctx->hB);
            ++it;
            if (std::isnan(err)) break;
            continue;
        }
        if (it == 0 && ahead) {
            // The first two Picard iterations, the second enqueued before the
            // first one's read-back, so the device does not wait for the
            // host between them.  The first iteration's finish opens the
            // second's gate when it has not converged and its Krylov solve is
            // done; otherwise every kernel of the second returns at once and
            // the host goes on as without it.  The second iteration writes
            // nothing the first one's continuation reads: its own Krylov
            // history, its own pinned state mirror, and the state and work
            // vectors only once the gate is open.
            const bool rel = P.picard_mode == WSPM_PICARD_REL;
            double* dv = jac ? ctx->dinv : nullptr;
            KrylovRun<double> run1(ctx, P.method, ctx->Av, dv, ctx->bv, ctx->hA, ctx->hB, P.tol, P.atol, P.maxiter,
                                   P.restart, 0, true, P.precond);
            int sg = 0;
            do_assemble(ctx, Ks, ctx->hA, ctx->Cn, nullptr, ctx->thm, ctx->thn, dt, ctx->Av, ctx->bv, dv, nullptr,
                        nullptr, mode, ctx->hB, P.tol, P.atol, P.maxiter, P.restart, ctx->ell, &sg);
            run1.start_in = FoldIn{ctx->partials, ctx->prow, sg};
            run1.enqueue(run1.first_chunk());
            const PicardOp pop1{ctx->sp, ctx->hB, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx), rel};
            picard_enqueue(ctx, pop1, ctx->hstate, ctx->gate, P.picard_tol, 1);
            // the second iteration: h_m = hB, its iterate into hA
            KrylovRun<double> run2(ctx, P.method, ctx->Av, dv, ctx->bv, ctx->hB, ctx->hA, P.tol, P.atol, P.maxiter,
                                   P.restart, 0, true, P.precond);
            run2.hist = khist_ahead(ctx, P.maxiter);
            run2.gate = ctx->gate;
            do_assemble(ctx, Ks, ctx->hB, ctx->Cn, nullptr, ctx->thm, ctx->thn, dt, ctx->Av, ctx->bv, dv, nullptr,
                        nullptr, mode, ctx->hA, P.tol, P.atol, P.maxiter, P.restart, ctx->ell, &sg, ctx->gate);
            run2.start_in = FoldIn{ctx->partials, ctx->prow, sg};
            run2.enqueue(run2.first_chunk());
            PicardOp pop2{ctx->sp, ctx->hA, ctx->hB, ctx->Cn, nullptr, ctx->thm, krec_of(ctx), rel};
            pop2.gate = ctx->gate;
            picard_enqueue(ctx, pop2, ctx->hstate_ahead);
            wait_published(ctx);   // the first iteration, in ctx->hstate
            const KState& s1 = ctx->hstate[0];
            const bool open = s1.picard_err >= P.picard_tol && s1.done;   // the finish's test (2 iterations allowed)
            wspm_krylov_stats st{};
            if (!open) {
                if (run1.more()) {
                    while (run1.more()) {
                        run1.enqueue(1);
                        poll(ctx, 0);
                    }
                    picard_poll(ctx, pop1);
                }
            }
            run1.stats(&st);
            kry += st.iters;
            err = ctx->hstate[0].picard_err;
            std::swap(ctx->hA, ctx->hB);
            ++it;
            if (!open) {
                if (std::isnan(err)) break;
                continue;
            }
            // the second iteration ran: its state, then on as usual
            wait_published(ctx);
            ctx->hstate[0] = ctx->hstate_ahead[0];
            if (run2.more()) {
                pop2.gate = nullptr;
                while (run2.more()) {
                    run2.enqueue(1);
                    poll(ctx, 0);
                }
                picard_poll(ctx, pop2);
            }
            run2.stats(&st);
            kry += st.iters;
            err = ctx->hstate[0].picard_err;
            std::swap(ctx->hA, ctx->hB);
            ++it;
            if (std::isnan(err)) break;
            continue;
        }
        if (mixed) ensure_mixed(ctx);   // the assembly's residual start copies h_m into hB
        // the FP64 Krylov solve (constructed here so that a finish-free
        // BiCGSTAB can take over the assembly start's finish)
        std::unique_ptr<KrylovRun<double>> krun;
        int start_grid = 0;
        const bool plain = !direct && !mixed;
        // CG / BiCGSTAB started by the assembly: the constructor enqueues nothing
        // (GMRES's scales its first basis vector, so it comes after the assembly)
        if (plain && !own_start && P.method != WSPM_GMRES) {
            krun.reset(new KrylovRun<double>(ctx, P.method, ctx->Av, jac ? ctx->dinv : nullptr, ctx->bv, ctx->hA,
                                             ctx->hB, P.tol, P.atol, P.maxiter, P.restart, 0, true, P.precond));
        }
        const bool defer = krun && krun->fold && P.method == WSPM_BICGSTAB;
        do_assemble(ctx, Ks, ctx->hA, ctx->Cn, nullptr, ctx->thm, ctx->thn, dt, ctx->Av, ctx->bv,
                    (jac && !direct && !mixed) ? ctx->dinv : nullptr, mixed ? ctx->A32 : nullptr,
                    (mixed && jac) ? ctx->dinv32 : nullptr,
                    mode, ctx->hB, P.tol, P.atol, P.maxiter, P.restart, ctx->ell, defer ? &start_grid : nullptr);
        if (defer) krun->start_in = FoldIn{ctx->partials, ctx->prow, start_grid};
        if (ctx->rich_neumann)   // src/solvers/nitrogen_solver.py:115-119
            launch_edge_list(ctx, 0, EDGE_NEUMANN, ctx->rich_flux, 0.0, nullptr, 0, nullptr, nullptr, ctx->bv,
                             nullptr);
        if (ctx->rich_freedrain)   // nitrogen_solver.py:121-125, K(h_m) from the nodal K
            launch_edge_list(ctx, 1, EDGE_FREEDRAIN_WATER, 0.0, 0.0, krec_of(ctx), 1, nullptr, nullptr, ctx->bv,
                             nullptr);
        wspm_krylov_stats st{};
        const PicardOp pop{ctx->sp, ctx->hB, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx),
                           P.picard_mode == WSPM_PICARD_REL};
        if (direct) {
            direct_solve(ctx, ctx->Av, ctx->bv, ctx->hB);
            picard_poll(ctx, pop);
            if (direct_info(ctx)) throw ApiFail{WSPM_EINVAL, "direct solve: singular matrix"};
        } else if (mixed) {
            const float* d32 = jac ? ctx->dinv32 : nullptr;
            if (own_start) {   // the refinement starts from h_m
                ++ctx->n_launch;
                HIPCHK(hipMemcpyAsync(ctx->hB, ctx->hA, ctx->n * sizeof(double), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            }
            bool started = !own_start, refined = false;
            if (jacobi_on(ctx, d32)) {
                // Jacobi sweeps and the Picard update enqueued behind the
                // residual start and read back once; if the refinement needs
                // more sweeps, the update is redone after them
                if (own_start)
                    launch_spmv<double>(ctx, ctx->Av, RefResidE{ctx->bv, ctx->hB, ctx->ref_r, P.tol, P.atol, 1}, 2,
                                        true);
                started = true;
                for (int k = 0; k < jac_sweeps(false); ++k)
                    jacobi_sweep(ctx, ctx->Av, ctx->A32, d32, ctx->bv, ctx->hB, P.tol, P.atol, 2, false);
                PicardOp popr = pop;
                popr.ref = states_of(ctx, 2);
                picard_poll(ctx, popr);
                refined = ctx->hstate[0].rdone != 0;
                if (ctx->hstate[0].rstall) ctx->jac_inner_off = true;   // a sweep stalled
                if (refined) st.iters = ctx->hstate[0].kry;
            }
            if (!refined) {
                mixed_solve(ctx, P.method, ctx->Av, jac ? ctx->dinv : nullptr, ctx->A32, d32, ctx->bv, ctx->hB,
                            P.tol, P.atol, P.maxiter, P.restart, P.inner_tol, 50, &st, started, true);
                picard_poll(ctx, pop);
            }
        } else {
            // Enqueue the expected number of Krylov iterations and the Picard
            // error behind them, then read both back with one synchronisation;
            // if the solve needed more iterations, finish it and redo the error.
            if (!krun)   // after the assembly and the edge terms
                krun.reset(new KrylovRun<double>(ctx, P.method, ctx->Av, jac ? ctx->dinv : nullptr, ctx->bv, ctx->hA,
                                                 ctx->hB, P.tol, P.atol, P.maxiter, P.restart, 0, !own_start,
                                                 P.precond));
            KrylovRun<double>& run = *krun;
            run.enqueue(run.first_chunk());
            picard_poll(ctx, pop);
            if (run.more()) {
                while (run.more()) {
                    run.enqueue(1);
                    poll(ctx, 0);
                }
                picard_poll(ctx, pop);
            }
            run.stats(&st);
        }
        kry += st.iters;
        err = ctx->hstate[0].picard_err;
        std::swap(ctx->hA, ctx->hB);
        ++it;
        if (std::isnan(err)) break;
    }
    ctx->n_picard += it;
    if (stats) {
        stats->picard_iters = it;
        stats->krylov_iters = kry;
        stats->converged = err < P.picard_tol;
        stats->reserved = 0;
        stats->err = err;
    }
    return WSPM_OK;
    API_END
}

int32_t wspm_richards_state(wspm_ctx* ctx, double* h, double* theta, double* theta_n) {
    if (!ctx) return WSPM_EINVAL;
    if (!ctx->rich_init) return fail(ctx, WSPM_ESTATE, "wspm_richards_init not called");
    API_BEGIN
    const size_t bytes = ctx->n * sizeof(double);
    if (h) HIPCHK(hipMemcpyAsync(h, ctx->hA, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (theta) HIPCHK(hipMemcpyAsync(theta, ctx->thm, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (theta_n) HIPCHK(hipMemcpyAsync(theta_n, ctx->thn, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

int32_t wspm_timing_enable(wspm_ctx* ctx, int32_t on) {
    if (!ctx) return WSPM_EINVAL;
    API_BEGIN
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (on < 0 || on > 7) return fail(ctx, WSPM_EINVAL, "timing: 0..7 (1 FP64 SpMV, 2 FP32 SpMV, +4 assembly)");
    if (on && !ctx->tflags) ctx->tflags = dalloc<int>(ctx, TIMING_SLOTS);
    if (on) HIPCHK(hipMemset(ctx->tflags, 0, TIMING_SLOTS * sizeof(int)));
    if ((on & 4) && !ctx->asm_tflags) ctx->asm_tflags = dalloc<int>(ctx, TIMING_SLOTS);
    if (on & 4) HIPCHK(hipMemset(ctx->asm_tflags, 0, TIMING_SLOTS * sizeof(int)));
    ctx->timing = on != 0;
    ctx->timing_fp32 = (on & 3) == 2;
    ctx->timing_asm = (on & 4) != 0;
    ctx->ev_used = 0;
    ctx->asm_ev_used = 0;
    return WSPM_OK;
    API_END
}

int32_t wspm_timing_read_asm(wspm_ctx* ctx, int64_t* launches, double* total_ms, int64_t* rows) {
    if (!ctx || !launches || !total_ms) return WSPM_EINVAL;
    API_BEGIN
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t pairs = ctx->asm_ev_used / 2;
    std::vector<int> flags(pairs, 0);
    if (pairs) HIPCHK(hipMemcpy(flags.data(), ctx->asm_tflags, pairs * sizeof(int), hipMemcpyDeviceToHost));
    double tot = 0.0;
    int64_t worked = 0, nrows = 0;
    for (size_t k = 0; k < pairs; ++k) {
        if (!flags[k]) continue;
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->asm_ev[2 * k], ctx->asm_ev[2 * k + 1]));
        tot += ms;
        ++worked;
        nrows += flags[k];
    }
    *launches = worked;
    *total_ms = tot;
    if (rows) *rows = nrows;
    ctx->asm_ev_used = 0;
    if (ctx->asm_tflags) HIPCHK(hipMemset(ctx->asm_tflags, 0, TIMING_SLOTS * sizeof(int)));
    return WSPM_OK;
    API_END
}

// Timed SpMV launches with at least min_rows rows: count, summed ms, rows.
static void timing_sum(wspm_ctx* ctx, int64_t min_rows, int64_t* launches, double* total_ms, int64_t* rows) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t pairs = ctx->ev_used / 2;
    std::vector<int> flags(pairs, 0);
    if (pairs) HIPCHK(hipMemcpy(flags.data(), ctx->tflags, pairs * sizeof(int), hipMemcpyDeviceToHost));
    double tot = 0.0;
    int64_t worked = 0, nrows = 0;
    for (size_t k = 0; k < pairs; ++k) {
        if (!flags[k] || flags[k] < min_rows) continue;
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev_pool[2 * k], ctx->ev_pool[2 * k + 1]));
        tot += ms;
        ++worked;
        nrows += flags[k];
    }
    *launches = worked;
    *total_ms = tot;
    if (rows) *rows = nrows;
}

int32_t wspm_timing_read_over(wspm_ctx* ctx, int64_t min_rows, int64_t* launches, double* total_ms,
                              int64_t* rows) {
    if (!ctx || !launches || !total_ms) return WSPM_EINVAL;
    API_BEGIN
    timing_sum(ctx, min_rows, launches, total_ms, rows);
    return WSPM_OK;
    API_END
}

int32_t wspm_launch_count(const wspm_ctx* ctx, int64_t* launches, int64_t* picard_rounds) {
    if (!ctx || !launches || !picard_rounds) return WSPM_EINVAL;
    *launches = ctx->n_launch;
    *picard_rounds = ctx->n_picard;
    return WSPM_OK;
}

int32_t wspm_timing_read(wspm_ctx* ctx, int64_t* launches, double* total_ms, int64_t* rows) {
    if (!ctx || !launches || !total_ms) return WSPM_EINVAL;
    API_BEGIN
    timing_sum(ctx, 1, launches, total_ms, rows);
    ctx->ev_used = 0;
    if (ctx->tflags) HIPCHK(hipMemset(ctx->tflags, 0, TIMING_SLOTS * sizeof(int)));
    return WSPM_OK;
    API_END
}


// ---------------------------------------------------------------------------
// Transport: edge lists, flux/dispersion fields, species systems
// ---------------------------------------------------------------------------
int32_t wspm_edge_set(wspm_ctx* ctx, int32_t list, int32_t n, const int32_t* nodes) {
    if (!ctx) return WSPM_EINVAL;
    if (list != WSPM_EDGE_INLET && list != WSPM_EDGE_OUTLET) return fail(ctx, WSPM_EINVAL, "unknown edge list");
    if (n < 0 || (n > 0 && !nodes)) return fail(ctx, WSPM_EINVAL, "bad edge node array");
    if (ctx->dim != 2) return fail(ctx, WSPM_EINVAL, "edge terms are 2D");
    const HostPattern& P = ctx->hp;
    auto pos = [&](int i, int j) -> int {
        const auto b = P.cols.begin() + P.row_ptr[i], e = P.cols.begin() + P.row_ptr[i + 1];
        const auto it = std::lower_bound(b, e, j);
        return (it != e && *it == j) ? (int)(it - P.cols.begin()) : -1;
    };
    std::vector<int32_t> nd(nodes, nodes + n), sp(n, -1), sn(n, -1);
    std::vector<uint8_t> seen(ctx->n, 0);
    for (int k = 0; k < n; ++k) {
        if (nd[k] < 0 || nd[k] >= ctx->n) return fail(ctx, WSPM_EINVAL, "edge node out of range");
        if (seen[nd[k]]) return fail(ctx, WSPM_EINVAL, "edge node listed twice");
        seen[nd[k]] = 1;
    }
    // a pair not joined by a mesh edge has no matrix entries in the pattern:
    // the list then takes right-hand-side terms only (Neumann, free drainage
    // of water), and the matrix terms (Cauchy, free drainage of nitrogen)
    // refuse it
    std::string unjoined;
    for (int k = 0; k < n; ++k) {
        if (k > 0) sp[k] = pos(nd[k], nd[k - 1]);
        if (k + 1 < n) sn[k] = pos(nd[k], nd[k + 1]);
        if (unjoined.empty() && k + 1 < n && sn[k] < 0)
            unjoined = std::to_string(nd[k]) + " and " + std::to_string(nd[k + 1]);
    }
    API_BEGIN
    auto& L = ctx->edges[list];
    if (n > L.cnt || !L.nodes) {
        L.nodes = dalloc<int32_t>(ctx, std::max(n, 1));
        L.slot_prev = dalloc<int32_t>(ctx, std::max(n, 1));
        L.slot_next = dalloc<int32_t>(ctx, std::max(n, 1));
    }
    if (n > 0) {
        HIPCHK(hipMemcpy(L.nodes, nd.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(L.slot_prev, sp.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(L.slot_next, sn.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    L.cnt = n;
    L.unjoined = unjoined;
    return WSPM_OK;
    API_END
}

int32_t wspm_richards_edges(wspm_ctx* ctx, int32_t inlet_neumann, double neumann_flux, int32_t outlet_freedrain) {
    if (!ctx) return WSPM_EINVAL;
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    ctx->rich_neumann = inlet_neumann != 0;
    ctx->rich_flux = neumann_flux;
    ctx->rich_freedrain = outlet_freedrain != 0;
    ++ctx->graph_epoch;
    return WSPM_OK;
}

int32_t wspm_transport_fields(wspm_ctx* ctx, const double* phi, const double* h, const double* Ks_elem, double DL,
                              double DT, double Dm, double* qx, double* qz, double* absq, double* dxx, double* dxz,
                              double* dzz, double* theta) {
    if (!ctx) return WSPM_EINVAL;
    if (!phi || !h || !qx || !qz || !absq || !dxx || !dxz || !dzz || !theta)
        return fail(ctx, WSPM_EINVAL, "null argument");
    if (ctx->dim != 2) return fail(ctx, WSPM_EINVAL, "transport is 2D");
    API_BEGIN
    FieldArgs a{};
    a.n = ctx->n;
    a.inc_ptr = ctx->inc_ptr;
    a.inc = ctx->inc;
    a.conn4 = ctx->conn4;
    a.nrec = nrec_of(ctx);
    a.h = h;
    a.ks_e = Ks_elem;
    a.sp = make_soil(WSPM_SOIL_VAN_GENUCHTEN, phi);
    a.DL = DL; a.DT = DT; a.Dm = Dm;
    a.qx = qx; a.qz = qz; a.aq = absq; a.dxx = dxx; a.dxz = dxz; a.dzz = dzz; a.th = theta;
    HIPCHK(launch_transport_fields(a, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

int32_t wspm_assemble_transport(wspm_ctx* ctx, const wspm_transport_params* tp, const double* theta,
                                const double* qx, const double* qz, const double* dxx, const double* dxz,
                                const double* dzz, const double* K, const double* c_n, const double* c_src,
                                double* A_vals, double* b, double* dinv) {
    if (!ctx) return WSPM_EINVAL;
    if (!tp || !theta || !qx || !qz || !dxx || !dxz || !dzz || !c_n || !A_vals || !b)
        return fail(ctx, WSPM_EINVAL, "null argument");
    if (ctx->dim != 2) return fail(ctx, WSPM_EINVAL, "transport is 2D");
    if (!(tp->dt > 0.0)) return fail(ctx, WSPM_EINVAL, "dt must be positive");
    if (ctx->max_row > 16) return fail(ctx, WSPM_EINVAL, "rows longer than 16 entries");
    if (tp->outlet_freedrain && !K) return fail(ctx, WSPM_EINVAL, "free drainage needs K");
    API_BEGIN
    TrArgs a{};
    a.n = ctx->n;
    a.row_ptr = ctx->row_ptr; a.cols = ctx->cols; a.diag_pos = ctx->diag_pos;
    a.inc_ptr = ctx->inc_ptr; a.inc = ctx->inc; a.conn4 = ctx->conn4; a.nrec = nrec_of(ctx);
    a.th = theta; a.qx = qx; a.qz = qz; a.dxx = dxx; a.dxz = dxz; a.dzz = dzz;
    a.c_n = c_n;
    a.c_src = c_src;
    a.mu = tp->mu; a.rho_kd = tp->rho_kd; a.mu_src = tp->mu_src;
    a.inv_dt = 1.0 / tp->dt;
    a.lumped = tp->lumped != 0;
    a.tq = ctx->tq;
    a.A = A_vals; a.b = b; a.dinv = dinv;
    HIPCHK(launch_assemble_transport(a, ctx->max_row, ctx->stream));
    if (tp->inlet_cauchy)   // src/models/boundary.py:217-277
        launch_edge_list(ctx, WSPM_EDGE_INLET, EDGE_CAUCHY, tp->cauchy_flux, tp->c_inlet, nullptr, 0, nullptr,
                         A_vals, b, dinv);
    if (tp->outlet_freedrain)   // boundary.py:336-421
        launch_edge_list(ctx, WSPM_EDGE_OUTLET, EDGE_FREEDRAIN_N, 0.0, 0.0, K, 1, theta, A_vals, nullptr, dinv);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the implicit step
// ---------------------------------------------------------------------------
// Position of the transposed entry of every CSR slot (the pattern is
// structurally symmetric), built once per context.
static void ensure_tpos(wspm_ctx* ctx) {
    if (ctx->adj.tpos) return;
    const HostPattern& P = ctx->hp;
    std::vector<int32_t> tp(P.cols.size());
    for (int i = 0; i < ctx->n; ++i)
        for (int k = P.row_ptr[i]; k < P.row_ptr[i + 1]; ++k) {
            const int j = P.cols[k];
            const auto b = P.cols.begin() + P.row_ptr[j], e = P.cols.begin() + P.row_ptr[j + 1];
            tp[k] = (int32_t)(std::lower_bound(b, e, i) - P.cols.begin());
        }
    ctx->adj.tpos = dupload(ctx, tp.data(), tp.size());
}

static void adj_reduce(wspm_ctx* ctx, const double* a, int len, double* out) {
    const AdjSum op{a, len, out};
    const int g = std::max(1, std::min((len + BLOCK - 1) / BLOCK, 1024));
    LAUNCH(k_vec<AdjSum>, dim3(g), dim3(BLOCK), 0, ctx->stream, len, op, red_of(ctx, 0),
                       EnsWalk{nullptr, nullptr, 0});
    HIPCHK(hipGetLastError());
    LAUNCH((k_finish<2, AdjSum>), dim3(1), dim3(BLOCK), 0, ctx->stream, op, ctx->partials, ctx->prow, g,
                       ctx->states);
    HIPCHK(hipGetLastError());
}

int32_t wspm_adj_step(wspm_ctx* ctx, const wspm_adj_params* prm, const double* Ks_elem, const double* h,
                      const double* h_prev, const int32_t* obs, const double* dldy, const double* lam_next,
                      const double* dldh_in, double* lam, double* grad_logks, double* grad_an,
                      wspm_krylov_stats* stats) {
    if (!ctx) return WSPM_EINVAL;
    if (!prm || !Ks_elem || !h || !h_prev || !lam || !grad_logks || !grad_an)
        return fail(ctx, WSPM_EINVAL, "null argument");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    if (ctx->dim != 2) return fail(ctx, WSPM_EINVAL, "the adjoint is 2D");
    if (ctx->max_row > 16) return fail(ctx, WSPM_EINVAL, "rows longer than 16 entries");
    if (!(prm->dt > 0.0)) return fail(ctx, WSPM_EINVAL, "dt must be positive");
    if (prm->n_obs < 0 || (prm->n_obs > 0 && (!obs || !dldy))) return fail(ctx, WSPM_EINVAL, "bad observations");
    if (prm->maxiter < 1) return fail(ctx, WSPM_EINVAL, "maxiter must be >= 1");
    API_BEGIN
    const int n = ctx->n, E = ctx->E;
    auto& W = ctx->adj;
    if (!W.nodal) {
        W.nodal = dalloc<double>(ctx, 9 * (size_t)n);
        W.ml = dalloc<double>(ctx, n);
        W.J = dalloc<double>(ctx, ctx->nnz);
        W.JT = dalloc<double>(ctx, ctx->nnz);
        W.dinv = dalloc<double>(ctx, n);
        W.rhs = dalloc<double>(ctx, n);
        W.dldh = dalloc<double>(ctx, n);
        W.ge = dalloc<double>(ctx, 2 * (size_t)E);
        W.gn = dalloc<double>(ctx, 2 * (size_t)n);
    }
    ensure_tpos(ctx);
    AdjArgs a{};
    a.n = n;
    a.E = E;
    a.row_ptr = ctx->row_ptr;
    a.diag_pos = ctx->diag_pos;
    a.inc_ptr = ctx->inc_ptr;
    a.inc = ctx->inc;
    a.conn4 = ctx->conn4;
    a.nrec = nrec_of(ctx);
    a.bc_flag = ctx->n_bc > 0 ? ctx->bc_flag : nullptr;
    a.ks_e = Ks_elem;
    a.h = h;
    a.hp = h_prev;
    double p[6];
    for (int k = 0; k < 6; ++k) p[k] = prm->phi[k];
    p[5] = 1.0;
    a.sp = make_soil(WSPM_SOIL_VAN_GENUCHTEN, p);
    adj_tie_m(a.sp);
    a.inv_dt = 1.0 / prm->dt;
    a.qc = ctx->qc;
    double* nd = W.nodal;
    a.nd = AdjNodal{nd, nd + n, nd + 2 * (size_t)n, nd + 3 * (size_t)n, nd + 4 * (size_t)n, nd + 5 * (size_t)n,
                    nd + 6 * (size_t)n, nd + 7 * (size_t)n, nd + 8 * (size_t)n};
    a.ml = W.ml;
    HIPCHK(launch_adj_mass(a, W.ml, ctx->stream));
    HIPCHK(launch_adj_nodal(a, ctx->stream));
    HIPCHK(launch_adj_jacobian(a, ctx->max_row, W.J, W.JT, W.dinv, W.tpos, ctx->nnz, ctx->stream));
    if (dldh_in) HIPCHK(hipMemcpyAsync(W.dldh, dldh_in, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    else HIPCHK(hipMemsetAsync(W.dldh, 0, n * sizeof(double), ctx->stream));
    if (prm->n_obs > 0) HIPCHK(launch_adj_obs(a, prm->n_obs, obs, dldy, W.dldh, grad_an, ctx->stream));
    HIPCHK(launch_adj_rhs(a, W.dldh, lam_next, W.rhs, ctx->stream));
    if (lam_next) HIPCHK(hipMemcpyAsync(lam, lam_next, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    else HIPCHK(hipMemsetAsync(lam, 0, n * sizeof(double), ctx->stream));
    wspm_krylov_stats st{};
    krylov_solve<double>(ctx, WSPM_BICGSTAB, W.JT, W.dinv, W.rhs, lam, lam, prm->tol, 0.0, prm->maxiter, 30, 0, &st);
    HIPCHK(launch_adj_params(a, lam, grad_logks, W.ge, W.gn, ctx->stream));
    adj_reduce(ctx, W.ge, E, grad_an);
    adj_reduce(ctx, W.gn, n, grad_an);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (stats) *stats = st;
    return WSPM_OK;
    API_END
}

int32_t wspm_adj_transport_step(wspm_ctx* ctx, const wspm_adj_tr_params* prm, const double* Ks_elem,
                                const double* h, const double* h_next, const double* c, const double* c_prev,
                                const int32_t* obs, const double* dldy, const double* lam_next, double* lam,
                                double* dldh, double* grad_logks, double* grad_an, wspm_krylov_stats* stats) {
    if (!ctx) return WSPM_EINVAL;
    if (!prm || !Ks_elem || !h || !c || !c_prev || !lam || !dldh || !grad_logks || !grad_an)
        return fail(ctx, WSPM_EINVAL, "null argument");
    if ((h_next == nullptr) != (lam_next == nullptr))
        return fail(ctx, WSPM_EINVAL, "h_next and lam_next go together (both NULL at the last step)");
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    if (ctx->dim != 2) return fail(ctx, WSPM_EINVAL, "the adjoint is 2D");
    if (ctx->max_row > 16) return fail(ctx, WSPM_EINVAL, "rows longer than 16 entries");
    if (!(prm->dt > 0.0)) return fail(ctx, WSPM_EINVAL, "dt must be positive");
    if (prm->n_obs < 0 || (prm->n_obs > 0 && (!obs || !dldy))) return fail(ctx, WSPM_EINVAL, "bad observations");
    if (prm->maxiter < 1) return fail(ctx, WSPM_EINVAL, "maxiter must be >= 1");
    if (prm->cauchy_flux != 0.0 && ctx->edges[WSPM_EDGE_INLET].cnt < 2)
        return fail(ctx, WSPM_EINVAL, "a Cauchy flux needs the inlet edge list (wspm_edge_set)");
    API_BEGIN
    const int n = ctx->n, E = ctx->E;
    const size_t N = (size_t)n;
    auto& W = ctx->adjtr;
    if (!W.f) {
        W.f = dalloc<double>(ctx, 7 * N);
        W.fnext = dalloc<double>(ctx, 7 * N);
        W.mth = dalloc<double>(ctx, N);
        W.mT = dalloc<double>(ctx, N);
        W.A = dalloc<double>(ctx, ctx->nnz);
        W.AT = dalloc<double>(ctx, ctx->nnz);
        W.dinv = dalloc<double>(ctx, N);
        W.b = dalloc<double>(ctx, N);
        W.rhs = dalloc<double>(ctx, N);
        W.gf = dalloc<double>(ctx, 6 * N);
        W.w = dalloc<double>(ctx, 3 * N);
        W.gn = dalloc<double>(ctx, 2 * N);
    }
    ensure_tpos(ctx);
    double p[6];
    for (int k = 0; k < 6; ++k) p[k] = prm->phi[k];
    p[5] = 1.0;                                   // Ks enters through Ks_elem
    SoilParams sp = make_soil(WSPM_SOIL_VAN_GENUCHTEN, p);
    adj_tie_m(sp);
    // node fields of h_k (and of h_{k+1} for the next step's species mass)
    auto fields = [&](const double* hh, double* out) {
        FieldArgs f{};
        f.n = n;
        f.inc_ptr = ctx->inc_ptr; f.inc = ctx->inc; f.conn4 = ctx->conn4; f.nrec = nrec_of(ctx);
        f.h = hh; f.ks_e = Ks_elem; f.sp = sp;
        f.DL = prm->DL; f.DT = prm->DT; f.Dm = prm->Dm;
        f.qx = out; f.qz = out + N; f.aq = out + 2 * N; f.dxx = out + 3 * N; f.dxz = out + 4 * N;
        f.dzz = out + 5 * N; f.th = out + 6 * N;
        HIPCHK(launch_transport_fields(f, ctx->stream));
    };
    fields(h, W.f);
    const double* F = W.f;
    TrAdjArgs a{};
    a.n = n; a.E = E;
    a.inc_ptr = ctx->inc_ptr; a.inc = ctx->inc; a.conn4 = ctx->conn4; a.nrec = nrec_of(ctx);
    a.tq = ctx->tq;
    a.ks_e = Ks_elem;
    a.h = h;
    a.sp = sp;
    a.DL = prm->DL; a.DT = prm->DT; a.Dm = prm->Dm;
    a.inv_dt = 1.0 / prm->dt;
    for (int k = 0; k < 3; ++k) a.mu[k] = prm->mu[k];
    a.rho_kd = prm->rho_kd;
    a.qx = F; a.qz = F + N; a.th = F + 6 * N;
    a.c = c; a.cp = c_prev; a.lam = lam;
    if (h_next) {
        fields(h_next, W.fnext);
        HIPCHK(launch_tr_adj_mass(a, W.fnext + 6 * N, W.mth, W.mT, ctx->stream));
    }
    wspm_krylov_stats tot{};
    for (int s = 2; s >= 0; --s) {
        double* ls = lam + s * N;
        HIPCHK(launch_tr_adj_rhs(a, s, W.mth, W.mT, lam_next ? lam_next + s * N : nullptr,
                                 s < 2 ? lam + (s + 1) * N : nullptr, W.rhs, ctx->stream));
        if (s == 2 && prm->n_obs > 0) HIPCHK(launch_tr_adj_obs(prm->n_obs, obs, dldy, W.rhs, ctx->stream));
        // A_s of the forward step (k_assemble_transport + the Cauchy inlet)
        TrArgs t{};
        t.n = n;
        t.row_ptr = ctx->row_ptr; t.cols = ctx->cols; t.diag_pos = ctx->diag_pos;
        t.inc_ptr = ctx->inc_ptr; t.inc = ctx->inc; t.conn4 = ctx->conn4; t.nrec = nrec_of(ctx);
        t.th = F + 6 * N; t.qx = F; t.qz = F + N; t.dxx = F + 3 * N; t.dxz = F + 4 * N; t.dzz = F + 5 * N;
        t.c_n = c_prev + s * N;
        t.c_src = s > 0 ? c + (s - 1) * N : nullptr;
        t.mu = prm->mu[s];
        t.rho_kd = s == 0 ? prm->rho_kd : 0.0;
        t.mu_src = s > 0 ? prm->mu[s - 1] : 0.0;
        t.inv_dt = a.inv_dt;
        t.lumped = 1;
        t.tq = ctx->tq;
        t.A = W.A; t.b = W.b; t.dinv = W.dinv;
        HIPCHK(launch_assemble_transport(t, ctx->max_row, ctx->stream));
        if (prm->cauchy_flux != 0.0)
            launch_edge_list(ctx, WSPM_EDGE_INLET, EDGE_CAUCHY, prm->cauchy_flux, 0.0, nullptr, 0, nullptr, W.A,
                             W.b, W.dinv);
        HIPCHK(launch_transpose_values(W.A, ctx->adj.tpos, W.AT, ctx->nnz, ctx->stream));
        if (lam_next) HIPCHK(hipMemcpyAsync(ls, lam_next + s * N, N * sizeof(double), hipMemcpyDeviceToDevice,
                                            ctx->stream));
        else HIPCHK(hipMemsetAsync(ls, 0, N * sizeof(double), ctx->stream));
        wspm_krylov_stats st{};
        krylov_solve<double>(ctx, WSPM_BICGSTAB, W.AT, W.dinv, W.rhs, ls, ls, prm->tol, 0.0, prm->maxiter, 30, 0,
                             &st);
        tot.iters += st.iters;
        if (st.code != 0) tot.code = st.code;
        tot.res_norm = std::max(tot.res_norm, st.res_norm);
        tot.b_norm = std::max(tot.b_norm, st.b_norm);
    }
    HIPCHK(launch_tr_adj_fields(a, W.gf, ctx->stream));
    HIPCHK(launch_tr_adj_chain(a, W.gf, W.w, dldh, W.gn, grad_logks, ctx->stream));
    adj_reduce(ctx, W.gn, n, grad_an);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (stats) *stats = tot;
    return WSPM_OK;
    API_END
}

// ---------------------------------------------------------------------------
// Ensemble of independent instances
// ---------------------------------------------------------------------------
// Rows per ensemble table entry: the register assembly runs one thread per row.
static constexpr int ENS_ROWS = BLOCK;
// Workgroups of an ensemble launch (they stride over the listed entries).
static constexpr int ENS_GRID = 2048;

// Ensemble launch shapes.  The operator format follows from the meshes
// alone: ELL when the rows allow it (ell_shape_ok), else CSR at the row
// width wspm_ctx_create chose (WSPM_LANES overrides it).  Every context that
// holds the same instances therefore sums each row in the same order, and
// results are repeatable run after run and across groupings of the
// instances.  The tuner times only what changes no bit of the results: the
// number of workgroups that walk the entry lists.  Every listed entry (256
// rows of one instance) is worked by one workgroup and writes its own
// partial-row column, so the grid decides who works an entry, not what is
// summed.  Candidates 8192 ... 256 workgroups (at most one per entry),
// timed on the whole operator (zero values) with the time-step format: 2
// warm-up and 10 timed launches each; a smaller grid is kept only when it is
// more than 3 % faster.  (The solves' launches walk short lists -- 8 % of the
// entries on average over a 256-instance time step -- where a large grid
// keeps every entry on a workgroup of its own; on the whole operator the
// grids are within a few per cent of each other, and a smaller grid that won
// that probe by noise cost 2-5 % per time step.)  WSPM_ENS_GRID fixes the
// grid instead.
static void ens_autotune(wspm_ctx* ctx) {
    const char* env = std::getenv("WSPM_LANES");
    if (env) {
        const int l = std::atoi(env);
        if (l == 2 || l == 4 || l == 8 || l == 16 || l == 32 || l == 64) ctx->lanes = l;
    }
    const char* g = std::getenv("WSPM_ENS_GRID");
    if (g && std::atoi(g) > 0) {
        ctx->ens_grid = std::max(1, std::min(std::atoi(g), ctx->n_tab));
        ctx->ens_grid_cands = 0;
        return;
    }
    double* av = nullptr;
    double* xy = nullptr;
    const size_t cap = op_cap(ctx);
    HIPCHK(hipMalloc(&av, cap * sizeof(double)));
    hipError_t e = hipMalloc(&xy, 2 * (size_t)ctx->n * sizeof(double));
    if (e != hipSuccess) {
        (void)hipFree(av);
        HIPCHK(e);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        HIPCHK(hipMemsetAsync(av, 0, cap * sizeof(double), ctx->stream));
        HIPCHK(hipMemsetAsync(xy, 0, 2 * (size_t)ctx->n * sizeof(double), ctx->stream));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const bool ell = ell_shape_ok(ctx);
        if (ell) ensure_ell(ctx);
        const EllScope scope(ctx, ell);
        const EpiPlain<double> epi{xy, xy + ctx->n};
        int best = std::min(ENS_GRID, ctx->n_tab), cands = 0;
        float best_ms = 1e30f;
        int last = -1;
        for (int grid = 8192; grid >= 256; grid /= 2) {
            const int gg = std::min(grid, ctx->n_tab);
            if (gg == last) continue;               // the table has fewer entries
            last = gg;
            ctx->ens_grid = gg;
            for (int r = 0; r < 2; ++r) launch_spmv<double>(ctx, av, epi, 0);
            HIPCHK(hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < 10; ++r) launch_spmv<double>(ctx, av, epi, 0);
            HIPCHK(hipEventRecord(e1, ctx->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            ++cands;
            if (ms < best_ms * 0.97f) {
                best_ms = ms;
                best = gg;
            }
        }
        ctx->ens_grid = best;
        ctx->ens_grid_cands = cands;
        ctx->ens_tune_us = best_ms * 100.0f;   // per launch, microseconds
    } catch (const Err&) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(av);
        (void)hipFree(xy);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(av);
    (void)hipFree(xy);
}

int32_t wspm_ens_create(int32_t dim, int32_t n_inst, const int32_t* node_counts, const int32_t* elem_counts,
                        const int32_t* conn, const double* coords, void* stream, wspm_ctx** out) {
    if (!out || !node_counts || !elem_counts || !conn || !coords) {
        g_create_error = "null argument";
        return WSPM_EINVAL;
    }
    *out = nullptr;
    if (dim != 2) {
        g_create_error = "ensembles are 2D (triangle meshes)";
        return WSPM_EINVAL;
    }
    if (n_inst < 1) {
        g_create_error = "n_inst must be >= 1";
        return WSPM_EINVAL;
    }
    const int K = dim + 1;
    std::vector<int> noff(n_inst + 1, 0), eoff(n_inst + 1, 0);
    for (int i = 0; i < n_inst; ++i) {
        if (node_counts[i] < 1 || elem_counts[i] < 1) {
            g_create_error = "every instance needs nodes and elements";
            return WSPM_EINVAL;
        }
        const int64_t nn = (int64_t)noff[i] + node_counts[i], ne = (int64_t)eoff[i] + elem_counts[i];
        if (nn > INT32_MAX / 2 || ne > INT32_MAX / 4) {
            g_create_error = "ensemble too large for 32-bit indices";
            return WSPM_EINVAL;
        }
        noff[i + 1] = (int)nn;
        eoff[i + 1] = (int)ne;
    }
    // instance-local connectivity -> global node numbers
    std::vector<int32_t> gconn((size_t)eoff[n_inst] * K);
    for (int i = 0; i < n_inst; ++i) {
        for (int64_t q = (int64_t)eoff[i] * K; q < (int64_t)eoff[i + 1] * K; ++q) {
            const int32_t v = conn[q];
            if (v < 0 || v >= node_counts[i]) {
                g_create_error = "instance " + std::to_string(i) + ": vertex index out of range";
                return WSPM_EINVAL;
            }
            gconn[q] = v + noff[i];
        }
    }
    wspm_ctx* ctx = nullptr;
    int rc = wspm_ctx_create(dim, noff[n_inst], eoff[n_inst], gconn.data(), coords, stream, &ctx);
    if (rc != WSPM_OK) return rc;
    if (!ctx->asm_reg) {
        g_create_error = "ensemble rows must have at most 8 entries (register assembly)";
        wspm_ctx_destroy(ctx);
        return WSPM_EINVAL;
    }
    try {
        std::vector<RowTab> tab;
        std::vector<int> iblk(n_inst + 1, 0);
        for (int i = 0; i < n_inst; ++i) {
            iblk[i] = (int)tab.size();
            for (int r = noff[i]; r < noff[i + 1]; r += ENS_ROWS)
                tab.push_back(make_int4(r, std::min(r + ENS_ROWS, noff[i + 1]), i,
                                        r == noff[i] ? noff[i + 1] - noff[i] : 0));
        }
        iblk[n_inst] = (int)tab.size();
        ctx->n_inst = n_inst;
        ctx->n_tab = (int)tab.size();
        ctx->tab = dupload(ctx, tab.data(), tab.size());
        ctx->htab = tab;
        ctx->iblk = dupload(ctx, iblk.data(), iblk.size());
        if (ctx->n_tab > ctx->prow) {
            ctx->prow = ctx->n_tab;
            ctx->partials = dalloc<double>(ctx, (size_t)ctx->prow * MAX_DOTS);
        }
        ctx->states = dalloc<KState>(ctx, 3 * (size_t)n_inst);
        HIPCHK(hipMemset(ctx->states, 0, 3 * (size_t)n_inst * sizeof(KState)));
        ctx->ens_status = dalloc<int>(ctx, 4);
        ctx->ens_out = dalloc<int>(ctx, 5 * (size_t)n_inst + 2);  // 3 int + 1 double each
        ctx->kact = dalloc<int>(ctx, 1 + (size_t)ctx->n_tab);
        ctx->kact_alt = dalloc<int>(ctx, 1 + (size_t)ctx->n_tab);
        HIPCHK(hipMemset(ctx->kact_alt, 0, sizeof(int)));
        ctx->ens_list_kernel = env_is("WSPM_ENS_LIST", "kernel");
        ctx->pact = dalloc<int>(ctx, 1 + (size_t)ctx->n_tab);
        HIPCHK(hipMemset(ctx->kact, 0, sizeof(int)));
        HIPCHK(hipMemset(ctx->pact, 0, sizeof(int)));
        {
            const char* ce = std::getenv("WSPM_ENS_COMPACT");
            ctx->ens_compact_every = std::max(1, ce ? std::atoi(ce) : 1);
        }
        void* hp = nullptr;
        HIPCHK(hipHostMalloc(&hp, 4 * sizeof(int), hipHostMallocDefault));
        ctx->host_allocs.push_back(hp);
        ctx->hstatus = static_cast<int*>(hp);
        std::memset(ctx->hstatus, 0, 4 * sizeof(int));
        ens_autotune(ctx);
        HIPCHK(hipStreamSynchronize(ctx->stream));
    } catch (const Err&) {
        g_create_error = ctx->err;
        free_ctx(ctx);
        return WSPM_EHIP;
    }
    *out = ctx;
    return WSPM_OK;
}

int32_t wspm_ens_richards_init(wspm_ctx* ctx, const wspm_richards_params* prm, const double* phi,
                               const double* Ks_elem, const double* h0) {
    if (!ctx) return WSPM_EINVAL;
    if (!ctx->tab) return fail(ctx, WSPM_EINVAL, "not an ensemble context");
    if (!prm || !h0) return fail(ctx, WSPM_EINVAL, "null argument");
    int rc = check_richards_params(ctx, prm);
    if (rc) return rc;
    if (prm->method == WSPM_GMRES) return fail(ctx, WSPM_EINVAL, "ensembles solve with CG or BiCGSTAB");
    if (prm->precond > WSPM_PRECOND_JACOBI) return fail(ctx, WSPM_EINVAL, "ensembles take no preconditioner or Jacobi");
    if (prm->precision == WSPM_MIXED && !(prm->inner_tol > 0.0 && prm->inner_tol < 1.0))
        return fail(ctx, WSPM_EINVAL, "inner_tol must be in (0, 1)");
    API_BEGIN
    richards_alloc(ctx, prm, Ks_elem, h0);
    const int ni = ctx->n_inst;
    std::vector<KState> hs(ni);
    std::memset(hs.data(), 0, ni * sizeof(KState));
    for (int i = 0; i < ni; ++i) {
        double p[6];
        for (int k = 0; k < 6; ++k) p[k] = phi ? phi[6 * (size_t)i + k] : prm->phi[k];
        if (prm->soil_model == WSPM_SOIL_GARDNER) p[4] = p[5] = 0.0;
        if (Ks_elem) p[prm->soil_model == WSPM_SOIL_GARDNER ? 3 : 5] = 1.0;  // nodal model gives Kr
        hs[i].sp = make_soil(prm->soil_model, p);
        hs[i].picard_tol = prm->picard_tol;
        hs[i].picard_maxit = prm->picard_maxit;
        hs[i].picard_rel = prm->picard_mode == WSPM_PICARD_REL;
        hs[i].done = 1;
        hs[i].pdone = 1;
    }
    // hs is pageable and local: the stream is synchronised below, before it goes
    HIPCHK(hipMemcpyAsync(states_of(ctx, 0), hs.data(), ni * sizeof(KState), hipMemcpyHostToDevice, ctx->stream));
    launch_vec(ctx, SoilOp{ctx->sp, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx), 1}, 0, LIST_ALL);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nrec_dirty = false;
    ctx->rich_init = true;
    ctx->has_ks = Ks_elem != nullptr;
    return WSPM_OK;
    API_END
}

// One time step of every instance: the Picard loop of
// src/solvers/richards_solver_v2.py:141-168 per instance.  Each Picard
// iteration is one assembly (+ solver start) over the instances still
// iterating, their Krylov solves (instances drop out as they converge), and
// the Picard error per instance; the host reads back three counters per
// iteration chunk.
// Mixed-precision linear solves of one ensemble Picard iteration: the
// assembly left r = b - A x (x = h_m) and |r|, |b| per instance.  Sweeps of
// FP32 Krylov corrections on the instances still refining (device-compacted
// lists; each instance's own FP64 residual test sets rdone), each followed by
// the FP64 update and residual, until every instance's FP64 residual meets
// the Krylov tolerance.  The host reads the count of instances still refining
// once per ENS_SWEEPS sweeps: the corrections use the true FP64 residual, so
// an inner solve cut at ENS_INNER iterations only costs another sweep, and
// instances that are done drop out of the lists on the device.  One sweep of
// up to two inner iterations per read-back measured fastest on the
// 256-instance ensemble at inner tolerance 1e-4 (64.9k-65.4k instance-steps/s,
// against 63.9k-64.4k for three, 62.3k for one and 63.6k for four);
// WSPM_ENS_SWEEPS and WSPM_ENS_INNER override them.
static constexpr int ENS_SWEEPS = 1;
static constexpr int ENS_INNER = 2;


static void ens_mixed_solve(wspm_ctx* ctx, bool jac) {
    const wspm_richards_params& P = ctx->prm;
    while (true) {
        const int sweeps = env_int("WSPM_ENS_SWEEPS", ENS_SWEEPS), inner = env_int("WSPM_ENS_INNER", ENS_INNER);
        // fc: the lists are rebuilt by the finishes before them (the
        // assembly's finish and every residual finish leave the instances
        // still refining; the start's finish those whose inner solve runs;
        // the inner solve's last finish those refined in this sweep)
        const bool fc = fused_compact(ctx);
        for (int sw = 0; sw < sweeps; ++sw) {
            if (!fc) ens_compact(ctx, LIST_K);        // instances still refining (done clear)
            // the inner solve starts from zero: scaling and solver start in one
            // pass, without the start's SpMV
            KrylovRun<float> run(ctx, P.method, ctx->A32, jac ? ctx->dinv32 : nullptr, ctx->ref_r32, ctx->ref_d32,
                                 ctx->ref_d32, P.inner_tol, 0.0, P.maxiter, P.restart, 0, true);
            if (fc) run.last_mode = 2;
            if (P.method == WSPM_BICGSTAB)
                launch_vec(ctx, RefStartE<BicgInit<float>>{{run.v, P.inner_tol, 0.0, P.maxiter}, ctx->ref_r,
                                                           ctx->ref_r32, ctx->ref_d32, 0.0}, 0, LIST_K, fc);
            else
                launch_vec(ctx, RefStartE<CgInit<float>>{{run.v, P.inner_tol, 0.0, P.maxiter}, ctx->ref_r,
                                                         ctx->ref_r32, ctx->ref_d32, 0.0}, 0, LIST_K, fc);
            if (!fc) ens_compact(ctx, LIST_K);
            run.enqueue(std::min(inner, P.maxiter));
            if (!fc) ens_compact(ctx, LIST_K, 2);     // the instances refined in this sweep
            launch_vec(ctx, RefCorrectE{ctx->hB, ctx->ref_d32, 0.0}, 0, LIST_K);
            launch_spmv<double>(ctx, ctx->Av, RefResidE{ctx->bv, ctx->hB, ctx->ref_r, P.tol, P.atol, 0}, 0, true,
                                LIST_K, nullptr, true, fc);
        }
        poll_ens(ctx, 0);                             // hstatus[0]: instances whose refinement goes on
        if (ctx->hstatus[0] == 0) break;
    }
}

static void ens_step(wspm_ctx* ctx, double dt, wspm_step_stats* stats) {
    const wspm_richards_params& P = ctx->prm;
    const bool jac = P.precond == WSPM_PRECOND_JACOBI;
    const double* Ks = ctx->has_ks ? ctx->Ks : nullptr;
    KState* st0 = states_of(ctx, 0);
    const int ni = ctx->n_inst;
    LAUNCH(k_ens_step_begin, dim3((ni + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, st0, ni);
    HIPCHK(hipGetLastError());
    ens_compact(ctx, LIST_P);
    ++ctx->n_launch;
    HIPCHK(hipMemcpyAsync(ctx->thn, ctx->thm, ctx->n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    const bool mixed = P.precision == WSPM_MIXED;
    const int mode = mixed ? ASM_INIT_RESID : P.method == WSPM_BICGSTAB ? ASM_INIT_BICG : ASM_INIT_CG;
    const EllScope ell_scope(ctx, ctx->ell);
    int it = 0;
    while (true) {
        if (mixed) {
            ensure_mixed(ctx);
            // the assembly's residual start also copies h_m into the refinement's iterate hB
            // (mixed precision reads the FP32 inverse diagonal only)
            do_assemble(ctx, Ks, ctx->hA, ctx->Cn, nullptr, ctx->thm, ctx->thn, dt, ctx->Av, ctx->bv, nullptr,
                        ctx->A32, jac ? ctx->dinv32 : nullptr, mode, ctx->hB, P.tol, P.atol, P.maxiter, P.restart,
                        ctx->ell);
            const PicardOp pop{ctx->sp, ctx->hB, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx),
                               P.picard_mode == WSPM_PICARD_REL, 1};
            if (jacobi_on(ctx, jac ? ctx->dinv32 : nullptr)) {
                // Jacobi sweeps on the instances still refining, then the
                // Picard update of those refined (the others skip it), read
                // back once; more sweeps and updates while any instance refines
                const bool fc = fused_compact(ctx);
                const int ns = jac_sweeps(true);
                while (true) {
                    for (int k = 0; k < ns; ++k) {
                        jacobi_sweep(ctx, ctx->Av, ctx->A32, ctx->dinv32, ctx->bv, ctx->hB, P.tol, P.atol, 0, fc);
                        if (!fc) ens_compact(ctx, LIST_K);
                    }
                    launch_vec(ctx, pop, 0, LIST_P);
                    poll_ens(ctx, 0);
                    if (ctx->hstatus[3] > 0) ctx->jac_inner_off = true;   // a sweep stalled
                    if (ctx->hstatus[0] == 0) break;
                    if (!jacobi_on(ctx, ctx->dinv32)) {
                        ens_mixed_solve(ctx, jac);   // Krylov sweeps for the rest
                        launch_vec(ctx, pop, 0, LIST_P);
                        poll_ens(ctx, 0);
                        break;
                    }
                }
            } else {
                ens_mixed_solve(ctx, jac);
                launch_vec(ctx, pop, 0, LIST_P);
                poll_ens(ctx, 0);
            }
            ens_compact(ctx, LIST_P);
            ++it;
            ++ctx->n_picard;
            if (ctx->hstatus[1] == 0 || it >= P.picard_maxit) break;
            continue;
        }
        const bool graphs = ctx->use_graphs && ctx->fast_poll && !ctx->timing;
        // the constructor enqueues nothing (started): the assembly starts the solves
        KrylovRun<double> run(ctx, P.method, ctx->Av, jac ? ctx->dinv : nullptr, ctx->bv, ctx->hA, ctx->hB,
                              P.tol, P.atol, P.maxiter, P.restart, 0, true);
        const PicardOp pop{ctx->sp, ctx->hB, ctx->hA, ctx->Cn, nullptr, ctx->thm, krec_of(ctx),
                           P.picard_mode == WSPM_PICARD_REL, 1};
        const int K = std::min(std::max(1, ctx->last_iters[P.method]), P.maxiter);
        auto first = [&] {
            do_assemble(ctx, Ks, ctx->hA, ctx->Cn, nullptr, ctx->thm, ctx->thn, dt, ctx->Av, ctx->bv,
                        jac ? ctx->dinv : nullptr, nullptr, nullptr, mode, ctx->hB, P.tol, P.atol, P.maxiter,
                        P.restart, ctx->ell);
            run.enqueue(K);
            launch_vec(ctx, pop, 0, LIST_P);
            ens_status_publish(ctx, 0);
        };
        if (graphs) {
            const hipGraphExec_t ex = step_graph(ctx, GRAPH_ENS_PICARD, dt, K, first);
            run.enq = K;
            ++ctx->n_launch;
            HIPCHK(hipGraphLaunch(ex, ctx->stream));
            wait_published(ctx);
        } else {
            first();
            if (ctx->fast_poll) wait_published(ctx);
            else {
                HIPCHK(hipMemcpyAsync(ctx->hstatus, ctx->ens_status, 4 * sizeof(int), hipMemcpyDeviceToHost,
                                      ctx->stream));
                HIPCHK(hipStreamSynchronize(ctx->stream));
            }
        }
        while (ctx->hstatus[0] > 0 && run.enq < P.maxiter) {
            if (graphs && run.enq + 2 <= P.maxiter) {
                // two more Krylov iterations, the Picard error, the counts
                const int enq0 = run.enq;
                // the compactions inside the chunk depend on the iteration count
                const hipGraphExec_t ex = step_graph(ctx, GRAPH_ENS_CHUNK, dt,
                                                     2 | (enq0 % ctx->ens_compact_every) << 8, [&] {
                    run.enqueue(2);
                    launch_vec(ctx, pop, 0, LIST_P);
                    ens_status_publish(ctx, 0);
                });
                run.enq = enq0 + 2;
                ++ctx->n_launch;
                HIPCHK(hipGraphLaunch(ex, ctx->stream));
                wait_published(ctx);
                continue;
            }
            run.enqueue(2);
            launch_vec(ctx, pop, 0, LIST_P);
            poll_ens(ctx, 0);
        }
        ens_compact(ctx, LIST_P);
        ctx->last_iters[P.method] = std::max(1, ctx->hstatus[2]);
        ++it;
        ++ctx->n_picard;
        if (ctx->hstatus[1] == 0 || it >= P.picard_maxit) break;
    }
    if (stats) {
        int* ip = ctx->ens_out;
        int* ik = ip + ni;
        int* ic = ik + ni;
        double* er = reinterpret_cast<double*>(ctx->ens_out + 3 * (size_t)ni + (ni & 1));
        LAUNCH(k_ens_collect, dim3((ni + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, st0, ni,
                           ip, ik, ic, er);
        HIPCHK(hipGetLastError());
        std::vector<int> hi(3 * (size_t)ni);
        std::vector<double> he(ni);
        HIPCHK(hipMemcpyAsync(hi.data(), ip, 3 * (size_t)ni * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(he.data(), er, ni * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        int pmax = 0, conv = 1;
        int64_t ksum = 0;
        double emax = 0.0;
        for (int i = 0; i < ni; ++i) {
            pmax = std::max(pmax, hi[i]);
            ksum += hi[ni + i];
            conv &= hi[2 * ni + i];
            emax = std::isnan(he[i]) || std::isnan(emax) ? NAN : std::max(emax, he[i]);
        }
        stats->picard_iters = pmax;
        stats->krylov_iters = (int32_t)std::min<int64_t>(ksum, INT32_MAX);
        stats->converged = conv;
        stats->reserved = 0;
        stats->err = emax;
    }
}

int32_t wspm_ens_step_stats(wspm_ctx* ctx, int32_t* picard_iters, int32_t* krylov_iters, int32_t* converged,
                            double* err) {
    if (!ctx) return WSPM_EINVAL;
    if (!ctx->tab) return fail(ctx, WSPM_EINVAL, "not an ensemble context");
    API_BEGIN
    const int ni = ctx->n_inst;
    int* ip = ctx->ens_out;
    int* ik = ip + ni;
    int* ic = ik + ni;
    double* er = reinterpret_cast<double*>(ctx->ens_out + 3 * (size_t)ni + (ni & 1));
    LAUNCH(k_ens_collect, dim3((ni + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream,
                       states_of(ctx, 0), ni, ip, ik, ic, er);
    HIPCHK(hipGetLastError());
    if (picard_iters) HIPCHK(hipMemcpyAsync(picard_iters, ip, ni * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (krylov_iters) HIPCHK(hipMemcpyAsync(krylov_iters, ik, ni * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (converged) HIPCHK(hipMemcpyAsync(converged, ic, ni * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (err) HIPCHK(hipMemcpyAsync(err, er, ni * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

int32_t wspm_ens_shape(const wspm_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return WSPM_EINVAL;
    if (!ctx->tab) return fail(const_cast<wspm_ctx*>(ctx), WSPM_EINVAL, "not an ensemble context");
    out[0] = ell_shape_ok(ctx) ? 1 : 0;
    out[1] = ell_shape_ok(ctx) ? ctx->max_row : ctx->lanes;
    out[2] = ctx->ens_grid;
    out[3] = ENS_ROWS;
    out[4] = ctx->n_tab;
    out[5] = ctx->ens_grid_cands;
    out[6] = (int32_t)std::lround(ctx->ens_tune_us * 1000.0f);
    out[7] = ctx->jac_inner_off ? 1 : 0;
    return WSPM_OK;
}

// ---------------------------------------------------------------------------
// Diagnostics
// ---------------------------------------------------------------------------
static int diag_check(wspm_ctx* ctx) {
    if (ctx->tab) return fail(ctx, WSPM_EINVAL, "not available on an ensemble context");
    if (ctx->dim != 2) return fail(ctx, WSPM_EINVAL, "2D only, as the reference");
    if (ctx->E < 1) return fail(ctx, WSPM_EINVAL, "empty mesh");
    return WSPM_OK;
}

static DiagArgs diag_args(wspm_ctx* ctx) {
    DiagArgs a{};
    a.E = ctx->E;
    a.conn4 = ctx->conn4;
    a.nrec = nrec_of(ctx);
    a.partials = ctx->partials;
    a.prow = ctx->prow;
    a.grid = std::max(1, std::min((ctx->E + BLOCK - 1) / BLOCK, ctx->prow));
    return a;
}

int32_t wspm_domain_integral(wspm_ctx* ctx, const double* f, double* out) {
    if (!ctx) return WSPM_EINVAL;
    if (!f || !out) return fail(ctx, WSPM_EINVAL, "null argument");
    if (int rc = diag_check(ctx)) return rc;
    API_BEGIN
    if (!ctx->diag_out) ctx->diag_out = dalloc<double>(ctx, 4);
    HIPCHK(launch_domain_integral(diag_args(ctx), f, ctx->diag_out, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, ctx->diag_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

int32_t wspm_element_size(wspm_ctx* ctx, double* size) {
    if (!ctx) return WSPM_EINVAL;
    if (!size) return fail(ctx, WSPM_EINVAL, "null argument");
    if (int rc = diag_check(ctx)) return rc;
    API_BEGIN
    HIPCHK(launch_element_size(diag_args(ctx), size, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

int32_t wspm_peclet_courant(wspm_ctx* ctx, const double* qx, const double* qz, const double* theta,
                            const double* dxx, const double* dzz, double dt, int32_t all_elements,
                            double* pe_elem, double* cr_elem, double* out) {
    if (!ctx) return WSPM_EINVAL;
    if (!qx || !qz || !theta || !dxx || !dzz || !out) return fail(ctx, WSPM_EINVAL, "null argument");
    if (int rc = diag_check(ctx)) return rc;
    API_BEGIN
    if (!ctx->diag_out) ctx->diag_out = dalloc<double>(ctx, 4);
    const PeCrArgs p{qx, qz, theta, dxx, dzz, dt, pe_elem, cr_elem, all_elements != 0};
    HIPCHK(launch_peclet_courant(diag_args(ctx), p, ctx->diag_out, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, ctx->diag_out, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WSPM_OK;
    API_END
}

}  // extern "C"
