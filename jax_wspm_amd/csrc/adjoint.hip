// Discrete adjoint kernels for gfx950 (see adjoint.h).
//
// Soil derivatives come from forward-mode dual numbers pushed through the
// same van Genuchten-Mualem expression the forward kernels evaluate
// (soil_eval in wspm_device.h), so they are the derivatives of the discrete
// forward model, not of a separately derived formula.  Every kernel is a
// gather (row, node or element owns its output); sums have fixed orders.
#include "adjoint.h"

namespace wspm {

// value + derivatives with respect to (h, alpha, n)
struct D3 {
    double v, d[3];
};
__device__ inline D3 dconst(double v) { return D3{v, {0.0, 0.0, 0.0}}; }
__device__ inline D3 operator+(const D3& a, const D3& b) {
    return D3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ inline D3 operator-(const D3& a, const D3& b) {
    return D3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ inline D3 operator*(const D3& a, const D3& b) {
    D3 r{a.v * b.v, {}};
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
__device__ inline D3 operator/(const D3& a, const D3& b) {
    const double q = a.v / b.v;
    D3 r{q, {}};
    for (int k = 0; k < 3; ++k) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
    return r;
}
__device__ inline D3 operator-(const D3& a) { return D3{-a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
// a^b for a > 0: d(a^b) = a^b (b' ln a + b a'/a)
__device__ inline D3 dpow(const D3& a, const D3& b) {
    const double p = pow(a.v, b.v);
    const double la = log(a.v);
    D3 r{p, {}};
    for (int k = 0; k < 3; ++k) r.d[k] = p * (b.d[k] * la + b.v * a.d[k] / a.v);
    return r;
}
__device__ inline D3 dsqrt(const D3& a) {
    const double s = sqrt(a.v);
    D3 r{s, {}};
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] / (2.0 * s);
    return r;
}

// theta and Kr (Ks = 1) of soil_eval's van Genuchten branch, with m = 1 - 1/n.
__device__ inline void soil_dual(const SoilParams& sp, double hv, D3& th, D3& kr) {
    const double ts = sp.p[1], tr = sp.p[2];
    if (!(hv < 0.0)) {
        th = dconst(ts);
        kr = dconst(1.0);
        return;
    }
    const D3 h{hv, {1.0, 0.0, 0.0}};
    const D3 alpha{sp.p[0], {0.0, 1.0, 0.0}};
    const D3 n{sp.p[3], {0.0, 0.0, 1.0}};
    const D3 one = dconst(1.0);
    const D3 m = one - one / n;
    const D3 ah = alpha * (-h);
    const D3 ahn = dpow(ah, n);
    const D3 base = one + ahn;
    const D3 bm = dpow(base, -m);
    th = dconst(ts - tr) * bm + dconst(tr);
    const D3 se = (th - dconst(tr)) / dconst(ts - tr);
    const D3 inner = one - dpow(one - dpow(se, one / m), m);
    kr = dsqrt(se) * (inner * inner);
}

void adj_tie_m(SoilParams& sp) { sp.p[4] = 1.0 - 1.0 / sp.p[3]; }

__global__ __launch_bounds__(BLOCK) void k_adj_nodal(AdjArgs a) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    D3 th, kr;
    soil_dual(a.sp, a.h[i], th, kr);
    a.nd.kr[i] = kr.v;
    a.nd.dkr_dh[i] = kr.d[0];
    a.nd.dkr_da[i] = kr.d[1];
    a.nd.dkr_dn[i] = kr.d[2];
    a.nd.dth_dh[i] = th.d[0];
    a.nd.dth_da[i] = th.d[1];
    a.nd.dth_dn[i] = th.d[2];
    soil_dual(a.sp, a.hp[i], th, kr);
    a.nd.dthp_da[i] = th.d[1];
    a.nd.dthp_dn[i] = th.d[2];
}

// P1 triangle geometry (as the forward kernels): G[d][k], det J.
__device__ inline double adj_geometry(const double (&px)[3], const double (&py)[3], double (&Gx)[3], double (&Gz)[3]) {
    const double J00 = px[1] - px[0], J01 = py[1] - py[0];
    const double J10 = px[2] - px[0], J11 = py[2] - py[0];
    const double det = J00 * J11 - J01 * J10;
    const double id = 1.0 / det;
    Gx[1] = J11 * id;
    Gz[1] = -J10 * id;
    Gx[2] = -J01 * id;
    Gz[2] = J00 * id;
    Gx[0] = -(Gx[1] + Gx[2]);
    Gz[0] = -(Gz[1] + Gz[2]);
    return det;
}

__device__ inline void load_element(const AdjArgs& a, int e, int (&nd)[3], double (&px)[3], double (&py)[3]) {
    const int4 cn = a.conn4[e];
    nd[0] = cn.x;
    nd[1] = cn.y;
    nd[2] = cn.z;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double4 r = a.nrec[nd[c]];
        px[c] = r.x;
        py[c] = r.y;
    }
}

// Global lumped mass ml_i = sum over incidences of det J ml[a].
__global__ __launch_bounds__(BLOCK) void k_adj_mass(AdjArgs a, double* ml) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    double m = 0.0;
    for (int q = a.inc_ptr[i]; q < a.inc_ptr[i + 1]; ++q) {
        const int2 ic = a.inc[q];
        int nd[3];
        double px[3], py[3], Gx[3], Gz[3];
        load_element(a, ic.x >> 2, nd, px, py);
        const double det = adj_geometry(px, py, Gx, Gz);
        m += det * a.qc.ml[ic.x & 3];
    }
    ml[i] = m;
}

template <int MAXR>
__global__ __launch_bounds__(BLOCK) void k_adj_jacobian(AdjArgs a, double* J, double* dinv) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const int rs = a.row_ptr[i];
    const int rl = a.row_ptr[i + 1] - rs;
    const int dg = a.diag_pos[i] - rs;
    double v[MAXR];
#pragma unroll
    for (int s = 0; s < MAXR; ++s) v[s] = 0.0;
    if (a.bc_flag != nullptr && a.bc_flag[i]) {
#pragma unroll
        for (int s = 0; s < MAXR; ++s) v[s] = (s == dg) ? 1.0 : 0.0;
    } else {
        double msum = 0.0;
        for (int q = a.inc_ptr[i]; q < a.inc_ptr[i + 1]; ++q) {
            const int2 ic = a.inc[q];
            const int e = ic.x >> 2, la = ic.x & 3;
            int nd[3];
            double px[3], py[3], Gx[3], Gz[3];
            load_element(a, e, nd, px, py);
            const double det = adj_geometry(px, py, Gx, Gz);
            double hv[3], kr[3], dkr[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                hv[c] = a.h[nd[c]];
                kr[c] = a.nd.kr[nd[c]];
                dkr[c] = a.nd.dkr_dh[nd[c]];
            }
            const double ks = a.ks_e[e];
            double kbar = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) kbar += a.qc.ck[c] * kr[c];
            kbar *= ks;
            double gg[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) gg[c] = Gx[la] * Gx[c] + Gz[la] * Gz[c];
            const double sa = gg[0] * hv[0] + gg[1] * hv[1] + gg[2] * hv[2] + Gz[la];
            const uint32_t w = (uint32_t)ic.y;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double jac = det * (kbar * gg[c] + ks * a.qc.ck[c] * dkr[c] * sa);
                const uint32_t sl = (w >> (8 * c)) & 0xffu;
#pragma unroll
                for (int s2 = 0; s2 < MAXR; ++s2) v[s2] += (sl == (uint32_t)s2) ? jac : 0.0;
            }
            msum += det * a.qc.ml[la];
        }
        const double md = a.inv_dt * msum * a.nd.dth_dh[i];
#pragma unroll
        for (int s = 0; s < MAXR; ++s) v[s] += (s == dg) ? md : 0.0;
    }
    double dgv = 0.0;
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
        if (s < rl) J[rs + s] = v[s];
        if (s == dg) dgv = v[s];
    }
    dinv[i] = 1.0 / (dgv != 0.0 ? dgv : 1.0);   // J^T has the same diagonal
}

__global__ __launch_bounds__(BLOCK) void k_adj_transpose(const double* __restrict__ J, const int32_t* __restrict__ tpos,
                                                         double* JT, int64_t nnz) {
    for (int64_t k = blockIdx.x * (int64_t)BLOCK + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * BLOCK)
        JT[k] = J[tpos[k]];
}

__global__ __launch_bounds__(BLOCK) void k_adj_rhs(AdjArgs a, const double* dldh, const double* lam_next,
                                                   double* rhs) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    double r = dldh ? -dldh[i] : 0.0;
    const bool dir = a.bc_flag != nullptr && a.bc_flag[i];
    if (lam_next && !dir) r += a.inv_dt * a.ml[i] * a.nd.dth_dh[i] * lam_next[i];
    rhs[i] = r;
}

// Element terms: t_e = det sum_a lam_a s_a over non-Dirichlet vertices;
// d/dlogKs_e = Kbar_e t_e; d/d(alpha, n) through Kbar: Ks_e sum_c ck_c dKr_c t_e.
__global__ __launch_bounds__(BLOCK) void k_adj_elem(AdjArgs a, const double* lam, double* grad_logks, double* ge) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= a.E) return;
    int nd[3];
    double px[3], py[3], Gx[3], Gz[3];
    load_element(a, e, nd, px, py);
    const double det = adj_geometry(px, py, Gx, Gz);
    double hv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) hv[c] = a.h[nd[c]];
    double t = 0.0;
#pragma unroll
    for (int la = 0; la < 3; ++la) {
        const bool dir = a.bc_flag != nullptr && a.bc_flag[nd[la]];
        if (dir) continue;
        double sa = Gz[la];
#pragma unroll
        for (int c = 0; c < 3; ++c) sa += (Gx[la] * Gx[c] + Gz[la] * Gz[c]) * hv[c];
        t += lam[nd[la]] * sa;
    }
    t *= det;
    const double ks = a.ks_e[e];
    double kbar = 0.0, ka = 0.0, kn = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        kbar += a.qc.ck[c] * a.nd.kr[nd[c]];
        ka += a.qc.ck[c] * a.nd.dkr_da[nd[c]];
        kn += a.qc.ck[c] * a.nd.dkr_dn[nd[c]];
    }
    grad_logks[e] += ks * kbar * t;
    ge[e] = ks * ka * t;
    ge[a.E + e] = ks * kn * t;
}

// Node terms of the storage: ml_i lam_i (dtheta(h_k) - dtheta(h_{k-1})) / dt.
__global__ __launch_bounds__(BLOCK) void k_adj_node(AdjArgs a, const double* lam, double* gn) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const bool dir = a.bc_flag != nullptr && a.bc_flag[i];
    const double w = dir ? 0.0 : a.inv_dt * a.ml[i] * lam[i];
    gn[i] = w * (a.nd.dth_da[i] - a.nd.dthp_da[i]);
    gn[a.n + i] = w * (a.nd.dth_dn[i] - a.nd.dthp_dn[i]);
}

// One thread: fixed order over the observations.
__global__ void k_adj_obs(AdjArgs a, int n_obs, const int32_t* obs, const double* dldy, double* dldh, double* gan) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double sa = 0.0, sn = 0.0;
    for (int j = 0; j < n_obs; ++j) {
        const int i = obs[j];
        dldh[i] += dldy[j] * a.nd.dth_dh[i];
        sa += dldy[j] * a.nd.dth_da[i];
        sn += dldy[j] * a.nd.dth_dn[i];
    }
    gan[0] += sa;
    gan[1] += sn;
}

static dim3 grid_of(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

hipError_t launch_adj_nodal(const AdjArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_nodal, grid_of(a.n), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj_mass(const AdjArgs& a, double* ml, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_mass, grid_of(a.n), dim3(BLOCK), 0, s, a, ml);
    return hipGetLastError();
}

hipError_t launch_adj_jacobian(const AdjArgs& a, int max_row, double* J, double* JT, double* dinv,
                               const int32_t* tpos, int64_t nnz, hipStream_t s) {
    if (max_row <= 8) hipLaunchKernelGGL(k_adj_jacobian<8>, grid_of(a.n), dim3(BLOCK), 0, s, a, J, dinv);
    else if (max_row <= 16) hipLaunchKernelGGL(k_adj_jacobian<16>, grid_of(a.n), dim3(BLOCK), 0, s, a, J, dinv);
    else return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned g = (unsigned)std::min<int64_t>((nnz + BLOCK - 1) / BLOCK, 8192);
    hipLaunchKernelGGL(k_adj_transpose, dim3(g), dim3(BLOCK), 0, s, J, tpos, JT, nnz);
    return hipGetLastError();
}

hipError_t launch_adj_rhs(const AdjArgs& a, const double* dldh, const double* lam_next, double* rhs, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_rhs, grid_of(a.n), dim3(BLOCK), 0, s, a, dldh, lam_next, rhs);
    return hipGetLastError();
}

hipError_t launch_adj_params(const AdjArgs& a, const double* lam, double* grad_logks, double* ge, double* gn,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_adj_elem, grid_of(a.E), dim3(BLOCK), 0, s, a, lam, grad_logks, ge);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_node, grid_of(a.n), dim3(BLOCK), 0, s, a, lam, gn);
    return hipGetLastError();
}

hipError_t launch_adj_obs(const AdjArgs& a, int n_obs, const int32_t* obs, const double* dldy, double* dldh,
                          double* gan, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_obs, dim3(1), dim3(64), 0, s, a, n_obs, obs, dldy, dldh, gan);
    return hipGetLastError();
}

}  // namespace wspm
