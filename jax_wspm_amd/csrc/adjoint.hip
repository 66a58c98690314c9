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
// log(a), d = a' / a; exp(a), d = e^a a'
__device__ inline D3 dlog(const D3& a) {
    D3 r{log(a.v), {}};
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] / a.v;
    return r;
}
__device__ inline D3 dexp(const D3& a) {
    const double e = exp(a.v);
    D3 r{e, {}};
    for (int k = 0; k < 3; ++k) r.d[k] = e * a.d[k];
    return r;
}
// log1p(a), d = a' / (1 + a); expm1(a), d = e^a a'
__device__ inline D3 dlog1p(const D3& a) {
    D3 r{log1p(a.v), {}};
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] / (1.0 + a.v);
    return r;
}
__device__ inline D3 dexpm1(const D3& a) {
    const double e = expm1(a.v);
    D3 r{e, {}};
    for (int k = 0; k < 3; ++k) r.d[k] = (e + 1.0) * a.d[k];
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
    // the forward's steps: s = (a|h|)^n, L = log(1+s) and Q = log(1+1/s)
    // from one log1p, Se = exp(-m L), 1 - (1 - Se^(1/m))^m = -expm1(-m Q)
    const D3 ah = alpha * (-h);
    const D3 nl = n * dlog(ah);
    const D3 ahn = dexp(nl);
    const bool big = ahn.v >= 1.0;
    const D3 lp = dlog1p(big ? one / ahn : ahn);
    const D3 L = big ? nl + lp : lp;
    const D3 Q = big ? lp : lp - nl;
    const D3 se = dexp(-(m * L));
    th = dconst(ts - tr) * se + dconst(tr);
    const D3 inner = -dexpm1(-(m * Q));
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

template <class Args>
__device__ inline void load_element(const Args& a, int e, int (&nd)[3], double (&px)[3], double (&py)[3]) {
    const int4 cn = a.conn4[e];
    nd[0] = cn.x;
    nd[1] = cn.y;
    nd[2] = cn.z;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double2 r = a.nrec.xy(nd[c]);
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
            double jac[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) jac[c] = det * (kbar * gg[c] + ks * a.qc.ck[c] * dkr[c] * sa);
            add_slots<3, MAXR>(v, (uint32_t)ic.y, jac);
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

// ---------------------------------------------------------------------------
// Transport adjoint (nitrogen species chain)
// ---------------------------------------------------------------------------

// Lumped species mass at step k+1: per incidence (e, a) of node i,
// sum_c m_ac = det (-sum_{c,u} U[a][c][u] theta_u - rho_kd sum_c T[a][c])
// (the forward's m, k_assemble_transport).
__global__ __launch_bounds__(BLOCK) void k_tr_adj_mass(TrAdjArgs a, const double* th, double* mth, double* mT) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const TrQuad& Q = a.tq;
    double s0 = 0.0, s1 = 0.0;
    for (int q = a.inc_ptr[i]; q < a.inc_ptr[i + 1]; ++q) {
        const int2 ic = a.inc[q];
        const int la = ic.x & 3;
        int nd[3];
        double px[3], py[3], Gx[3], Gz[3];
        load_element(a, ic.x >> 2, nd, px, py);
        const double det = adj_geometry(px, py, Gx, Gz);
        double u = 0.0, t = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            t += Q.T[la][c];
#pragma unroll
            for (int v = 0; v < 3; ++v) u += Q.U[la][c][v] * th[nd[v]];
        }
        s0 += -u * det;
        s1 += -t * det;
    }
    mth[i] = s0;
    mT[i] = s1;
}

// rhs_i = (mth_i + rho_kd_s mT_i) lam_next_i / dt
//         - mu_s sum_{(e,b) at i} det sum_{a,c} lam_up_a U[a][c][b] theta_c
// The second term is -(dR_{s+1}/dc_s)^T lam_up: R_{s+1} carries
// +mu_s det sum_{c,u} U[a][c][u] theta_c c_s,u (minus the source of b).
__global__ __launch_bounds__(BLOCK) void k_tr_adj_rhs(TrAdjArgs a, int sp, const double* mth, const double* mT,
                                                      const double* lam_next, const double* lam_up, double* rhs) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    double r = 0.0;
    if (lam_next) r = a.inv_dt * (mth[i] + (sp == 0 ? a.rho_kd : 0.0) * mT[i]) * lam_next[i];
    if (lam_up) {
        const TrQuad& Q = a.tq;
        double s = 0.0;
        for (int q = a.inc_ptr[i]; q < a.inc_ptr[i + 1]; ++q) {
            const int2 ic = a.inc[q];
            const int lb = ic.x & 3;
            int nd[3];
            double px[3], py[3], Gx[3], Gz[3];
            load_element(a, ic.x >> 2, nd, px, py);
            const double det = adj_geometry(px, py, Gx, Gz);
            double t = 0.0;
#pragma unroll
            for (int aa = 0; aa < 3; ++aa)
#pragma unroll
                for (int c = 0; c < 3; ++c) t += lam_up[nd[aa]] * Q.U[aa][c][lb] * a.th[nd[c]];
            s += det * t;
        }
        r += -a.mu[sp] * s;
    }
    rhs[i] = r;
}

__global__ void k_tr_adj_obs(int n_obs, const int32_t* obs, const double* dldy, double* rhs) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int j = 0; j < n_obs; ++j) rhs[obs[j]] += dldy[j];   // serial: repeated nodes add up
}

// Field gradients at node j: g_f[j] = -sum_s sum_{(e,v) at j} d(phi_s,e)/d f_v,
// phi_s,e = sum_a mu_a R_a^e.  Per element (G, det as the forward):
//   d/dDxx_v = -det ck_v (mu.Gx)(Gx.c),  d/dDzz_v = -det ck_v (mu.Gz)(Gz.c),
//   d/dDxz_v = -det ck_v ((mu.Gx)(Gz.c) + (mu.Gz)(Gx.c)),
//   d/dqx_v  = -det (sum_a mu_a T[a][v]) (Gx.c),  d/dqz_v likewise with Gz,
//   d/dtheta_v = det ( mu_s sum_{a,c} mu_a c_c U[a][c][v]
//                     - sum_a mu_a (c_a - cp_a) sum_c U[a][c][v] / dt
//                     + mu_{s-1} sum_{a,u} mu_a U[a][v][u] c_{s-1,u} ).
__global__ __launch_bounds__(BLOCK) void k_tr_adj_fields(TrAdjArgs a, double* gf) {
    const int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= a.n) return;
    const TrQuad& Q = a.tq;
    const size_t n = (size_t)a.n;
    double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = a.inc_ptr[j]; q < a.inc_ptr[j + 1]; ++q) {
        const int2 ic = a.inc[q];
        const int v = ic.x & 3;
        int nd[3];
        double px[3], py[3], Gx[3], Gz[3];
        load_element(a, ic.x >> 2, nd, px, py);
        const double det = adj_geometry(px, py, Gx, Gz);
        for (int sp = 0; sp < 3; ++sp) {
            double m[3], c[3], cp[3], cs[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[k] = a.lam[sp * n + nd[k]];
                c[k] = a.c[sp * n + nd[k]];
                cp[k] = a.cp[sp * n + nd[k]];
                cs[k] = sp > 0 ? a.c[(sp - 1) * n + nd[k]] : 0.0;
            }
            double mGx = 0.0, mGz = 0.0, cGx = 0.0, cGz = 0.0, mT = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                mGx += m[k] * Gx[k];
                mGz += m[k] * Gz[k];
                cGx += c[k] * Gx[k];
                cGz += c[k] * Gz[k];
                mT += m[k] * Q.T[k][v];
            }
            double rx = 0.0, dm = 0.0, sr = 0.0;
#pragma unroll
            for (int aa = 0; aa < 3; ++aa) {
                double uc = 0.0, us = 0.0, u1 = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    uc += Q.U[aa][k][v] * c[k];
                    us += Q.U[aa][v][k] * cs[k];
                    u1 += Q.U[aa][k][v];
                }
                rx += m[aa] * uc;
                dm += m[aa] * (c[aa] - cp[aa]) * u1;
                sr += m[aa] * us;
            }
            const double ck = Q.ck[v];
            g[0] -= det * (a.mu[sp] * rx - a.inv_dt * dm + (sp > 0 ? a.mu[sp - 1] : 0.0) * sr);
            g[1] -= -det * mT * cGx;
            g[2] -= -det * mT * cGz;
            g[3] -= -det * ck * mGx * cGx;
            g[4] -= -det * ck * (mGx * cGz + mGz * cGx);
            g[5] -= -det * ck * mGz * cGz;
        }
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) gf[f * n + j] = g[f];
}

// Node gradient of h averaged over the incident elements (b/c form), the
// mean incident Ks, and the incidence count (as k_transport_fields).
__device__ inline void tr_node_grad(const TrAdjArgs& a, int i, double& gx, double& gz, double& ksn, double& cnt) {
    double sx = 0.0, sz = 0.0, ks = 0.0;
    const int q0 = a.inc_ptr[i], q1 = a.inc_ptr[i + 1];
    for (int q = q0; q < q1; ++q) {
        const int e = a.inc[q].x >> 2;
        ks += a.ks_e[e];
        int nd[3];
        double x[3], y[3];
        load_element(a, e, nd, x, y);
        const double b1 = y[1] - y[2], b2 = y[2] - y[0], b3 = y[0] - y[1];
        const double c1 = x[2] - x[1], c2 = x[0] - x[2], c3 = x[1] - x[0];
        const double aire = fabs(c3 * b2 - c2 * b3) / 2;
        sx += (b1 * a.h[nd[0]] + b2 * a.h[nd[1]] + b3 * a.h[nd[2]]) / (2 * aire);
        sz += (c1 * a.h[nd[0]] + c2 * a.h[nd[1]] + c3 * a.h[nd[2]]) / (2 * aire);
    }
    cnt = (double)(q1 - q0);
    gx = sx / cnt;
    gz = sz / cnt;
    ksn = ks / cnt;
}

// Per node: carry g_f through D(q, theta) (dual numbers over (qx, qz, theta)
// on the forward's expressions), then q = -K grad h, K = Kr(h) Ks_node and
// theta(h) (soil_dual over (h, alpha, n)).  Writes the local part of dL/dh,
// the alpha / n terms, and for the neighbour pass w = (G_gx / cnt, G_gz / cnt,
// G_K Kr / cnt).
__global__ __launch_bounds__(BLOCK) void k_tr_adj_chain(TrAdjArgs a, const double* gf, double* w, double* dldh,
                                                        double* gn) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const size_t n = (size_t)a.n;
    double gx, gz, ksn, cnt;
    tr_node_grad(a, i, gx, gz, ksn, cnt);
    const D3 qx{a.qx[i], {1.0, 0.0, 0.0}}, qz{a.qz[i], {0.0, 1.0, 0.0}}, th{a.th[i], {0.0, 0.0, 1.0}};
    const D3 aq = dsqrt(qx * qx + qz * qz);
    const double ts = a.sp.p[1];
    const D3 tau = dpow(th, dconst(7.0 / 3.0)) / dconst(ts * ts);
    const D3 DLd = dconst(a.DL), DTd = dconst(a.DT), Dmd = dconst(a.Dm);
    const D3 dxx = (DLd * qx * qx + DTd * qz * qz) / aq + th * Dmd * tau;
    const D3 dzz = (DLd * qz * qz + DTd * qx * qx) / aq + th * Dmd * tau;
    const D3 dxz = dconst(a.DL - a.DT) * qx * qz / aq;
    const double gth = gf[i], gqx = gf[n + i], gqz = gf[2 * n + i];
    const double gxx = gf[3 * n + i], gxz = gf[4 * n + i], gzz = gf[5 * n + i];
    const double Gqx = gqx + gxx * dxx.d[0] + gxz * dxz.d[0] + gzz * dzz.d[0];
    const double Gqz = gqz + gxx * dxx.d[1] + gxz * dxz.d[1] + gzz * dzz.d[1];
    const double Gth = gth + gxx * dxx.d[2] + gxz * dxz.d[2] + gzz * dzz.d[2];
    D3 sth, skr;
    soil_dual(a.sp, a.h[i], sth, skr);
    double C, Kr, thv;
    soil_eval(a.sp, a.h[i], C, Kr, thv);
    const double K = Kr * ksn;
    const double GK = -Gqx * gx - Gqz * (gz + 1.0);
    dldh[i] = Gth * sth.d[0] + GK * ksn * skr.d[0];
    gn[i] = Gth * sth.d[1] + GK * ksn * skr.d[1];
    gn[n + i] = Gth * sth.d[2] + GK * ksn * skr.d[2];
    w[i] = -K * Gqx / cnt;
    w[n + i] = -K * Gqz / cnt;
    w[2 * n + i] = GK * Kr / cnt;
}

// dldh_j += sum_{(e,b) at j} sum_{i in e} (w_gx,i b_b + w_gz,i c_b) / (2 |area_e|)
__global__ __launch_bounds__(BLOCK) void k_tr_adj_dh(TrAdjArgs a, const double* w, double* dldh) {
    const int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= a.n) return;
    const size_t n = (size_t)a.n;
    double s = 0.0;
    for (int q = a.inc_ptr[j]; q < a.inc_ptr[j + 1]; ++q) {
        const int2 ic = a.inc[q];
        const int lb = ic.x & 3;
        int nd[3];
        double x[3], y[3];
        load_element(a, ic.x >> 2, nd, x, y);
        const double b[3] = {y[1] - y[2], y[2] - y[0], y[0] - y[1]};
        const double c[3] = {x[2] - x[1], x[0] - x[2], x[1] - x[0]};
        const double aire = fabs(c[2] * b[1] - c[1] * b[2]) / 2;
        double wx = 0.0, wz = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            wx += w[nd[k]];
            wz += w[n + nd[k]];
        }
        s += (wx * b[lb] + wz * c[lb]) / (2 * aire);
    }
    dldh[j] += s;
}

// grad_logks[e] += Ks_e sum_{i in e} G_K,i Kr_i / cnt_i
__global__ __launch_bounds__(BLOCK) void k_tr_adj_ks(TrAdjArgs a, const double* w, double* grad_logks) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= a.E) return;
    const int4 cn = a.conn4[e];
    const double* wk = w + 2 * (size_t)a.n;
    grad_logks[e] += a.ks_e[e] * (wk[cn.x] + wk[cn.y] + wk[cn.z]);
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

hipError_t launch_transpose_values(const double* J, const int32_t* tpos, double* JT, int64_t nnz, hipStream_t s) {
    const unsigned g = (unsigned)std::min<int64_t>((nnz + BLOCK - 1) / BLOCK, 8192);
    hipLaunchKernelGGL(k_adj_transpose, dim3(g), dim3(BLOCK), 0, s, J, tpos, JT, nnz);
    return hipGetLastError();
}

hipError_t launch_tr_adj_mass(const TrAdjArgs& a, const double* th_next, double* mth, double* mT, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_adj_mass, grid_of(a.n), dim3(BLOCK), 0, s, a, th_next, mth, mT);
    return hipGetLastError();
}

hipError_t launch_tr_adj_rhs(const TrAdjArgs& a, int species, const double* mth, const double* mT,
                             const double* lam_next, const double* lam_up, double* rhs, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_adj_rhs, grid_of(a.n), dim3(BLOCK), 0, s, a, species, mth, mT, lam_next, lam_up, rhs);
    return hipGetLastError();
}

hipError_t launch_tr_adj_obs(int n_obs, const int32_t* obs, const double* dldy, double* rhs, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_adj_obs, dim3(1), dim3(64), 0, s, n_obs, obs, dldy, rhs);
    return hipGetLastError();
}

hipError_t launch_tr_adj_fields(const TrAdjArgs& a, double* gf, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_adj_fields, grid_of(a.n), dim3(BLOCK), 0, s, a, gf);
    return hipGetLastError();
}

hipError_t launch_tr_adj_chain(const TrAdjArgs& a, const double* gf, double* w, double* dldh, double* gn,
                               double* grad_logks, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_adj_chain, grid_of(a.n), dim3(BLOCK), 0, s, a, gf, w, dldh, gn);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tr_adj_dh, grid_of(a.n), dim3(BLOCK), 0, s, a, w, dldh);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tr_adj_ks, grid_of(a.E), dim3(BLOCK), 0, s, a, w, grad_logks);
    return hipGetLastError();
}

}  // namespace wspm
