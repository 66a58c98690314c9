// Solute / nitrogen transport kernels for gfx950 (see transport.h).
//
// All three kernels are gathers over the mesh context's fixed structures
// (CSR pattern, node -> (element, vertex) incidences, packed connectivity and
// node records); none scatters, so every sum has a fixed order and results
// are repeatable.  They are HBM/latency-bound integer + FP64 work with no
// matrix-shaped contraction.
#include "transport.h"

#include "assemble.h"

namespace wspm {

void compute_tr_quad(TrQuad& q) {
    // the reference's literal rule (src/numerics/gauss.py:31-37)
    const double pts[3][2] = {{0.166666666666667, 0.166666666666667},
                              {0.166666666666667, 0.666666666666667},
                              {0.666666666666667, 0.166666666666667}};
    const double w = 0.166666666666667;
    for (int a = 0; a < 3; ++a) {
        q.ck[a] = 0.0;
        for (int c = 0; c < 3; ++c) {
            q.T[a][c] = 0.0;
            for (int v = 0; v < 3; ++v) q.U[a][c][v] = 0.0;
        }
    }
    for (int p = 0; p < 3; ++p) {
        const double N[3] = {1.0 - pts[p][0] - pts[p][1], pts[p][0], pts[p][1]};
        for (int a = 0; a < 3; ++a) {
            q.ck[a] += w * N[a];
            for (int c = 0; c < 3; ++c) {
                q.T[a][c] += w * N[a] * N[c];
                for (int v = 0; v < 3; ++v) q.U[a][c][v] += w * N[a] * N[c] * N[v];
            }
        }
    }
}

// P1 triangle: global shape derivatives G[d][k] and det J (signed), with
// J = dN_ref X as src/numerics/assembly_solute.py:47-51.
__device__ inline double tri_geometry(const double (&px)[3], const double (&py)[3], double (&Gx)[3],
                                      double (&Gz)[3]) {
    const double J00 = px[1] - px[0], J01 = py[1] - py[0];
    const double J10 = px[2] - px[0], J11 = py[2] - py[0];
    const double det = J00 * J11 - J01 * J10;
    const double id = 1.0 / det;
    Gx[1] = J11 * id;
    Gx[2] = -J01 * id;
    Gz[1] = -J10 * id;
    Gz[2] = J00 * id;
    Gx[0] = -(Gx[1] + Gx[2]);
    Gz[0] = -(Gz[1] + Gz[2]);
    return det;
}

// ---------------------------------------------------------------------------
// Flux / dispersion at the nodes: mean of the incident elements' gradients
// (b/c form over |area|, transport_utilities.py:150-171), then q = -K grad h
// (+ gravity in z) and the dispersion tensor (:173-189).  With a per-element
// Ks field (not in the reference, whose Ks is one number) the node's K is the
// soil model's K times the mean Ks of its incident elements.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_transport_fields(FieldArgs a) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    double sx = 0.0, sz = 0.0, ks = 0.0;
    const int q0 = a.inc_ptr[i], q1 = a.inc_ptr[i + 1];
    for (int q = q0; q < q1; ++q) {
        const int e = a.inc[q].x >> 2;
        if (a.ks_e) ks += a.ks_e[e];
        const int4 cn = a.conn4[e];
        const int nd[3] = {cn.x, cn.y, cn.z};
        double x[3], y[3], hv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 r = a.nrec.xy(nd[c]);
            x[c] = r.x;
            y[c] = r.y;
            hv[c] = a.h[nd[c]];
        }
        const double b1 = y[1] - y[2], b2 = y[2] - y[0], b3 = y[0] - y[1];
        const double c1 = x[2] - x[1], c2 = x[0] - x[2], c3 = x[1] - x[0];
        const double aire = fabs(c3 * b2 - c2 * b3) / 2;
        sx += (b1 * hv[0] + b2 * hv[1] + b3 * hv[2]) / (2 * aire);
        sz += (c1 * hv[0] + c2 * hv[1] + c3 * hv[2]) / (2 * aire);
    }
    const double cnt = (double)(q1 - q0);
    const double gx = sx / cnt, gz = sz / cnt;
    double C, K, th;
    soil_eval(a.sp, a.h[i], C, K, th);
    if (a.ks_e) K *= ks / cnt;
    const double qx = -K * gx;
    const double qz = -K * (gz + 1.0);
    const double aq = sqrt(qx * qx + qz * qz);
    const double ts = a.sp.p[1];
    const double tau = pow(th, 7.0 / 3.0) / (ts * ts);
    a.qx[i] = qx;
    a.qz[i] = qz;
    a.aq[i] = aq;
    a.dxx[i] = (a.DL * qx * qx + a.DT * qz * qz) / aq + th * a.Dm * tau;
    a.dzz[i] = (a.DL * qz * qz + a.DT * qx * qx) / aq + th * a.Dm * tau;
    a.dxz[i] = (a.DL - a.DT) * qx * qz / aq;
    a.th[i] = th;
}

// ---------------------------------------------------------------------------
// Species system, one thread per row.  For incidence (e, a) the row receives
// row a of the element's stiffness (dispersion + advection + mu theta N N^T),
// of its mass (-(theta + rho Kd) N N^T: lumped into the diagonal or kept
// consistent) and its source entry (-mu_src theta c_src N_a); b = M c_n / dt
// + source.  Slots follow the select pattern of the Richards kernel.
// ---------------------------------------------------------------------------
template <int MAXR>
__global__ __launch_bounds__(BLOCK) void k_assemble_transport(TrArgs a) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const int rs = a.row_ptr[i];
    const int rl = a.row_ptr[i + 1] - rs;
    const int dg = a.diag_pos[i] - rs;
    double v[MAXR];
#pragma unroll
    for (int s = 0; s < MAXR; ++s) v[s] = 0.0;
    double msum = 0.0, bm = 0.0, bsrc = 0.0;
    const TrQuad& Q = a.tq;
    for (int q = a.inc_ptr[i]; q < a.inc_ptr[i + 1]; ++q) {
        const int2 ic = a.inc[q];
        const int la = ic.x & 3;
        const int4 cn = a.conn4[ic.x >> 2];
        const int nd[3] = {cn.x, cn.y, cn.z};
        double px[3], py[3], th[3], qx[3], qz[3], dxx[3], dxz[3], dzz[3], cs[3], cnv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 r = a.nrec.xy(nd[c]);
            px[c] = r.x;
            py[c] = r.y;
            th[c] = a.th[nd[c]];
            qx[c] = a.qx[nd[c]];
            qz[c] = a.qz[nd[c]];
            dxx[c] = a.dxx[nd[c]];
            dxz[c] = a.dxz[nd[c]];
            dzz[c] = a.dzz[nd[c]];
            cs[c] = a.c_src ? a.c_src[nd[c]] : 0.0;
            cnv[c] = a.lumped ? 0.0 : a.c_n[nd[c]];
        }
        double Gx[3], Gz[3];
        const double det = tri_geometry(px, py, Gx, Gz);
        double Dxx = 0.0, Dxz = 0.0, Dzz = 0.0, Tqx = 0.0, Tqz = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Dxx += Q.ck[c] * dxx[c];
            Dxz += Q.ck[c] * dxz[c];
            Dzz += Q.ck[c] * dzz[c];
            Tqx += Q.T[la][c] * qx[c];
            Tqz += Q.T[la][c] * qz[c];
        }
        const uint32_t w = (uint32_t)ic.y;
        double adds[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double uth = 0.0;
#pragma unroll
            for (int u = 0; u < 3; ++u) uth += Q.U[la][c][u] * th[u];
            double k = -(Dxx * Gx[la] * Gx[c] + Dxz * (Gx[la] * Gz[c] + Gz[la] * Gx[c]) + Dzz * Gz[la] * Gz[c])
                       - Tqx * Gx[c] - Tqz * Gz[c] + a.mu * uth;
            k *= det;
            const double m = (-uth - a.rho_kd * Q.T[la][c]) * det;
            double add = k;
            if (a.lumped) {
                msum += m;
            } else {
                add += a.inv_dt * m;
                bm += m * cnv[c];
            }
            adds[c] = add;
        }
        add_slots<3, MAXR>(v, w, adds);
        if (a.c_src) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int u = 0; u < 3; ++u) s += Q.U[la][c][u] * th[c] * cs[u];
            bsrc += -a.mu_src * s * det;
        }
    }
    double bi;
    if (a.lumped) {
#pragma unroll
        for (int s = 0; s < MAXR; ++s) v[s] += (s == dg) ? a.inv_dt * msum : 0.0;
        bi = a.inv_dt * (msum * a.c_n[i]) + bsrc;
    } else {
        bi = a.inv_dt * bm + bsrc;
    }
    double dgv = 0.0;
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
        if (s < rl) a.A[rs + s] = v[s];
        if (s == dg) dgv = v[s];
    }
    a.b[i] = bi;
    if (a.dinv) a.dinv[i] = 1.0 / (dgv != 0.0 ? dgv : 1.0);
}

// ---------------------------------------------------------------------------
// Edge terms: 2-point Gauss on [x_I, x_J] (jac = (x_J - x_I) / 2, weights 1,
// N = ((1 - xi)/2, (1 + xi)/2) at xi = -+1/sqrt(3)).
// ---------------------------------------------------------------------------
// Row-l (l = 0 for I, 1 for J) contributions of edge (I, J): m[0] to (row,
// row) and m[1] to (row, other); returns the b contribution.
__device__ inline double edge_row(const EdgeArgs& a, int I, int J, int l, double (&m)[2]) {
    const double xi[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
    const double jac = (a.nrec.xy(J).x - a.nrec.xy(I).x) / 2;
    m[0] = m[1] = 0.0;
    double bl = 0.0;
    for (int j = 0; j < 2; ++j) {
        const double N[2] = {(1 - xi[j]) / 2, (1 + xi[j]) / 2};
        switch (a.kind) {
            case EDGE_NEUMANN:
                bl = bl - jac * 1.0 * a.flux * N[l];
                break;
            case EDGE_FREEDRAIN_WATER: {
                const double kavg = N[0] * a.K[(size_t)I * a.kstride] + N[1] * a.K[(size_t)J * a.kstride];
                bl = bl + (-jac * 1.0 * kavg * N[l]);
                break;
            }
            case EDGE_CAUCHY: {
                const double wgt = jac * 1.0 * a.flux;
                m[0] += wgt * N[l] * N[l];
                m[1] += wgt * N[l] * N[1 - l];
                break;
            }
            default: {  // EDGE_FREEDRAIN_N
                const double kavg = N[0] * a.K[(size_t)I * a.kstride] + N[1] * a.K[(size_t)J * a.kstride];
                const double tavg = N[0] * a.th[I] + N[1] * a.th[J];
                const double wgt = jac * 1.0 * (-kavg / tavg);
                m[0] += wgt * N[l] * N[l];
                m[1] += wgt * N[l] * N[1 - l];
                break;
            }
        }
    }
    if (a.kind == EDGE_CAUCHY) {
        const double N0[2] = {(1 - xi[0]) / 2, (1 + xi[0]) / 2};
        const double N1[2] = {(1 - xi[1]) / 2, (1 + xi[1]) / 2};
        bl = jac * a.c_in * a.flux * (1.0 * N0[l] + 1.0 * N1[l]);
    }
    return bl;
}

__global__ __launch_bounds__(BLOCK) void k_edges(EdgeArgs a) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= a.cnt) return;
    const int i = a.nodes[k];
    const bool mat = a.kind == EDGE_CAUCHY || a.kind == EDGE_FREEDRAIN_N;
    double bi = a.b ? a.b[i] : 0.0;
    if (k > 0) {   // edge (nodes[k-1], i): i is its second node
        double m[2];
        bi += edge_row(a, a.nodes[k - 1], i, 1, m);
        if (mat) {
            a.A[a.diag_pos[i]] += m[0];
            a.A[a.slot_prev[k]] += m[1];
        }
    }
    if (k + 1 < a.cnt) {   // edge (i, nodes[k+1]): i is its first node
        double m[2];
        bi += edge_row(a, i, a.nodes[k + 1], 0, m);
        if (mat) {
            a.A[a.diag_pos[i]] += m[0];
            a.A[a.slot_next[k]] += m[1];
        }
    }
    if (a.b) a.b[i] = bi;
    if (mat && a.dinv) {
        const double d = a.A[a.diag_pos[i]];
        a.dinv[i] = 1.0 / (d != 0.0 ? d : 1.0);
    }
}

hipError_t launch_transport_fields(const FieldArgs& a, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_transport_fields, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_assemble_transport(const TrArgs& a, int max_row, hipStream_t s) {
    const dim3 g((a.n + BLOCK - 1) / BLOCK), b(BLOCK);
    if (a.n == 0) return hipSuccess;
    if (max_row <= 8) hipLaunchKernelGGL(k_assemble_transport<8>, g, b, 0, s, a);
    else if (max_row <= 16) hipLaunchKernelGGL(k_assemble_transport<16>, g, b, 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_edges(const EdgeArgs& a, hipStream_t s) {
    if (a.cnt > 1) hipLaunchKernelGGL(k_edges, dim3((a.cnt + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

}  // namespace wspm
