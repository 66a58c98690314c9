// Diagnostics kernels for gfx950 (see diagnostics.h).
//
// One thread per element over a fixed grid (grid-stride), the element's three
// node records gathered through the packed connectivity.  Sums use the
// deterministic two-phase reduction of wspm_device.h (workgroup partials,
// then one workgroup adding them in workgroup order); maxima are exact in any
// order.  Both passes are a few bytes per element: HBM/latency-bound.
#include "diagnostics.h"

namespace wspm {

namespace {

// jnp.max / jnp.maximum semantics: NaN wins.
__device__ inline double nan_max(double a, double b) {
    if (a != a || b != b) return __builtin_nan("");
    return a > b ? a : b;
}
__device__ inline double nan_min(double a, double b) {
    if (a != a || b != b) return __builtin_nan("");
    return a < b ? a : b;
}

__device__ inline void gather_xy(const DiagArgs& a, int e, int (&nd)[3], double (&px)[3], double (&py)[3]) {
    const int4 c = a.conn4[e];
    nd[0] = c.x;
    nd[1] = c.y;
    nd[2] = c.z;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double2 r = a.nrec.xy(nd[k]);
        px[k] = r.x;
        py[k] = r.y;
    }
}

// Smallest of |v0 - v2|, |v1 - v0|, |v2 - v1| (vertices minus their roll by
// one, transport_utilities.py:199-202).
__device__ inline double min_edge(const double (&px)[3], const double (&py)[3]) {
    // separately rounded squares and sum, as jnp.sum(edges**2) evaluates
    // them: no fused multiply-add here
#pragma clang fp contract(off)
    double h = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int j = (k + 2) % 3;
        const double dx = px[k] - px[j], dy = py[k] - py[j];
        const double l = sqrt(dx * dx + dy * dy);
        h = k == 0 ? l : nan_min(h, l);
    }
    return h;
}

// Pe and Cr of element e (transport_utilities.py:220-235).
__device__ inline void pe_cr(const DiagArgs& a, const PeCrArgs& p, int e, double& pe, double& cr) {
    int nd[3];
    double px[3], py[3];
    gather_xy(a, e, nd, px, py);
    const double h = min_edge(px, py);
    double pex = 0.0, pez = 0.0, crx = 0.0, crz = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int i = nd[k];
        const double th = p.th[i];
        const double vx = fabs(p.qx[i] / th), vz = fabs(p.qz[i] / th);
        const double px_ = (vx * h) / (2.0 * p.dxx[i]), pz_ = (vz * h) / (2.0 * p.dzz[i]);
        const double cx = (vx * p.dt) / h, cz = (vz * p.dt) / h;
        pex = k == 0 ? px_ : nan_max(pex, px_);
        pez = k == 0 ? pz_ : nan_max(pez, pz_);
        crx = k == 0 ? cx : nan_max(crx, cx);
        crz = k == 0 ? cz : nan_max(crz, cz);
    }
    pe = nan_max(pex, pez);
    cr = nan_max(crx, crz);
}

// Workgroup maximum of two values (NaN wins); the result is valid in thread 0.
__device__ inline void block_max2(double& m0, double& m1) {
    __shared__ double sh[2 * NWAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m0 = nan_max(m0, __shfl_xor(m0, o, 64));
        m1 = nan_max(m1, __shfl_xor(m1, o, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sh[wid] = m0;
        sh[NWAVE + wid] = m1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NWAVE; ++w) {
            m0 = nan_max(m0, sh[w]);
            m1 = nan_max(m1, sh[NWAVE + w]);
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_domain_integral(DiagArgs a, const double* __restrict__ f) {
    // points (1/6,1/6), (2/3,1/6), (1/6,2/3), weights 1/3 (checkpoint :30-35)
    const double gp[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}};
    const double w = 1.0 / 3.0;
    double acc[1] = {0.0};
    for (int e = blockIdx.x * BLOCK + threadIdx.x; e < a.E; e += a.grid * BLOCK) {
        int nd[3];
        double px[3], py[3];
        gather_xy(a, e, nd, px, py);
        const double fv[3] = {f[nd[0]], f[nd[1]], f[nd[2]]};
        const double J00 = px[1] - px[0], J01 = px[2] - px[0];
        const double J10 = py[1] - py[0], J11 = py[2] - py[0];
        const double det = fabs(J00 * J11 - J01 * J10);
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double N0 = 1.0 - gp[q][0] - gp[q][1], N1 = gp[q][0], N2 = gp[q][1];
            s += w * (fv[0] * N0 + fv[1] * N1 + fv[2] * N2);
        }
        acc[0] += s * det;
    }
    store_partials<1>(acc, a.partials, a.prow);
}

__global__ __launch_bounds__(BLOCK) void k_sum_finish(const double* __restrict__ partials, int nblk,
                                                       double* out) {
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < nblk; b += BLOCK) acc[0] += partials[b];
    __shared__ double sh[NWAVE];
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) out[0] = acc[0];
}

__global__ __launch_bounds__(BLOCK) void k_element_size(DiagArgs a, double* __restrict__ size) {
    for (int e = blockIdx.x * BLOCK + threadIdx.x; e < a.E; e += a.grid * BLOCK) {
        int nd[3];
        double px[3], py[3];
        gather_xy(a, e, nd, px, py);
        size[e] = min_edge(px, py);
    }
}

__global__ __launch_bounds__(BLOCK) void k_peclet_courant(DiagArgs a, PeCrArgs p) {
    double m0 = -__builtin_inf(), m1 = -__builtin_inf();
    for (int e = blockIdx.x * BLOCK + threadIdx.x; e < a.E; e += a.grid * BLOCK) {
        double pe, cr;
        pe_cr(a, p, e, pe, cr);
        if (p.pe) p.pe[e] = pe;
        if (p.cr) p.cr[e] = cr;
        m0 = nan_max(m0, pe);
        m1 = nan_max(m1, cr);
    }
    block_max2(m0, m1);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = m0;
        a.partials[a.prow + blockIdx.x] = m1;
    }
}

// The reference collects the vmapped (Pe[E], Cr[E]) pair with
// [pc[0] for pc in Pe_Cr] and [pc[1] for pc in Pe_Cr] (transport_utilities.py:
// 237-240): iterating the pair yields the two arrays, so its "Pe" is
// max(Pe_0, Cr_0) and its "Cr" is max(Pe_1, Cr_1) (index 1 clamped to E - 1,
// as JAX clamps).  all_elements gives the maxima over the mesh instead.
__global__ __launch_bounds__(BLOCK) void k_pecr_finish(DiagArgs a, PeCrArgs p, int nblk, double* out) {
    if (!p.all_elements) {
        if (threadIdx.x == 0) {
            double pe0, cr0, pe1, cr1;
            pe_cr(a, p, 0, pe0, cr0);
            pe_cr(a, p, a.E > 1 ? 1 : 0, pe1, cr1);
            out[0] = nan_max(pe0, cr0);
            out[1] = nan_max(pe1, cr1);
        }
        return;
    }
    double m0 = -__builtin_inf(), m1 = -__builtin_inf();
    for (int b = threadIdx.x; b < nblk; b += BLOCK) {
        m0 = nan_max(m0, a.partials[b]);
        m1 = nan_max(m1, a.partials[a.prow + b]);
    }
    block_max2(m0, m1);
    if (threadIdx.x == 0) {
        out[0] = m0;
        out[1] = m1;
    }
}

}  // namespace

hipError_t launch_domain_integral(const DiagArgs& a, const double* f, double* out, hipStream_t s) {
    if (a.grid < 1 || a.grid > a.prow) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_domain_integral, dim3(a.grid), dim3(BLOCK), 0, s, a, f);
    hipLaunchKernelGGL(k_sum_finish, dim3(1), dim3(BLOCK), 0, s, a.partials, a.grid, out);
    return hipGetLastError();
}

hipError_t launch_element_size(const DiagArgs& a, double* size, hipStream_t s) {
    if (a.grid < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_element_size, dim3(a.grid), dim3(BLOCK), 0, s, a, size);
    return hipGetLastError();
}

hipError_t launch_peclet_courant(const DiagArgs& a, const PeCrArgs& p, double* out, hipStream_t s) {
    if (a.grid < 1 || a.grid > a.prow || a.E < 1) return hipErrorInvalidValue;
    if (p.all_elements || p.pe || p.cr)
        hipLaunchKernelGGL(k_peclet_courant, dim3(a.grid), dim3(BLOCK), 0, s, a, p);
    hipLaunchKernelGGL(k_pecr_finish, dim3(1), dim3(BLOCK), 0, s, a, p, a.grid, out);
    return hipGetLastError();
}

}  // namespace wspm
