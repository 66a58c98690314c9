// Diagnostics over the mesh context: domain integrals (mass balance) and the
// element Peclet / Courant numbers of the transport fields (diagnostics.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wspm_device.h"

namespace wspm {

struct DiagArgs {
    int E;
    const int4* conn4;        // padded connectivity
    NodeRecs nrec;      // node coordinates (xy)
    double* partials;         // column-major [2][prow]
    int prow;
    int grid;                 // workgroups of the element pass (<= prow)
};

// out[0] = sum_e |det J_e| * sum_q w_q f(x_q), the 3-point rule with weights
// 1/3 of richards_mass_balance-checkpoint.py:25-65 (twice the integral).
// out: device double.
hipError_t launch_domain_integral(const DiagArgs& a, const double* f, double* out, hipStream_t s);

// Smallest edge length of every triangle (transport_utilities.py:196-204).
hipError_t launch_element_size(const DiagArgs& a, double* size, hipStream_t s);

// Element Peclet / Courant numbers (transport_utilities.py:206-242).
// pe, cr: per-element outputs (nullable).  out[0..1] (device): with
// all_elements, the maxima over the mesh; otherwise the reference's
// (max(Pe_0, Cr_0), max(Pe_1, Cr_1)).
struct PeCrArgs {
    const double *qx, *qz, *th, *dxx, *dzz;
    double dt;
    double *pe, *cr;
    int all_elements;
};
hipError_t launch_peclet_courant(const DiagArgs& a, const PeCrArgs& p, double* out, hipStream_t s);

}  // namespace wspm
