// ao_device.hpp — Cartesian Gaussian AO value and gradient at one point.
//
// Follows crates/calchem-dft/src/ao_grid.rs:182-230 (eval_aos_grad_filtered):
//   * the shell contributes nothing when r^2 > r2_cutoff = (40 + 4 l) / alpha_min;
//   * radial sums R0 = sum_p c~_p N_p e^{-a_p r^2}, R1 = -2 sum_p a_p c~_p N_p e^{-a_p r^2},
//     and the shell is skipped when |R0| and |R1| are both below 1e-18;
//   * phi = s x^lx y^ly z^lz R0, grad phi = s (d(x^lx y^ly z^lz) R0 + x^lx y^ly z^lz r R1);
//   * the AO is dropped (all four values zero) unless |phi| or a gradient
//     component exceeds 1e-16.
#pragma once
#include "common.hpp"

namespace chx {

__device__ __forceinline__ void mono_d(double x, int l, double& p, double& d) {
    // x^l and d/dx x^l for l <= 3 (the reference's pow_d0_d1_d2 branches).
    switch (l) {
    case 0: p = 1.0; d = 0.0; break;
    case 1: p = x; d = 1.0; break;
    case 2: p = x * x; d = 2.0 * x; break;
    default: p = x * x * x; d = 3.0 * x * x; break;
    }
}

struct AoVal {
    double v, gx, gy, gz;
};

__device__ __forceinline__ AoVal ao_eval(const DevShell* __restrict__ shells, const double2* __restrict__ prims,
                                         const DevAo& ao, double x, double y, double z) {
    AoVal o{0.0, 0.0, 0.0, 0.0};
    const DevShell s = shells[ao.shell];
    const double dx = x - s.cx, dy = y - s.cy, dz = z - s.cz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > s.r2cut) return o;
    double r0 = 0.0, r1 = 0.0;
    for (int p = 0; p < s.n_prims; ++p) {
        const double2 pr = prims[s.prim_start + p];  // (exponent, weighted norm)
        const double e = pr.y * exp(-pr.x * r2);
        r0 += e;
        r1 -= 2.0 * pr.x * e;
    }
    if (fabs(r0) < 1e-18 && fabs(r1) < 1e-18) return o;
    double px, dpx, py, dpy, pz, dpz;
    mono_d(dx, ao.lx, px, dpx);
    mono_d(dy, ao.ly, py, dpy);
    mono_d(dz, ao.lz, pz, dpz);
    const double m = px * py * pz;
    const double sc = ao.scale;
    const double v = sc * m * r0;
    const double gx = sc * (dpx * py * pz * r0 + m * dx * r1);
    const double gy = sc * (px * dpy * pz * r0 + m * dy * r1);
    const double gz = sc * (px * py * dpz * r0 + m * dz * r1);
    if (fabs(v) > 1e-16 || fabs(gx) > 1e-16 || fabs(gy) > 1e-16 || fabs(gz) > 1e-16) {
        o.v = v;
        o.gx = gx;
        o.gy = gy;
        o.gz = gz;
    }
    return o;
}

}  // namespace chx
