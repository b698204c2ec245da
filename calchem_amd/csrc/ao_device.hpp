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
    // x^l and d/dx x^l for l <= 3 (the reference's pow_d0_d1_d2 cases),
    // by selects so that lanes with different l do not diverge.
    const double x2 = x * x;
    p = l == 0 ? 1.0 : (l == 1 ? x : (l == 2 ? x2 : x2 * x));
    d = l == 0 ? 0.0 : (l == 1 ? 1.0 : (l == 2 ? 2.0 * x : 3.0 * x * x));
}

struct AoVal {
    double v, gx, gy, gz;
};

// x^l and d/dx x^l for l <= 3 in straight-line code (selects, no branches):
// q = x^(l-1) for l >= 1, p = x q (1 for l = 0), d = l q.
__device__ __forceinline__ void mono_sel(double x, int l, double& p, double& d) {
    const double q = (l & 2) ? ((l & 1) ? x * x : x) : 1.0;
    p = l ? x * q : 1.0;
    d = static_cast<double>(l) * q;
}

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

// The same evaluation split in two: the radial sums of one shell at one point
// (zero when the shell is cut off or both sums fall below 1e-18), then any
// Cartesian component of that shell from them.
__device__ __forceinline__ double2 shell_radial(const DevShell& s, const double2* __restrict__ prims, double x,
                                                double y, double z) {
    const double dx = x - s.cx, dy = y - s.cy, dz = z - s.cz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    double r0 = 0.0, r1 = 0.0;
    if (r2 <= s.r2cut) {
        for (int p = 0; p < s.n_prims; ++p) {
            const double2 pr = prims[s.prim_start + p];
            const double e = pr.y * exp(-pr.x * r2);
            r0 += e;
            r1 -= 2.0 * pr.x * e;
        }
        if (fabs(r0) < 1e-18 && fabs(r1) < 1e-18) r0 = r1 = 0.0;
    }
    return make_double2(r0, r1);
}

// e^{-t} for t >= 0, table-driven (P. T. P. Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15 (1989) 144-157): t = (64 m + j) ln2/64 + r with
// |r| <= ln2/128 (Cody-Waite reduction with ln2/64 in two parts; the first
// product is exact under the FMA), e^{-t} = 2^-m 2^{-j/64} e^{-r}, e^{-r}
// from its degree-5 Taylor polynomial (truncation below 4e-17 relative).
// tab holds 2^{-j/64}, j = 0..63 (kExpNegTab, staged in LDS by the caller).
// Within 2 ulp of the correctly rounded value (20 M samples on [0, 708]);
// 0 for t > 708, where e^{-t} is below the smallest normal double.
__device__ constexpr double kExpNegTab[64] = {
    0x1.0000000000000p+0, 0x1.fa7c1819e90d8p-1, 0x1.f50765b6e4540p-1, 0x1.efa1bee615a27p-1,
    0x1.ea4afa2a490dap-1, 0x1.e502ee78b3ff6p-1, 0x1.dfc97337b9b5fp-1, 0x1.da9e603db3285p-1,
    0x1.d5818dcfba487p-1, 0x1.d072d4a07897cp-1, 0x1.cb720dcef9069p-1, 0x1.c67f12e57d14bp-1,
    0x1.c199bdd85529cp-1, 0x1.bcc1e904bc1d2p-1, 0x1.b7f76f2fb5e47p-1, 0x1.b33a2b84f15fbp-1,
    0x1.ae89f995ad3adp-1, 0x1.a9e6b5579fdbfp-1, 0x1.a5503b23e255dp-1, 0x1.a0c667b5de565p-1,
    0x1.9c49182a3f090p-1, 0x1.97d829fde4e50p-1, 0x1.93737b0cdc5e5p-1, 0x1.8f1ae99157736p-1,
    0x1.8ace5422aa0dbp-1, 0x1.868d99b4492edp-1, 0x1.82589994cce13p-1, 0x1.7e2f336cf4e62p-1,
    0x1.7a11473eb0187p-1, 0x1.75feb564267c9p-1, 0x1.71f75e8ec5f74p-1, 0x1.6dfb23c651a2fp-1,
    0x1.6a09e667f3bcdp-1, 0x1.6623882552225p-1, 0x1.6247eb03a5585p-1, 0x1.5e76f15ad2148p-1,
    0x1.5ab07dd485429p-1, 0x1.56f4736b527dap-1, 0x1.5342b569d4f82p-1, 0x1.4f9b2769d2ca7p-1,
    0x1.4bfdad5362a27p-1, 0x1.486a2b5c13cd0p-1, 0x1.44e086061892dp-1, 0x1.4160a21f72e2ap-1,
    0x1.3dea64c123422p-1, 0x1.3a7db34e59ff7p-1, 0x1.371a7373aa9cbp-1, 0x1.33c08b26416ffp-1,
    0x1.306fe0a31b715p-1, 0x1.2d285a6e4030bp-1, 0x1.29e9df51fdee1p-1, 0x1.26b4565e27cddp-1,
    0x1.2387a6e756238p-1, 0x1.2063b88628cd6p-1, 0x1.1d4873168b9aap-1, 0x1.1a35beb6fcb75p-1,
    0x1.172b83c7d517bp-1, 0x1.1429aaea92de0p-1, 0x1.11301d0125b51p-1, 0x1.0e3ec32d3d1a2p-1,
    0x1.0b5586cf9890fp-1, 0x1.0874518759bc8p-1, 0x1.059b0d3158574p-1, 0x1.02c9a3e778061p-1,
};

__device__ __forceinline__ double exp_neg_tab(double t, const double* __restrict__ tab) {
    const double kd = rint(t * 92.33248261689366);  // 64 / ln 2
    const double rh = fma(-kd, 0.010830424696249145, t);
    const double r = fma(-kd, 3.623510646634843e-19, rh);
    const int k = static_cast<int>(kd);
    double p = fma(r, -1.0 / 120.0, 1.0 / 24.0);
    p = fma(r, p, -1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(r, p, -1.0);
    p = fma(r, p, 1.0);
    const double v = ldexp(tab[k & 63] * p, -(k >> 6));
    return t <= 708.0 ? v : 0.0;
}

// shell_radial with the table-driven exponential.
__device__ __forceinline__ double2 shell_radial_tab(const DevShell& s, const double2* __restrict__ prims, double x,
                                                    double y, double z, const double* __restrict__ tab) {
    const double dx = x - s.cx, dy = y - s.cy, dz = z - s.cz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    double r0 = 0.0, r1 = 0.0;
    if (r2 <= s.r2cut) {
        for (int p = 0; p < s.n_prims; ++p) {
            const double2 pr = prims[s.prim_start + p];
            const double e = pr.y * exp_neg_tab(pr.x * r2, tab);
            r0 += e;
            r1 -= 2.0 * pr.x * e;
        }
        if (fabs(r0) < 1e-18 && fabs(r1) < 1e-18) r0 = r1 = 0.0;
    }
    return make_double2(r0, r1);
}

// Straight-line: the unscaled values u are formed first and the drop test
// |scale| max|u| > 1e-16 (the same test on the scaled values, since scaling
// by |scale| is monotonic) selects the scale or zero.  Zero radial sums give
// zero values by themselves.
__device__ __forceinline__ AoVal ao_from_radial(const DevShell& s, const DevAo& ao, double x, double y, double z,
                                                double2 rr) {
    const double dx = x - s.cx, dy = y - s.cy, dz = z - s.cz;
    double px, dpx, py, dpy, pz, dpz;
    mono_sel(dx, ao.lx, px, dpx);
    mono_sel(dy, ao.ly, py, dpy);
    mono_sel(dz, ao.lz, pz, dpz);
    const double pyz = py * pz, m = px * pyz, mr1 = m * rr.y;
    const double u = m * rr.x;
    const double ux = fma(dpx * pyz, rr.x, mr1 * dx);
    const double uy = fma(px * dpy * pz, rr.x, mr1 * dy);
    const double uz = fma(px * py * dpz, rr.x, mr1 * dz);
    const double mx = fmax(fmax(fabs(u), fabs(ux)), fmax(fabs(uy), fabs(uz)));
    const double sc = fabs(ao.scale) * mx > 1e-16 ? ao.scale : 0.0;
    return AoVal{sc * u, sc * ux, sc * uy, sc * uz};
}

// ao_from_radial for a row held in LDS.
__device__ __forceinline__ AoVal ao_from_radial_row(const DevShell& s, const RowL& r, double x, double y, double z,
                                                    double2 rr) {
    DevAo ao;
    ao.scale = r.scale;
    ao.shell = r.sh;
    ao.lx = r.lmn & 3;
    ao.ly = (r.lmn >> 2) & 3;
    ao.lz = (r.lmn >> 4) & 3;
    return ao_from_radial(s, ao, x, y, z, rr);
}

// FP32 variants for the mixed build's V stage.  Offsets from the shell
// centre are formed in FP64 and rounded once, so r^2 carries FP32 relative
// error only; e^{-x} then has absolute error below ~1e-7 x e^{-x} <= 4e-8.
// e^{-x} uses the hardware base-2 exponential (v_exp_f32).

__device__ __forceinline__ float exp_neg_f(float x) { return __builtin_amdgcn_exp2f(-x * 1.4426950408889634f); }

// mono_sel in FP32.
__device__ __forceinline__ void mono_self(float x, int l, float& p, float& d) {
    const float q = (l & 2) ? ((l & 1) ? x * x : x) : 1.0f;
    p = l ? x * q : 1.0f;
    d = static_cast<float>(l) * q;
}

// For the mixed build's V stage, the AO values only enter through
// P = phi and Z = w0 phi + w1 dphi/dx + w2 dphi/dy + w3 dphi/dz, with
// (w0, w1, w2, w3) = wv of the point.  With phi = s m R0 and
// grad phi = s (grad(m) R0 + m R1 d) (m = x^lx y^ly z^lz, d = r - centre):
//   Z = s (m A + R0 B),  A = w0 R0 + R1 (w1 dx + w2 dy + w3 dz),
//   B = w1 dm/dx + w2 dm/dy + w3 dm/dz,
// so A is formed once per (shell, point) beside the radial sums, and only m
// and B per row.  No per-AO 1e-16 drop test here: what it removes is below
// FP32 resolution of these sums.
// Radial pair (R0, A) of one shell at one point, FP32, with the cutoff and
// 1e-18 rules of shell_radial.
__device__ __forceinline__ float2 shell_radial_za_f(const DevShell& s, const double2* __restrict__ prims, double x,
                                                    double y, double z, const float4& w) {
    const float dx = static_cast<float>(x - s.cx), dy = static_cast<float>(y - s.cy),
                dz = static_cast<float>(z - s.cz);
    const float r2 = dx * dx + dy * dy + dz * dz;
    float r0 = 0.0f, r1 = 0.0f;
    if (r2 <= static_cast<float>(s.r2cut)) {
        for (int p = 0; p < s.n_prims; ++p) {
            const double2 pr = prims[s.prim_start + p];
            const float al = static_cast<float>(pr.x);
            const float e = static_cast<float>(pr.y) * exp_neg_f(al * r2);
            r0 += e;
            r1 -= 2.0f * al * e;
        }
        if (fabsf(r0) < 1e-18f && fabsf(r1) < 1e-18f) r0 = r1 = 0.0f;
    }
    return make_float2(r0, fmaf(w.x, r0, r1 * fmaf(w.y, dx, fmaf(w.z, dy, w.w * dz))));
}

// (P, Z) of one row at one point from its shell's (R0, A).
__device__ __forceinline__ float2 pz_from_radial_row_f(const DevShell& s, const RowL& r, double x, double y, double z,
                                                       float2 ra, const float4& w) {
    const float dx = static_cast<float>(x - s.cx), dy = static_cast<float>(y - s.cy),
                dz = static_cast<float>(z - s.cz);
    float px, dpx, py, dpy, pz, dpz;
    mono_self(dx, r.lmn & 3, px, dpx);
    mono_self(dy, (r.lmn >> 2) & 3, py, dpy);
    mono_self(dz, (r.lmn >> 4) & 3, pz, dpz);
    const float pyz = py * pz, m = px * pyz;
    const float b = fmaf(w.y * dpx, pyz, px * fmaf(w.z * dpy, pz, w.w * py * dpz));
    const float sc = static_cast<float>(r.scale);
    return make_float2(sc * (m * ra.x), sc * fmaf(m, ra.y, ra.x * b));
}

// For the nuclear gradient (eval_aos_hess, ao_grid.rs:234-299): radial sums
// R0, R1, R2 = 4 sum_p a_p^2 c~_p N_p e^{-a_p r^2}, zero when the shell is
// cut off or all three fall below 1e-18; no per-AO drop rule.
struct Rad3 {
    double r0, r1, r2;
};

__device__ __forceinline__ Rad3 shell_radial3(const DevShell& s, const double2* __restrict__ prims, double x,
                                              double y, double z, const double* __restrict__ tab) {
    const double dx = x - s.cx, dy = y - s.cy, dz = z - s.cz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    Rad3 o{0.0, 0.0, 0.0};
    if (r2 <= s.r2cut) {
        for (int p = 0; p < s.n_prims; ++p) {
            const double2 pr = prims[s.prim_start + p];
            const double e = pr.y * exp_neg_tab(pr.x * r2, tab);
            o.r0 += e;
            o.r1 -= 2.0 * pr.x * e;
            o.r2 += 4.0 * pr.x * pr.x * e;
        }
        if (fabs(o.r0) < 1e-18 && fabs(o.r1) < 1e-18 && fabs(o.r2) < 1e-18) o.r0 = o.r1 = o.r2 = 0.0;
    }
    return o;
}

// x^l and its first two derivatives, l <= 3 (pow_d0_d1_d2, ao_grid.rs:80-95).
__device__ __forceinline__ void mono_d2(double x, int l, double& p, double& d, double& d2) {
    const double x2 = x * x;
    p = l == 0 ? 1.0 : (l == 1 ? x : (l == 2 ? x2 : x2 * x));
    d = l == 0 ? 0.0 : (l == 1 ? 1.0 : (l == 2 ? 2.0 * x : 3.0 * x * x));
    d2 = l <= 1 ? 0.0 : (l == 2 ? 2.0 : 6.0 * x);
}

// Value, gradient and Hessian (xx, xy, xz, yy, yz, zz) of one row.
struct AoHess {
    double v, g[3], h[6];
};

__device__ __forceinline__ AoHess ao_hess_row(const DevShell& s, const RowL& r, double x, double y, double z,
                                              const Rad3& rr) {
    AoHess o{};
    const double dx = x - s.cx, dy = y - s.cy, dz = z - s.cz;
    double px, dpx, d2px, py, dpy, d2py, pz, dpz, d2pz;
    mono_d2(dx, r.lmn & 3, px, dpx, d2px);
    mono_d2(dy, (r.lmn >> 2) & 3, py, dpy, d2py);
    mono_d2(dz, (r.lmn >> 4) & 3, pz, dpz, d2pz);
    const double m = px * py * pz, sc = r.scale;
    const double dm0 = dpx * py * pz, dm1 = px * dpy * pz, dm2 = px * py * dpz;
    o.v = sc * m * rr.r0;
    o.g[0] = sc * (dm0 * rr.r0 + m * dx * rr.r1);
    o.g[1] = sc * (dm1 * rr.r0 + m * dy * rr.r1);
    o.g[2] = sc * (dm2 * rr.r0 + m * dz * rr.r1);
    o.h[0] = sc * (d2px * py * pz * rr.r0 + (2.0 * dm0 * dx + m) * rr.r1 + m * dx * dx * rr.r2);
    o.h[1] = sc * (dpx * dpy * pz * rr.r0 + (dm0 * dy + dm1 * dx) * rr.r1 + m * dx * dy * rr.r2);
    o.h[2] = sc * (dpx * py * dpz * rr.r0 + (dm0 * dz + dm2 * dx) * rr.r1 + m * dx * dz * rr.r2);
    o.h[3] = sc * (px * d2py * pz * rr.r0 + (2.0 * dm1 * dy + m) * rr.r1 + m * dy * dy * rr.r2);
    o.h[4] = sc * (px * dpy * dpz * rr.r0 + (dm1 * dz + dm2 * dy) * rr.r1 + m * dy * dz * rr.r2);
    o.h[5] = sc * (px * py * d2pz * rr.r0 + (2.0 * dm2 * dz + m) * rr.r1 + m * dz * dz * rr.r2);
    return o;
}

// Value and gradient of one row from (R0, R1) of its shell, without the
// per-AO drop rule (the gradient's AO values, eval_aos_hess, have none):
// the v and g of ao_hess_row.
__device__ __forceinline__ AoVal ao_grad_row(const DevShell& s, const RowL& r, double x, double y, double z,
                                             const Rad3& rr) {
    const double dx = x - s.cx, dy = y - s.cy, dz = z - s.cz;
    double px, dpx, d2px, py, dpy, d2py, pz, dpz, d2pz;
    mono_d2(dx, r.lmn & 3, px, dpx, d2px);
    mono_d2(dy, (r.lmn >> 2) & 3, py, dpy, d2py);
    mono_d2(dz, (r.lmn >> 4) & 3, pz, dpz, d2pz);
    const double m = px * py * pz, sc = r.scale;
    const double dm0 = dpx * py * pz, dm1 = px * dpy * pz, dm2 = px * py * dpz;
    return AoVal{sc * m * rr.r0, sc * (dm0 * rr.r0 + m * dx * rr.r1), sc * (dm1 * rr.r0 + m * dy * rr.r1),
                 sc * (dm2 * rr.r0 + m * dz * rr.r1)};
}

// Stage the rows, shells (with block-local primitive offsets) and primitives
// of a block or group into LDS.  rows: (global AO, local shell); shl:
// (global shell, local primitive offset).
__device__ __forceinline__ void stage_meta(const int2* __restrict__ rows, int m, int mp,
                                           const int2* __restrict__ shl, int nsh,
                                           const DevShell* __restrict__ shells, const double2* __restrict__ prims,
                                           const DevAo* __restrict__ aos, RowL* sRow, DevShell* sSh,
                                           double2* sPrim, int tid, int nthreads) {
    for (int r = tid; r < mp; r += nthreads) {
        RowL o{0.0, 0, -1};
        if (r < m) {
            const int2 e = rows[r];
            const DevAo ao = aos[e.x];
            o.scale = ao.scale;
            o.lmn = ao.lx | (ao.ly << 2) | (ao.lz << 4);
            o.sh = e.y;
        }
        sRow[r] = o;
    }
    for (int s = tid; s < nsh; s += nthreads) {
        const int2 e = shl[s];
        DevShell d = shells[e.x];
        for (int p = 0; p < d.n_prims; ++p) sPrim[e.y + p] = prims[d.prim_start + p];
        d.prim_start = e.y;
        sSh[s] = d;
    }
}

}  // namespace chx
