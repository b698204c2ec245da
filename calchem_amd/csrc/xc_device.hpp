// xc_device.hpp — SVWN5 / BLYP / PBE / B3LYP (+ B3LYP5, PBE0) exchange-correlation
// functionals evaluated on the GPU in FP64.
//
// Semantics follow crates/calchem-dft/src/xc.rs:
//   eval_xc_spin dispatch and mixing coefficients      xc.rs:159-223
//   Slater (Dirac) exchange, per spin                    xc.rs:229-251
//   VWN (1980) correlation, VWN5 and VWN_RPA parameters  xc.rs:257-381
//   Becke 88 gradient correction                         xc.rs:387-424
//   LYP in Miehlich et al. (1989) form                   xc.rs:433-547
//   PBE exchange (spin scaling) and PBE correlation on PW92   xc.rs:556-828
// Inputs are the spin invariants (rho_a, rho_b, gamma_aa, gamma_ab, gamma_bb);
// the output holds f_xc and its five first partial derivatives.
#pragma once
#include <hip/hip_runtime.h>

namespace chx {

enum XcId : int { XC_SVWN5 = 0, XC_BLYP = 1, XC_PBE = 2, XC_B3LYP = 3, XC_B3LYP5 = 4, XC_PBE0 = 5 };

struct XcOut {
    double f, va, vb, gaa, gab, gbb;
};

constexpr double kRhoCut = 1e-18;                        // xc.rs:13
constexpr double kPi = 3.14159265358979323846;
constexpr double kSlaterCx = 0.930525736349100025;       // (3/4)(6/pi)^(1/3)
constexpr double kCbrt2 = 1.2599210498948731648;
constexpr double kLypK0 = 36.462398978764777098;         // 2^(11/3) * 0.3 (3 pi^2)^(2/3)
constexpr double kPbeCxu = -0.73855876638202240588;      // -(3/4)(3/pi)^(1/3)
constexpr double kThreePi2 = 29.608813203268075857;      // 3 pi^2
constexpr double kPbeCt = 15.755920349483144659;         // (16/pi)(3 pi^2)^(1/3)
constexpr double kRsK = 0.23873241463784300365;          // 3 / (4 pi)

__device__ __forceinline__ double clamp_pm1(double x) { return fmin(1.0, fmax(-1.0, x)); }

__device__ __forceinline__ XcOut xc_zero() { return XcOut{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }

__device__ __forceinline__ void xc_axpy(XcOut& acc, double s, const XcOut& x) {
    acc.f += s * x.f;
    acc.va += s * x.va;
    acc.vb += s * x.vb;
    acc.gaa += s * x.gaa;
    acc.gab += s * x.gab;
    acc.gbb += s * x.gbb;
}

// ---- Dirac/Slater exchange: e_x(rho_s) = -Cx rho_s^{4/3} for each spin. ----
__device__ __forceinline__ XcOut dirac_exchange(double ra, double rb) {
    XcOut o = xc_zero();
    if (ra >= kRhoCut) {
        double c = cbrt(ra);
        o.f -= kSlaterCx * ra * c;
        o.va = -(4.0 / 3.0) * kSlaterCx * c;
    }
    if (rb >= kRhoCut) {
        double c = cbrt(rb);
        o.f -= kSlaterCx * rb * c;
        o.vb = -(4.0 / 3.0) * kSlaterCx * c;
    }
    return o;
}

// ---- VWN interpolation in x = sqrt(rs) ----
struct VwnFit {
    double A, x0, b, c;
};
struct VwnSet {
    VwnFit para, ferro, alpha;
};
// VWN5 (xc.rs:270-289) and VWN_RPA (xc.rs:291-310) fits.
__device__ constexpr VwnSet kVwn5 = {{0.0310907, -0.10498, 3.72744, 12.9352},
                                     {0.01554535, -0.32500, 7.06042, 18.0578},
                                     {-0.016886863940389905, -0.00475840, 1.13107, 13.0045}};
__device__ constexpr VwnSet kVwnRpa = {{0.0310907, -0.409286, 13.0720, 42.7198},
                                       {0.01554535, -0.743294, 20.1231, 101.578},
                                       {-0.016886863940389905, -0.228344, 1.06835, 11.4813}};

// eps(x) of one VWN channel and rs * d eps / d rs (exact Padé form).
__device__ __forceinline__ void vwn_fit(double x, const VwnFit& p, double& eps, double& rs_deps) {
    const double X = x * x + p.b * x + p.c;
    const double X0 = p.x0 * p.x0 + p.b * p.x0 + p.c;
    const double Q = sqrt(4.0 * p.c - p.b * p.b);
    const double at = atan(Q / (2.0 * x + p.b));
    const double dx0 = x - p.x0;
    const double lead = log(x * x / X) + (2.0 * p.b / Q) * at;
    const double tail = log(dx0 * dx0 / X) + (2.0 * (p.b + 2.0 * p.x0) / Q) * at;
    eps = p.A * (lead - (p.b * p.x0 / X0) * tail);
    rs_deps = p.A * (p.c * dx0 - p.b * x * p.x0) / (dx0 * X);
}

// Spin interpolation f(zeta), f'(zeta) and 1/f''(0).
__device__ __forceinline__ void spin_f(double zeta, double& fz, double& dfz, double& inv_fdd0) {
    const double c = kCbrt2 - 1.0;
    const double zp = fmax(1.0 + zeta, 0.0), zm = fmax(1.0 - zeta, 0.0);
    const double zp3 = cbrt(zp), zm3 = cbrt(zm);
    fz = (zp * zp3 + zm * zm3 - 2.0) / (2.0 * c);
    dfz = (4.0 / (6.0 * c)) * (zp3 - zm3);
    inv_fdd0 = (9.0 * c) / 4.0;
}

__device__ __forceinline__ XcOut vwn_correlation(double ra, double rb, const VwnSet& P) {
    XcOut o = xc_zero();
    const double rho = ra + rb;
    if (rho < kRhoCut) return o;
    const double rs = cbrt(kRsK / rho);
    const double x = sqrt(rs);
    const double zeta = clamp_pm1((ra - rb) / rho);
    double ep, dp;
    vwn_fit(x, P.para, ep, dp);
    // zeta = 0 (every point of a restricted density): f(0) = f'(0) = 0, so the
    // ferromagnetic and spin-stiffness channels add exactly nothing
    double ec = ep, rsdec = dp, dz = 0.0;
    if (zeta != 0.0) {
        double ef, df, ea, da, fz, dfz, ifdd;
        vwn_fit(x, P.ferro, ef, df);
        vwn_fit(x, P.alpha, ea, da);
        spin_f(zeta, fz, dfz, ifdd);
        const double z3 = zeta * zeta * zeta, z4 = z3 * zeta, w4 = 1.0 - z4;
        const double dfe = ef - ep, dfd = df - dp;
        ec = ep + ea * fz * ifdd * w4 + dfe * fz * z4;
        rsdec = dp + da * fz * ifdd * w4 + dfd * fz * z4;
        dz = ea * ifdd * (dfz * w4 - 4.0 * z3 * fz) + dfe * (dfz * z4 + 4.0 * z3 * fz);
    }
    const double base = ec - rsdec / 3.0;
    o.f = rho * ec;
    o.va = base + (1.0 - zeta) * dz;
    o.vb = base - (1.0 + zeta) * dz;
    return o;
}

// ---- Becke 88 gradient correction for one spin channel ----
__device__ __forceinline__ void b88_spin(double r, double g, double& f, double& vr, double& vg) {
    f = vr = vg = 0.0;
    if (r < kRhoCut) return;
    constexpr double beta = 0.0042;
    const double r13 = cbrt(r), r43 = r * r13;
    const double x2 = fmax(g, 0.0) / (r43 * r43);
    const double x = sqrt(x2);
    const double ash = asinh(x);
    const double d = 1.0 + 6.0 * beta * x * ash;
    const double gfac = -beta * x2 / d;
    // h = g'(x)/x, regular at x = 0
    const double h = -beta * (2.0 + 6.0 * beta * x * ash - 6.0 * beta * x2 / sqrt(1.0 + x2)) / (d * d);
    f = r43 * gfac;
    vr = (4.0 / 3.0) * r13 * (gfac - x2 * h);
    vg = 0.5 * h / r43;
}

__device__ __forceinline__ XcOut b88_correction(double ra, double rb, double gaa, double gbb) {
    XcOut o = xc_zero();
    double fa, fb;
    b88_spin(ra, gaa, fa, o.va, o.gaa);
    b88_spin(rb, gbb, fb, o.vb, o.gbb);
    o.f = fa + fb;
    return o;
}

// ---- LYP (Miehlich 1989 gradient-only form) ----
// Bracket S(ra, rb) = S_TF + A_aa g_aa + A_ab g_ab + A_bb g_bb and dS/d(ra);
// ra83, rb83 = ra^(8/3), rb^(8/3).
__device__ __forceinline__ void lyp_bracket(double ra, double rb, double ra83, double rb83, double dl, double ddl,
                                            double gaa, double gab, double gbb, double& S, double& dS,
                                            double& Aaa, double& Aab, double& Abb) {
    const double rho = ra + rb, rho2 = rho * rho;
    const double stf = kLypK0 * ra * rb * (ra83 + rb83);
    const double dstf = kLypK0 * ((11.0 / 3.0) * ra83 * rb + rb * rb83);
    const double k47 = (47.0 - 7.0 * dl) / 18.0;
    const double cg = ra * rb * k47 - (2.0 / 3.0) * rho2;
    const double dcg = rb * k47 - ra * rb * (7.0 * ddl) / 18.0 - (4.0 / 3.0) * rho;
    const double k25 = 2.5 - dl / 18.0;
    const double k11 = (dl - 11.0) / 9.0;
    const double qa = ra * ra * rb / rho;
    const double dqa = (2.0 * ra * rb * rho - ra * ra * rb) / rho2;
    const double ea = -ra * rb * k25 - k11 * qa + ((2.0 / 3.0) * rho2 - rb * rb);
    const double dea = -rb * k25 + ra * rb * (ddl / 18.0) - (ddl / 9.0) * qa - k11 * dqa + (4.0 / 3.0) * rho;
    const double qb = ra * rb * rb / rho;
    const double dqb = (rb * rb * rho - ra * rb * rb) / rho2;
    const double eb = -ra * rb * k25 - k11 * qb + ((2.0 / 3.0) * rho2 - ra * ra);
    const double deb =
        -rb * k25 + ra * rb * (ddl / 18.0) - (ddl / 9.0) * qb - k11 * dqb + (4.0 / 3.0) * rho - 2.0 * ra;
    Aaa = cg + ea;
    Aab = 2.0 * cg;
    Abb = cg + eb;
    S = stf + Aaa * gaa + Aab * gab + Abb * gbb;
    dS = dstf + (dcg + dea) * gaa + 2.0 * dcg * gab + (dcg + deb) * gbb;
}

__device__ __forceinline__ XcOut lyp_correlation(double ra, double rb, double gaa, double gab, double gbb) {
    XcOut o = xc_zero();
    const double rho = ra + rb;
    if (rho < kRhoCut) return o;
    constexpr double A = 0.04918, B = 0.132, C = 0.2533, D = 0.349;
    // powers of one third through cbrt (cheaper than pow)
    const double c = cbrt(rho), rm13 = 1.0 / c;
    const double den = 1.0 + D * rm13;
    const double g0 = -4.0 * A / (rho * den);
    const double dg0 = -g0 * (1.0 + (2.0 / 3.0) * D * rm13) / (rho * den);
    o.f = g0 * ra * rb;
    o.va = dg0 * ra * rb + g0 * rb;
    o.vb = dg0 * ra * rb + g0 * ra;
    if (C * rm13 > 500.0) return o;  // omega underflows: only the local term survives
    const double om = exp(-C * rm13) / den / (rho * rho * rho * c * c);  // rho^(-11/3)
    const double dl = C * rm13 + D * rm13 / den;
    const double dom = om * (dl - 11.0) / (3.0 * rho);
    const double ddl = -(1.0 / 3.0) * (rm13 / rho) * (C + D / (den * den));
    double S, dSa, Aaa, Aab, Abb, S2, dSb, t0, t1, t2;
    const double ca = ra > 0.0 ? cbrt(ra) : 0.0, cb = rb > 0.0 ? cbrt(rb) : 0.0;
    const double ra83 = ra * ra * ca * ca, rb83 = rb * rb * cb * cb;
    lyp_bracket(ra, rb, ra83, rb83, dl, ddl, gaa, gab, gbb, S, dSa, Aaa, Aab, Abb);
    lyp_bracket(rb, ra, rb83, ra83, dl, ddl, gbb, gab, gaa, S2, dSb, t0, t1, t2);
    const double ab = A * B;
    o.f -= ab * om * S;
    o.va -= ab * (dom * S + om * dSa);
    o.vb -= ab * (dom * S + om * dSb);
    o.gaa = -ab * om * Aaa;
    o.gab = -ab * om * Aab;
    o.gbb = -ab * om * Abb;
    return o;
}

// ---- PBE exchange via spin scaling f(rho_s, g_ss) = f_u(2 rho_s, 4 g_ss) / 2 ----
__device__ __forceinline__ void pbe_x_spin(double rs_, double gss, double& f, double& vr, double& vg) {
    f = vr = vg = 0.0;
    if (rs_ < kRhoCut) return;
    constexpr double kappa = 0.804, mu = 0.2195149727645171;
    const double ru = 2.0 * rs_, gu = 4.0 * fmax(gss, 0.0);
    const double r13 = cbrt(ru);
    const double exl = kPbeCxu * ru * r13;
    const double kf = cbrt(kThreePi2 * ru);
    const double d2 = 4.0 * kf * kf * ru * ru;
    const double s2 = gu / d2;
    const double q = 1.0 + (mu / kappa) * s2;
    const double fx = 1.0 + kappa - kappa / q;
    const double dfx = mu / (q * q);
    f = 0.5 * exl * fx;
    vr = (4.0 / 3.0) * kPbeCxu * r13 * fx - exl * dfx * (8.0 / 3.0) * s2 / ru;
    vg = 2.0 * exl * dfx / d2;
}

__device__ __forceinline__ XcOut pbe_exchange(double ra, double rb, double gaa, double gbb) {
    XcOut o = xc_zero();
    double fa, fb;
    pbe_x_spin(ra, gaa, fa, o.va, o.gaa);
    pbe_x_spin(rb, gbb, fb, o.vb, o.gbb);
    o.f = fa + fb;
    return o;
}

struct Pw92Fit {
    double A, a1, b1, b2, b3, b4;
};
__device__ constexpr Pw92Fit kPwUnpol = {0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
__device__ constexpr Pw92Fit kPwPol = {0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517};
__device__ constexpr Pw92Fit kPwAlpha = {0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671};

// G(rs) = -2A(1 + a1 rs) ln(1 + 1/Q1), Q1 = 2A(b1 x + b2 x^2 + b3 x^3 + b4 x^4).
__device__ __forceinline__ void pw92(double rs, const Pw92Fit& p, double& G, double& dG) {
    const double x = sqrt(rs);
    const double q0 = -2.0 * p.A * (1.0 + p.a1 * rs);
    const double q1 = 2.0 * p.A * (x * (p.b1 + x * (p.b2 + x * (p.b3 + x * p.b4))));
    const double dq1 = p.A * (p.b1 / x + 2.0 * p.b2 + 3.0 * p.b3 * x + 4.0 * p.b4 * rs);
    const double ln = log(1.0 + 1.0 / q1);
    G = q0 * ln;
    dG = -2.0 * p.A * p.a1 * ln - q0 * dq1 / (q1 * (q1 + 1.0));
}

__device__ __forceinline__ XcOut pbe_correlation(double ra, double rb, double gaa, double gab, double gbb) {
    XcOut o = xc_zero();
    const double rho = ra + rb;
    if (rho < kRhoCut) return o;
    constexpr double beta = 0.06672455060314922, gam = 0.0310906908696549;
    const double gt = fmax(gaa + 2.0 * gab + gbb, 0.0);
    const double rs = cbrt(kRsK / rho);
    const double drs = -rs / (3.0 * rho);
    const double zeta = clamp_pm1((ra - rb) / rho);
    double e0, de0;
    pw92(rs, kPwUnpol, e0, de0);
    // zeta = 0: f(0) = f'(0) = 0, the polarized and spin-stiffness fits add
    // exactly nothing (and phi = 1 below)
    double ec = e0, dec_rs = de0, dec_z = 0.0;
    if (zeta != 0.0) {
        double e1, de1, ma, dma, fz, dfz, ifdd;
        pw92(rs, kPwPol, e1, de1);
        pw92(rs, kPwAlpha, ma, dma);
        spin_f(zeta, fz, dfz, ifdd);
        const double ac = -ma, dac = -dma;
        const double z3 = zeta * zeta * zeta, z4 = z3 * zeta, w4 = 1.0 - z4;
        ec = e0 + ac * fz * ifdd * w4 + (e1 - e0) * fz * z4;
        dec_rs = de0 + dac * fz * ifdd * w4 + (de1 - de0) * fz * z4;
        dec_z = ac * ifdd * (dfz * w4 - 4.0 * z3 * fz) + (e1 - e0) * (dfz * z4 + 4.0 * z3 * fz);
    }
    const double dec_r = dec_rs * drs;
    if (gt < 1e-36) {
        o.f = rho * ec;
        o.va = ec + rho * dec_r + (1.0 - zeta) * dec_z;
        o.vb = ec + rho * dec_r - (1.0 + zeta) * dec_z;
        return o;
    }
    const double b0 = beta / gam;
    const double zp = fmax(1.0 + zeta, 1e-24), zm = fmax(1.0 - zeta, 1e-24);
    const double czp = zeta != 0.0 ? cbrt(zp) : 1.0, czm = zeta != 0.0 ? cbrt(zm) : 1.0;
    const double zp23 = czp * czp, zm23 = czm * czm;
    const double phi = 0.5 * (zp23 + zm23), phi2 = phi * phi, phi3 = phi2 * phi;
    const double omz_dphi = ((1.0 - zeta) / czp - zm23) / 3.0;   // (1 - zeta) dphi/dzeta
    const double opz_dphi = ((1.0 + zeta) / czm - zp23) / 3.0;   // -(1 + zeta) dphi/dzeta
    const double dt = kPbeCt * phi2 * (rho * rho * cbrt(rho));
    const double t2 = gt / dt;
    const double w = -ec / (gam * phi3);
    double Af = 0.0, dAf = 0.0;
    if (w <= 100.0) {
        const double em1 = expm1(w);
        Af = b0 / em1;
        dAf = -Af * exp(w) / em1;
    }
    const double u = Af * t2;
    const double dpoly = 1.0 + u + u * u;
    const double zarg = 1.0 + b0 * t2 * (1.0 + u) / dpoly;
    const double lnz = log(zarg);
    const double H = gam * phi3 * lnz;
    const double dZ_t2 = b0 * (1.0 + 2.0 * u) / (dpoly * dpoly);
    const double dZ_A = -b0 * t2 * t2 * u * (2.0 + u) / (dpoly * dpoly);
    const double dH_t2 = gam * phi3 * dZ_t2 / zarg;
    const double dH_w = gam * phi3 * dZ_A * dAf / zarg;
    const double dH_r = dH_t2 * (-(7.0 / 3.0) * t2 / rho) + dH_w * (-dec_r / (gam * phi3));
    const double dH_phi = 3.0 * gam * phi2 * lnz + dH_w * (-3.0 * w / phi) + dH_t2 * (-2.0 * t2 / phi);
    const double dH_ec = -dH_w / (gam * phi3);
    const double e_tot = ec + H;
    const double common = e_tot + rho * (dec_r + dH_r);
    o.f = rho * e_tot;
    o.va = common + (1.0 - zeta) * dec_z * (1.0 + dH_ec) + dH_phi * omz_dphi;
    o.vb = common - (1.0 + zeta) * dec_z * (1.0 + dH_ec) + dH_phi * opz_dphi;
    const double vg = rho * dH_t2 / dt;
    o.gaa = vg;
    o.gab = 2.0 * vg;
    o.gbb = vg;
    return o;
}

// eval_xc_spin (xc.rs:159-223).
__device__ __forceinline__ XcOut eval_xc(int func, double ra_in, double rb_in, double gaa_in, double gab, double gbb_in) {
    const double ra = fmax(ra_in, 0.0), rb = fmax(rb_in, 0.0);
    if (ra + rb < kRhoCut) return xc_zero();
    const double gaa = fmax(gaa_in, 0.0), gbb = fmax(gbb_in, 0.0);
    XcOut o = xc_zero();
    switch (func) {
    case XC_SVWN5:
        xc_axpy(o, 1.0, dirac_exchange(ra, rb));
        xc_axpy(o, 1.0, vwn_correlation(ra, rb, kVwn5));
        break;
    case XC_BLYP:
        xc_axpy(o, 1.0, dirac_exchange(ra, rb));
        xc_axpy(o, 1.0, b88_correction(ra, rb, gaa, gbb));
        xc_axpy(o, 1.0, lyp_correlation(ra, rb, gaa, gab, gbb));
        break;
    case XC_PBE:
        xc_axpy(o, 1.0, pbe_exchange(ra, rb, gaa, gbb));
        xc_axpy(o, 1.0, pbe_correlation(ra, rb, gaa, gab, gbb));
        break;
    case XC_B3LYP:
    case XC_B3LYP5:
        xc_axpy(o, 0.80, dirac_exchange(ra, rb));
        xc_axpy(o, 0.72, b88_correction(ra, rb, gaa, gbb));
        xc_axpy(o, 0.19, vwn_correlation(ra, rb, func == XC_B3LYP ? kVwnRpa : kVwn5));
        xc_axpy(o, 0.81, lyp_correlation(ra, rb, gaa, gab, gbb));
        break;
    case XC_PBE0:
        xc_axpy(o, 0.75, pbe_exchange(ra, rb, gaa, gbb));
        xc_axpy(o, 1.00, pbe_correlation(ra, rb, gaa, gab, gbb));
        break;
    default:
        break;
    }
    return o;
}

}  // namespace chx
