// int_device.hpp — device side of the Gaussian integral engine: the Boys
// function, McMurchie-Davidson Hermite expansion coefficients E and Hermite
// Coulomb integrals R, and the one-electron and electron-repulsion class
// kernels (one template instance per angular-momentum class, every loop of
// compile-time length so that the working arrays live in registers).
//
// What is computed is fixed by the reference's integrals
// (crates/calchem-int/src/one_electron.rs:7-168, eri.rs:276-1018: S, T,
// V_ne and (ab|cd) over contracted Cartesian Gaussians whose components are
// each normalised); the reference reaches them by Obara-Saika / HGP
// recursions, this engine by Hermite expansions (McMurchie and Davidson,
// J. Comput. Phys. 26, 218 (1978); notation of Helgaker, Jorgensen and
// Olsen, Molecular Electronic-Structure Theory, ch. 9).
#pragma once
#include <hip/hip_runtime.h>

namespace chx {
namespace ints {

constexpr int kMaxL = 2;  // s, p and d shells (the bases here stop at d)

__host__ __device__ constexpr int tri_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
// Index of the Hermite / Cartesian triple (t, u, v) among all triples, by total first.
__host__ __device__ constexpr int hidx(int t, int u, int v) {
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
// Cartesian component order of a shell (crates/calchem-basis/src/lib.rs:135-160).
__host__ __device__ constexpr int cart_x(int l, int c) {
    return l == 0 ? 0 : (l == 1 ? (c == 0 ? 1 : 0) : (c == 0 ? 2 : (c == 1 || c == 2 ? 1 : 0)));
}
__host__ __device__ constexpr int cart_y(int l, int c) {
    return l == 0 ? 0 : (l == 1 ? (c == 1 ? 1 : 0) : (c == 3 ? 2 : (c == 1 || c == 4 ? 1 : 0)));
}
__host__ __device__ constexpr int cart_z(int l, int c) {
    return l == 0 ? 0 : (l == 1 ? (c == 2 ? 1 : 0) : (c == 5 ? 2 : (c == 2 || c == 4 ? 1 : 0)));
}
// Class of a shell pair with la >= lb: (0,0) 0, (1,0) 1, (1,1) 2, (2,0) 3, (2,1) 4, (2,2) 5.
__host__ __device__ constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }
constexpr int kPairClasses = 6;

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi52 = 34.98683665524972497;  // 2 pi^(5/2)

// One primitive pair of a shell pair (a, b): p = alpha + beta, the product
// centre P, P - A, P - B and K = c_a c_b exp(-alpha beta / p |A - B|^2)
// (c: contraction coefficient times N(alpha; L00)).
struct PrimPair {
    double p, inv2p, px, py, pz, pax, pay, paz, pbx, pby, pbz, k;
};

// A shell pair with la >= lb (shells swapped where needed) and its primitive pairs.
struct ShellPairDev {
    int prim_begin, n_prim;
    int ao_a, ao_b;   // first AO of shell a / b
    int la, lb;
    double q;         // Schwarz factor max sqrt((mu nu|mu nu))
};

// F_m(T), m = 0..M: the series e^{-T} sum_k (2T)^k / ((2M+1)(2M+3)..(2M+2k+1))
// for F_M below T = 30 with downward recursion, else F_0 = sqrt(pi/T) erf(sqrt T) / 2
// and upward recursion (stable there: (2m+1)/(2T) < 1).
template <int M>
__device__ __forceinline__ void boys(double T, double (&F)[M + 1]) {
    const double et = exp(-T);
    if (T < 30.0) {
        double term = 1.0 / (2 * M + 1), sum = term;
        for (int k = 1; k < 160; ++k) {
            term *= 2.0 * T / (double)(2 * M + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        F[M] = et * sum;
#pragma unroll
        for (int m = M; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + et) / (double)(2 * m - 1);
    } else {
        F[0] = 0.5 * sqrt(kPi / T) * erf(sqrt(T));
#pragma unroll
        for (int m = 0; m < M; ++m) F[m + 1] = ((double)(2 * m + 1) * F[m] - et) / (2.0 * T);
    }
}

// Hermite expansion coefficients of one Cartesian direction:
// E[i][j][t] for i <= LA, j <= LB, t <= i + j (E^{00}_0 = 1; the Gaussian
// prefactor K is kept apart).  E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + PA E^{ij}_t + (t+1) E^{ij}_{t+1},
// and likewise for j with PB.
template <int LA, int LB>
struct HermE {
    double e[LA + 1][LB + 1][LA + LB + 1];
};

template <int LA, int LB>
__device__ __forceinline__ void herm_e(double pa, double pb, double inv2p, HermE<LA, LB>& E) {
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= LA + LB; ++t) E.e[i][j][t] = 0.0;
    E.e[0][0][0] = 1.0;
#pragma unroll
    for (int i = 1; i <= LA; ++i)
#pragma unroll
        for (int t = 0; t <= LA + LB; ++t)
            if (t <= i) {
                double v = pa * E.e[i - 1][0][t];
                if (t >= 1) v += inv2p * E.e[i - 1][0][t - 1];
                if (t + 1 <= i - 1) v += (double)(t + 1) * E.e[i - 1][0][t + 1];
                E.e[i][0][t] = v;
            }
#pragma unroll
    for (int j = 1; j <= LB; ++j)
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int t = 0; t <= LA + LB; ++t)
                if (t <= i + j) {
                    double v = pb * E.e[i][j - 1][t];
                    if (t >= 1) v += inv2p * E.e[i][j - 1][t - 1];
                    if (t + 1 <= i + j - 1) v += (double)(t + 1) * E.e[i][j - 1][t + 1];
                    E.e[i][j][t] = v;
                }
}

// Hermite Coulomb integrals R_{tuv} = R^{(0)}_{tuv}(alpha, X, Y, Z) for
// t + u + v <= L from R^{(n)}_{000} = (-2 alpha)^n F_n:
// R^{(n)}_{t+1,u,v} = t R^{(n+1)}_{t-1,u,v} + X R^{(n+1)}_{tuv} (and u, v alike).
template <int L>
__device__ __forceinline__ void herm_r(double m2a, double X, double Y, double Z, const double (&F)[L + 1],
                                       double (&R)[tri_count(L)]) {
    double pw[L + 1];
    pw[0] = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * m2a;
    R[0] = pw[L] * F[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
        double cur[tri_count(L)];
        cur[0] = pw[n] * F[n];
#pragma unroll
        for (int t = 0; t <= L; ++t)
#pragma unroll
            for (int u = 0; u <= L; ++u)
#pragma unroll
                for (int v = 0; v <= L; ++v) {
                    if (t + u + v == 0 || t + u + v > L - n) continue;
                    double val;
                    if (t > 0) {
                        val = X * R[hidx(t - 1, u, v)];
                        if (t > 1) val += (double)(t - 1) * R[hidx(t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * R[hidx(t, u - 1, v)];
                        if (u > 1) val += (double)(u - 1) * R[hidx(t, u - 2, v)];
                    } else {
                        val = Z * R[hidx(t, u, v - 1)];
                        if (v > 1) val += (double)(v - 1) * R[hidx(t, u, v - 2)];
                    }
                    cur[hidx(t, u, v)] = val;
                }
#pragma unroll
        for (int k = 0; k < tri_count(L); ++k)
            if (k < tri_count(L - n)) R[k] = cur[k];
    }
}

// ---------------------------------------------------------------------------
// One-electron integrals of one class (LA >= LB): one thread per
// (shell pair, component pair); S, T and V_ne = -sum_C Z_C (mu|1/r_C|nu)
// written to both triangles of the row-major n x n matrices.
// ---------------------------------------------------------------------------
struct OneArgs {
    const ShellPairDev* pairs;
    const int* list;         // pair ids of this class
    int n_list;
    const PrimPair* prims;
    const double* beta;      // per primitive pair: the exponent of shell b
    const double* ao_scale;  // component scale per AO
    const double4* atoms;    // (x, y, z, Z)
    int n_atoms, n_ao;
    double *S, *T, *V;
};

template <int LA, int LB>
__global__ __launch_bounds__(128) void k_one_electron(OneArgs a) {
    constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.n_list * NA * NB) return;
    const int comp = tid % (NA * NB), pi = a.list[tid / (NA * NB)];
    const int ia = comp / NB, ib = comp % NB;
    const ShellPairDev sp = a.pairs[pi];
    // a shell with itself: (ia, ib) and (ib, ia) are one matrix element pair; ia >= ib writes both
    if (sp.ao_a == sp.ao_b && ia < ib) return;
    int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
#pragma unroll
    for (int c = 0; c < NA; ++c)
        if (c == ia) {
            ax = cart_x(LA, c);
            ay = cart_y(LA, c);
            az = cart_z(LA, c);
        }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c == ib) {
            bx = cart_x(LB, c);
            by = cart_y(LB, c);
            bz = cart_z(LB, c);
        }
    double s_acc = 0.0, t_acc = 0.0, v_acc = 0.0;
    for (int k = 0; k < sp.n_prim; ++k) {
        const PrimPair pp = a.prims[sp.prim_begin + k];
        const double beta = a.beta[sp.prim_begin + k];
        HermE<LA, LB + 2> ex, ey, ez;
        herm_e<LA, LB + 2>(pp.pax, pp.pbx, pp.inv2p, ex);
        herm_e<LA, LB + 2>(pp.pay, pp.pby, pp.inv2p, ey);
        herm_e<LA, LB + 2>(pp.paz, pp.pbz, pp.inv2p, ez);
        // this thread's component: Hermite slices E_t and the overlaps with b's power j - 2, j, j + 2
        double hx[L + 1], hy[L + 1], hz[L + 1], s0[3], sm[3], sp2[3];
        const int lb3[3] = {bx, by, bz};
#pragma unroll
        for (int t = 0; t <= L; ++t) hx[t] = hy[t] = hz[t] = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) s0[q] = sm[q] = sp2[q] = 0.0;
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) {
#pragma unroll
                for (int t = 0; t <= L; ++t) {
                    if (t > i + j) continue;
                    if (i == ax && j == bx) hx[t] = ex.e[i][j][t];
                    if (i == ay && j == by) hy[t] = ey.e[i][j][t];
                    if (i == az && j == bz) hz[t] = ez.e[i][j][t];
                }
                if (i == ax && j == bx) { s0[0] = ex.e[i][j][0]; sp2[0] = ex.e[i][j + 2][0]; if (j >= 2) sm[0] = ex.e[i][j - 2][0]; }
                if (i == ay && j == by) { s0[1] = ey.e[i][j][0]; sp2[1] = ey.e[i][j + 2][0]; if (j >= 2) sm[1] = ey.e[i][j - 2][0]; }
                if (i == az && j == bz) { s0[2] = ez.e[i][j][0]; sp2[2] = ez.e[i][j + 2][0]; if (j >= 2) sm[2] = ez.e[i][j - 2][0]; }
            }
        // d^2/dq^2 of q^j e^{-beta q^2}: j(j-1) q^{j-2} - 2 beta (2j+1) q^j + 4 beta^2 q^{j+2}
        double d2[3];
#pragma unroll
        for (int q = 0; q < 3; ++q)
            d2[q] = (double)(lb3[q] * (lb3[q] - 1)) * sm[q] - 2.0 * beta * (2 * lb3[q] + 1) * s0[q] +
                    4.0 * beta * beta * sp2[q];
        const double ovl = pp.k * (kPi / pp.p) * sqrt(kPi / pp.p);
        s_acc += ovl * s0[0] * s0[1] * s0[2];
        t_acc += ovl * -0.5 * (d2[0] * s0[1] * s0[2] + s0[0] * d2[1] * s0[2] + s0[0] * s0[1] * d2[2]);
        double vn = 0.0;
        for (int c = 0; c < a.n_atoms; ++c) {
            const double4 at = a.atoms[c];
            const double X = pp.px - at.x, Y = pp.py - at.y, Z = pp.pz - at.z;
            double F[L + 1], R[tri_count(L)];
            boys<L>(pp.p * (X * X + Y * Y + Z * Z), F);
            herm_r<L>(-2.0 * pp.p, X, Y, Z, F, R);
            double sum = 0.0;
#pragma unroll
            for (int t = 0; t <= L; ++t)
#pragma unroll
                for (int u = 0; u <= L; ++u)
#pragma unroll
                    for (int v = 0; v <= L; ++v)
                        if (t + u + v <= L) sum += hx[t] * hy[u] * hz[v] * R[hidx(t, u, v)];
            vn -= at.w * sum;
        }
        v_acc += pp.k * (2.0 * kPi / pp.p) * vn;
    }
    const int mu = sp.ao_a + ia, nu = sp.ao_b + ib;
    const double sc = a.ao_scale[mu] * a.ao_scale[nu];
    const size_t o1 = (size_t)mu * a.n_ao + nu, o2 = (size_t)nu * a.n_ao + mu;
    a.S[o1] = sc * s_acc;
    a.S[o2] = sc * s_acc;
    a.T[o1] = sc * t_acc;
    a.T[o2] = sc * t_acc;
    a.V[o1] = sc * v_acc;
    a.V[o2] = sc * v_acc;
}

// ---------------------------------------------------------------------------
// Electron repulsion columns of one class pair (bra LA >= LB, ket LC >= LD):
// one thread per (bra pair, ket pair, ket component); it forms (mu nu|lam sig)
// for every bra component pair and writes them at the packed index
// mu (mu + 1) / 2 + nu (mu >= nu) of output column col_off(ket) + ket component.
// DIAG: ket pair = bra pair and only (mu nu|mu nu) is written, to diag[pair index].
// ---------------------------------------------------------------------------
struct EriArgs {
    const ShellPairDev* pairs;
    const int* bra;  // bra pair ids of class (LA, LB)
    int n_bra;
    const int* ket;      // ket pair ids of class (LC, LD)
    const int* ket_col;  // first output column of each ket pair (n_c n_d columns each)
    int n_ket;
    const PrimPair* prims;
    const double* ao_scale;
    double* out;     // column-major, leading dimension ld (packed pair index)
    long long ld;
    double schwarz;  // quartets with q_bra q_ket below this are written as zeros
};

template <int LA, int LB, int LC, int LD, bool DIAG>
__global__ __launch_bounds__(128) void k_eri(EriArgs a) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n_work = DIAG ? (long long)a.n_bra * NC * ND : (long long)a.n_bra * a.n_ket * NC * ND;
    if (tid >= n_work) return;
    const int kc = (int)(tid % (NC * ND));
    const long long rest = tid / (NC * ND);
    const int bi = (int)(rest % a.n_bra);
    const int bra_id = a.bra[bi];
    const int ki = DIAG ? 0 : (int)(rest / a.n_bra);
    const int ket_id = DIAG ? bra_id : a.ket[ki];
    const ShellPairDev B = a.pairs[bra_id], K = a.pairs[ket_id];
    double out[NA * NB];
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) out[k] = 0.0;
    if (DIAG || B.q * K.q >= a.schwarz) {
        for (int kq = 0; kq < K.n_prim; ++kq) {
            const PrimPair qq = a.prims[K.prim_begin + kq];
            HermE<LC, LD> fx, fy, fz;
            herm_e<LC, LD>(qq.pax, qq.pbx, qq.inv2p, fx);
            herm_e<LC, LD>(qq.pay, qq.pby, qq.inv2p, fy);
            herm_e<LC, LD>(qq.paz, qq.pbz, qq.inv2p, fz);
            // this thread's ket component, with the (-1)^tau signs folded in
            double gx[LCD + 1], gy[LCD + 1], gz[LCD + 1];
#pragma unroll
            for (int t = 0; t <= LCD; ++t) gx[t] = gy[t] = gz[t] = 0.0;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int cd = 0; cd < ND; ++cd)
                    if (cc * ND + cd == kc) {
#pragma unroll
                        for (int t = 0; t <= LCD; ++t) {
                            const double sg = (t & 1) ? -1.0 : 1.0;
                            if (t <= cart_x(LC, cc) + cart_x(LD, cd)) gx[t] = sg * fx.e[cart_x(LC, cc)][cart_x(LD, cd)][t];
                            if (t <= cart_y(LC, cc) + cart_y(LD, cd)) gy[t] = sg * fy.e[cart_y(LC, cc)][cart_y(LD, cd)][t];
                            if (t <= cart_z(LC, cc) + cart_z(LD, cd)) gz[t] = sg * fz.e[cart_z(LC, cc)][cart_z(LD, cd)][t];
                        }
                    }
            for (int kp = 0; kp < B.n_prim; ++kp) {
                const PrimPair pp = a.prims[B.prim_begin + kp];
                const double pq = pp.p + qq.p, alpha = pp.p * qq.p / pq;
                const double X = pp.px - qq.px, Y = pp.py - qq.py, Z = pp.pz - qq.pz;
                double F[L + 1], R[tri_count(L)];
                boys<L>(alpha * (X * X + Y * Y + Z * Z), F);
                herm_r<L>(-2.0 * alpha, X, Y, Z, F, R);
                // ket contracted with R: Xh[tuv] over the bra Hermite triples
                double Xh[tri_count(LAB)];
#pragma unroll
                for (int t = 0; t <= LAB; ++t)
#pragma unroll
                    for (int u = 0; u <= LAB; ++u)
#pragma unroll
                        for (int v = 0; v <= LAB; ++v) {
                            if (t + u + v > LAB) continue;
                            double s = 0.0;
#pragma unroll
                            for (int t2 = 0; t2 <= LCD; ++t2)
#pragma unroll
                                for (int u2 = 0; u2 <= LCD; ++u2)
#pragma unroll
                                    for (int v2 = 0; v2 <= LCD; ++v2)
                                        if (t2 + u2 + v2 <= LCD)
                                            s += gx[t2] * gy[u2] * gz[v2] * R[hidx(t + t2, u + u2, v + v2)];
                            Xh[hidx(t, u, v)] = s;
                        }
                HermE<LA, LB> ex, ey, ez;
                herm_e<LA, LB>(pp.pax, pp.pbx, pp.inv2p, ex);
                herm_e<LA, LB>(pp.pay, pp.pby, pp.inv2p, ey);
                herm_e<LA, LB>(pp.paz, pp.pbz, pp.inv2p, ez);
                const double pref = kTwoPi52 / (pp.p * qq.p * sqrt(pq)) * pp.k * qq.k;
#pragma unroll
                for (int ca = 0; ca < NA; ++ca)
#pragma unroll
                    for (int cb = 0; cb < NB; ++cb) {
                        const int ax = cart_x(LA, ca), ay = cart_y(LA, ca), az = cart_z(LA, ca);
                        const int bx = cart_x(LB, cb), by = cart_y(LB, cb), bz = cart_z(LB, cb);
                        double s = 0.0;
#pragma unroll
                        for (int t = 0; t <= LAB; ++t)
#pragma unroll
                            for (int u = 0; u <= LAB; ++u)
#pragma unroll
                                for (int v = 0; v <= LAB; ++v)
                                    if (t <= ax + bx && u <= ay + by && v <= az + bz)
                                        s += ex.e[ax][bx][t] * ey.e[ay][by][u] * ez.e[az][bz][v] * Xh[hidx(t, u, v)];
                        out[ca * NB + cb] += pref * s;
                    }
            }
        }
    }
    const int lc = kc / ND, ld = kc % ND;
    const double sck = a.ao_scale[K.ao_a + lc] * a.ao_scale[K.ao_b + ld];
    if (DIAG) {
        // a pair of one shell with itself: (lc, ld) and (ld, lc) are one element; lc >= ld writes it
        if (B.ao_a == B.ao_b && lc < ld) return;
        const int mu = B.ao_a + lc, nu = B.ao_b + ld;
        const int hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
#pragma unroll
        for (int k = 0; k < NA * NB; ++k)
            if (k == kc) a.out[(long long)hi * (hi + 1) / 2 + lo] = out[k] * sck * sck;
        return;
    }
    double* col = a.out + (long long)(a.ket_col[ki] + kc) * a.ld;
#pragma unroll
    for (int ca = 0; ca < NA; ++ca)
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            const int mu = B.ao_a + ca, nu = B.ao_b + cb;
            const int hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
            col[(long long)hi * (hi + 1) / 2 + lo] = out[ca * NB + cb] * sck * a.ao_scale[mu] * a.ao_scale[nu];
        }
}

}  // namespace ints
}  // namespace chx
