/*
 * oracle.c — CPU FP64 restatement of CalChem's DFT exchange-correlation grid
 * quadrature (E_xc, V_xc and the integrated electron count N).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the calchem_amd
 * HIP path; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  Nothing in the product links or calls it.
 *
 * Reference semantics followed (paths relative to the CalChem tree):
 *   basis normalisation   crates/calchem-basis/src/normalization.rs:4-97,
 *                         crates/calchem-basis/src/lib.rs:135-160,462-520
 *   Lebedev rules         crates/calchem-dft/src/lebedev.rs:69-250
 *   molecular grid        crates/calchem-dft/src/grid.rs:58-173,268-419,732-828
 *   AO evaluation         crates/calchem-dft/src/ao_grid.rs:24-95,182-230
 *   functionals           crates/calchem-dft/src/xc.rs:159-828
 *   per-point RKS Vxc     crates/calchem-dft/src/vxc.rs:99-384
 *   cached BLAS-3 RKS     crates/calchem-dft/src/vxc.rs:429-716,801-985
 *   per-point UKS Vxc     crates/calchem-dft/src/vxc.rs:1141-1470
 *   overlap (for Tr PS)   crates/calchem-int/src/one_electron.rs (Obara-Saika)
 * N = sum_g w_g rho_g has no reference counterpart; it is pinned by N = Tr(PS).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   Lebedev: scipy.integrate.lebedev_rule point sets, monomial exactness up to
 *     the rule order, int exp(x) dOmega = 4 pi sinh 1 (lebedev.rs:279-325).
 *   Grid: closed-form Gaussian integrals (grid.rs:838-857), 3-centre sum
 *     (grid.rs:860-888), Becke partition of unity, diffuse normalisation.
 *   Basis/AO: <phi|phi> = 1 by quadrature and by the analytic overlap,
 *     AO gradients vs finite differences (ao_grid.rs:308-379), a brute-force
 *     numpy AO evaluation.
 *   Functionals: mpmath fixtures written from the literature formulas
 *     (tests/golden/gen_xc_fixtures.py), derivative FD (xc.rs:1212-1238),
 *     LYP one-electron zero (xc.rs:1240-1247).
 *   Vxc: V = dE/dP by finite differences (vxc.rs:2388-2434), symmetry,
 *     per-point == cached (vxc.rs:2604-2649), RKS == UKS(P/2,P/2)
 *     (vxc.rs:2436-2448), N ~ Tr(PS).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846
static const double ANG2BOHR = 1.8897261245650618; /* calchem-model/src/lib.rs:4 */
static const double RHO_CUTOFF = 1e-18;            /* xc.rs:13 */
static const double DENSITY_SCREEN = 1e-18;        /* vxc.rs:23 */
#define BATCH 128                                  /* vxc.rs:96 */

/* Integer power by binary exponentiation (the expansion Rust's powi uses). */
static double ipow(double x, int n) {
    double r = 1.0;
    int neg = n < 0;
    unsigned k = neg ? (unsigned)(-n) : (unsigned)n;
    for (;;) {
        if (k & 1u) r *= x;
        k >>= 1;
        if (!k) break;
        x *= x;
    }
    return neg ? 1.0 / r : r;
}

/* ------------------------------------------------------------------------ */
/* Basis normalisation: normalization.rs:4-97                               */
/* ------------------------------------------------------------------------ */
static double dfact_odd(int k) { /* (k)!! for odd k, 1 for k <= 0 */
    double a = 1.0;
    if (k <= 0) return 1.0;
    while (k > 0) { a *= (double)k; k -= 2; }
    return a;
}

double or_primitive_norm(double alpha, int l, int m, int n) {
    int L = l + m + n;
    double radial = pow(2.0 * alpha / OR_PI, 0.75);
    double num = ipow(4.0 * alpha, L);
    double den = dfact_odd(2 * l - 1) * dfact_odd(2 * m - 1) * dfact_odd(2 * n - 1);
    return radial * sqrt(num / den);
}

static double prim_self_overlap(double ai, double aj, int l, int m, int n) {
    double g = ai + aj;
    double ni = or_primitive_norm(ai, l, m, n), nj = or_primitive_norm(aj, l, m, n);
    double radial = pow(OR_PI / g, 1.5);
    double ang = 1.0;
    if (l > 0) ang *= dfact_odd(2 * l - 1) / ipow(2.0 * g, l);
    if (m > 0) ang *= dfact_odd(2 * m - 1) / ipow(2.0 * g, m);
    if (n > 0) ang *= dfact_odd(2 * n - 1) / ipow(2.0 * g, n);
    return ni * nj * radial * ang;
}

/* Contraction coefficients scaled so <phi|phi> = 1 for lmn = (L,0,0). */
int or_normalized_contraction(int n_prim, const double *e, const double *c, int L, double *out) {
    double s = 0.0;
    if (n_prim <= 0) return -1;
    for (int i = 0; i < n_prim; ++i)
        if (!(e[i] > 0.0) || !isfinite(e[i]) || !isfinite(c[i])) return -1;
    for (int i = 0; i < n_prim; ++i)
        for (int j = 0; j < n_prim; ++j) s += c[i] * c[j] * prim_self_overlap(e[i], e[j], L, 0, 0);
    if (!(s > 0.0) || !isfinite(s)) return -2;
    double sc = 1.0 / sqrt(s);
    for (int i = 0; i < n_prim; ++i) out[i] = c[i] * sc;
    return 0;
}

/* Cartesian component order of a shell: calchem-basis/src/lib.rs:135-160. */
static const int COMP_L0[1][3] = {{0, 0, 0}};
static const int COMP_L1[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
static const int COMP_L2[6][3] = {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
static const int COMP_L3[10][3] = {{3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1},
                                   {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}};
static const int *comp_of(int l, int c) {
    switch (l) {
    case 0: return COMP_L0[c];
    case 1: return COMP_L1[c];
    case 2: return COMP_L2[c];
    default: return COMP_L3[c];
    }
}
static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

int or_n_ao(const or_basis_t *b) {
    int n = 0;
    for (int s = 0; s < b->n_shells; ++s) n += ncart(b->l[s]);
    return n;
}

/* PrecomputedAoShell (ao_grid.rs:24-71). */
typedef struct {
    double center[3];
    double r2_cutoff, r_cutoff;
    int l, ao_offset, n_prim, nfn;
    double *alpha, *wnorm; /* exponent, c~ * N(alpha; L00) */
    double scale[10];      /* N(alpha0; lmn) / N(alpha0; L00) */
} shell_t;

typedef struct {
    int n_shells, n_ao;
    shell_t *sh;
    double *store;
} shells_t;

static int shells_build(const or_basis_t *b, shells_t *out) {
    int tot = 0;
    out->n_shells = b->n_shells;
    out->sh = (shell_t *)calloc((size_t)(b->n_shells ? b->n_shells : 1), sizeof(shell_t));
    for (int s = 0; s < b->n_shells; ++s) tot += b->n_prim[s];
    out->store = (double *)calloc((size_t)(2 * tot + 1), sizeof(double));
    int off = 0, ao = 0;
    for (int s = 0; s < b->n_shells; ++s) {
        shell_t *S = &out->sh[s];
        int np = b->n_prim[s], L = b->l[s];
        const double *e = b->exps + b->prim_start[s], *c = b->coefs + b->prim_start[s];
        S->alpha = out->store + off;
        S->wnorm = out->store + off + np;
        off += 2 * np;
        double *cn = (double *)malloc(sizeof(double) * (size_t)np);
        if (or_normalized_contraction(np, e, c, L, cn) != 0) { free(cn); return -1; }
        double amin = INFINITY;
        for (int p = 0; p < np; ++p) {
            S->alpha[p] = e[p];
            S->wnorm[p] = cn[p] * or_primitive_norm(e[p], L, 0, 0);
            if (e[p] < amin) amin = e[p];
        }
        free(cn);
        if (amin < 1e-12) amin = 1e-12;
        S->r2_cutoff = (40.0 + 4.0 * L) / amin;
        S->r_cutoff = sqrt(S->r2_cutoff);
        S->l = L;
        S->n_prim = np;
        S->nfn = ncart(L);
        S->ao_offset = ao;
        ao += S->nfn;
        for (int k = 0; k < 3; ++k) S->center[k] = b->center[3 * s + k];
        double lead = or_primitive_norm(e[0], L, 0, 0);
        for (int ic = 0; ic < S->nfn; ++ic) {
            const int *lmn = comp_of(L, ic);
            S->scale[ic] = (lead == 0.0) ? 1.0 : or_primitive_norm(e[0], lmn[0], lmn[1], lmn[2]) / lead;
        }
    }
    out->n_ao = ao;
    return 0;
}
static void shells_free(shells_t *s) { free(s->sh); free(s->store); }

/* pow_d0_d1 (ao_grid.rs:80-95) */
static inline void powd(double x, int l, double *p, double *d) {
    switch (l) {
    case 0: *p = 1.0; *d = 0.0; break;
    case 1: *p = x; *d = 1.0; break;
    case 2: *p = x * x; *d = 2.0 * x; break;
    case 3: *p = x * x * x; *d = 3.0 * x * x; break;
    default: *p = ipow(x, l); *d = (double)l * ipow(x, l - 1);
    }
}

/* pow_d0_d1_d2 (ao_grid.rs:80-95): x^l and its first two derivatives. */
static inline void pow012(double x, int l, double *p, double *d, double *d2) {
    switch (l) {
    case 0: *p = 1.0; *d = 0.0; *d2 = 0.0; break;
    case 1: *p = x; *d = 1.0; *d2 = 0.0; break;
    case 2: *p = x * x; *d = 2.0 * x; *d2 = 2.0; break;
    case 3: *p = x * x * x; *d = 3.0 * x * x; *d2 = 6.0 * x; break;
    default:
        *p = ipow(x, l); *d = (double)l * ipow(x, l - 1); *d2 = (double)l * (l - 1) * ipow(x, l - 2);
    }
}

/* One AO in the active list of a point. */
typedef struct { int mu; double val, g[3]; } aoval_t;

/* eval_aos_grad_filtered (ao_grid.rs:182-230) over the listed shells. */
static int eval_aos(const shells_t *S, const int *list, int nlist, const double r[3], aoval_t *out) {
    int n = 0;
    for (int k = 0; k < nlist; ++k) {
        const shell_t *sh = &S->sh[list ? list[k] : k];
        double dx = r[0] - sh->center[0], dy = r[1] - sh->center[1], dz = r[2] - sh->center[2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > sh->r2_cutoff) continue;
        double r0 = 0.0, r1 = 0.0;
        for (int p = 0; p < sh->n_prim; ++p) {
            double e = sh->wnorm[p] * exp(-sh->alpha[p] * r2);
            r0 += e;
            r1 -= 2.0 * sh->alpha[p] * e;
        }
        if (fabs(r0) < 1e-18 && fabs(r1) < 1e-18) continue;
        for (int ic = 0; ic < sh->nfn; ++ic) {
            const int *c = comp_of(sh->l, ic);
            double s = sh->scale[ic], px, dpx, py, dpy, pz, dpz;
            powd(dx, c[0], &px, &dpx);
            powd(dy, c[1], &py, &dpy);
            powd(dz, c[2], &pz, &dpz);
            double m = px * py * pz;
            double val = s * m * r0;
            double gx = s * (dpx * py * pz * r0 + m * dx * r1);
            double gy = s * (px * dpy * pz * r0 + m * dy * r1);
            double gz = s * (px * py * dpz * r0 + m * dz * r1);
            if (fabs(val) > 1e-16 || fabs(gx) > 1e-16 || fabs(gy) > 1e-16 || fabs(gz) > 1e-16) {
                out[n].mu = sh->ao_offset + ic;
                out[n].val = val;
                out[n].g[0] = gx;
                out[n].g[1] = gy;
                out[n].g[2] = gz;
                ++n;
            }
        }
    }
    return n;
}

void or_eval_aos(const or_basis_t *b, const double r[3], double *val, double *grad) {
    shells_t S;
    if (shells_build(b, &S) != 0) return;
    aoval_t *buf = (aoval_t *)malloc(sizeof(aoval_t) * (size_t)(S.n_ao + 1));
    memset(val, 0, sizeof(double) * (size_t)S.n_ao);
    memset(grad, 0, sizeof(double) * 3 * (size_t)S.n_ao);
    int n = eval_aos(&S, NULL, S.n_shells, r, buf);
    for (int k = 0; k < n; ++k) {
        val[buf[k].mu] = buf[k].val;
        for (int q = 0; q < 3; ++q) grad[3 * buf[k].mu + q] = buf[k].g[q];
    }
    free(buf);
    shells_free(&S);
}

/* ------------------------------------------------------------------------ */
/* Lebedev-Laikov rules: lebedev.rs:69-250 (orbit generators + parameters   */
/* of Lebedev & Laikov, Doklady Math. 59, 477 (1999)).                      */
/* ------------------------------------------------------------------------ */
typedef struct { int kind; double a, b, v; } orbit_t;

static const orbit_t LEB26[] = {{1, 0, 0, 0.4761904761904762e-1}, {2, 0, 0, 0.3809523809523810e-1},
                                {3, 0, 0, 0.3214285714285714e-1}};
static const orbit_t LEB50[] = {{1, 0, 0, 0.1269841269841270e-1}, {2, 0, 0, 0.2257495590828924e-1},
                                {3, 0, 0, 0.2109375000000000e-1},
                                {4, 0.3015113445777636, 0, 0.2017333553791887e-1}};
static const orbit_t LEB110[] = {{1, 0, 0, 0.3828270494937162e-2}, {3, 0, 0, 0.9793737512487512e-2},
                                 {4, 0.1851156353447362, 0, 0.8211737283191111e-2},
                                 {4, 0.6904210483822922, 0, 0.9942814891178103e-2},
                                 {4, 0.3956894730559419, 0, 0.9595471336070963e-2},
                                 {5, 0.4783690288121502, 0, 0.9694996361663028e-2}};
static const orbit_t LEB194[] = {{1, 0, 0, 0.1782340447244611e-2}, {2, 0, 0, 0.5716905949977102e-2},
                                 {3, 0, 0, 0.5573383178848738e-2},
                                 {4, 0.6712973442695226, 0, 0.5608704082587997e-2},
                                 {4, 0.2892465627575439, 0, 0.5158237711805383e-2},
                                 {4, 0.4446933178717437, 0, 0.5518771467273614e-2},
                                 {4, 0.1299335447650067, 0, 0.4106777028169394e-2},
                                 {5, 0.3457702197611283, 0, 0.5051846064614808e-2},
                                 {6, 0.1590417105383530, 0.8360360154824589, 0.5530248916233094e-2}};
static const orbit_t LEB302[] = {{1, 0, 0, 0.8545911725128148e-3}, {3, 0, 0, 0.3599119285025571e-2},
                                 {4, 0.3515640345570105, 0, 0.3449788424305883e-2},
                                 {4, 0.6566329410219612, 0, 0.3604822601419882e-2},
                                 {4, 0.4729054132581005, 0, 0.3576729661743367e-2},
                                 {4, 0.9618308522614784e-1, 0, 0.2352101413689164e-2},
                                 {4, 0.2219645236294178, 0, 0.3108953122413675e-2},
                                 {4, 0.7011766416089545, 0, 0.3650045807677255e-2},
                                 {5, 0.2644152887060663, 0, 0.2982344963171804e-2},
                                 {5, 0.5718955891878961, 0, 0.3600820932216460e-2},
                                 {6, 0.2510034751770465, 0.8000727494073952, 0.3571540554273387e-2},
                                 {6, 0.1233548532583327, 0.4127724083168531, 0.3392312205006170e-2}};
static const orbit_t LEB590[] = {
    {1, 0, 0, 0.3095121295306187e-3}, {3, 0, 0, 0.1852379698597489e-2},
    {4, 0.7040954938227469, 0, 0.1871790639277744e-2}, {4, 0.6807744066455243, 0, 0.1858812585438317e-2},
    {4, 0.6372546939258752, 0, 0.1852028828296213e-2}, {4, 0.5044419707800358, 0, 0.1846715956151242e-2},
    {4, 0.4215761784010967, 0, 0.1818471778162769e-2}, {4, 0.3317920736472123, 0, 0.1749564657281154e-2},
    {4, 0.2384736701421887, 0, 0.1617210647254411e-2}, {4, 0.1459036449157763, 0, 0.1384737234851692e-2},
    {4, 0.06095034115507196, 0, 0.9764331165051050e-3}, {5, 0.6116843442009876, 0, 0.1857161196774078e-2},
    {5, 0.3964755348199858, 0, 0.1705153996395864e-2}, {5, 0.1724782009907724, 0, 0.1300321685886048e-2},
    {6, 0.5610263808622060, 0.3518280927733519, 0.1842866472905286e-2},
    {6, 0.4742392842551980, 0.2634716655937950, 0.1802658934377451e-2},
    {6, 0.5984126497885380, 0.1816640840360209, 0.1849830560443660e-2},
    {6, 0.3791035407695563, 0.1720795225656878, 0.1713904507106709e-2},
    {6, 0.2778673190586244, 0.08213021581932511, 0.1555213603396808e-2},
    {6, 0.5033564271075117, 0.08999205842074875, 0.1802239128008525e-2}};

static int put(double *o, int n, double x, double y, double z, double w) {
    o[4 * n] = x; o[4 * n + 1] = y; o[4 * n + 2] = z; o[4 * n + 3] = w;
    return n + 1;
}
static int put_octants(double *o, int n, double x, double y, double z, double w) {
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2) n = put(o, n, sx * x, sy * y, sz * z, w);
    return n;
}
static int expand_orbit(double *o, int n, const orbit_t *ob) {
    double w = 4.0 * OR_PI * ob->v, a = ob->a, b = ob->b;
    switch (ob->kind) {
    case 1:
        for (int s = -1; s <= 1; s += 2) {
            n = put(o, n, s, 0, 0, w);
            n = put(o, n, 0, s, 0, w);
            n = put(o, n, 0, 0, s, w);
        }
        break;
    case 2: {
        double s = sqrt(0.5);
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                double s1 = i ? s : -s, s2 = j ? s : -s;
                n = put(o, n, s1, s2, 0, w);
                n = put(o, n, s1, 0, s2, w);
                n = put(o, n, 0, s1, s2, w);
            }
        break;
    }
    case 3: {
        double s = sqrt(1.0 / 3.0);
        n = put_octants(o, n, s, s, s, w);
        break;
    }
    case 4: {
        double m = sqrt(1.0 - 2.0 * a * a);
        n = put_octants(o, n, a, a, m, w);
        n = put_octants(o, n, a, m, a, w);
        n = put_octants(o, n, m, a, a, w);
        break;
    }
    case 5: {
        double q = sqrt(1.0 - a * a);
        double ut[2][2] = {{a, q}, {q, a}};
        for (int k = 0; k < 2; ++k)
            for (int su = -1; su <= 1; su += 2)
                for (int st = -1; st <= 1; st += 2) {
                    double u = su * ut[k][0], t = st * ut[k][1];
                    n = put(o, n, u, t, 0, w);
                    n = put(o, n, u, 0, t, w);
                    n = put(o, n, 0, u, t, w);
                }
        break;
    }
    case 6: {
        double c = sqrt(1.0 - a * a - b * b);
        double perm[6][3] = {{a, b, c}, {b, a, c}, {c, a, b}, {a, c, b}, {c, b, a}, {b, c, a}};
        for (int k = 0; k < 6; ++k) n = put_octants(o, n, perm[k][0], perm[k][1], perm[k][2], w);
        break;
    }
    }
    return n;
}

int or_lebedev(int n_ang, double *xyzw) {
    const orbit_t *t;
    int no;
    switch (n_ang) {
    case 26: t = LEB26; no = 3; break;
    case 50: t = LEB50; no = 4; break;
    case 110: t = LEB110; no = 6; break;
    case 194: t = LEB194; no = 9; break;
    case 302: t = LEB302; no = 12; break;
    case 590: t = LEB590; no = 20; break;
    default: return -1;
    }
    int n = 0;
    for (int k = 0; k < no; ++k) n = expand_orbit(xyzw, n, &t[k]);
    return n;
}

/* ------------------------------------------------------------------------ */
/* Molecular grid: grid.rs:58-173 (radii, radial counts, pruning),          */
/* 268-419 (Becke k=3 fast partition), 732-828 (point generation).          */
/* ------------------------------------------------------------------------ */
double or_bragg_slater_bohr(int z) {
    double a;
    if (z == 1 || z == 2) a = 0.35;
    else if (z == 3) a = 1.45;
    else if (z == 4) a = 1.05;
    else if (z == 5) a = 0.85;
    else if (z == 6) a = 0.70;
    else if (z == 7) a = 0.65;
    else if (z == 8) a = 0.60;
    else if (z == 9) a = 0.50;
    else if (z == 10) a = 0.45;
    else if (z == 11) a = 1.80;
    else if (z == 12) a = 1.50;
    else if (z == 13) a = 1.25;
    else if (z == 14) a = 1.10;
    else if (z >= 15 && z <= 18) a = 1.00;
    else if (z == 19) a = 2.20;
    else if (z == 20) a = 1.80;
    else if (z >= 21 && z <= 30) a = 1.40;
    else if (z == 31) a = 1.30;
    else if (z == 32) a = 1.25;
    else if (z >= 33 && z <= 35) a = 1.15;
    else if (z == 36) a = 1.10;
    else if (z == 37) a = 2.35;
    else if (z == 38) a = 2.00;
    else if (z >= 39 && z <= 49) a = 1.55;
    else if (z == 50 || z == 51) a = 1.45;
    else if (z == 52 || z == 53) a = 1.40;
    else if (z == 54) a = 1.35;
    else if (z == 55) a = 2.60;
    else if (z == 56) a = 2.15;
    else a = 1.75;
    return a * ANG2BOHR;
}

static double becke_asym(int zc, int zd) {
    if (zc == zd) return 0.0;
    double chi = or_bragg_slater_bohr(zc) / or_bragg_slater_bohr(zd);
    double u = (chi - 1.0) / (chi + 1.0);
    double a = u / (u * u - 1.0);
    return a < -0.5 ? -0.5 : (a > 0.5 ? 0.5 : a);
}

static int radial_count(int z, int level) {
    switch (level) {
    case OR_COARSE: return z <= 2 ? 25 : (z <= 18 ? 35 : 45);
    case OR_MEDIUM: return z <= 2 ? 35 : (z <= 18 ? 50 : 65);
    case OR_FINE: return z <= 2 ? 55 : (z <= 18 ? 75 : 90);
    default: return 200;
    }
}

static int pruned_nang(double r, double rbs, int level) {
    double ratio = r / (rbs > 0.3 ? rbs : 0.3);
    switch (level) {
    case OR_COARSE: return ratio < 0.25 ? 26 : ratio < 0.75 ? 50 : ratio < 4.5 ? 110 : 50;
    case OR_MEDIUM:
        return ratio < 0.20 ? 26 : ratio < 0.55 ? 50 : ratio < 1.10 ? 110 : ratio < 4.50 ? 194 : 110;
    case OR_FINE:
        return ratio < 0.18 ? 50 : ratio < 0.50 ? 110 : ratio < 0.95 ? 194 : ratio < 5.00 ? 302 : 194;
    default: return 590;
    }
}

typedef struct {
    int n;
    const double *xyz;
    double *inv_rij, *a_ij;
} becke_tab_t;

static void becke_tab_init(becke_tab_t *t, int n, const int *z, const double *xyz) {
    t->n = n;
    t->xyz = xyz;
    t->inv_rij = (double *)calloc((size_t)n * (size_t)n + 1, sizeof(double));
    t->a_ij = (double *)calloc((size_t)n * (size_t)n + 1, sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                   dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            double rij = sqrt(dx * dx + dy * dy + dz * dz);
            double inv = rij >= 1e-14 ? 1.0 / rij : 0.0;
            double a = becke_asym(z[i], z[j]);
            t->inv_rij[i * n + j] = t->inv_rij[j * n + i] = inv;
            t->a_ij[i * n + j] = a;
            t->a_ij[j * n + i] = -a;
        }
}
static void becke_tab_free(becke_tab_t *t) { free(t->inv_rij); free(t->a_ij); }

static inline double becke_step(double nu) {
    if (nu <= -1.0) return 1.0;
    if (nu >= 1.0) return 0.0;
    double f1 = 0.5 * nu * (3.0 - nu * nu);
    double f2 = 0.5 * f1 * (3.0 - f1 * f1);
    double f3 = 0.5 * f2 * (3.0 - f2 * f2);
    return 0.5 * (1.0 - f3);
}
static inline double clamp1(double x) { return x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x); }

/* becke_partition_weight_fast (grid.rs:338-419). scratch: 2n doubles + n ints. */
static double becke_fast(const becke_tab_t *t, double *dist, double *s_hb, int *active, int host,
                         const double r[3]) {
    int n = t->n;
    if (n <= 1) return 1.0;
    for (int a = 0; a < n; ++a) {
        double dx = r[0] - t->xyz[3 * a], dy = r[1] - t->xyz[3 * a + 1], dz = r[2] - t->xyz[3 * a + 2];
        dist[a] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    double rh = dist[host], ph = 1.0;
    int nact = 0;
    for (int b = 0; b < n; ++b) {
        if (b == host) continue;
        double inv = t->inv_rij[host * n + b];
        if (inv == 0.0) continue;
        double mu = clamp1((rh - dist[b]) * inv);
        double nu = mu + t->a_ij[host * n + b] * (1.0 - mu * mu);
        double s = becke_step(nu);
        ph *= s;
        if (ph < 1e-20) return 0.0;
        s_hb[b] = s;
        if (1.0 - s > 1e-17) active[nact++] = b;
    }
    if (nact == 0) return 1.0;
    double sum = ph;
    for (int k = 0; k < nact; ++k) {
        int a = active[k];
        double ra = dist[a], pa = 1.0 - s_hb[a];
        if (pa < 1e-20) continue;
        for (int b = 0; b < n; ++b) {
            if (b == host || b == a) continue;
            double inv = t->inv_rij[a * n + b];
            if (inv == 0.0) continue;
            double mu = clamp1((ra - dist[b]) * inv);
            double nu = mu + t->a_ij[a * n + b] * (1.0 - mu * mu);
            pa *= becke_step(nu);
            if (pa < 1e-20) break;
        }
        sum += pa;
    }
    return sum <= 0.0 ? 0.0 : ph / sum;
}

double or_becke_weight_fast(int n_atoms, const int *z, const double *xyz, int host, const double r[3]) {
    becke_tab_t t;
    becke_tab_init(&t, n_atoms, z, xyz);
    double *d = (double *)malloc(sizeof(double) * 2 * (size_t)(n_atoms + 1));
    int *act = (int *)malloc(sizeof(int) * (size_t)(n_atoms + 1));
    double w = becke_fast(&t, d, d + n_atoms + 1, act, host, r);
    free(d);
    free(act);
    becke_tab_free(&t);
    return w;
}

long or_build_grid(int n_atoms, const int *z, const double *xyz, int level, int n_rad_x, int n_ang_x,
                   int radial_min, double *out, long max_out) {
    if (n_atoms <= 0 || !z || !xyz) return -1;
    if (level == OR_LEVEL_EXPLICIT && n_rad_x <= 0) return -1;
    double *leb[6];
    int lebn[6];
    const int lebsizes[6] = {26, 50, 110, 194, 302, 590};
    for (int k = 0; k < 6; ++k) {
        leb[k] = (double *)malloc(sizeof(double) * 4 * (size_t)lebsizes[k]);
        lebn[k] = or_lebedev(lebsizes[k], leb[k]);
    }
    becke_tab_t t;
    becke_tab_init(&t, n_atoms, z, xyz);
    double *d = (double *)malloc(sizeof(double) * 2 * (size_t)(n_atoms + 1));
    int *act = (int *)malloc(sizeof(int) * (size_t)(n_atoms + 1));
    long cnt = 0;
    int bad = 0;
    for (int A = 0; A < n_atoms && !bad; ++A) {
        double rbs = or_bragg_slater_bohr(z[A]);
        int nr;
        if (level == OR_LEVEL_EXPLICIT) nr = n_rad_x;
        else {
            int base_min = (level == OR_FINE) ? 99 : 0;
            nr = radial_count(z[A], level);
            if (radial_min > nr) nr = radial_min;
            if (base_min > nr) nr = base_min;
        }
        double m = (double)(nr + 1);
        for (int i = 1; i <= nr; ++i) {
            double ti = (double)i, mt = m - ti;
            double r = rbs * (ti * ti) / (mt * mt);
            double wr = 2.0 * ipow(rbs, 3) * m * ipow(ti, 5) / ipow(mt, 7);
            int na = (level == OR_LEVEL_EXPLICIT) ? n_ang_x : pruned_nang(r, rbs, level);
            int li = -1;
            for (int k = 0; k < 6; ++k)
                if (lebsizes[k] == na) li = k;
            if (li < 0) { bad = 1; break; }
            for (int g = 0; g < lebn[li]; ++g) {
                const double *u = leb[li] + 4 * g;
                double c[3] = {xyz[3 * A] + r * u[0], xyz[3 * A + 1] + r * u[1], xyz[3 * A + 2] + r * u[2]};
                double raw = wr * u[3];
                double bw = becke_fast(&t, d, d + n_atoms + 1, act, A, c);
                double wt = raw * bw;
                if (wt > 1e-18) {
                    if (out && cnt < max_out) {
                        double *o = out + 6 * cnt;
                        o[0] = c[0]; o[1] = c[1]; o[2] = c[2]; o[3] = wt; o[4] = raw; o[5] = (double)A;
                    }
                    ++cnt;
                }
            }
        }
    }
    for (int k = 0; k < 6; ++k) free(leb[k]);
    free(d);
    free(act);
    becke_tab_free(&t);
    if (bad) return -1;
    if (!out || cnt > max_out) return -cnt;
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* XC functionals: xc.rs:159-828                                            */
/* ------------------------------------------------------------------------ */
typedef struct { double f, va, vb, gaa, gab, gbb; } xc_t;
static void xc_add(xc_t *o, double s, xc_t x) {
    o->f += s * x.f; o->va += s * x.va; o->vb += s * x.vb;
    o->gaa += s * x.gaa; o->gab += s * x.gab; o->gbb += s * x.gbb;
}

static void slater1(double r, double *f, double *v) { /* xc.rs:229-240 */
    if (r < RHO_CUTOFF) { *f = 0; *v = 0; return; }
    double cx = 0.75 * pow(6.0 / OR_PI, 1.0 / 3.0), r13 = cbrt(r);
    *f = -cx * r * r13;
    *v = -(4.0 / 3.0) * cx * r13;
}
static xc_t slater_x(double ra, double rb) {
    xc_t o = {0, 0, 0, 0, 0, 0};
    double fa, va, fb, vb;
    slater1(ra, &fa, &va);
    slater1(rb, &fb, &vb);
    o.f = fa + fb; o.va = va; o.vb = vb;
    return o;
}

typedef struct { double a, x0, b, c; } vwnch_t;
typedef struct { vwnch_t para, ferro, stiff; } vwn_t;
static const vwn_t VWN5 = {{0.0310907, -0.10498, 3.72744, 12.9352},
                           {0.01554535, -0.32500, 7.06042, 18.0578},
                           {-0.016886863940389905, -0.00475840, 1.13107, 13.0045}};
static const vwn_t VWNRPA = {{0.0310907, -0.409286, 13.0720, 42.7198},
                             {0.01554535, -0.743294, 20.1231, 101.578},
                             {-0.016886863940389905, -0.228344, 1.06835, 11.4813}};

static void vwn_ch(double x, const vwnch_t *c, double *eps, double *rsd) { /* xc.rs:316-331 */
    double xx = x * x + c->b * x + c->c;
    double xx0 = c->x0 * c->x0 + c->b * c->x0 + c->c;
    double q = sqrt(4.0 * c->c - c->b * c->b);
    double at = atan(q / (2.0 * x + c->b));
    double xm = x - c->x0;
    *eps = c->a * (log(x * x / xx) + (2.0 * c->b / q) * at -
                   (c->b * c->x0 / xx0) * (log(xm * xm / xx) + (2.0 * (c->b + 2.0 * c->x0) / q) * at));
    *rsd = c->a * (c->c * xm - c->b * x * c->x0) / (xm * xx);
}

static void spin_interp(double z, double *fz, double *dfz, double *fdd0) { /* xc.rs:333-345 */
    double c2 = cbrt(2.0) - 1.0, den = 2.0 * c2;
    double zp = 1.0 + z > 0 ? 1.0 + z : 0.0, zm = 1.0 - z > 0 ? 1.0 - z : 0.0;
    double zp13 = cbrt(zp), zm13 = cbrt(zm);
    *fz = (zp * zp13 + zm * zm13 - 2.0) / den;
    *dfz = (4.0 / (3.0 * den)) * (zp13 - zm13);
    *fdd0 = 4.0 / (9.0 * c2);
}

static xc_t vwn_c(double ra, double rb, const vwn_t *p) { /* xc.rs:347-381 */
    xc_t o = {0, 0, 0, 0, 0, 0};
    double rho = ra + rb;
    if (rho < RHO_CUTOFF) return o;
    double rs = cbrt(3.0 / (4.0 * OR_PI * rho)), x = sqrt(rs);
    double zeta = clamp1((ra - rb) / rho);
    double ep, rdep, ef, rdef, ac, rdac, fz, dfz, fdd0;
    vwn_ch(x, &p->para, &ep, &rdep);
    vwn_ch(x, &p->ferro, &ef, &rdef);
    vwn_ch(x, &p->stiff, &ac, &rdac);
    spin_interp(zeta, &fz, &dfz, &fdd0);
    double z2 = zeta * zeta, z3 = z2 * zeta, z4 = z2 * z2, omz4 = 1.0 - z4, inv = 1.0 / fdd0;
    double dec = ef - ep, rddec = rdef - rdep;
    double ec = ep + ac * fz * inv * omz4 + dec * fz * z4;
    double rsdec = rdep + rdac * fz * inv * omz4 + rddec * fz * z4;
    double decdz = ac * inv * (dfz * omz4 - 4.0 * z3 * fz) + dec * (dfz * z4 + 4.0 * z3 * fz);
    double base = ec - (1.0 / 3.0) * rsdec;
    o.f = rho * ec;
    o.va = base + (1.0 - zeta) * decdz;
    o.vb = base - (1.0 + zeta) * decdz;
    return o;
}

static void b88_1(double r, double g, double *f, double *vr, double *vg) { /* xc.rs:387-412 */
    if (r < RHO_CUTOFF) { *f = *vr = *vg = 0; return; }
    const double BETA = 0.0042;
    double r13 = cbrt(r), r43 = r * r13;
    double s2 = (g > 0 ? g : 0.0) / (r43 * r43), s = sqrt(s2), as = asinh(s);
    double den = 1.0 + 6.0 * BETA * s * as;
    double gg = -BETA * s2 / den;
    double sq = sqrt(1.0 + s2);
    double h = -BETA * (2.0 + 6.0 * BETA * s * as - 6.0 * BETA * s2 / sq) / (den * den);
    *f = r43 * gg;
    *vr = (4.0 / 3.0) * r13 * (gg - s2 * h);
    *vg = 0.5 * h / r43;
}
static xc_t b88_x(double ra, double rb, double gaa, double gbb) {
    xc_t o = {0, 0, 0, 0, 0, 0};
    double fa, va, ga, fb, vb, gb;
    b88_1(ra, gaa, &fa, &va, &ga);
    b88_1(rb, gbb, &fb, &vb, &gb);
    o.f = fa + fb; o.va = va; o.vb = vb; o.gaa = ga; o.gbb = gb;
    return o;
}

/* Miehlich bracket S and dS/d(ra): xc.rs:433-491 */
static void lyp_bracket(double ra, double rb, double delta, double ddelta, double gaa, double gab,
                        double gbb, double *S, double *dS, double *Aaa, double *Aab, double *Abb) {
    double rho = ra + rb, rho2 = rho * rho;
    double cf = 0.3 * pow(3.0 * OR_PI * OR_PI, 2.0 / 3.0);
    double k0 = pow(2.0, 11.0 / 3.0) * cf;
    double ra83 = ra > 0 ? pow(ra, 8.0 / 3.0) : 0.0, rb83 = rb > 0 ? pow(rb, 8.0 / 3.0) : 0.0;
    double stf = k0 * ra * rb * (ra83 + rb83);
    double dstf = k0 * ((11.0 / 3.0) * ra83 * rb + rb * rb83);
    double cg = ra * rb * (47.0 - 7.0 * delta) / 18.0 - (2.0 / 3.0) * rho2;
    double dcg = rb * (47.0 - 7.0 * delta) / 18.0 + ra * rb * (-7.0 * ddelta) / 18.0 - (4.0 / 3.0) * rho;
    double qaa = ra * ra * rb / rho, dqaa = (2.0 * ra * rb * rho - ra * ra * rb) / rho2;
    double eaa = -ra * rb * (2.5 - delta / 18.0) - ((delta - 11.0) / 9.0) * qaa + ((2.0 / 3.0) * rho2 - rb * rb);
    double deaa = -rb * (2.5 - delta / 18.0) + ra * rb * (ddelta / 18.0) - (ddelta / 9.0) * qaa -
                  ((delta - 11.0) / 9.0) * dqaa + (4.0 / 3.0) * rho;
    double qbb = ra * rb * rb / rho, dqbb = (rb * rb * rho - ra * rb * rb) / rho2;
    double ebb = -ra * rb * (2.5 - delta / 18.0) - ((delta - 11.0) / 9.0) * qbb + ((2.0 / 3.0) * rho2 - ra * ra);
    double debb = -rb * (2.5 - delta / 18.0) + ra * rb * (ddelta / 18.0) - (ddelta / 9.0) * qbb -
                  ((delta - 11.0) / 9.0) * dqbb + (4.0 / 3.0) * rho - 2.0 * ra;
    *Aaa = cg + eaa;
    *Aab = 2.0 * cg;
    *Abb = cg + ebb;
    *S = stf + *Aaa * gaa + *Aab * gab + *Abb * gbb;
    *dS = dstf + (dcg + deaa) * gaa + (2.0 * dcg) * gab + (dcg + debb) * gbb;
}

static xc_t lyp_c(double ra, double rb, double gaa, double gab, double gbb) { /* xc.rs:493-547 */
    xc_t o = {0, 0, 0, 0, 0, 0};
    double rho = ra + rb;
    if (rho < RHO_CUTOFF) return o;
    const double A = 0.04918, B = 0.132, C = 0.2533, D = 0.349;
    double r13 = pow(rho, -1.0 / 3.0);
    double den = 1.0 + D * r13;
    double g0 = -4.0 * A / (rho * den);
    double dg0 = -g0 * (1.0 + (2.0 / 3.0) * D * r13) / (rho * den);
    if (C * r13 > 500.0) {
        o.f = g0 * ra * rb;
        o.va = dg0 * ra * rb + g0 * rb;
        o.vb = dg0 * ra * rb + g0 * ra;
        return o;
    }
    double t0 = g0 * ra * rb, dt0a = dg0 * ra * rb + g0 * rb, dt0b = dg0 * ra * rb + g0 * ra;
    double om = exp(-C * r13) / den * pow(rho, -11.0 / 3.0);
    double delta = C * r13 + D * r13 / den;
    double dom = om * (delta - 11.0) / (3.0 * rho);
    double r43i = r13 / rho;
    double ddelta = -(1.0 / 3.0) * r43i * (C + D / (den * den));
    double S, dSa, Aaa, Aab, Abb, Ss, dSb, t1, t2, t3;
    lyp_bracket(ra, rb, delta, ddelta, gaa, gab, gbb, &S, &dSa, &Aaa, &Aab, &Abb);
    lyp_bracket(rb, ra, delta, ddelta, gbb, gab, gaa, &Ss, &dSb, &t1, &t2, &t3);
    double ab = A * B;
    o.f = t0 - ab * om * S;
    o.va = dt0a - ab * (dom * S + om * dSa);
    o.vb = dt0b - ab * (dom * S + om * dSb);
    o.gaa = -ab * om * Aaa;
    o.gab = -ab * om * Aab;
    o.gbb = -ab * om * Abb;
    return o;
}

static void pbex_1(double rs_, double gss, double *f, double *vr, double *vg) { /* xc.rs:575-606 */
    const double kappa = 0.804, mu = 0.2195149727645171;
    if (rs_ < RHO_CUTOFF) { *f = *vr = *vg = 0; return; }
    double ru = 2.0 * rs_, gu = 4.0 * (gss > 0 ? gss : 0.0);
    double cxu = -0.75 * pow(3.0 / OR_PI, 1.0 / 3.0);
    double r13 = cbrt(ru), r43 = ru * r13, exl = cxu * r43;
    double kf = cbrt(3.0 * OR_PI * OR_PI * ru);
    double dens2 = 4.0 * kf * kf * ru * ru, s2 = gu / dens2;
    double opm = 1.0 + (mu / kappa) * s2;
    double fx = 1.0 + kappa - kappa / opm, dfx = mu / (opm * opm);
    double fu = exl * fx;
    double dfr = (4.0 / 3.0) * cxu * r13 * fx + exl * dfx * (-(8.0 / 3.0) * s2 / ru);
    double dfg = exl * dfx / dens2;
    *f = 0.5 * fu;
    *vr = dfr;
    *vg = 2.0 * dfg;
}
static xc_t pbe_x(double ra, double rb, double gaa, double gbb) {
    xc_t o = {0, 0, 0, 0, 0, 0};
    double fa, va, ga, fb, vb, gb;
    pbex_1(ra, gaa, &fa, &va, &ga);
    pbex_1(rb, gbb, &fb, &vb, &gb);
    o.f = fa + fb; o.va = va; o.vb = vb; o.gaa = ga; o.gbb = gb;
    return o;
}

typedef struct { double a, a1, b1, b2, b3, b4; } pw_t;
static const pw_t PW_G0 = {0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
static const pw_t PW_G1 = {0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517};
static const pw_t PW_GA = {0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671};

static void pw92g(double rs, const pw_t *c, double *g, double *dg) { /* xc.rs:681-695 */
    double x = sqrt(rs), x2 = rs, x3 = x2 * x, x4 = x2 * x2;
    double q0 = -2.0 * c->a * (1.0 + c->a1 * rs);
    double q1 = 2.0 * c->a * (c->b1 * x + c->b2 * x2 + c->b3 * x3 + c->b4 * x4);
    double dq1 = c->a * (c->b1 / x + 2.0 * c->b2 + 3.0 * c->b3 * x + 4.0 * c->b4 * x2);
    double ln = log(1.0 + 1.0 / q1);
    *g = q0 * ln;
    *dg = -2.0 * c->a * c->a1 * ln - q0 * dq1 / (q1 * (q1 + 1.0));
}

static xc_t pbe_c(double ra, double rb, double gaa, double gab, double gbb) { /* xc.rs:712-828 */
    const double beta = 0.06672455060314922, gam = 0.0310906908696549;
    xc_t o = {0, 0, 0, 0, 0, 0};
    double rho = ra + rb;
    if (rho < RHO_CUTOFF) return o;
    double gt = gaa + 2.0 * gab + gbb;
    if (gt < 0) gt = 0;
    double rs = cbrt(3.0 / (4.0 * OR_PI * rho)), drs = -rs / (3.0 * rho);
    double zeta = clamp1((ra - rb) / rho);
    double e0, de0, e1, de1, mac, dmac, fz, dfz, fdd0;
    pw92g(rs, &PW_G0, &e0, &de0);
    pw92g(rs, &PW_G1, &e1, &de1);
    pw92g(rs, &PW_GA, &mac, &dmac);
    double ac = -mac, dac = -dmac;
    spin_interp(zeta, &fz, &dfz, &fdd0);
    double inv = 1.0 / fdd0, z2 = zeta * zeta, z3 = z2 * zeta, z4 = z2 * z2, omz4 = 1.0 - z4;
    double ec = e0 + ac * fz * inv * omz4 + (e1 - e0) * fz * z4;
    double decrs = de0 + dac * fz * inv * omz4 + (de1 - de0) * fz * z4;
    double decr = decrs * drs;
    double decz = ac * inv * (dfz * omz4 - 4.0 * z3 * fz) + (e1 - e0) * (dfz * z4 + 4.0 * z3 * fz);
    if (gt < 1e-36) {
        o.f = rho * ec;
        o.va = ec + rho * decr + (1.0 - zeta) * decz;
        o.vb = ec + rho * decr - (1.0 + zeta) * decz;
        return o;
    }
    double b0 = beta / gam;
    double zp = 1.0 + zeta > 1e-24 ? 1.0 + zeta : 1e-24, zm = 1.0 - zeta > 1e-24 ? 1.0 - zeta : 1e-24;
    double zp23 = ipow(cbrt(zp), 2), zm23 = ipow(cbrt(zm), 2);
    double phi = 0.5 * (zp23 + zm23), phi2 = phi * phi, phi3 = phi2 * phi;
    double omzdphi = ((1.0 - zeta) * (1.0 / cbrt(zp)) - zm23) / 3.0;
    double mopzdphi = ((1.0 + zeta) * (1.0 / cbrt(zm)) - zp23) / 3.0;
    double ct = (16.0 / OR_PI) * cbrt(3.0 * OR_PI * OR_PI);
    double rho73 = rho * rho * cbrt(rho);
    double dent2 = ct * phi2 * rho73, t2 = gt / dent2;
    double w = -ec / (gam * phi3);
    double af, daf;
    if (w > 100.0) { af = 0; daf = 0; }
    else {
        double em1 = expm1(w), ew = exp(w);
        af = b0 / em1;
        daf = -af * ew / em1;
    }
    double u = af * t2, dp = 1.0 + u + u * u, qr = (1.0 + u) / dp;
    double zarg = 1.0 + b0 * t2 * qr, lnz = log(zarg);
    double h = gam * phi3 * lnz;
    double dzdt2 = b0 * (1.0 + 2.0 * u) / (dp * dp);
    double dzda = -b0 * t2 * t2 * u * (2.0 + u) / (dp * dp);
    double dhdt2 = gam * phi3 * dzdt2 / zarg;
    double dhdw = gam * phi3 * (dzda * daf) / zarg;
    double dhdr = dhdt2 * (-(7.0 / 3.0) * t2 / rho) + dhdw * (-decr / (gam * phi3));
    double dhdphi = 3.0 * gam * phi2 * lnz + dhdw * (-3.0 * w / phi) + dhdt2 * (-2.0 * t2 / phi);
    double dhdec = dhdw * (-1.0 / (gam * phi3));
    double te = ec + h;
    double common = te + rho * (decr + dhdr);
    o.f = rho * te;
    o.va = common + (1.0 - zeta) * decz * (1.0 + dhdec) + dhdphi * omzdphi;
    o.vb = common - (1.0 + zeta) * decz * (1.0 + dhdec) + dhdphi * mopzdphi;
    double vgt = rho * dhdt2 / dent2;
    o.gaa = vgt;
    o.gab = 2.0 * vgt;
    o.gbb = vgt;
    return o;
}

static xc_t eval_xc(int func, double rha, double rhb, double gaa_, double gab, double gbb_) { /* xc.rs:159-223 */
    xc_t o = {0, 0, 0, 0, 0, 0};
    double ra = rha > 0 ? rha : 0.0, rb = rhb > 0 ? rhb : 0.0;
    if (ra + rb < RHO_CUTOFF) return o;
    double gaa = gaa_ > 0 ? gaa_ : 0.0, gbb = gbb_ > 0 ? gbb_ : 0.0;
    switch (func) {
    case OR_SVWN5:
        o = slater_x(ra, rb);
        xc_add(&o, 1.0, vwn_c(ra, rb, &VWN5));
        break;
    case OR_BLYP:
        o = slater_x(ra, rb);
        xc_add(&o, 1.0, b88_x(ra, rb, gaa, gbb));
        xc_add(&o, 1.0, lyp_c(ra, rb, gaa, gab, gbb));
        break;
    case OR_PBE:
        o = pbe_x(ra, rb, gaa, gbb);
        xc_add(&o, 1.0, pbe_c(ra, rb, gaa, gab, gbb));
        break;
    case OR_B3LYP:
    case OR_B3LYP5:
        xc_add(&o, 0.80, slater_x(ra, rb));
        xc_add(&o, 0.72, b88_x(ra, rb, gaa, gbb));
        xc_add(&o, 0.19, vwn_c(ra, rb, func == OR_B3LYP ? &VWNRPA : &VWN5));
        xc_add(&o, 0.81, lyp_c(ra, rb, gaa, gab, gbb));
        break;
    case OR_PBE0:
        xc_add(&o, 0.75, pbe_x(ra, rb, gaa, gbb));
        xc_add(&o, 1.00, pbe_c(ra, rb, gaa, gab, gbb));
        break;
    default: break;
    }
    return o;
}

void or_eval_xc(int func, double ra, double rb, double gaa, double gab, double gbb, double *out) {
    xc_t o = eval_xc(func, ra, rb, gaa, gab, gbb);
    out[0] = o.f; out[1] = o.va; out[2] = o.vb; out[3] = o.gaa; out[4] = o.gab; out[5] = o.gbb;
}

/* ------------------------------------------------------------------------ */
/* Batch screening (vxc.rs:112-165, 443-493).                               */
/* ------------------------------------------------------------------------ */
typedef struct {
    long begin, end;   /* raw point range of the batch */
    int nshell, m;     /* candidate shells, AOs */
    int *shell_list;   /* candidate shell indices (ascending) */
} batch_t;

static int screen_batch(const shells_t *S, const double *pts, long begin, long end, int *shell_list,
                        int *m_out) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    int has = 0;
    for (long g = begin; g < end; ++g) {
        const double *p = pts + 4 * g;
        if (fabs(p[3]) < 1e-24) continue;
        has = 1;
        for (int q = 0; q < 3; ++q) {
            if (p[q] < mn[q]) mn[q] = p[q];
            if (p[q] > mx[q]) mx[q] = p[q];
        }
    }
    *m_out = 0;
    if (!has) return 0;
    double c[3] = {0.5 * (mn[0] + mx[0]), 0.5 * (mn[1] + mx[1]), 0.5 * (mn[2] + mx[2])};
    double dx = mx[0] - c[0], dy = mx[1] - c[1], dz = mx[2] - c[2];
    double rbox = sqrt(dx * dx + dy * dy + dz * dz);
    int ns = 0, m = 0;
    for (int s = 0; s < S->n_shells; ++s) {
        const shell_t *sh = &S->sh[s];
        double ex = sh->center[0] - c[0], ey = sh->center[1] - c[1], ez = sh->center[2] - c[2];
        double d2 = ex * ex + ey * ey + ez * ez, th = rbox + sh->r_cutoff;
        if (d2 <= th * th) {
            shell_list[ns++] = s;
            m += sh->nfn;
        }
    }
    *m_out = m;
    return ns;
}

/* ------------------------------------------------------------------------ */
/* Threading helper: static partition of [0, n_items) into n_threads runs.  */
/* ------------------------------------------------------------------------ */
typedef struct {
    void (*fn)(void *ctx, long b, long e, int tid);
    void *ctx;
    long b, e;
    int tid;
} job_t;
static void *job_run(void *a) {
    job_t *j = (job_t *)a;
    j->fn(j->ctx, j->b, j->e, j->tid);
    return NULL;
}
static void parallel_for(long n, int nt, void (*fn)(void *, long, long, int), void *ctx) {
    if (nt < 1) nt = 1;
    if (nt > n) nt = (int)(n > 0 ? n : 1);
    pthread_t th[256];
    job_t jb[256];
    if (nt > 256) nt = 256;
    long chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        jb[t].fn = fn; jb[t].ctx = ctx; jb[t].tid = t;
        jb[t].b = t * chunk < n ? t * chunk : n;
        jb[t].e = (t + 1) * chunk < n ? (t + 1) * chunk : n;
        if (t) pthread_create(&th[t], NULL, job_run, &jb[t]);
    }
    job_run(&jb[0]);
    for (int t = 1; t < nt; ++t) pthread_join(th[t], NULL);
}

static inline double kahan_add(double *s, double *c, double x) {
    double t = x - *c, nx = *s + t;
    *c = (nx - *s) - t;
    *s = nx;
    return nx;
}

/* ------------------------------------------------------------------------ */
/* Per-point RKS path: accumulate_rks_points_with (vxc.rs:99-276).          */
/* Threads take contiguous point chunks; each keeps a lower-triangle V.     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const shells_t *S;
    const double *pts, *P;
    long n_pts;
    int func, n, uks;
    const double *Pb;
    double *vloc, *vlocb; /* [nt][n*n] */
    double *e, *ne;       /* [nt] */
} pp_ctx_t;

static void rks_pp_worker(void *vctx, long cb, long ce, int tid) {
    pp_ctx_t *X = (pp_ctx_t *)vctx;
    const shells_t *S = X->S;
    int n = X->n;
    long g0 = cb * BATCH, g1 = ce * BATCH < X->n_pts ? ce * BATCH : X->n_pts;
    double *V = X->vloc + (size_t)tid * n * n;
    int *slist = (int *)malloc(sizeof(int) * (size_t)(S->n_shells + 1));
    int *ao = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *loc = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    aoval_t *act = (aoval_t *)malloc(sizeof(aoval_t) * (size_t)(n + 1));
    int *li = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    double *phip = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *z = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *psub = NULL, *vsub = NULL;
    size_t cap = 0;
    double e = 0.0, ec = 0.0, ne = 0.0;
    for (long b0 = g0; b0 < g1; b0 += BATCH) {
        long b1 = b0 + BATCH < g1 ? b0 + BATCH : g1;
        int m, ns = screen_batch(S, X->pts, b0, b1, slist, &m);
        if (ns == 0 || m == 0) continue;
        if ((size_t)m * m > cap) {
            cap = (size_t)m * m;
            psub = (double *)realloc(psub, sizeof(double) * cap);
            vsub = (double *)realloc(vsub, sizeof(double) * cap);
        }
        int k = 0;
        for (int s = 0; s < ns; ++s)
            for (int c = 0; c < S->sh[slist[s]].nfn; ++c) ao[k++] = S->sh[slist[s]].ao_offset + c;
        for (int i = 0; i < m; ++i) loc[ao[i]] = i;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) psub[(size_t)i * m + j] = X->P[(size_t)ao[i] * n + ao[j]];
        memset(vsub, 0, sizeof(double) * (size_t)m * m);
        for (long g = b0; g < b1; ++g) {
            const double *p = X->pts + 4 * g;
            double w = p[3];
            if (fabs(w) < 1e-24) continue;
            int na = eval_aos(S, slist, ns, p, act);
            if (na == 0) continue;
            for (int a = 0; a < na; ++a) li[a] = loc[act[a].mu];
            double rho = 0.0, gr[3] = {0, 0, 0};
            for (int a = 0; a < na; ++a) {
                const double *row = psub + (size_t)li[a] * m;
                double s = 0.0;
                for (int b = 0; b < na; ++b) s += row[li[b]] * act[b].val;
                phip[a] = s;
                rho += act[a].val * s;
                for (int q = 0; q < 3; ++q) gr[q] += 2.0 * act[a].g[q] * s;
            }
            ne += w * rho;
            if (rho < DENSITY_SCREEN) continue;
            double rs = 0.5 * rho, gss = 0.25 * (gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
            xc_t o = eval_xc(X->func, rs, rs, gss, gss, gss);
            kahan_add(&e, &ec, w * o.f);
            double us = o.gaa + 0.5 * o.gab;
            double u[3] = {us * gr[0], us * gr[1], us * gr[2]};
            double hw = 0.5 * w * o.va;
            for (int a = 0; a < na; ++a)
                z[a] = hw * act[a].val + w * (u[0] * act[a].g[0] + u[1] * act[a].g[1] + u[2] * act[a].g[2]);
            for (int a = 0; a < na; ++a) {
                double *row = vsub + (size_t)li[a] * m;
                for (int b = 0; b <= a; ++b) row[li[b]] += z[a] * act[b].val + act[a].val * z[b];
            }
        }
        for (int i = 0; i < m; ++i)
            for (int j = 0; j <= i; ++j) V[(size_t)ao[i] * n + ao[j]] += vsub[(size_t)i * m + j];
    }
    X->e[tid] = e;
    X->ne[tid] = ne;
    free(slist); free(ao); free(loc); free(act); free(li); free(phip); free(z); free(psub); free(vsub);
}

static void reduce_sym(int n, int nt, const double *vloc, double *V) {
    memset(V, 0, sizeof(double) * (size_t)n * n);
    for (int t = 0; t < nt; ++t) {
        const double *src = vloc + (size_t)t * n * n;
        for (size_t k = 0; k < (size_t)n * n; ++k) V[k] += src[k];
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) V[(size_t)j * n + i] = V[(size_t)i * n + j];
}

int or_vxc_rks_perpoint(const or_basis_t *b, long n_pts, const double *pts, int func, const double *P,
                        int nt, double *V, double *exc, double *nelec) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    int n = S.n_ao;
    if (nt < 1) nt = 1;
    long nb = (n_pts + BATCH - 1) / BATCH;
    if (nt > nb) nt = (int)(nb > 0 ? nb : 1);
    pp_ctx_t X;
    memset(&X, 0, sizeof X);
    X.S = &S; X.pts = pts; X.P = P; X.n_pts = n_pts; X.func = func; X.n = n;
    X.vloc = (double *)calloc((size_t)nt * n * n + 1, sizeof(double));
    X.e = (double *)calloc((size_t)nt, sizeof(double));
    X.ne = (double *)calloc((size_t)nt, sizeof(double));
    if (n_pts > 0 && n > 0) parallel_for(nb, nt, rks_pp_worker, &X);
    double e = 0.0, ne = 0.0;
    for (int t = 0; t < nt; ++t) { e += X.e[t]; ne += X.ne[t]; }
    reduce_sym(n, nt, X.vloc, V);
    *exc = e;
    if (nelec) *nelec = ne;
    free(X.vloc); free(X.e); free(X.ne);
    shells_free(&S);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Cached-batch BLAS-3 path (vxc.rs:429-716, 801-985): per batch of 128     */
/* points, Phi/dPhi (m x b) row-major, T = P_sub Phi, rho/grad, functional, */
/* Z, V_sub = Z Phi^T, scatter V_sub + V_sub^T into the lower triangle.     */
/* The cache is rebuilt per batch here; the arithmetic is the same.         */
/* ------------------------------------------------------------------------ */
typedef struct {
    const shells_t *S;
    const double *pts, *P;
    long n_pts, bbeg;
    int func, n;
    double *vloc, *e, *ne, *fl;
} ca_ctx_t;

static void rks_cached_worker(void *vctx, long cb, long ce, int tid) {
    ca_ctx_t *X = (ca_ctx_t *)vctx;
    const shells_t *S = X->S;
    int n = X->n;
    double *V = X->vloc + (size_t)tid * n * n;
    int *slist = (int *)malloc(sizeof(int) * (size_t)(S->n_shells + 1));
    int *ao = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *loc = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    aoval_t *act = (aoval_t *)malloc(sizeof(aoval_t) * (size_t)(n + 1));
    double *psub = NULL, *vsub = NULL, *phi = NULL, *dx = NULL, *dy = NULL, *dz = NULL, *T = NULL, *Z = NULL;
    size_t capmm = 0, capmb = 0;
    double w[BATCH], rho[BATCH], gx[BATCH], gy[BATCH], gz[BATCH], wv[BATCH], ux[BATCH], uy[BATCH], uz[BATCH];
    double e = 0.0, ec = 0.0, ne = 0.0, fl = 0.0;
    for (long bi = X->bbeg + cb; bi < X->bbeg + ce; ++bi) {
        long b0 = bi * BATCH, b1 = b0 + BATCH < X->n_pts ? b0 + BATCH : X->n_pts;
        int m, ns = screen_batch(S, X->pts, b0, b1, slist, &m);
        if (ns == 0 || m == 0) continue;
        int nb = 0;
        long idx[BATCH];
        for (long g = b0; g < b1; ++g)
            if (fabs(X->pts[4 * g + 3]) >= 1e-24) idx[nb++] = g;
        if (nb == 0) continue;
        if ((size_t)m * m > capmm) {
            capmm = (size_t)m * m;
            psub = (double *)realloc(psub, sizeof(double) * capmm);
            vsub = (double *)realloc(vsub, sizeof(double) * capmm);
        }
        if ((size_t)m * BATCH > capmb) {
            capmb = (size_t)m * BATCH;
            phi = (double *)realloc(phi, sizeof(double) * capmb);
            dx = (double *)realloc(dx, sizeof(double) * capmb);
            dy = (double *)realloc(dy, sizeof(double) * capmb);
            dz = (double *)realloc(dz, sizeof(double) * capmb);
            T = (double *)realloc(T, sizeof(double) * capmb);
            Z = (double *)realloc(Z, sizeof(double) * capmb);
        }
        int k = 0;
        for (int s = 0; s < ns; ++s)
            for (int c = 0; c < S->sh[slist[s]].nfn; ++c) ao[k++] = S->sh[slist[s]].ao_offset + c;
        for (int i = 0; i < m; ++i) loc[ao[i]] = i;
        size_t mb = (size_t)m * nb;
        memset(phi, 0, sizeof(double) * mb);
        memset(dx, 0, sizeof(double) * mb);
        memset(dy, 0, sizeof(double) * mb);
        memset(dz, 0, sizeof(double) * mb);
        for (int g = 0; g < nb; ++g) {
            const double *p = X->pts + 4 * idx[g];
            w[g] = p[3];
            int na = eval_aos(S, slist, ns, p, act);
            for (int a = 0; a < na; ++a) {
                size_t o = (size_t)loc[act[a].mu] * nb + g;
                phi[o] = act[a].val; dx[o] = act[a].g[0]; dy[o] = act[a].g[1]; dz[o] = act[a].g[2];
            }
        }
        /* 1. gather P_sub */
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) psub[(size_t)i * m + j] = X->P[(size_t)ao[i] * n + ao[j]];
        /* 2. T = P_sub * Phi */
        memset(T, 0, sizeof(double) * mb);
        for (int i = 0; i < m; ++i) {
            double *Ti = T + (size_t)i * nb;
            for (int j = 0; j < m; ++j) {
                double a = psub[(size_t)i * m + j];
                const double *Pj = phi + (size_t)j * nb;
                for (int g = 0; g < nb; ++g) Ti[g] += a * Pj[g];
            }
        }
        /* 3. rho and grad rho */
        for (int g = 0; g < nb; ++g) rho[g] = gx[g] = gy[g] = gz[g] = 0.0;
        for (int i = 0; i < m; ++i) {
            size_t o = (size_t)i * nb;
            for (int g = 0; g < nb; ++g) {
                double t = T[o + g];
                rho[g] += phi[o + g] * t;
                gx[g] += dx[o + g] * t;
                gy[g] += dy[o + g] * t;
                gz[g] += dz[o + g] * t;
            }
        }
        /* 4. functional */
        int any = 0;
        for (int g = 0; g < nb; ++g) {
            ne += w[g] * rho[g];
            if (rho[g] < DENSITY_SCREEN) { wv[g] = ux[g] = uy[g] = uz[g] = 0.0; continue; }
            any = 1;
            double Gx = 2.0 * gx[g], Gy = 2.0 * gy[g], Gz = 2.0 * gz[g];
            double rs = 0.5 * rho[g], gss = 0.25 * (Gx * Gx + Gy * Gy + Gz * Gz);
            xc_t o = eval_xc(X->func, rs, rs, gss, gss, gss);
            kahan_add(&e, &ec, w[g] * o.f);
            wv[g] = 0.5 * w[g] * o.va;
            double us = w[g] * (o.gaa + 0.5 * o.gab);
            ux[g] = us * Gx; uy[g] = us * Gy; uz[g] = us * Gz;
        }
        fl += 4.0 * (double)m * m * nb;
        if (!any) continue;
        /* 5. Z */
        for (int i = 0; i < m; ++i) {
            size_t o = (size_t)i * nb;
            for (int g = 0; g < nb; ++g)
                Z[o + g] = wv[g] * phi[o + g] + ux[g] * dx[o + g] + uy[g] * dy[o + g] + uz[g] * dz[o + g];
        }
        /* 6. V_sub = Z * Phi^T */
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double *zi = Z + (size_t)i * nb, *pj = phi + (size_t)j * nb;
                double s = 0.0;
                for (int g = 0; g < nb; ++g) s += zi[g] * pj[g];
                vsub[(size_t)i * m + j] = s;
            }
        /* 7. scatter V_sub + V_sub^T into the lower triangle */
        for (int i = 0; i < m; ++i)
            for (int j = 0; j <= i; ++j)
                V[(size_t)ao[i] * n + ao[j]] += vsub[(size_t)i * m + j] + vsub[(size_t)j * m + i];
    }
    X->e[tid] = e;
    X->ne[tid] = ne;
    X->fl[tid] = fl;
    free(slist); free(ao); free(loc); free(act);
    free(psub); free(vsub); free(phi); free(dx); free(dy); free(dz); free(T); free(Z);
}

long or_vxc_rks_cached_range(const or_basis_t *b, long n_pts, const double *pts, int func, const double *P,
                             int nt, long bbeg, long bend, double *V, double *exc, double *nelec,
                             double *flops) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    int n = S.n_ao;
    long nbt = (n_pts + BATCH - 1) / BATCH;
    if (bbeg < 0) bbeg = 0;
    if (bend > nbt || bend < 0) bend = nbt;
    long nb = bend > bbeg ? bend - bbeg : 0;
    if (nt < 1) nt = 1;
    if (nt > nb) nt = (int)(nb > 0 ? nb : 1);
    ca_ctx_t X;
    memset(&X, 0, sizeof X);
    X.S = &S; X.pts = pts; X.P = P; X.n_pts = n_pts; X.func = func; X.n = n; X.bbeg = bbeg;
    X.vloc = (double *)calloc((size_t)nt * n * n + 1, sizeof(double));
    X.e = (double *)calloc((size_t)nt, sizeof(double));
    X.ne = (double *)calloc((size_t)nt, sizeof(double));
    X.fl = (double *)calloc((size_t)nt, sizeof(double));
    if (nb > 0 && n > 0) parallel_for(nb, nt, rks_cached_worker, &X);
    double e = 0.0, ne = 0.0, fl = 0.0;
    for (int t = 0; t < nt; ++t) { e += X.e[t]; ne += X.ne[t]; fl += X.fl[t]; }
    reduce_sym(n, nt, X.vloc, V);
    *exc = e;
    if (nelec) *nelec = ne;
    if (flops) *flops = fl;
    free(X.vloc); free(X.e); free(X.ne); free(X.fl);
    shells_free(&S);
    return nb;
}

int or_vxc_rks_cached(const or_basis_t *b, long n_pts, const double *pts, int func, const double *P, int nt,
                      double *V, double *exc, double *nelec) {
    long r = or_vxc_rks_cached_range(b, n_pts, pts, func, P, nt, 0, -1, V, exc, nelec, NULL);
    return r < 0 ? (int)r : 0;
}

/* ------------------------------------------------------------------------ */
/* Per-point UKS path: accumulate_uks_points (vxc.rs:1141-1348).            */
/* ------------------------------------------------------------------------ */
static void uks_pp_worker(void *vctx, long cb, long ce, int tid) {
    pp_ctx_t *X = (pp_ctx_t *)vctx;
    const shells_t *S = X->S;
    int n = X->n;
    long g0 = cb * BATCH, g1 = ce * BATCH < X->n_pts ? ce * BATCH : X->n_pts;
    double *Va = X->vloc + (size_t)tid * n * n, *Vb = X->vlocb + (size_t)tid * n * n;
    int *slist = (int *)malloc(sizeof(int) * (size_t)(S->n_shells + 1));
    int *ao = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *loc = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    aoval_t *act = (aoval_t *)malloc(sizeof(aoval_t) * (size_t)(n + 1));
    int *li = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    double *za = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *zb = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *pa = NULL, *pb = NULL, *va = NULL, *vb = NULL;
    size_t cap = 0;
    double e = 0.0, ec = 0.0, ne = 0.0;
    for (long b0 = g0; b0 < g1; b0 += BATCH) {
        long b1 = b0 + BATCH < g1 ? b0 + BATCH : g1;
        int m, ns = screen_batch(S, X->pts, b0, b1, slist, &m);
        if (ns == 0 || m == 0) continue;
        if ((size_t)m * m > cap) {
            cap = (size_t)m * m;
            pa = (double *)realloc(pa, sizeof(double) * cap);
            pb = (double *)realloc(pb, sizeof(double) * cap);
            va = (double *)realloc(va, sizeof(double) * cap);
            vb = (double *)realloc(vb, sizeof(double) * cap);
        }
        int k = 0;
        for (int s = 0; s < ns; ++s)
            for (int c = 0; c < S->sh[slist[s]].nfn; ++c) ao[k++] = S->sh[slist[s]].ao_offset + c;
        for (int i = 0; i < m; ++i) loc[ao[i]] = i;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                pa[(size_t)i * m + j] = X->P[(size_t)ao[i] * n + ao[j]];
                pb[(size_t)i * m + j] = X->Pb[(size_t)ao[i] * n + ao[j]];
            }
        memset(va, 0, sizeof(double) * (size_t)m * m);
        memset(vb, 0, sizeof(double) * (size_t)m * m);
        for (long g = b0; g < b1; ++g) {
            const double *p = X->pts + 4 * g;
            double w = p[3];
            if (fabs(w) < 1e-24) continue;
            int na = eval_aos(S, slist, ns, p, act);
            if (na == 0) continue;
            for (int a = 0; a < na; ++a) li[a] = loc[act[a].mu];
            double rha = 0, rhb = 0, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
            for (int a = 0; a < na; ++a) {
                const double *ra_ = pa + (size_t)li[a] * m, *rb_ = pb + (size_t)li[a] * m;
                double sa = 0, sb = 0;
                for (int bb = 0; bb < na; ++bb) {
                    sa += ra_[li[bb]] * act[bb].val;
                    sb += rb_[li[bb]] * act[bb].val;
                }
                rha += act[a].val * sa;
                rhb += act[a].val * sb;
                for (int q = 0; q < 3; ++q) {
                    ga[q] += 2.0 * act[a].g[q] * sa;
                    gb[q] += 2.0 * act[a].g[q] * sb;
                }
            }
            ne += w * (rha + rhb);
            double ra = rha > 0 ? rha : 0, rb = rhb > 0 ? rhb : 0;
            if (ra + rb < DENSITY_SCREEN) continue;
            double gaa = ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2];
            double gab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
            double gbb = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2];
            xc_t o = eval_xc(X->func, ra, rb, gaa, gab, gbb);
            kahan_add(&e, &ec, w * o.f);
            double ua[3], ub[3];
            for (int q = 0; q < 3; ++q) {
                ua[q] = 2.0 * o.gaa * ga[q] + o.gab * gb[q];
                ub[q] = 2.0 * o.gbb * gb[q] + o.gab * ga[q];
            }
            double hwa = 0.5 * w * o.va, hwb = 0.5 * w * o.vb;
            for (int a = 0; a < na; ++a) {
                const double *g = act[a].g;
                za[a] = hwa * act[a].val + w * (ua[0] * g[0] + ua[1] * g[1] + ua[2] * g[2]);
                zb[a] = hwb * act[a].val + w * (ub[0] * g[0] + ub[1] * g[1] + ub[2] * g[2]);
            }
            for (int a = 0; a < na; ++a) {
                double *rowa = va + (size_t)li[a] * m, *rowb = vb + (size_t)li[a] * m;
                for (int bb = 0; bb <= a; ++bb) {
                    rowa[li[bb]] += za[a] * act[bb].val + act[a].val * za[bb];
                    rowb[li[bb]] += zb[a] * act[bb].val + act[a].val * zb[bb];
                }
            }
        }
        for (int i = 0; i < m; ++i)
            for (int j = 0; j <= i; ++j) {
                Va[(size_t)ao[i] * n + ao[j]] += va[(size_t)i * m + j];
                Vb[(size_t)ao[i] * n + ao[j]] += vb[(size_t)i * m + j];
            }
    }
    X->e[tid] = e;
    X->ne[tid] = ne;
    free(slist); free(ao); free(loc); free(act); free(li); free(za); free(zb);
    free(pa); free(pb); free(va); free(vb);
}

int or_vxc_uks_perpoint(const or_basis_t *b, long n_pts, const double *pts, int func, const double *Pa,
                        const double *Pb, int nt, double *Va, double *Vb, double *exc, double *nelec) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    int n = S.n_ao;
    if (nt < 1) nt = 1;
    long nb = (n_pts + BATCH - 1) / BATCH;
    if (nt > nb) nt = (int)(nb > 0 ? nb : 1);
    pp_ctx_t X;
    memset(&X, 0, sizeof X);
    X.S = &S; X.pts = pts; X.P = Pa; X.Pb = Pb; X.n_pts = n_pts; X.func = func; X.n = n; X.uks = 1;
    X.vloc = (double *)calloc((size_t)nt * n * n + 1, sizeof(double));
    X.vlocb = (double *)calloc((size_t)nt * n * n + 1, sizeof(double));
    X.e = (double *)calloc((size_t)nt, sizeof(double));
    X.ne = (double *)calloc((size_t)nt, sizeof(double));
    if (n_pts > 0 && n > 0) parallel_for(nb, nt, uks_pp_worker, &X);
    double e = 0.0, ne = 0.0;
    for (int t = 0; t < nt; ++t) { e += X.e[t]; ne += X.ne[t]; }
    reduce_sym(n, nt, X.vloc, Va);
    reduce_sym(n, nt, X.vlocb, Vb);
    *exc = e;
    if (nelec) *nelec = ne;
    free(X.vloc); free(X.vlocb); free(X.e); free(X.ne);
    shells_free(&S);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* XC nuclear gradient (xc_nuclear_gradient_cpu / accumulate_xc_grad_points, */
/* vxc.rs:1769-1899; becke_weight_nuclear_gradient, grid.rs:423-584).        */
/* ------------------------------------------------------------------------ */
typedef struct { int mu, atom; double val, g[3], h[6]; } aohess_t; /* h: xx xy xz yy yz zz */

/* eval_aos_hess (ao_grid.rs:234-299): values, gradients and Hessians of every
 * AO of every shell within its cutoff; no per-AO drop rule.  shell_atom maps
 * shells to atoms. */
static int eval_aos_hess(const shells_t *S, const int *shell_atom, const double r[3], aohess_t *out) {
    int n = 0;
    for (int k = 0; k < S->n_shells; ++k) {
        const shell_t *sh = &S->sh[k];
        double dx = r[0] - sh->center[0], dy = r[1] - sh->center[1], dz = r[2] - sh->center[2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > sh->r2_cutoff) continue;
        double r0 = 0.0, r1 = 0.0, rr2 = 0.0;
        for (int p = 0; p < sh->n_prim; ++p) {
            double e = sh->wnorm[p] * exp(-sh->alpha[p] * r2);
            r0 += e;
            r1 -= 2.0 * sh->alpha[p] * e;
            rr2 += 4.0 * sh->alpha[p] * sh->alpha[p] * e;
        }
        if (fabs(r0) < 1e-18 && fabs(r1) < 1e-18 && fabs(rr2) < 1e-18) continue;
        for (int c = 0; c < sh->nfn; ++c) {
            const int *lmn = comp_of(sh->l, c);
            double px, dpx, d2px, py, dpy, d2py, pz, dpz, d2pz;
            pow012(dx, lmn[0], &px, &dpx, &d2px);
            pow012(dy, lmn[1], &py, &dpy, &d2py);
            pow012(dz, lmn[2], &pz, &dpz, &d2pz);
            double m = px * py * pz, sc = sh->scale[c];
            double dm[3] = {dpx * py * pz, px * dpy * pz, px * py * dpz};
            aohess_t *o = &out[n++];
            o->mu = sh->ao_offset + c;
            o->atom = shell_atom[k];
            o->val = sc * m * r0;
            o->g[0] = sc * (dm[0] * r0 + m * dx * r1);
            o->g[1] = sc * (dm[1] * r0 + m * dy * r1);
            o->g[2] = sc * (dm[2] * r0 + m * dz * r1);
            o->h[0] = sc * (d2px * py * pz * r0 + (2.0 * dm[0] * dx + m) * r1 + m * dx * dx * rr2);
            o->h[1] = sc * (dpx * dpy * pz * r0 + (dm[0] * dy + dm[1] * dx) * r1 + m * dx * dy * rr2);
            o->h[2] = sc * (dpx * py * dpz * r0 + (dm[0] * dz + dm[2] * dx) * r1 + m * dx * dz * rr2);
            o->h[3] = sc * (px * d2py * pz * r0 + (2.0 * dm[1] * dy + m) * r1 + m * dy * dy * rr2);
            o->h[4] = sc * (px * dpy * dpz * r0 + (dm[1] * dz + dm[2] * dy) * r1 + m * dy * dz * rr2);
            o->h[5] = sc * (px * py * d2pz * r0 + (2.0 * dm[2] * dz + m) * r1 + m * dz * dz * rr2);
        }
    }
    return n;
}

/* becke_step_and_deriv (grid.rs:129-140): s(nu) and ds/dnu. */
static void becke_step_deriv(double nu, double *s, double *ds) {
    double x = clamp1(nu);
    double f1 = 0.5 * x * (3.0 - x * x), d1 = 1.5 * (1.0 - x * x);
    double f2 = 0.5 * f1 * (3.0 - f1 * f1), d2 = 1.5 * (1.0 - f1 * f1);
    double f3 = 0.5 * f2 * (3.0 - f2 * f2), d3 = 1.5 * (1.0 - f2 * f2);
    *s = 0.5 * (1.0 - f3);
    *ds = -0.5 * d3 * d2 * d1;
}

typedef struct {
    int n;
    const int *z;
    const double *xyz;
    double *vra, *dist, *smat, *dsdmu, *cellp; /* scratch per thread */
} bgrad_t;

/* becke_weight_nuclear_gradient (grid.rs:423-584): adds raw_w f_xc dw_host/dR_A. */
static void becke_grad(bgrad_t *B, int host, const double r[3], double raw_w, double fxc, double *grad) {
    int n = B->n;
    if (n <= 1 || fabs(fxc) < 1e-18 || raw_w == 0.0) return;
    for (int c = 0; c < n; ++c) {
        double dx = r[0] - B->xyz[3 * c], dy = r[1] - B->xyz[3 * c + 1], dz = r[2] - B->xyz[3 * c + 2];
        B->vra[3 * c] = dx; B->vra[3 * c + 1] = dy; B->vra[3 * c + 2] = dz;
        B->dist[c] = sqrt(dx * dx + dy * dy + dz * dz);
        B->cellp[c] = 1.0;
    }
    for (long k = 0; k < (long)n * n; ++k) { B->smat[k] = 1.0; B->dsdmu[k] = 0.0; }
    for (int c = 0; c < n; ++c)
        for (int d = c + 1; d < n; ++d) {
            const double *ac = B->xyz + 3 * c, *ad = B->xyz + 3 * d;
            double rcd = sqrt((ac[0] - ad[0]) * (ac[0] - ad[0]) + (ac[1] - ad[1]) * (ac[1] - ad[1]) +
                              (ac[2] - ad[2]) * (ac[2] - ad[2]));
            if (rcd < 1e-14) continue;
            double mu = clamp1((B->dist[c] - B->dist[d]) / rcd);
            double a = becke_asym(B->z[c], B->z[d]);
            double nu = mu + a * (1.0 - mu * mu), scd, dsdnu;
            becke_step_deriv(nu, &scd, &dsdnu);
            B->smat[c * n + d] = scd;
            B->smat[d * n + c] = 1.0 - scd;
            B->dsdmu[c * n + d] = B->dsdmu[d * n + c] = dsdnu * (1.0 - 2.0 * a * mu);
            B->cellp[c] *= scd;
            B->cellp[d] *= 1.0 - scd;
        }
    double zs = 0.0;
    for (int c = 0; c < n; ++c) zs += B->cellp[c];
    if (zs <= 1e-30) return;
    double wn = B->cellp[host] / zs, pref = raw_w * fxc / zs;
    double hacc[3] = {0, 0, 0};
    for (int a = 0; a < n; ++a) {
        if (a == host) continue;
        double ra = B->dist[a];
        if (ra < 1e-14) continue;
        const double *pa = B->xyz + 3 * a;
        double dra[3] = {-B->vra[3 * a] / ra, -B->vra[3 * a + 1] / ra, -B->vra[3 * a + 2] / ra};
        double dph[3] = {0, 0, 0}, dzs[3] = {0, 0, 0};
        for (int b = 0; b < n; ++b) {
            if (b == a) continue;
            const double *pb = B->xyz + 3 * b;
            double rv[3] = {pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
            double rab = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
            if (rab < 1e-14) continue;
            double mu = clamp1((ra - B->dist[b]) / rab), inv = 1.0 / rab;
            double dmu[3];
            for (int q = 0; q < 3; ++q) dmu[q] = (dra[q] - mu * rv[q] * inv) * inv;
            double slope = B->dsdmu[a * n + b];
            if (fabs(slope) < 1e-30) continue;
            double pab, pba;
            { /* prod over d != a, skip b of s_ad (and b's row skipping a) */
                double sv = B->smat[a * n + b];
                if (sv > 1e-12) pab = B->cellp[a] / sv;
                else { pab = 1.0; for (int d = 0; d < n; ++d) if (d != a && d != b) pab *= B->smat[a * n + d]; }
                sv = B->smat[b * n + a];
                if (sv > 1e-12) pba = B->cellp[b] / sv;
                else { pba = 1.0; for (int d = 0; d < n; ++d) if (d != b && d != a) pba *= B->smat[b * n + d]; }
            }
            for (int q = 0; q < 3; ++q) {
                double dsab = slope * dmu[q];
                double ta = pab * dsab, tb = -pba * dsab;
                dzs[q] += ta + tb;
                if (b == host) dph[q] += tb;
            }
        }
        for (int q = 0; q < 3; ++q) {
            double dwf = pref * (dph[q] - wn * dzs[q]);
            grad[3 * a + q] += dwf;
            hacc[q] -= dwf;
        }
    }
    for (int q = 0; q < 3; ++q) grad[3 * host + q] += hacc[q];
}

typedef struct {
    const shells_t *S;
    const int *shell_atom;
    const double *pts6, *Pa, *Pb;
    long n_pts;
    int func, n, n_atoms;
    const int *z;
    const double *xyz;
    double *gpart; /* [nt][n_atoms*3] */
} grad_ctx_t;

static void grad_worker(void *vctx, long cb, long ce, int tid) {
    grad_ctx_t *X = (grad_ctx_t *)vctx;
    const int n = X->n, na = X->n_atoms;
    long g0 = cb * BATCH, g1 = ce * BATCH < X->n_pts ? ce * BATCH : X->n_pts;
    double *grad = X->gpart + (size_t)tid * na * 3;
    aohess_t *act = (aohess_t *)malloc(sizeof(aohess_t) * (size_t)(n + 1));
    double *sa = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *sb = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *gpa = (double *)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    double *gpb = (double *)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    bgrad_t B;
    B.n = na; B.z = X->z; B.xyz = X->xyz;
    B.vra = (double *)malloc(sizeof(double) * 3 * (size_t)na);
    B.dist = (double *)malloc(sizeof(double) * (size_t)na);
    B.cellp = (double *)malloc(sizeof(double) * (size_t)na);
    B.smat = (double *)malloc(sizeof(double) * (size_t)na * na);
    B.dsdmu = (double *)malloc(sizeof(double) * (size_t)na * na);
    static const int hq[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}}; /* hess_row(q) */
    for (long g = g0; g < g1; ++g) {
        const double *p = X->pts6 + 6 * g;
        const double w = p[3], raw = p[4];
        const int host = (int)p[5];
        if (fabs(raw) < 1e-24) continue;
        int na_ = eval_aos_hess(X->S, X->shell_atom, p, act);
        if (na_ == 0) continue;
        double rha = 0, rhb = 0, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
        for (int i = 0; i < na_; ++i) {
            const double *ra_ = X->Pa + (size_t)act[i].mu * n, *rb_ = X->Pb + (size_t)act[i].mu * n;
            double s_a = 0, s_b = 0, va[3] = {0, 0, 0}, vb[3] = {0, 0, 0};
            for (int j = 0; j < na_; ++j) {
                double pa_ = ra_[act[j].mu], pb_ = rb_[act[j].mu];
                s_a += pa_ * act[j].val;
                s_b += pb_ * act[j].val;
                for (int q = 0; q < 3; ++q) { va[q] += pa_ * act[j].g[q]; vb[q] += pb_ * act[j].g[q]; }
            }
            sa[i] = s_a; sb[i] = s_b;
            for (int q = 0; q < 3; ++q) { gpa[3 * i + q] = va[q]; gpb[3 * i + q] = vb[q]; }
            rha += act[i].val * s_a;
            rhb += act[i].val * s_b;
            for (int q = 0; q < 3; ++q) { ga[q] += 2.0 * act[i].g[q] * s_a; gb[q] += 2.0 * act[i].g[q] * s_b; }
        }
        double ra = rha > 0 ? rha : 0, rb = rhb > 0 ? rhb : 0;
        if (ra + rb < DENSITY_SCREEN) continue;
        double gaa = ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2];
        double gab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        double gbb = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2];
        xc_t o = eval_xc(X->func, ra, rb, gaa, gab, gbb);
        if (fabs(w) > 1e-24) {
            double ua[3], ub[3], sum[3] = {0, 0, 0};
            for (int q = 0; q < 3; ++q) {
                ua[q] = 2.0 * o.gaa * ga[q] + o.gab * gb[q];
                ub[q] = 2.0 * o.gbb * gb[q] + o.gab * ga[q];
            }
            for (int i = 0; i < na_; ++i) {
                const double *G = act[i].g, *H = act[i].h;
                double uga = ua[0] * gpa[3 * i] + ua[1] * gpa[3 * i + 1] + ua[2] * gpa[3 * i + 2];
                double ugb = ub[0] * gpb[3 * i] + ub[1] * gpb[3 * i + 1] + ub[2] * gpb[3 * i + 2];
                for (int q = 0; q < 3; ++q) {
                    double hua = ua[0] * H[hq[q][0]] + ua[1] * H[hq[q][1]] + ua[2] * H[hq[q][2]];
                    double hub = ub[0] * H[hq[q][0]] + ub[1] * H[hq[q][1]] + ub[2] * H[hq[q][2]];
                    double ta = o.va * G[q] * sa[i] + hua * sa[i] + G[q] * uga;
                    double tb = o.vb * G[q] * sb[i] + hub * sb[i] + G[q] * ugb;
                    double f = -2.0 * w * (ta + tb);
                    grad[3 * act[i].atom + q] += f;
                    sum[q] += f;
                }
            }
            for (int q = 0; q < 3; ++q) grad[3 * host + q] -= sum[q];
        }
        becke_grad(&B, host, p, raw, o.f, grad);
    }
    free(act); free(sa); free(sb); free(gpa); free(gpb);
    free(B.vra); free(B.dist); free(B.cellp); free(B.smat); free(B.dsdmu);
}

int or_xc_gradient(const or_basis_t *b, int n_atoms, const int *z, const double *xyz_bohr, long n_pts,
                   const double *pts6, int func, const double *Pa, const double *Pb, int nt, double *grad) {
    if (!b || n_atoms <= 0 || !z || !xyz_bohr || !pts6 || !Pa || !Pb || !grad) return -1;
    for (int k = 0; k < 3 * n_atoms; ++k) grad[k] = 0.0;
    if (n_atoms <= 1 || n_pts <= 0) return 0;
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    if (nt < 1) nt = 1;
    long nb = (n_pts + BATCH - 1) / BATCH;
    if (nt > nb) nt = (int)nb;
    grad_ctx_t X;
    X.S = &S; X.shell_atom = b->atom; X.pts6 = pts6; X.Pa = Pa; X.Pb = Pb; X.n_pts = n_pts;
    X.func = func; X.n = S.n_ao; X.n_atoms = n_atoms; X.z = z; X.xyz = xyz_bohr;
    X.gpart = (double *)calloc((size_t)nt * n_atoms * 3, sizeof(double));
    parallel_for(nb, nt, grad_worker, &X);
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 3 * n_atoms; ++k) grad[k] += X.gpart[(size_t)t * n_atoms * 3 + k];
    free(X.gpart);
    shells_free(&S);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Analytic overlap of contracted Cartesian Gaussians (Obara-Saika 1D).     */
/* ------------------------------------------------------------------------ */
static double os1d(int la, int lb, double PA, double PB, double p) {
    /* E(i,j) via S_{i+1,j} = PA S_ij + (i S_{i-1,j} + j S_{i,j-1}) / (2p) */
    double s[5][5];
    memset(s, 0, sizeof s);
    s[0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            if (i > 0) {
                double v = PA * s[i - 1][j];
                if (i > 1) v += (i - 1) * s[i - 2][j] / (2.0 * p);
                if (j > 0) v += j * s[i - 1][j - 1] / (2.0 * p);
                s[i][j] = v;
            } else {
                double v = PB * s[i][j - 1];
                if (j > 1) v += (j - 1) * s[i][j - 2] / (2.0 * p);
                s[i][j] = v;
            }
        }
    return s[la][lb];
}

int or_overlap(const or_basis_t *b, double *Smat) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    int n = S.n_ao;
    for (int s1 = 0; s1 < S.n_shells; ++s1)
        for (int s2 = 0; s2 < S.n_shells; ++s2) {
            const shell_t *A = &S.sh[s1], *B = &S.sh[s2];
            double AB2 = 0;
            for (int q = 0; q < 3; ++q) AB2 += (A->center[q] - B->center[q]) * (A->center[q] - B->center[q]);
            for (int ia = 0; ia < A->nfn; ++ia)
                for (int ib = 0; ib < B->nfn; ++ib) {
                    const int *la = comp_of(A->l, ia), *lb = comp_of(B->l, ib);
                    double tot = 0.0;
                    for (int pa = 0; pa < A->n_prim; ++pa)
                        for (int pb = 0; pb < B->n_prim; ++pb) {
                            double a = A->alpha[pa], bb = B->alpha[pb], p = a + bb;
                            double pre = exp(-a * bb / p * AB2) * pow(OR_PI / p, 1.5);
                            double v = pre;
                            for (int q = 0; q < 3; ++q) {
                                double Pq = (a * A->center[q] + bb * B->center[q]) / p;
                                v *= os1d(la[q], lb[q], Pq - A->center[q], Pq - B->center[q], p);
                            }
                            tot += A->wnorm[pa] * B->wnorm[pb] * v;
                        }
                    Smat[(size_t)(A->ao_offset + ia) * n + B->ao_offset + ib] = tot * A->scale[ia] * B->scale[ib];
                }
        }
    shells_free(&S);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-point rho and grad rho over all shells (the quantities the reference */
/* GPU backend returns from calchem_cuda_vxc_eval_rho, cuda_kernels.cu:     */
/* 7452-7494; formula of vxc.rs:205-217).  out: [n_pts*4].                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    const shells_t *S;
    const double *pts, *P;
    int n;
    double *out;
} rho_ctx_t;

static void rho_worker(void *vctx, long b, long e, int tid) {
    rho_ctx_t *X = (rho_ctx_t *)vctx;
    int n = X->n;
    aoval_t *act = (aoval_t *)malloc(sizeof(aoval_t) * (size_t)(n + 1));
    for (long g = b; g < e; ++g) {
        const double *p = X->pts + 4 * g;
        int na = eval_aos(X->S, NULL, X->S->n_shells, p, act);
        double rho = 0, gr[3] = {0, 0, 0};
        for (int a = 0; a < na; ++a) {
            const double *row = X->P + (size_t)act[a].mu * n;
            double s = 0;
            for (int c = 0; c < na; ++c) s += row[act[c].mu] * act[c].val;
            rho += act[a].val * s;
            for (int q = 0; q < 3; ++q) gr[q] += 2.0 * act[a].g[q] * s;
        }
        X->out[4 * g] = rho;
        X->out[4 * g + 1] = gr[0];
        X->out[4 * g + 2] = gr[1];
        X->out[4 * g + 3] = gr[2];
    }
    free(act);
    (void)tid;
}

int or_rho_points(const or_basis_t *b, long n_pts, const double *pts, const double *P, int nt, double *out) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    rho_ctx_t X = {&S, pts, P, S.n_ao, out};
    parallel_for(n_pts, nt, rho_worker, &X);
    shells_free(&S);
    return 0;
}

/* Dense AO values and gradients at many points (test helper):            */
/* val [n_pts*n_ao], grad [n_pts*n_ao*3], zero where the value rule drops. */
int or_ao_dense(const or_basis_t *b, long n_pts, const double *pts, double *val, double *grad) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    int n = S.n_ao;
    aoval_t *act = (aoval_t *)malloc(sizeof(aoval_t) * (size_t)(n + 1));
    memset(val, 0, sizeof(double) * (size_t)n_pts * n);
    memset(grad, 0, sizeof(double) * 3 * (size_t)n_pts * n);
    for (long g = 0; g < n_pts; ++g) {
        int na = eval_aos(&S, NULL, S.n_shells, pts + 4 * g, act);
        for (int a = 0; a < na; ++a) {
            size_t o = (size_t)g * n + act[a].mu;
            val[o] = act[a].val;
            for (int q = 0; q < 3; ++q) grad[3 * o + q] = act[a].g[q];
        }
    }
    free(act);
    shells_free(&S);
    return 0;
}

/* ======================================================================== */
/* Integrals of the SCF Hamiltonian (the SCF row beside the Vxc build).     */
/*   Boys function F_m(T)                         boys.rs:109-192           */
/*   overlap / kinetic / nuclear attraction        one_electron.rs:7-168,   */
/*                                                 primitive.rs:44-174      */
/*   electron repulsion (ab|cd): Obara-Saika VRR to [e0|f0], Head-Gordon-    */
/*   Pople HRR to (ab|cd) after contraction        eri.rs:276-1018          */
/* Contracted functions as in the AO code: c~_p N(alpha_p; L00) times the   */
/* component scale N(alpha0; lmn)/N(alpha0; L00) (each Cartesian component  */
/* normalised, calchem-basis/src/lib.rs:200-202).                           */
/* ======================================================================== */

/* F_m(T) = int_0^1 u^{2m} exp(-T u^2) du for m = 0..mmax, in long double:  */
/* the series e^{-T} sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)) for F_M and   */
/* downward recursion below T = 45; F_0 = sqrt(pi/T) erf(sqrt T) / 2 and     */
/* upward recursion above.                                                  */
void or_boys(int mmax, double T, double *F) {
    long double t = T, et = expl(-t), Fm[64];
    if (mmax > 63) mmax = 63;
    if (T < 45.0) {
        long double term = 1.0L / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 4000; ++k) {
            term *= 2.0L * t / (long double)(2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1e-22L * sum) break;
        }
        Fm[mmax] = et * sum;
        for (int m = mmax; m > 0; --m) Fm[m - 1] = (2.0L * t * Fm[m] + et) / (long double)(2 * m - 1);
    } else {
        Fm[0] = 0.5L * sqrtl(3.14159265358979323846264338327950288L / t) * erfl(sqrtl(t));
        for (int m = 0; m < mmax; ++m) Fm[m + 1] = ((long double)(2 * m + 1) * Fm[m] - et) / (2.0L * t);
    }
    for (int m = 0; m <= mmax; ++m) F[m] = (double)Fm[m];
}

/* Cartesian triples (x, y, z) with x + y + z <= L, indexed by total n first. */
static int tri_index(int x, int y, int z) {
    int n = x + y + z, r = n - x;
    (void)y;
    return n * (n + 1) * (n + 2) / 6 + r * (r + 1) / 2 + z;
}
static int tri_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

/* Kinetic 1D factor <d/dx a | d/dx b> in the units of os1d. */
static double kin1d(int i, int j, double a, double b, double PA, double PB, double p) {
    double v = 4.0 * a * b * os1d(i + 1, j + 1, PA, PB, p);
    if (i > 0) v -= 2.0 * b * i * os1d(i - 1, j + 1, PA, PB, p);
    if (j > 0) v -= 2.0 * a * j * os1d(i + 1, j - 1, PA, PB, p);
    if (i > 0 && j > 0) v += (double)(i * j) * os1d(i - 1, j - 1, PA, PB, p);
    return v;
}

/* Nuclear attraction (a|1/r_C|b) of one primitive pair by Obara-Saika VRR on
 * Theta^(m)(e, 0) and HRR (a, b+1_i) = (a+1_i, b) + AB_i (a, b). */
static double nuc_hrr(const double *th0, const int a[3], const int b[3], const double AB[3]) {
    int i = b[0] > 0 ? 0 : (b[1] > 0 ? 1 : (b[2] > 0 ? 2 : -1));
    if (i < 0) return th0[tri_index(a[0], a[1], a[2])];
    int a1[3] = {a[0], a[1], a[2]}, b1[3] = {b[0], b[1], b[2]};
    a1[i] += 1;
    b1[i] -= 1;
    return nuc_hrr(th0, a1, b1, AB) + AB[i] * nuc_hrr(th0, a, b1, AB);
}

int or_one_electron(const or_basis_t *b, int n_atoms, const int *z, const double *xyz, double *Smat,
                    double *Tmat, double *Vmat) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    const int n = S.n_ao;
    for (int s1 = 0; s1 < S.n_shells; ++s1)
        for (int s2 = 0; s2 < S.n_shells; ++s2) {
            const shell_t *A = &S.sh[s1], *B = &S.sh[s2];
            const int L = A->l + B->l, ne = tri_count(L);
            double AB[3], AB2 = 0.0;
            for (int q = 0; q < 3; ++q) {
                AB[q] = A->center[q] - B->center[q];
                AB2 += AB[q] * AB[q];
            }
            double *th = (double *)malloc(sizeof(double) * (size_t)(L + 1) * ne);
            double F[64];
            for (int ia = 0; ia < A->nfn; ++ia)
                for (int ib = 0; ib < B->nfn; ++ib) {
                    const int *la = comp_of(A->l, ia), *lb = comp_of(B->l, ib);
                    double ts = 0.0, tt = 0.0, tv = 0.0;
                    for (int pa = 0; pa < A->n_prim; ++pa)
                        for (int pb = 0; pb < B->n_prim; ++pb) {
                            const double a = A->alpha[pa], bb = B->alpha[pb], p = a + bb;
                            const double Kab = exp(-a * bb / p * AB2);
                            const double cw = A->wnorm[pa] * B->wnorm[pb];
                            double P[3], PA[3], PB[3], s1[3], k1[3];
                            for (int q = 0; q < 3; ++q) {
                                P[q] = (a * A->center[q] + bb * B->center[q]) / p;
                                PA[q] = P[q] - A->center[q];
                                PB[q] = P[q] - B->center[q];
                                s1[q] = os1d(la[q], lb[q], PA[q], PB[q], p);
                                k1[q] = kin1d(la[q], lb[q], a, bb, PA[q], PB[q], p);
                            }
                            const double pre = Kab * pow(OR_PI / p, 1.5);
                            ts += cw * pre * s1[0] * s1[1] * s1[2];
                            tt += cw * pre * 0.5 * (k1[0] * s1[1] * s1[2] + s1[0] * k1[1] * s1[2] + s1[0] * s1[1] * k1[2]);
                            for (int c = 0; c < n_atoms; ++c) {
                                double PC[3], PC2 = 0.0;
                                for (int q = 0; q < 3; ++q) {
                                    PC[q] = P[q] - xyz[3 * c + q];
                                    PC2 += PC[q] * PC[q];
                                }
                                or_boys(L, p * PC2, F);
                                /* th[m * ne + e] */
                                for (int m = 0; m <= L; ++m) th[(size_t)m * ne] = 2.0 * OR_PI / p * Kab * F[m];
                                for (int tot = 1; tot <= L; ++tot)
                                    for (int x = tot; x >= 0; --x)
                                        for (int y = tot - x; y >= 0; --y) {
                                            int e[3] = {x, y, tot - x - y};
                                            int i = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
                                            int e1[3] = {e[0], e[1], e[2]};
                                            e1[i] -= 1;
                                            const int ie = tri_index(e[0], e[1], e[2]), i1 = tri_index(e1[0], e1[1], e1[2]);
                                            int i2 = -1;
                                            if (e1[i] > 0) {
                                                int e2[3] = {e1[0], e1[1], e1[2]};
                                                e2[i] -= 1;
                                                i2 = tri_index(e2[0], e2[1], e2[2]);
                                            }
                                            for (int m = 0; m <= L - tot; ++m) {
                                                double v = PA[i] * th[(size_t)m * ne + i1] - PC[i] * th[(size_t)(m + 1) * ne + i1];
                                                if (i2 >= 0)
                                                    v += e1[i] / (2.0 * p) * (th[(size_t)m * ne + i2] - th[(size_t)(m + 1) * ne + i2]);
                                                th[(size_t)m * ne + ie] = v;
                                            }
                                        }
                                tv -= z[c] * cw * nuc_hrr(th, la, lb, AB);
                            }
                        }
                    const size_t o = (size_t)(A->ao_offset + ia) * n + B->ao_offset + ib;
                    const double sc = A->scale[ia] * B->scale[ib];
                    if (Smat) Smat[o] = ts * sc;
                    if (Tmat) Tmat[o] = tt * sc;
                    if (Vmat) Vmat[o] = tv * sc;
                }
            free(th);
        }
    shells_free(&S);
    return 0;
}

/* Electron repulsion: contracted [e0|f0] over the primitive quartets, then HRR. */
typedef struct {
    const double *C; /* [ne * nf] contracted [e0|f0] */
    int nf;
    double AB[3], CD[3];
} eri_hrr_t;

static double eri_hrr_bra(const eri_hrr_t *h, const int a[3], const int b[3], int jf) {
    int i = b[0] > 0 ? 0 : (b[1] > 0 ? 1 : (b[2] > 0 ? 2 : -1));
    if (i < 0) return h->C[(size_t)tri_index(a[0], a[1], a[2]) * h->nf + jf];
    int a1[3] = {a[0], a[1], a[2]}, b1[3] = {b[0], b[1], b[2]};
    a1[i] += 1;
    b1[i] -= 1;
    return eri_hrr_bra(h, a1, b1, jf) + h->AB[i] * eri_hrr_bra(h, a, b1, jf);
}

static double eri_hrr(const eri_hrr_t *h, const int a[3], const int b[3], const int c[3], const int d[3]) {
    int i = d[0] > 0 ? 0 : (d[1] > 0 ? 1 : (d[2] > 0 ? 2 : -1));
    if (i < 0) return eri_hrr_bra(h, a, b, tri_index(c[0], c[1], c[2]));
    int c1[3] = {c[0], c[1], c[2]}, d1[3] = {d[0], d[1], d[2]};
    c1[i] += 1;
    d1[i] -= 1;
    return eri_hrr(h, a, b, c1, d1) + h->CD[i] * eri_hrr(h, a, b, c, d1);
}

/* out[nfa*nfb*nfc*nfd] = (ab|cd) of one shell quartet, component order a, b, c, d. */
static void eri_quartet(const shell_t *A, const shell_t *B, const shell_t *Cs, const shell_t *D, double *out) {
    const int Le = A->l + B->l, Lf = Cs->l + D->l, M = Le + Lf;
    const int ne = tri_count(Le), nf = tri_count(Lf);
    double *X = (double *)malloc(sizeof(double) * (size_t)(M + 1) * ne * nf);
    double *Cc = (double *)calloc((size_t)ne * nf, sizeof(double));
    double AB[3], CD[3], AB2 = 0.0, CD2 = 0.0, F[64];
    for (int q = 0; q < 3; ++q) {
        AB[q] = A->center[q] - B->center[q];
        CD[q] = Cs->center[q] - D->center[q];
        AB2 += AB[q] * AB[q];
        CD2 += CD[q] * CD[q];
    }
#define XI(m, ie, jf) X[((size_t)(m) * ne + (ie)) * nf + (jf)]
    for (int pa = 0; pa < A->n_prim; ++pa)
        for (int pb = 0; pb < B->n_prim; ++pb) {
            const double a = A->alpha[pa], b = B->alpha[pb], p = a + b;
            double P[3], PA[3];
            for (int q = 0; q < 3; ++q) {
                P[q] = (a * A->center[q] + b * B->center[q]) / p;
                PA[q] = P[q] - A->center[q];
            }
            const double Kab = exp(-a * b / p * AB2) * A->wnorm[pa] * B->wnorm[pb];
            for (int pc = 0; pc < Cs->n_prim; ++pc)
                for (int pd = 0; pd < D->n_prim; ++pd) {
                    const double c = Cs->alpha[pc], d = D->alpha[pd], qq = c + d;
                    double Q[3], QC[3], W[3], WP[3], WQ[3], PQ2 = 0.0;
                    for (int k = 0; k < 3; ++k) {
                        Q[k] = (c * Cs->center[k] + d * D->center[k]) / qq;
                        QC[k] = Q[k] - Cs->center[k];
                        W[k] = (p * P[k] + qq * Q[k]) / (p + qq);
                        WP[k] = W[k] - P[k];
                        WQ[k] = W[k] - Q[k];
                        PQ2 += (P[k] - Q[k]) * (P[k] - Q[k]);
                    }
                    const double rho = p * qq / (p + qq);
                    const double Kcd = exp(-c * d / qq * CD2) * Cs->wnorm[pc] * D->wnorm[pd];
                    const double pre = 2.0 * pow(OR_PI, 2.5) / (p * qq * sqrt(p + qq)) * Kab * Kcd;
                    or_boys(M, rho * PQ2, F);
                    for (int m = 0; m <= M; ++m) XI(m, 0, 0) = pre * F[m];
                    /* [e0|00]^(m) */
                    for (int tot = 1; tot <= Le; ++tot)
                        for (int x = tot; x >= 0; --x)
                            for (int y = tot - x; y >= 0; --y) {
                                int e[3] = {x, y, tot - x - y};
                                int i = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
                                int e1[3] = {e[0], e[1], e[2]};
                                e1[i] -= 1;
                                const int ie = tri_index(e[0], e[1], e[2]), i1 = tri_index(e1[0], e1[1], e1[2]);
                                int i2 = -1;
                                if (e1[i] > 0) {
                                    int e2[3] = {e1[0], e1[1], e1[2]};
                                    e2[i] -= 1;
                                    i2 = tri_index(e2[0], e2[1], e2[2]);
                                }
                                for (int m = 0; m <= M - tot; ++m) {
                                    double v = PA[i] * XI(m, i1, 0) + WP[i] * XI(m + 1, i1, 0);
                                    if (i2 >= 0) v += e1[i] / (2.0 * p) * (XI(m, i2, 0) - rho / p * XI(m + 1, i2, 0));
                                    XI(m, ie, 0) = v;
                                }
                            }
                    /* [e0|f0]^(m), f built up with e at every total */
                    for (int ftot = 1; ftot <= Lf; ++ftot)
                        for (int fx = ftot; fx >= 0; --fx)
                            for (int fy = ftot - fx; fy >= 0; --fy) {
                                int f[3] = {fx, fy, ftot - fx - fy};
                                int i = f[0] > 0 ? 0 : (f[1] > 0 ? 1 : 2);
                                int f1[3] = {f[0], f[1], f[2]};
                                f1[i] -= 1;
                                const int jf = tri_index(f[0], f[1], f[2]), j1 = tri_index(f1[0], f1[1], f1[2]);
                                int j2 = -1;
                                if (f1[i] > 0) {
                                    int f2[3] = {f1[0], f1[1], f1[2]};
                                    f2[i] -= 1;
                                    j2 = tri_index(f2[0], f2[1], f2[2]);
                                }
                                for (int etot = 0; etot <= Le; ++etot)
                                    for (int x = etot; x >= 0; --x)
                                        for (int y = etot - x; y >= 0; --y) {
                                            int e[3] = {x, y, etot - x - y};
                                            const int ie = tri_index(e[0], e[1], e[2]);
                                            int ie1 = -1;
                                            if (e[i] > 0) {
                                                int e1[3] = {e[0], e[1], e[2]};
                                                e1[i] -= 1;
                                                ie1 = tri_index(e1[0], e1[1], e1[2]);
                                            }
                                            for (int m = 0; m <= M - ftot - etot; ++m) {
                                                double v = QC[i] * XI(m, ie, j1) + WQ[i] * XI(m + 1, ie, j1);
                                                if (j2 >= 0)
                                                    v += f1[i] / (2.0 * qq) * (XI(m, ie, j2) - rho / qq * XI(m + 1, ie, j2));
                                                if (ie1 >= 0) v += e[i] / (2.0 * (p + qq)) * XI(m + 1, ie1, j1);
                                                XI(m, ie, jf) = v;
                                            }
                                        }
                            }
                    for (int k = 0; k < ne * nf; ++k) Cc[k] += X[k];
                }
        }
#undef XI
    eri_hrr_t h;
    h.C = Cc;
    h.nf = nf;
    for (int q = 0; q < 3; ++q) {
        h.AB[q] = AB[q];
        h.CD[q] = CD[q];
    }
    int o = 0;
    for (int ia = 0; ia < A->nfn; ++ia)
        for (int ib = 0; ib < B->nfn; ++ib)
            for (int ic = 0; ic < Cs->nfn; ++ic)
                for (int id = 0; id < D->nfn; ++id)
                    out[o++] = eri_hrr(&h, comp_of(A->l, ia), comp_of(B->l, ib), comp_of(Cs->l, ic),
                                       comp_of(D->l, id)) *
                               A->scale[ia] * B->scale[ib] * Cs->scale[ic] * D->scale[id];
    free(X);
    free(Cc);
}

typedef struct {
    const shells_t *S;
    const int *quart; /* unique quartets (s1 >= s2, s3 >= s4, pair12 >= pair34), 4 ints each */
    double *out;
} eri_ctx_t;

static void eri_worker(void *vctx, long b, long e, int tid) {
    eri_ctx_t *X = (eri_ctx_t *)vctx;
    const int n = X->S->n_ao;
    double *buf = (double *)malloc(sizeof(double) * 10000);
    for (long k = b; k < e; ++k) {
        const int *q = X->quart + 4 * k;
        const shell_t *A = &X->S->sh[q[0]], *B = &X->S->sh[q[1]], *C = &X->S->sh[q[2]], *D = &X->S->sh[q[3]];
        eri_quartet(A, B, C, D, buf);
        int o = 0;
        for (int ia = 0; ia < A->nfn; ++ia)
            for (int ib = 0; ib < B->nfn; ++ib)
                for (int ic = 0; ic < C->nfn; ++ic)
                    for (int id = 0; id < D->nfn; ++id) {
                        const size_t i = A->ao_offset + ia, j = B->ao_offset + ib, kk = C->ao_offset + ic,
                                     l = D->ao_offset + id;
                        const double v = buf[o++];
                        const size_t N = n;
#define PUT(p, q, r, s) X->out[(((p) * N + (q)) * N + (r)) * N + (s)] = v
                        PUT(i, j, kk, l); PUT(j, i, kk, l); PUT(i, j, l, kk); PUT(j, i, l, kk);
                        PUT(kk, l, i, j); PUT(l, kk, i, j); PUT(kk, l, j, i); PUT(l, kk, j, i);
#undef PUT
                    }
    }
    free(buf);
}

/* Dense (ij|kl), out[n^4], from the unique shell quartets, all 8 permutations written. */
int or_eri_full(const or_basis_t *b, int n_threads, double *out) {
    shells_t S;
    if (shells_build(b, &S) != 0) return -1;
    const int ns = S.n_shells;
    long nq = 0;
    int *quart = (int *)malloc(sizeof(int) * 4 * (size_t)((long)ns * (ns + 1) / 2 * ((long)ns * (ns + 1) / 2 + 1) / 2 + 1));
    for (int s1 = 0; s1 < ns; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2)
            for (int s3 = 0; s3 <= s1; ++s3)
                for (int s4 = 0; s4 <= (s3 == s1 ? s2 : s3); ++s4) {
                    int *q = quart + 4 * nq++;
                    q[0] = s1; q[1] = s2; q[2] = s3; q[3] = s4;
                }
    eri_ctx_t ctx = {&S, quart, out};
    parallel_for(nq, n_threads, eri_worker, &ctx);
    free(quart);
    shells_free(&S);
    return 0;
}
