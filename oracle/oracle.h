/*
 * oracle.h — CPU FP64 restatement of CalChem's DFT exchange-correlation
 * grid quadrature (Vxc / E_xc / N).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * calchem_amd HIP path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Every function follows a reference file:line (paths relative to the
 * CalChem source tree) and is pinned by the checks listed in oracle.c.
 */
#ifndef CALCHEM_ORACLE_H
#define CALCHEM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* XC functional ids (crates/calchem-dft/src/xc.rs:16-35). */
enum { OR_SVWN5 = 0, OR_BLYP = 1, OR_PBE = 2, OR_B3LYP = 3, OR_B3LYP5 = 4, OR_PBE0 = 5 };

/* Grid levels (crates/calchem-dft/src/grid.rs:13-20); OR_LEVEL_EXPLICIT
 * means fixed (n_rad, n_ang) with no angular pruning. */
enum { OR_COARSE = 0, OR_MEDIUM = 1, OR_FINE = 2, OR_ULTRAFINE = 3, OR_LEVEL_EXPLICIT = 4 };

/* Raw (unnormalised) shell description handed in by the caller.  Shell s
 * has n_prim[s] primitives starting at prim_start[s] in exps/coefs. */
typedef struct {
    int n_shells;
    const int *atom;        /* [n_shells] */
    const double *center;   /* [n_shells*3] bohr */
    const int *l;           /* [n_shells] 0..3 */
    const int *n_prim;      /* [n_shells] */
    const int *prim_start;  /* [n_shells] */
    const double *exps;     /* [total prims] */
    const double *coefs;    /* [total prims] raw contraction coefficients */
} or_basis_t;

double or_primitive_norm(double alpha, int l, int m, int n);
int or_normalized_contraction(int n_prim, const double *exps, const double *coefs, int L,
                              double *out);
int or_n_ao(const or_basis_t *b);

int or_lebedev(int n_ang, double *xyzw /* [n_ang*4] */);

/* Build the molecular grid.  out: [max_out*6] rows (x, y, z, w, raw_w, host).
 * Returns the number of points written, or -(needed) when max_out is too
 * small, or -1 on bad arguments. */
long or_build_grid(int n_atoms, const int *z, const double *xyz_bohr, int level, int n_rad,
                   int n_ang, int radial_min, double *out, long max_out);
double or_becke_weight_fast(int n_atoms, const int *z, const double *xyz_bohr, int host,
                            const double r[3]);
double or_bragg_slater_bohr(int z);

/* AO values and gradients at one point over all shells.  val/grad rows are
 * written for every AO (zero where the reference drops it). */
void or_eval_aos(const or_basis_t *b, const double r[3], double *val, double *grad /* [n_ao*3] */);

/* out[6] = f, v_rho_a, v_rho_b, v_gamma_aa, v_gamma_ab, v_gamma_bb. */
void or_eval_xc(int func, double ra, double rb, double gaa, double gab, double gbb, double *out);

/* Per-point RKS path (vxc.rs:99-384).  pts: [n_pts*4] (x, y, z, w).
 * V: [n_ao*n_ao] full symmetric.  Returns 0 or a negative error. */
int or_vxc_rks_perpoint(const or_basis_t *b, long n_pts, const double *pts, int func,
                        const double *P, int n_threads, double *V, double *exc, double *nelec);

/* Cached-batch BLAS-3 RKS path (vxc.rs:429-985, CPU branch). */
int or_vxc_rks_cached(const or_basis_t *b, long n_pts, const double *pts, int func,
                      const double *P, int n_threads, double *V, double *exc, double *nelec);

/* Same as or_vxc_rks_cached restricted to batches [batch_begin, batch_end)
 * of the 128-point batching; used to time a bounded CPU sample. Returns
 * the number of batches processed. */
long or_vxc_rks_cached_range(const or_basis_t *b, long n_pts, const double *pts, int func,
                             const double *P, int n_threads, long batch_begin, long batch_end,
                             double *V, double *exc, double *nelec, double *flops);

/* XC nuclear gradient dE_xc/dR_A (xc_nuclear_gradient_cpu semantics:
 * accumulate_xc_grad_points, vxc.rs:1769-1899, and
 * becke_weight_nuclear_gradient, grid.rs:423-584) for spin densities Pa, Pb
 * (RKS: Pa = Pb = P/2).  pts6: [n_pts*6] grid rows (x, y, z, w, raw_w,
 * host).  The basis' shell atom indices name the atoms.  grad: [n_atoms*3]. */
int or_xc_gradient(const or_basis_t *b, int n_atoms, const int *z, const double *xyz_bohr, long n_pts,
                   const double *pts6, int func, const double *Pa, const double *Pb, int nt, double *grad);

/* Per-point UKS path (vxc.rs:1141-1470). */
int or_vxc_uks_perpoint(const or_basis_t *b, long n_pts, const double *pts, int func,
                        const double *Pa, const double *Pb, int n_threads, double *Va,
                        double *Vb, double *exc, double *nelec);

/* rho and grad rho at each point (x, y, z, w rows), out [n_pts*4]. */
int or_rho_points(const or_basis_t *b, long n_pts, const double *pts, const double *P, int n_threads,
                  double *out);

/* Dense AO values [n_pts*n_ao] and gradients [n_pts*n_ao*3] (x, y, z, w rows). */
int or_ao_dense(const or_basis_t *b, long n_pts, const double *pts, double *val, double *grad);

/* Analytic overlap S (Obara-Saika 1D recursion), [n_ao*n_ao]. */
int or_overlap(const or_basis_t *b, double *S);

/* Boys function F_m(T), m = 0..mmax (<= 63), long-double series / recursion. */
void or_boys(int mmax, double T, double *F);
/* Overlap S, kinetic T and nuclear attraction V = -sum_C Z_C (mu|1/r_C|nu)
 * ([n_ao*n_ao] each, any may be NULL) over the n_atoms charges z at xyz (bohr):
 * one_electron.rs:7-168 (Obara-Saika). */
int or_one_electron(const or_basis_t *b, int n_atoms, const int *z, const double *xyz, double *S, double *T,
                    double *V);
/* Dense electron repulsion integrals (ij|kl), out[n_ao^4]: eri.rs:276-1018
 * (Obara-Saika VRR, Head-Gordon-Pople HRR). */
int or_eri_full(const or_basis_t *b, int n_threads, double *out);

#ifdef __cplusplus
}
#endif
#endif
