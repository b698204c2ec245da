/*
 * calchem_hip.h — C ABI of the MI355X (gfx950) DFT exchange-correlation
 * quadrature engine: the per-SCF-iteration Vxc matrix, E_xc and integrated
 * electron count N of CalChem's RKS path.
 *
 * Conventions (kept from the reference GPU backend): plain pointers and
 * sizes, row-major FP64 matrices, caller-chosen 64-bit handles, int return
 * codes 0 ok, -1 bad arguments, -2 shape or handle mismatch, -3 no usable
 * device, -4 allocation failure, -5 launch or copy failure, -6 size limit
 * exceeded, -7 unsupported option.  Reference paths are relative to the
 * CalChem source tree.
 */
#ifndef CALCHEM_HIP_H
#define CALCHEM_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Record layouts of crates/calchem-cuda/src/lib.rs:46-85 (AoShellGpu,
 * AoPrimGpu, AoCompGpu, GridPtGpu), byte for byte. */
typedef struct {
    double center[3];      /* bohr */
    double r2_cutoff;      /* (40 + 4 l) / alpha_min */
    int32_t ao_offset;
    int32_t n_funcs;
    int32_t prim_start;
    int32_t n_prims;
    int32_t comp_start;
    int32_t atom_index;
} CalchemAoShell;

typedef struct {
    double exponent;
    double weighted_norm;  /* normalised contraction coefficient x N(alpha; L,0,0) */
} CalchemAoPrim;

typedef struct {
    double scale;          /* N(alpha_0; lx,ly,lz) / N(alpha_0; L,0,0) */
    int32_t lx, ly, lz, _pad;
} CalchemAoComp;

typedef struct {
    double x, y, z;
    int32_t batch_id;
    int32_t host_atom;
} CalchemGridPt;

/* Functional ids (crates/calchem-dft/src/xc.rs:16-35). */
enum {
    CALCHEM_XC_SVWN5 = 0,
    CALCHEM_XC_BLYP = 1,
    CALCHEM_XC_PBE = 2,
    CALCHEM_XC_B3LYP = 3,
    CALCHEM_XC_B3LYP5 = 4,
    CALCHEM_XC_PBE0 = 5
};

/* Precision of the two contractions (AO values, phi*P and phi^T (w v) phi).
 * The functional, per-point reductions and all sums across tiles are FP64
 * in both modes. */
enum { CALCHEM_PREC_FP64 = 0, CALCHEM_PREC_FP32 = 1 };

/* Grid levels (crates/calchem-dft/src/grid.rs:13-20); EXPLICIT = fixed
 * (n_rad, n_ang) per atom with no angular pruning. */
enum {
    CALCHEM_GRID_COARSE = 0,
    CALCHEM_GRID_MEDIUM = 1,
    CALCHEM_GRID_FINE = 2,
    CALCHEM_GRID_ULTRAFINE = 3,
    CALCHEM_GRID_EXPLICIT = 4
};

typedef struct {
    int32_t block_points;          /* max points per rho-stage block (default 320) */
    int32_t group_points;          /* target points per V-stage group (default 384) */
    size_t memory_limit_bytes;     /* 0 = no limit */
    double mixed_exact_threshold;  /* mixed builds keep AOs with max |phi| above this in FP64 (0 = 0.1) */
    int32_t rho_tile_points;       /* points per rho tile: 0 or 16 (other values: -1 / -7) */
    int32_t group_max_rows;        /* AO rows per register-resident V group: 0 = default (272), 192 or 272 */
    /* Per-geometry AO cache of the mixed build's rho stage (AO values and
     * gradients at every point, the leading rows FP64, the others FP32, about
     * 16 B per (significant AO, point)): 0 = when it fits a quarter of the
     * free device memory (and memory_limit_bytes), 1 = never, 2 = always (-3
     * / -6 when it cannot be had).  FP64 builds never use it. */
    int32_t ao_cache;
    /* Operand form of the mixed build's V-stage kernel for groups of up to 8
     * 16-row tiles: 0 or 1 = FP32 operands on v_mfma_f32_16x16x4f32 (the
     * default), 2 = each FP32 operand split into three BF16 terms on
     * v_mfma_f32_16x16x32_bf16 (the form the larger groups always use); which
     * is faster depends on the workload, so it is an autotuner choice.  Other
     * values: -1. */
    int32_t v_small_bf16;
} CalchemVxcOptions;

typedef struct {
    int64_t n_points_in, n_points, n_blocks, n_groups, n_rho_tiles, n_v_pairs;
    double flops_rho;                /* sum over rho blocks of 2 m_b^2 n_b, unpadded */
    double flops_v;                  /* sum over V groups of 2 m_g^2 n_g, unpadded */
    double flops_rho_executed_fp64;  /* with tile padding, as issued to MFMA */
    double flops_rho_executed_fp32;
    double flops_v_executed;
    int32_t max_m_rho, max_m_v, rho_points_fp64, rho_points_fp32;
    size_t device_bytes;
    size_t lds_bytes_per_block;      /* device limit used to size the rho tiles */
    int64_t profiled_builds;         /* builds timed since profiling was switched on */
    double ms_gather;                /* summed HIP-event times of those builds, per stage: */
    double ms_rho;                   /*   density-block gather, rho kernel(s), functional, */
    double ms_xc;                    /*   V kernel(s), V reduction, E_xc / N reduction */
    double ms_vxc;
    double ms_reduce_v;
    double ms_reduce_en;
    double flops_rho_fp64_rows;      /* part of flops_rho over the rows a mixed build keeps in FP64 */
    size_t ao_cache_bytes;           /* the mixed build's AO cache (0: none) */
} CalchemVxcStats;

/* ---- Device queries: crates/calchem-cuda/src/lib.rs:469-489 (is_cuda_available,
 *      CALCHEM_DISABLE_GPU) and cuda_kernels.cu exports
 *      calchem_cuda_is_available / calchem_cuda_get_memory_info /
 *      calchem_cuda_total_allocated_bytes.  CALCHEM_HIP_DISABLE=1 reports no device. */
int calchem_hip_is_available(void);
int calchem_hip_get_memory_info(size_t *free_bytes, size_t *total_bytes);
size_t calchem_hip_total_allocated_bytes(void);

/* ---- Reference-compatible GPU backend (functional on the host) ----
 * Replaces calchem_cuda_vxc_upload_grid (cuda_kernels.cu:7354-7450; Rust
 * binding crates/calchem-cuda/src/lib.rs:303-319, used by GpuVxcGrid::upload
 * lib.rs:1195-1260).  The batch CSR is validated; points are re-blocked
 * spatially inside the library. */
int calchem_hip_vxc_upload_grid(const CalchemAoShell *shells, int n_shells, const CalchemAoPrim *prims,
                                int n_prims, const CalchemAoComp *comps, int n_comps,
                                const CalchemGridPt *pts, int n_pts, const int *cand_offsets,
                                int n_cand_offsets, const int *cand_shells, int n_cand_shells, int n_ao,
                                uint64_t grid_id, size_t memory_limit_bytes);
/* Replaces calchem_cuda_vxc_eval_rho (cuda_kernels.cu:7452-7494; lib.rs:320-326,
 * GpuVxcGrid::eval_rho_grad lib.rs:1262-1283): rho_grad[n_pts*4] =
 * (rho, d rho/dx, d rho/dy, d rho/dz) per uploaded point for density P (n_ao^2). */
int calchem_hip_vxc_eval_rho(uint64_t grid_id, const double *P, double *rho_grad, int n_pts, int n_ao);
/* Replaces calchem_cuda_vxc_accum (cuda_kernels.cu:7496-7545; lib.rs:327-333,
 * GpuVxcGrid::accumulate_vxc lib.rs:1285-1306): wv[n_pts*4] = (w v_rho_a / 2,
 * w u_x, w u_y, w u_z) per point; V_out (n_ao^2) receives the full symmetric
 * Vxc (its lower triangle is what the reference returns). */
int calchem_hip_vxc_accum(uint64_t grid_id, const double *wv, double *V_out, int n_pts, int n_ao);
/* Replaces calchem_cuda_vxc_release (cuda_kernels.cu:7547-7562; lib.rs:334,
 * GpuVxcGrid::drop lib.rs:1184-1193).  Frees handles of either kind. */
int calchem_hip_vxc_release(uint64_t grid_id);

/* ---- Unrestricted build ----
 * Replaces UksVxcEngine::evaluate (crates/calchem-dft/src/vxc.rs:1662-1766)
 * and evaluate_vxc_uks (vxc.rs:1352-1470): spin densities Pa, Pb (n_ao^2,
 * row-major) in; Va, Vb (full symmetric), E_xc and N = sum_g w_g (rho_a + rho_b)
 * out.  Same plan as the restricted build; the first call allocates the
 * second spin's per-point buffers.  _device: device pointers, asynchronous
 * on `stream`, d_exc_nelec[2] = (E_xc, N).  Single-molecule plans only (-2
 * for a batched plan). */
int calchem_hip_vxc_build_uks(uint64_t handle, int xc, int precision, const double *Pa, const double *Pb,
                              double *Va_out, double *Vb_out, double *exc, double *nelec);
int calchem_hip_vxc_build_uks_device(uint64_t handle, int xc, int precision, const double *d_Pa,
                                     const double *d_Pb, double *d_Va, double *d_Vb, double *d_exc_nelec,
                                     void *stream);

/* ---- Many molecules in one plan ----
 * No single reference counterpart: the reference builds one molecule per
 * RksVxcEngine (vxc.rs:429-985); this puts n_mol independent molecules into
 * one plan so each build stage is one launch over all their blocks.
 * Molecule k owns shells [mol_shell_begin[k], mol_shell_begin[k+1]) (their
 * ao_offset local to the molecule; prim_start / comp_start index the shared
 * prims / comps arrays), points [mol_point_begin[k], mol_point_begin[k+1]) of
 * xyzw (x, y, z, w rows) and mol_n_ao[k] AOs.  Densities and Vxc matrices are
 * packed: molecule k's n_k x n_k row-major block follows those of molecules
 * 0..k-1.  exc / nelec receive n_mol values; d_exc_nelec[2k], [2k+1] =
 * (E_xc, N) of molecule k. */
int calchem_hip_vxc_create_batched(int n_mol, const int *mol_shell_begin, const int *mol_point_begin,
                                   const int *mol_n_ao, const CalchemAoShell *shells, const CalchemAoPrim *prims,
                                   int n_prims, const CalchemAoComp *comps, int n_comps, const double *xyzw,
                                   const CalchemVxcOptions *opts, uint64_t handle);
int calchem_hip_vxc_build_batched(uint64_t handle, int xc, int precision, const double *P_packed,
                                  double *V_packed, double *exc, double *nelec);
int calchem_hip_vxc_build_batched_device(uint64_t handle, int xc, int precision, const double *d_P_packed,
                                         double *d_V_packed, double *d_exc_nelec, void *stream);

/* ---- Captured builds ----
 * No reference counterpart (the reference launches kernels one by one per
 * batch): for callers that repeat the same build every SCF iteration.
 * graph_create runs the build once eagerly, then captures the whole build on
 * these device buffers (single-molecule or batched plan) into a HIP graph;
 * graph_launch replays it on `stream` (the caller rewrites d_P in place
 * between launches); graph_destroy frees it (release frees the rest). */
int calchem_hip_vxc_graph_create(uint64_t handle, int xc, int precision, const double *d_P, double *d_V,
                                 double *d_exc_nelec, uint64_t *graph_out);
int calchem_hip_vxc_graph_launch(uint64_t handle, uint64_t graph, void *stream);
int calchem_hip_vxc_graph_destroy(uint64_t handle, uint64_t graph);

/* ---- XC nuclear gradient ----
 * Replaces xc_nuclear_gradient_with_gpu (crates/calchem-dft/src/vxc.rs:2030-2334;
 * per-point semantics accumulate_xc_grad_points, vxc.rs:1769-1899, and
 * becke_weight_nuclear_gradient, crates/calchem-dft/src/grid.rs:423-584).
 * set_atoms gives a single-molecule plan its atoms (z, bohr; the shells'
 * atom_index name them) and, per point in the order handed to
 * calchem_hip_vxc_create, the raw (pre-Becke) weight and host atom
 * (MolecularGrid.points raw_weight / host_atom); up to 256 atoms.
 * xc_gradient returns grad[n_atoms*3] = dE_xc/dR_A at fixed density matrix:
 * unrestricted for spin densities Pa, Pb, or restricted with Pb = NULL and Pa
 * the total density. */
int calchem_hip_vxc_set_atoms(uint64_t handle, int n_atoms, const int *z, const double *xyz_bohr,
                              const double *raw_weight, const int *host_atom);
int calchem_hip_vxc_xc_gradient(uint64_t handle, int xc, const double *Pa, const double *Pb, double *grad);

/* ---- Fused build (functional on the device) ----
 * Replaces the per-geometry setup of RksVxcEngine::new_with_gpu
 * (crates/calchem-dft/src/vxc.rs:429-595): xyzw[n_pts*4] = (x, y, z, weight)
 * of MolecularGrid.points (grid.rs:33-51).  A handle already in use returns -2
 * (release it first: a plan is one geometry); so does create_batched. */
int calchem_hip_vxc_create(const CalchemAoShell *shells, int n_shells, const CalchemAoPrim *prims, int n_prims,
                           const CalchemAoComp *comps, int n_comps, const double *xyzw, int n_pts, int n_ao,
                           const CalchemVxcOptions *opts, uint64_t handle);
/* Replaces RksVxcEngine::evaluate (vxc.rs:720-985) for a spin-summed density
 * P (n_ao^2, host memory): V_out (n_ao^2, full symmetric), *exc = E_xc,
 * *nelec = sum_g w_g rho_g (may be NULL).  Synchronous. */
int calchem_hip_vxc_build(uint64_t handle, int xc, int precision, const double *P, double *V_out, double *exc,
                          double *nelec);
/* Same build on device pointers, enqueued on `stream` (a hipStream_t, NULL =
 * default stream) without synchronising: d_exc_nelec[0] = E_xc,
 * d_exc_nelec[1] = N.  Builds on one handle share its scratch: issue them on
 * one stream. */
int calchem_hip_vxc_build_device(uint64_t handle, int xc, int precision, const double *d_P, double *d_V,
                                 double *d_exc_nelec, void *stream);
/* Per-kernel HIP-event timing of subsequent builds (on = 1) and plan statistics;
 * no reference counterpart (the reference prints CALCHEM_PROFILE stage timers,
 * vxc.rs:571-578). */
int calchem_hip_vxc_set_profiling(uint64_t handle, int on);
int calchem_hip_vxc_stats(uint64_t handle, CalchemVxcStats *out);

/* ---- Molecular grid ----
 * Replaces MolecularGrid::build / build_for_basis (grid.rs:614-634,732-828)
 * and an explicit unpruned (n_rad x n_ang) grid.  Points, Becke weights and
 * the weight filter are computed on the GPU (the host lists radial shells).  out[max_out*6] rows = (x, y, z, weight, raw_weight, host_atom),
 * in the reference's atom / radial / angular order, weight > 1e-18 only.
 * With out == NULL returns the point count; otherwise the count written, or
 * -6 when max_out is too small. */
long calchem_hip_grid_build(int n_atoms, const int *z, const double *xyz_bohr, int level, int n_rad, int n_ang,
                            int radial_min, double *out, long max_out);
/* Streaming grids (MolecularGrid::streaming_tasks / generate_task_points,
 * grid.rs:685-782): the (atom, first radial shell, last radial shell) tasks of
 * 8 radial shells each, tasks[3*k..], and the points of one task, generated
 * and weighted on the device, rows as in calchem_hip_grid_build.  With a NULL
 * buffer each returns the count; -6 when max_* is too small.  The tasks'
 * points concatenated in task order are the points of calchem_hip_grid_build. */
long calchem_hip_grid_streaming_tasks(int n_atoms, const int *z, int level, int n_rad, int radial_min, int *tasks,
                                      long max_tasks);
long calchem_hip_grid_task_points(int n_atoms, const int *z, const double *xyz_bohr, int level, int n_rad, int n_ang,
                                  int radial_min, int atom, int rad_start, int rad_end, double *out, long max_out);

/* ---- Integrals of the SCF Hamiltonian ----
 * The SCF row beside the Vxc build (crates/calchem-scf/src/rks.rs:131-615):
 * one-electron integrals (compute_overlap / compute_kinetic / compute_nuclear,
 * crates/calchem-int/src/one_electron.rs:115-168) and J / K
 * (RksFockBuilder, crates/calchem-scf/src/fock_builder.rs; exact integrals in
 * the reference, direct_fock.rs:1028) from Cholesky-decomposed repulsion
 * integrals resident on the device.  Shells up to d (-7 above).
 * create: the basis in the Vxc records above (same shell / primitive /
 * component conventions) and the nuclei atoms_xyzz[n_atoms*4] = (x, y, z, Z)
 * in bohr; it also computes the repulsion diagonal.  A handle in use: -2. */
int calchem_hip_int_create(const CalchemAoShell *shells, int n_shells, const CalchemAoPrim *prims, int n_prims,
                           const CalchemAoComp *comps, int n_comps, int n_ao, int n_atoms, const double *atoms_xyzz,
                           uint64_t handle);
int calchem_hip_int_release(uint64_t handle);
/* S, T and V = -sum_C Z_C (mu|1/r_C|nu), each n_ao^2 row-major (host memory). */
int calchem_hip_int_one_electron(uint64_t handle, double *S, double *T, double *V);
/* (mu nu|mu nu) for every packed pair mu >= nu, index mu (mu + 1) / 2 + nu (host, n_ao (n_ao + 1) / 2). */
int calchem_hip_int_eri_diag(uint64_t handle, double *out);
/* Exact repulsion columns: for each shell pair (shell_a[k], shell_b[k]) its
 * n_a n_b columns (component of shell_a major), each holding
 * (mu nu|lam sig) over the packed pairs mu >= nu (host memory). */
int calchem_hip_int_eri_columns(uint64_t handle, int n_kets, const int *shell_a, const int *shell_b, double *out);
/* Pivoted Cholesky decomposition of the repulsion integrals down to residual
 * diagonal tau (max_rank 0: up to 25 n_ao vectors); *rank_out = vectors kept. */
int calchem_hip_int_cholesky(uint64_t handle, double tau, int max_rank, int *rank_out);
/* J[mu nu] = sum (mu nu|lam sig) P[lam sig] and K[mu nu] = sum (mu lam|nu sig) P[lam sig]
 * (n_ao^2 each, either may be NULL) for P = Cpos Cpos^T - Cneg Cneg^T, with
 * Cpos (n_ao x n_pos) and Cneg (n_ao x n_neg) row-major in host memory. */
int calchem_hip_int_jk(uint64_t handle, int n_pos, const double *Cpos, int n_neg, const double *Cneg, double *J,
                       double *K);

/* ---- The atom-pair network of the learned SCF guess ----
 * No reference counterpart: the network replaces sad_guess at
 * crates/calchem-scf/src/rks.rs:221-231 (SAD stays the fallback).  A
 * perceptron d_in -> d_hidden -> d_hidden -> d_out, SiLU on the hidden layers,
 * FP32 parameters in one array W1[h][d_in], b1[h], W2[h][h], b2[h],
 * W3[d_out][h], b3[d_out].  Rows are samples (row-major, host memory).
 * Loss: masked mean squared error sum mask (y - t)^2 / sum mask. */
int calchem_hip_mlp_create(int d_in, int d_hidden, int d_out, const float *params, uint64_t handle);
int calchem_hip_mlp_release(uint64_t handle);
int calchem_hip_mlp_get_params(uint64_t handle, float *params);
int calchem_hip_mlp_forward(uint64_t handle, int n, const float *x, float *y);
/* Loss and its gradient with respect to every parameter (grad may be NULL). */
int calchem_hip_mlp_loss_grad(uint64_t handle, int n, const float *x, const float *t, const float *mask, float *loss,
                              float *grad);
/* Training set resident on the device, then Adam over shuffled minibatches with
 * a cosine learning-rate decay; loss_out[epochs] = mean minibatch loss per epoch. */
int calchem_hip_mlp_set_data(uint64_t handle, int n, const float *x, const float *t, const float *mask);
int calchem_hip_mlp_train(uint64_t handle, int epochs, int batch, float lr, uint64_t seed, float *loss_out);

/* ---- The functional on caller points ----
 * No reference entry point: evaluates eval_xc_spin (crates/calchem-dft/src/xc.rs:159-223)
 * with the device code the builds use, at arbitrary spin invariants, so that
 * its branches can be checked against fixed values (the reference's own
 * checks at xc.rs:1212-1247 call eval_xc_spin directly).
 * in[n*5] = (rho_a, rho_b, gamma_aa, gamma_ab, gamma_bb) per point; w[n] the
 * weights (NULL = 1); out[n*6] = (f, df/drho_a, df/drho_b, df/dgamma_aa,
 * df/dgamma_ab, df/dgamma_bb); sums[2] = (sum w f, sum w (rho_a + rho_b))
 * summed in a fixed order, or NULL.  _device: device pointers, enqueued on
 * `stream` (a hipStream_t, NULL = default stream); the per-workgroup partial
 * sums of a call are stream-ordered allocations of that call.  The host form
 * is synchronous. */
int calchem_hip_xc_eval(int xc, long long n, const double *in, const double *w, double *out, double *sums);
int calchem_hip_xc_eval_device(int xc, long long n, const double *d_in, const double *d_w, double *d_out,
                               double *d_sums, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): calchem_hip_xc_eval_device without any allocation.  The
 * partial sums live in a fixed 1 MiB device array per GPU, so n is at most
 * 16,777,216 (-6 above).  Calls that overlap on one GPU share that array and
 * race: issue them on one stream. */
int calchem_hip_xc_eval_device_noalloc(int xc, long long n, const double *d_in, const double *d_w, double *d_out,
                                       double *d_sums, void *stream);

#ifdef __cplusplus
}
#endif
#endif
