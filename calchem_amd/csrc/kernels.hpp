// kernels.hpp — kernel argument records and host launchers of the Vxc engine.
#pragma once
#include "common.hpp"

namespace chx {

struct RhoArgs {
    const DevShell* shells;
    const double2* prims;   // (exponent, c~ N(alpha; L00))
    const DevAo* aos;
    const double4* pts;     // sorted (x, y, z, w)
    const int2* tiles;      // this launch's tiles: (block, first point within the block)
    const RhoBlock* blocks;
    const int2* block_rows; // per block row: (global AO, block-local shell) or (-1, 0)
    const int2* block_sh;   // per block shell: (global shell, block-local primitive offset)
    const uint4* meta;      // packed per-block records (RhoBlock::meta)
    const double* P64;      // gathered FP64 density blocks (RhoBlock::psub)
    const void* PS;         // gathered density blocks in the precision of the non-leading rows
    double4* rho_out;       // (rho, grad rho) per sorted point, or per caller point when rho_perm
    const int* perm;        // sorted point -> caller point
    int rho_perm;
    int seg[kXcds + 1];     // this launch's XCD segments (RhoClass::seg)
    // two-spin launches only: the beta density blocks and beta output
    const double* P64b;
    const void* PSb;
    double4* rho_out_b;
    // Per-geometry AO cache of the mixed build (null: AOs evaluated per
    // build).  aoc_off[t] is tile t's offset in 16-byte units; see
    // aoc_tile_bytes for the layout.
    void* aoc;
    const long long* aoc_off;
};

// AO cache of the mixed rho stage: per rho workgroup tile, per 16-point
// sub-tile, two planes over the block's 16-row tiles in order: phi, then
// (dphi/dx, dphi/dy, dphi/dz), each [v][lane] exactly as the rho kernel's
// lanes hold them (slot v of lane (j, q): row Mma<float>::row(q, v) of the
// row tile, point j), FP64 for the block's leading row tiles (rows kept FP64)
// and FP32 for the others.  The kernels read phi alone to stage the MFMA
// operand and the gradients (with phi) for the epilogue, so the first read
// touches a quarter of the bytes.
constexpr long long kAocPhi64 = 64LL * 4 * 8;    // bytes of one leading row tile's phi plane
constexpr long long kAocPhi32 = 64LL * 4 * 4;    // ... of an FP32 row tile
constexpr long long kAocGrad64 = 64LL * 4 * 24;  // bytes of one leading row tile's gradient plane
constexpr long long kAocGrad32 = 64LL * 4 * 12;  // ... of an FP32 row tile
__host__ __device__ inline long long aoc_phi_off(int it, int l16) {
    return it < l16 ? it * kAocPhi64 : l16 * kAocPhi64 + (it - l16) * kAocPhi32;
}
__host__ __device__ inline long long aoc_grad_off(int it, int l16, int t16) {
    return aoc_phi_off(t16, l16) + (it < l16 ? it * kAocGrad64 : l16 * kAocGrad64 + (it - l16) * kAocGrad32);
}
__host__ __device__ inline long long aoc_sub_bytes(int m_pad, int m64) {
    return aoc_grad_off(m_pad >> 4, m64 >> 4, m_pad >> 4);
}

// Functional-stage workgroups never straddle molecules: workgroup w covers
// sorted points [wg[w].x, wg[w].x + wg[w].y).
struct XcArgs {
    const double4* rho;     // per sorted point
    const double4* pts;     // sorted (x, y, z, w)
    const int2* wg;
    int n_wg;
    int func;
    double4* wv;            // per sorted point
    double* e_part;         // per workgroup
    double* n_part;
};

// Unrestricted functional stage: both spin densities in, both wv out.
struct XcUksArgs {
    const double4* rho_a;   // per sorted point (rho_a, grad rho_a)
    const double4* rho_b;
    const double4* pts;
    const int2* wg;
    int n_wg;
    int func;
    double4* wv_a;          // (w v_a / 2, w u_a), u_a = 2 v_gaa grad rho_a + v_gab grad rho_b
    double4* wv_b;
    double* e_part;
    double* n_part;
};

// XC nuclear gradient (grad_kernels.hip).
struct XcGradArgs {
    const double4* rho_a;   // per sorted point; total density when rho_b is null
    const double4* rho_b;   // null: restricted (spin densities = half the total)
    const double4* pts;
    int n_pts;
    int func;
    double4* wx_a;          // (w v_rho_a, w u_a)
    double4* wx_b;
    double* fxc;            // f_xc per point (0 where screened)
};

struct GradAoArgs {
    const DevShell* shells;
    const double2* prims;
    const DevAo* aos;
    const double4* pts;
    const int2* tiles;      // (block, first point within it)
    const RhoBlock* blocks;
    const int2* block_rows;
    const int2* block_sh;
    const int* shell_atom;  // atom of each global shell
    const double* PS;       // gathered FP64 density blocks of this pass
    const double4* wx;      // (w v_rho, w u) of this pass's spin, per sorted point
    const int* host;        // host atom per sorted point
    int n_atoms;
    double factor;          // -2 (per spin; restricted: total density, once)
    double* partial;        // [tile][3 n_atoms]
};

struct GradBeckeArgs {
    const double4* pts;
    const double* raw_w;    // per sorted point
    const int* host;
    const double* fxc;
    const double* xyz;      // [3 n_atoms] bohr
    const double* inv_r;    // [n_atoms^2] 1 / R_cd (0 for coincident atoms)
    const double* asym;     // [n_atoms^2] a_cd
    int n_atoms;
    int n_pts;
    int pts_per_wave;
    double* partial;        // [wave][3 n_atoms]
};

size_t grad_ao_lds_bytes(int max_mpad, int max_nsh, int max_prim);
hipError_t launch_xc_grad(const XcGradArgs& a, hipStream_t s);
hipError_t launch_grad_ao(const GradAoArgs& a, int n_tiles, size_t lds, hipStream_t s);
hipError_t launch_grad_becke(const GradBeckeArgs& a, int n_waves, hipStream_t s);
hipError_t launch_grad_sum(const double* part, long long rows, int width, long long rows_per_chunk, double* out,
                           hipStream_t s);
// Becke pair tables of a molecule (grid.hip): 1 / R_cd and the size
// adjustment a_cd (antisymmetric), row-major n x n.
void becke_pair_tables(int n_atoms, const int* z, const double* xyz, double* inv_r, double* asym);

// Launch class of rho tiles: blocks of similar size share one LDS budget.
// A launch class of rho tiles.  Its tiles are split into kXcds segments of
// about equal cost, segment x = [seg[x], seg[x + 1]) (class-local), each a
// run of whole blocks, costliest first; XCD x walks segment x.
struct RhoClass {
    int tile_begin, n_tiles;  // range in the class-ordered list of rho workgroups (kRhoSub x 16 points)
    int grad_tile_begin, grad_n_tiles;  // range in the 16-point tile list of the XC gradient's AO kernel
    int max_mpad, max_nsh, max_m64, max_prim;
    int seg[kXcds + 1];
    int block_begin, n_blocks;  // the class's blocks: range in the plan's class-ordered block list
    double cost;                // sum of its tiles' costs (m_pad^2 plus a fixed part per tile)
};

struct VArgs {
    const DevShell* shells;
    const double2* prims;
    const DevAo* aos;
    const double4* pts;     // sorted
    const VGroup* groups;
    const VPair* pairs;     // tile-pair kernel
    const int* group_list;  // register-resident kernel: groups it handles
    const int2* group_rows; // per group row: (global AO, group-local shell)
    const int2* group_sh;   // per group shell: (global shell, group-local primitive offset)
    const uint4* gmeta;     // packed per-group records (VGroup::meta)
    const double4* wv;      // per sorted point, or per caller point when wv_perm
    const int* perm;
    int wv_perm;
    double* scratch;
};

// Per build: copy P[row_i][row_j] of every block into its m_pad x m_pad
// slot (zero for padding rows), in FP64 and/or FP32.
struct GatherArgs {
    const RhoBlock* blocks;
    const int2* block_rows;
    const MolInfo* mols;
    const double* P;        // packed per-molecule density blocks
    double* out64;          // may be null
    float* out32;           // may be null
    int out64_leading_only; // write out64 only for the block's first m64 columns
    const int* block_list;  // workgroup w gathers block block_list[w] (null: block w)
};

// One (AO mu, group holding mu) entry of the V reduction: where row li(mu)
// of the group's lower tiles starts in scratch, the group's row list, and
// li | tile_shift << 16.
struct RedEntry {
    long long base;
    int rows;
    int li_ts;
};

struct ReduceArgs {
    const int* ao_gbegin;   // [n_ao + 1] CSR of the entries of each AO, groups ascending
    const RedEntry* entries;
    const int* group_ao;    // per group row: its global AO
    const int* ao_mol;      // molecule of each global AO
    const MolInfo* mols;
    int n_ao;               // global AO rows
    const double* scratch;
    double* V;              // packed per-molecule V blocks, full symmetric
    // With n_seg > 1 each row's entries are split over n_seg workgroups
    // (blockIdx.y), whose partial rows go to part[(seg n_ao + mu) row_stride
    // + nu] and are summed in segment order by a second kernel.
    int n_seg;
    int row_stride;
    double* part;
};

struct SigArgs {
    const DevShell* shells;
    const double2* prims;
    const DevAo* aos;
    const double4* pts;
    const int* cand_ao;     // [entry] AO index
    const int2* cand_pts;   // [entry] (first sorted point, count) of its block
    unsigned char* flag;    // [entry] 1 if the AO survives the value rule somewhere
    double* maxabs;         // [entry] max |phi| over the block's points
    long long n_entries;
};

// Rho kernel variants: FP64 throughout, or mixed (leading rows FP64, the
// rest FP32).
enum RhoMode { kRhoFp64 = 0, kRhoMixed = 1 };
size_t rho_lds_bytes(int mode, const RhoClass& c, int ns = 1, bool cached = false);
// Mixed launches read the AO cache when a.aoc is set.
hipError_t launch_rho(const RhoArgs& a, int mode, const RhoClass& c, hipStream_t s, int ns = 1);
// Fill a.aoc for one launch class (mixed layout): the AO evaluation of the
// mixed rho kernel, stored instead of contracted.
hipError_t launch_rho_fill_cache(const RhoArgs& a, const RhoClass& c, hipStream_t s);
hipError_t launch_xc(const XcArgs& a, hipStream_t s);
hipError_t launch_xc_uks(const XcUksArgs& a, hipStream_t s);
hipError_t launch_vxc(const VArgs& a, bool fp32, int n_pairs, hipStream_t s);
// Register-resident V kernel over groups of at most max_tiles x 16 rows
// (max_tiles = 17 or 12).
// max_tiles: 8, 12 or 17; bf16: the mixed kernel's BF16 three-term operand form (always for 12 and 17)
size_t vgroup_lds_bytes(bool fp32, int max_tiles, int max_nsh, int max_prim, bool bf16 = true);
hipError_t launch_vxc_group(const VArgs& a, bool fp32, int max_tiles, int n_groups, int max_nsh, int max_prim,
                            hipStream_t s, bool bf16 = true);
int reduce_v_waves(int max_mol_ao, size_t lds_limit);
hipError_t launch_reduce_v(const ReduceArgs& a, int waves, int max_mol_ao, hipStream_t s);
// Per molecule: (E_xc, N) = sums of its functional-stage partials, into out[2k], out[2k+1].
hipError_t launch_reduce_en(const double* e, const double* ne, const MolInfo* mols, int n_mol, double* out,
                            hipStream_t s);
hipError_t launch_to_float(const double* x, float* y, long long n, hipStream_t s);
hipError_t launch_fill_zero(double* x, long long n, hipStream_t s);
// max_mpad: the largest m_pad of the blocks (at most 16384).
hipError_t launch_gather_p(const GatherArgs& a, int n_blocks, int max_mpad, hipStream_t s);
hipError_t launch_significance(const SigArgs& a, hipStream_t s);

// Grid construction (grid.hip): points, Becke weights and the w > 1e-18
// filter on the device; build_grid_task restricts to one atom's radial shells
// rad_start..rad_end (task_atom < 0: the whole grid).
long build_grid_host(int n_atoms, const int* z, const double* xyz, int level, int n_rad, int n_ang,
                     int radial_min, double* out, long max_out, int* status);
long build_grid_task(int n_atoms, const int* z, const double* xyz, int level, int n_rad, int n_ang, int radial_min,
                     int task_atom, int rad_start, int rad_end, double* out, long max_out, int* status);
int atom_radial_count(int z, int level, int n_rad, int radial_min);

}  // namespace chx
