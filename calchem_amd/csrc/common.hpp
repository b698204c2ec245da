// common.hpp — shared host/device records of the calchem_hip Vxc engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace chx {

// Return codes of the C ABI (same meaning as the reference GPU backend,
// crates/calchem-cuda/src/cuda_kernels.cu:7372-7541): 0 ok, -1 bad arguments,
// -2 shape / handle mismatch, -3 no usable device, -4 allocation failure,
// -5 launch or copy failure, -6 size limit exceeded; -7 unsupported option.
enum Status : int {
    kOk = 0,
    kBadArgs = -1,
    kMismatch = -2,
    kNoDevice = -3,
    kAllocFail = -4,
    kLaunchFail = -5,
    kLimit = -6,
    kUnsupported = -7,
};

// One shell as the AO kernels read it (48 bytes).
struct DevShell {
    double cx, cy, cz, r2cut;
    int prim_start, n_prims, l, ao_offset;
};

// One AO row: its shell, Cartesian exponents and component scale.
struct DevAo {
    double scale;
    int shell, lx, ly, lz;
};

// rho-stage block: a spatially compact set of points and the AOs that are
// significant at any of them.  Rows whose |phi| exceeds the mixed-precision
// threshold somewhere in the block come first (m64 of them, rounded up to 16).
struct RhoBlock {
    int ao_begin;  // offset into the block row list
    int m;         // significant AOs
    int m_pad;     // m rounded up to the 16-row tile
    int m64;       // leading rows kept in FP64 by the mixed build
    int sh_begin;  // offset into the block shell list
    int n_sh;      // shells owning the rows
    int n_prim;    // primitives of those shells
    int mol;       // molecule of a batched plan (0 otherwise)
    long long psub;  // offset of the block's gathered density (m_pad x m_pad, row-major)
    int pt_begin;  // first sorted point
    int n_pts;
    // Offset, in 16-byte units, of the block's packed records: RowL[m_pad]
    // (padding rows with sh = -1), DevShell[n_sh] (prim_start block-local),
    // double2[n_prim]: what the rho kernel stages in LDS, in LDS order.
    long long meta;
};

// One molecule of a plan (a single-molecule plan has one).  Its AOs are the
// global rows [ao_base, ao_base + n_ao); its density and V blocks sit at
// p_off / v_off (elements) in the packed n_ao x n_ao row-major matrices; its
// sorted points are [pt_begin, pt_begin + n_pts) and its functional-stage
// workgroups [xc_wg_begin, xc_wg_begin + n_xc_wg).
struct MolInfo {
    int ao_base, n_ao;
    long long p_off;
    int pt_begin, n_pts;
    int xc_wg_begin, n_xc_wg;
};

// A row as the kernels keep it in LDS: component scale, Cartesian exponents
// (2 bits each) and the block-local shell (-1 for padding rows).
struct RowL {
    double scale;
    int lmn;
    int sh;
};


// V-stage group (a union of neighbouring rho blocks).  Its rows are the
// union of the blocks' AOs in ascending order; V tiles of (1 << tile_shift)
// rows go to scratch as lower tile pairs (I >= J), row-major inside a tile.
struct VGroup {
    int ao_begin, m, n_tiles, tile_shift;
    int pt_begin, n_pts;
    int sh_begin, n_sh;
    long long scratch_off;  // in doubles
    int n_prim, pad;
    // Register-resident groups: offset, in 16-byte units, of the packed
    // records RowL[16 n_tiles] (padding rows with sh = -1), DevShell[n_sh]
    // (prim_start group-local), double2[n_prim]; -1 otherwise.
    long long meta;
};

// One workgroup of the tile-pair V kernel: a 64x64 lower tile pair of a
// group too wide for the register-resident kernel.
struct VPair {
    int group, ti, tj, pad;
};

constexpr int kVTile = 64;       // tile of the tile-pair V kernel
constexpr int kVPts = 16;        // points per V-stage step (MFMA k extent 4 x 4)
constexpr int kThreads = 256;    // 4 wavefronts of 64
constexpr int kRhoPts = 16;      // points per rho-stage tile (MFMA N extent)
constexpr int kRhoSub = 2;       // 16-point tiles one rho workgroup takes in turn (same block)
constexpr int kDefaultBlockPoints = 320;  // rho-block size when the caller gives none (measured: 20 waters
                                           // 4 % faster mixed than 256, C8H18 and H2O within 1 %)
constexpr int kSmallGroupTiles = 8;  // V groups of at most this many 16-row tiles use the small-group kernels
constexpr int kGroupMaxTiles = 17;   // default 16-row tiles per register-resident V group (272 rows)
// Elements after the last gathered density block: the rho kernel's look-ahead
// loads may read up to 3 chunks of 16 past a row's end (values unused).
constexpr int kPsubTail = 64;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X, each with its own L2

inline bool hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "calchem_hip: %s failed: %s\n", what, hipGetErrorString(e));
        return false;
    }
    return true;
}

}  // namespace chx
