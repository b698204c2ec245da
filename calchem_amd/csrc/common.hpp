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
// significant at any of them.
struct RhoBlock {
    int ao_begin;  // offset into the block AO list
    int m;         // significant AOs
    int m_pad;     // m rounded up to the 32-row tile
    int pad;
};

// One workgroup of the rho kernel: up to kRhoPts consecutive sorted points
// of one block.
struct RhoTile {
    int block;
    int pt_begin;  // sorted point index
    int n_pts;
    int pad;
};

// V-stage group (a union of neighbouring rho blocks).
struct VGroup {
    int ao_begin, m, m_pad, n_tiles64;
    int pt_begin, n_pts;
    long long scratch_off;  // in doubles
};

// One workgroup of the V kernel: a 64x64 lower tile pair (I >= J) of a group.
struct VPair {
    int group, ti, tj, pad;
};

constexpr int kRhoPts = 32;     // points per rho workgroup
constexpr int kVTile = 64;      // V-stage AO tile
constexpr int kVPts = 16;       // points staged per V-stage step
constexpr int kThreads = 256;   // 4 wavefronts of 64

inline bool hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "calchem_hip: %s failed: %s\n", what, hipGetErrorString(e));
        return false;
    }
    return true;
}

}  // namespace chx
