// kernels.hpp — kernel argument records and host launchers of the Vxc engine.
#pragma once
#include "common.hpp"

namespace chx {

struct RhoArgs {
    const DevShell* shells;
    const double2* prims;   // (exponent, c~ N(alpha; L00))
    const DevAo* aos;
    const double4* pts;     // sorted (x, y, z, w)
    const RhoTile* tiles;
    const RhoBlock* blocks;
    const int* block_ao;    // global AO index per block row
    const void* P;          // n_ao x n_ao, double or float
    int n_ao;
    int func;
    double4* wv;            // [sorted point] (w v_rho_a / 2, w u grad rho)
    double4* rho_out;       // [caller point] (rho, grad rho) when func < 0
    const int* perm;        // sorted point -> caller point
    double* e_part;         // [tile]
    double* n_part;         // [tile]
};

struct VArgs {
    const DevShell* shells;
    const double2* prims;
    const DevAo* aos;
    const double4* pts;     // sorted
    const VGroup* groups;
    const VPair* pairs;
    const int* group_ao;
    const double4* wv;      // per sorted point, or per caller point when wv_perm
    const int* perm;
    int wv_perm;
    double* scratch;
};

struct ReduceArgs {
    const int* loc;         // [group][n_ao] group-local row of each AO, or -1
    const VGroup* groups;
    int n_groups;
    int n_ao;
    const double* scratch;
    double* V;              // n_ao x n_ao, full symmetric
};

struct SigArgs {
    const DevShell* shells;
    const double2* prims;
    const DevAo* aos;
    const double4* pts;
    const int* cand_ao;     // [entry] AO index
    const int2* cand_pts;   // [entry] (first sorted point, count) of its block
    unsigned char* flag;    // [entry] 1 if the AO survives the value rule somewhere
    long long n_entries;
};

size_t rho_lds_bytes(int prec_bytes, int pts, int max_mpad);
hipError_t launch_rho(const RhoArgs& a, bool fp32, int pts, bool xc, int n_tiles, int max_mpad, hipStream_t s);
hipError_t launch_vxc(const VArgs& a, bool fp32, int n_pairs, hipStream_t s);
hipError_t launch_reduce_v(const ReduceArgs& a, hipStream_t s);
hipError_t launch_reduce_en(const double* e, const double* ne, int n, double* out, hipStream_t s);
hipError_t launch_to_float(const double* x, float* y, long long n, hipStream_t s);
hipError_t launch_significance(const SigArgs& a, hipStream_t s);

// Grid construction (grid.hip).
long build_grid_host(int n_atoms, const int* z, const double* xyz, int level, int n_rad, int n_ang,
                     int radial_min, double* out, long max_out, int* status);

}  // namespace chx
