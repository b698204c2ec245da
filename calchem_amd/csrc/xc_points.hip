// xc_points.hip — the device functional (xc_device.hpp) on caller-supplied
// points, with the weighted sums E = sum_i w_i f_i and N = sum_i w_i (rho_a + rho_b).
//
// The build kernels evaluate the functional only at the densities a grid
// produces; this entry point evaluates it at any (rho_a, rho_b, gamma_aa,
// gamma_ab, gamma_bb), so that the branches of xc.rs (density cut-off,
// gamma clamp, LYP underflow, PBE large-w and small-gamma forms;
// crates/calchem-dft/src/xc.rs:159-223,505-516,744-781) can be checked point
// by point against fixed values.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>

#include "../../include/calchem_hip.h"
#include "common.hpp"
#include "xc_device.hpp"

namespace chx {
namespace {

constexpr int kPtThreads = 256;  // 4 waves
constexpr int kSumThreads = 256;

// Sum of (e, n) over the workgroup into part[2 * blockIdx.x + {0, 1}] in a fixed
// order (xor shuffles in a wave, then the waves in order): deterministic.
__device__ __forceinline__ void wg_pair_sum(double e, double n, double* part) {
    __shared__ double sw[2][kPtThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        e += __shfl_xor(e, o);
        n += __shfl_xor(n, o);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sw[0][wave] = e;
        sw[1][wave] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, tn = 0.0;
#pragma unroll
        for (int w = 0; w < kPtThreads / 64; ++w) {
            te += sw[0][w];
            tn += sw[1][w];
        }
        part[2 * blockIdx.x] = te;
        part[2 * blockIdx.x + 1] = tn;
    }
}

// One thread per point: out[6 i ..] = (f, df/drho_a, df/drho_b, df/dgamma_aa,
// df/dgamma_ab, df/dgamma_bb) of functional F at in[5 i ..]; with part != NULL
// the workgroup's sums of w f and w (rho_a + rho_b) (w = 1 without weights).
template <int F>
__global__ __launch_bounds__(kPtThreads) void k_xc_points(long long n, const double* __restrict__ in,
                                                          const double* __restrict__ w, double* __restrict__ out,
                                                          double* __restrict__ part) {
    const long long i = static_cast<long long>(blockIdx.x) * kPtThreads + threadIdx.x;
    double e = 0.0, ne = 0.0;
    if (i < n) {
        const double* p = in + 5 * i;
        const double ra = p[0], rb = p[1];
        const XcOut o = eval_xc(F, ra, rb, p[2], p[3], p[4]);
        double* q = out + 6 * i;
        q[0] = o.f;
        q[1] = o.va;
        q[2] = o.vb;
        q[3] = o.gaa;
        q[4] = o.gab;
        q[5] = o.gbb;
        const double wi = w ? w[i] : 1.0;
        e = wi * o.f;
        ne = wi * (ra + rb);
    }
    if (part) wg_pair_sum(e, ne, part);
}

// One workgroup: sums[0..1] = the sums of the n_wg partial pairs, each thread
// taking a strided run in order, then the threads in a fixed order.
__global__ __launch_bounds__(kSumThreads) void k_sum_pairs(const double* __restrict__ part, long long n_wg,
                                                           double* __restrict__ sums) {
    __shared__ double sw[2][kSumThreads / 64];
    double e = 0.0, n = 0.0;
    for (long long k = threadIdx.x; k < n_wg; k += kSumThreads) {
        e += part[2 * k];
        n += part[2 * k + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        e += __shfl_xor(e, o);
        n += __shfl_xor(n, o);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sw[0][wave] = e;
        sw[1][wave] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, tn = 0.0;
#pragma unroll
        for (int k = 0; k < kSumThreads / 64; ++k) {
            te += sw[0][k];
            tn += sw[1][k];
        }
        sums[0] = te;
        sums[1] = tn;
    }
}

const void* pick_points_kernel(int f) {
    switch (f) {
    case XC_SVWN5: return reinterpret_cast<const void*>(&k_xc_points<XC_SVWN5>);
    case XC_BLYP: return reinterpret_cast<const void*>(&k_xc_points<XC_BLYP>);
    case XC_PBE: return reinterpret_cast<const void*>(&k_xc_points<XC_PBE>);
    case XC_B3LYP: return reinterpret_cast<const void*>(&k_xc_points<XC_B3LYP>);
    case XC_B3LYP5: return reinterpret_cast<const void*>(&k_xc_points<XC_B3LYP5>);
    case XC_PBE0: return reinterpret_cast<const void*>(&k_xc_points<XC_PBE0>);
    default: return nullptr;
    }
}

bool device_present() {
    const char* dis = std::getenv("CALCHEM_HIP_DISABLE");
    if (dis && dis[0] == '1') return false;
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

// Enqueue the point kernel and, with d_sums, the sum over the partials in
// `part` (2 doubles per workgroup of 256 points) on stream s.
int enqueue_points(int xc, long long n, const double* d_in, const double* d_w, double* d_out, double* d_sums,
                   double* part, hipStream_t s) {
    const void* k = pick_points_kernel(xc);
    if (!k) return kUnsupported;
    const long long n_wg = (n + kPtThreads - 1) / kPtThreads;
    double* part_arg = d_sums ? part : nullptr;
    void* args[] = {&n, const_cast<double**>(&d_in), const_cast<double**>(&d_w), &d_out, &part_arg};
    if (hipLaunchKernel(k, dim3(static_cast<unsigned>(n_wg)), dim3(kPtThreads), args, 0, s) != hipSuccess)
        return kLaunchFail;
    if (d_sums) {
        hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kSumThreads), 0, s, part, n_wg, d_sums);
        if (hipGetLastError() != hipSuccess) return kLaunchFail;
    }
    return kOk;
}

// Largest point count per call: the grid's workgroup count stays below 2^31.
constexpr long long kMaxPoints = (1LL << 31) * kPtThreads - kPtThreads;

// Scratch of the entry point that allocates nothing: the per-workgroup partial
// sums of up to kNoAllocMaxWg workgroups (16,777,216 points), 1 MiB per GPU,
// held until the process ends.  Every call on a GPU shares that GPU's copy.
constexpr long long kNoAllocMaxWg = 65536;
constexpr long long kNoAllocMaxPoints = kNoAllocMaxWg * kPtThreads;
__device__ __align__(256) double scratch[2 * kNoAllocMaxWg];

// Device address of `scratch` per GPU, looked up once per device with that
// device current (the host-side name of the array is not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return kNoDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return kNoDevice;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return kOk;
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" {

int calchem_hip_xc_eval_device(int xc, long long n, const double* d_in, const double* d_w, double* d_out,
                               double* d_sums, void* stream) {
    if (n < 0 || (n > 0 && (!d_in || !d_out))) return kBadArgs;
    if (n > kMaxPoints) return kLimit;
    if (!pick_points_kernel(xc)) return kUnsupported;
    if (!device_present()) return kNoDevice;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {  // the sums of no partials are zeros
        if (d_sums) {
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kSumThreads), 0, s, nullptr, 0LL, d_sums);
            if (hipGetLastError() != hipSuccess) return kLaunchFail;
        }
        return kOk;
    }
    // The per-workgroup partials live in stream-ordered memory of this call:
    // calls on different streams never share them.
    double* part = nullptr;
    const long long n_wg = (n + kPtThreads - 1) / kPtThreads;
    if (d_sums && hipMallocAsync(reinterpret_cast<void**>(&part), sizeof(double) * 2 * n_wg, s) != hipSuccess)
        return kAllocFail;
    const int rc = enqueue_points(xc, n, d_in, d_w, d_out, d_sums, part, s);
    if (part && hipFreeAsync(part, s) != hipSuccess && rc == kOk) return kLaunchFail;
    return rc;
}

int calchem_hip_xc_eval_device_noalloc(int xc, long long n, const double* d_in, const double* d_w, double* d_out,
                                       double* d_sums, void* stream) {
    if (n < 0 || (n > 0 && (!d_in || !d_out))) return kBadArgs;
    if (n > kNoAllocMaxPoints) return kLimit;
    if (!pick_points_kernel(xc)) return kUnsupported;
    if (!device_present()) return kNoDevice;
    double* part = nullptr;
    const int rc = scratch_for_current_device(&part);
    if (rc != kOk) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {  // the sums of no partials are zeros
        if (d_sums) {
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kSumThreads), 0, s, nullptr, 0LL, d_sums);
            if (hipGetLastError() != hipSuccess) return kLaunchFail;
        }
        return kOk;
    }
    // Every partial that k_sum_pairs reads is written by this call's point
    // kernel first, so nothing needs clearing between calls.
    return enqueue_points(xc, n, d_in, d_w, d_out, d_sums, part, s);
}

int calchem_hip_xc_eval(int xc, long long n, const double* in, const double* w, double* out, double* sums) {
    if (n < 0 || (n > 0 && (!in || !out))) return kBadArgs;
    if (n > kMaxPoints) return kLimit;
    if (!pick_points_kernel(xc)) return kUnsupported;
    if (!device_present()) return kNoDevice;
    if (n == 0) {
        if (sums) sums[0] = sums[1] = 0.0;
        return kOk;
    }
    double *d_in = nullptr, *d_w = nullptr, *d_out = nullptr, *d_sums = nullptr;
    int rc = kOk;
    if (hipMalloc(&d_in, sizeof(double) * 5 * n) != hipSuccess || hipMalloc(&d_out, sizeof(double) * 6 * n) != hipSuccess ||
        (w && hipMalloc(&d_w, sizeof(double) * n) != hipSuccess) ||
        (sums && hipMalloc(&d_sums, sizeof(double) * 2) != hipSuccess))
        rc = kAllocFail;
    if (rc == kOk && (hipMemcpy(d_in, in, sizeof(double) * 5 * n, hipMemcpyHostToDevice) != hipSuccess ||
                      (w && hipMemcpy(d_w, w, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)))
        rc = kLaunchFail;
    if (rc == kOk) rc = calchem_hip_xc_eval_device(xc, n, d_in, d_w, d_out, d_sums, nullptr);
    if (rc == kOk && (hipDeviceSynchronize() != hipSuccess ||
                      hipMemcpy(out, d_out, sizeof(double) * 6 * n, hipMemcpyDeviceToHost) != hipSuccess ||
                      (sums && hipMemcpy(sums, d_sums, sizeof(double) * 2, hipMemcpyDeviceToHost) != hipSuccess)))
        rc = kLaunchFail;
    hipFree(d_in);
    hipFree(d_out);
    if (d_w) hipFree(d_w);
    if (d_sums) hipFree(d_sums);
    return rc;
}

}  // extern "C"
