// probe_mfma.hip — measured MFMA throughput ceiling of the device, for the
// measurement report (a diagnostic beside the datasheet peak; not part of
// the Vxc library's interface).
//
// Every wave runs `iters` rounds of four independent accumulation chains of
// the 16x16x4 MFMA the build kernels use (v_mfma_f32_16x16x4f32 or
// v_mfma_f64_16x16x4f64, 2048 flop each), with its operands in registers, so
// nothing but the matrix cores bounds it.  The accumulators are summed and
// stored at the end so that the work is kept.
#include <hip/hip_runtime.h>

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T>
struct ProbeOp;
template <>
struct ProbeOp<float> {
    using acc = v4f;
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct ProbeOp<double> {
    using acc = v4d;
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

constexpr int kProbeThreads = 256;

template <typename T>
__global__ __launch_bounds__(kProbeThreads) void k_probe_mfma(T* __restrict__ out, int iters, T seed) {
    using O = ProbeOp<T>;
    typename O::acc c0{}, c1{}, c2{}, c3{};
    // small operands keep every sum finite for any iteration count used here
    const T a = seed * static_cast<T>(1 + (threadIdx.x & 7)), b = seed * static_cast<T>(1 + (threadIdx.x >> 5));
    for (int i = 0; i < iters; ++i) {
        c0 = O::op(a, b, c0);
        c1 = O::op(b, a, c1);
        c2 = O::op(a, a, c2);
        c3 = O::op(b, b, c3);
    }
    const typename O::acc s = c0 + c1 + c2 + c3;
    out[static_cast<long long>(blockIdx.x) * kProbeThreads + threadIdx.x] = s[0] + s[1] + s[2] + s[3];
}

}  // namespace

extern "C" {

// Measured dense MFMA rate in TFLOP/s: precision 0 = FP64, 1 = FP32;
// n_wg workgroups of 256 threads, iters rounds of 4 MFMAs per wave, timed
// with events around one launch after one warm-up launch.  Returns 0, or -1
// on bad arguments, -3 without a device, -5 when a HIP call fails.
int calchem_probe_mfma_tflops(int precision, int n_wg, int iters, double* tflops) {
    if (!tflops || n_wg <= 0 || iters <= 0 || (precision != 0 && precision != 1)) return -1;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return -3;
    const size_t elem = precision == 0 ? sizeof(double) : sizeof(float);
    void* out = nullptr;
    if (hipMalloc(&out, elem * static_cast<size_t>(n_wg) * kProbeThreads) != hipSuccess) return -5;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = -5;
    auto launch = [&]() {
        if (precision == 0)
            hipLaunchKernelGGL(k_probe_mfma<double>, dim3(n_wg), dim3(kProbeThreads), 0, 0, static_cast<double*>(out),
                               iters, 1e-3);
        else
            hipLaunchKernelGGL(k_probe_mfma<float>, dim3(n_wg), dim3(kProbeThreads), 0, 0, static_cast<float*>(out),
                               iters, 1e-3f);
        return hipGetLastError();
    };
    float ms = 0.0f;
    if (rc == 0 && (launch() != hipSuccess || hipDeviceSynchronize() != hipSuccess)) rc = -5;
    if (rc == 0 && (hipEventRecord(e0, 0) != hipSuccess || launch() != hipSuccess ||
                    hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f))
        rc = -5;
    if (rc == 0) {
        const double flop = 2048.0 * 4.0 * iters * (kProbeThreads / 64) * static_cast<double>(n_wg);
        *tflops = flop / (ms * 1e-3) / 1e12;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(out);
    return rc;
}

}  // extern "C"
