// mlp.hip — the atom-pair network of the learned SCF guess: a multilayer
// perceptron d_in -> H -> H -> d_out (SiLU on the hidden layers, FP32
// weights), with its forward pass, backward pass and Adam update written as
// HIP kernels.  The learned guess replaces the SAD start of the SCF
// (crates/calchem-scf/src/rks.rs:221-231; SAD stays as the fallback).
//
// Layout: parameters are one FP32 array W1[H][d_in], b1[H], W2[H][H], b2[H],
// W3[d_out][H], b3[d_out]; activations are row-major [batch][features].
// Loss: masked mean squared error sum_m (y - t)^2 / sum_m 1 over the batch.
// Every reduction (weight gradients over the batch, the loss) is a fixed-order
// sum of partials: training is bit-reproducible for a seed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/calchem_hip.h"
#include "common.hpp"

namespace chx {
namespace mlp {
namespace {

constexpr int kT = 64;   // output tile (kT x kT) per workgroup
constexpr int kK = 16;   // reduction step
constexpr int kThr = 256;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

enum Epi { kStore = 0, kBiasSilu = 1, kBias = 2 };

// C[M x N] = A'[M x K] B'[K x N] over k in [z kchunk, (z+1) kchunk) (partial
// z at C + z M ldc), row-major; A' = A (M x K) or A^T (A stored K x M), B' = B
// (K x N) or B^T (B stored N x K).  Epilogue: plain store, or + bias and
// SiLU (pre-activation to Z, activation to C), or + bias.  Each thread owns a
// 4 x 4 block of the 64 x 64 tile; A' and B' tiles go through LDS.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(kThr) void k_sgemm(int M, int N, int K, int kchunk, const float* __restrict__ A, int lda,
                                                 const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                                 const float* __restrict__ bias, float* __restrict__ Z) {
    __shared__ float As[kK][kT + 4];
    __shared__ float Bs[kK][kT + 4];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    float acc[4][4] = {};
    for (int k0 = kb; k0 < ke; k0 += kK) {
        // stage A' (kT x kK) and B' (kK x kT): 1024 elements each, 4 per thread
        for (int e = threadIdx.x; e < kT * kK; e += kThr) {
            int mm, kk;
            if (TA) {  // A stored K x M: consecutive threads along m
                mm = e % kT;
                kk = e / kT;
            } else {   // A stored M x K: consecutive threads along k
                kk = e % kK;
                mm = e / kK;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.0f;
            if (gm < M && gk < ke) v = TA ? A[(long long)gk * lda + gm] : A[(long long)gm * lda + gk];
            As[kk][mm] = v;
            int nn, kk2;
            if (TB) {  // B stored N x K
                kk2 = e % kK;
                nn = e / kK;
            } else {   // B stored K x N
                nn = e % kT;
                kk2 = e / kT;
            }
            const int gn = n0 + nn, gk2 = k0 + kk2;
            float w = 0.0f;
            if (gn < N && gk2 < ke) w = TB ? B[(long long)gn * ldb + gk2] : B[(long long)gk2 * ldb + gn];
            Bs[kk2][nn] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    float* Cz = C + (long long)blockIdx.z * M * ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + ty + 16 * i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tx + 16 * j;
            if (gn >= N) continue;
            float v = acc[i][j];
            if (EPI == kBiasSilu) {
                v += bias[gn];
                Z[(long long)gm * ldc + gn] = v;
                v = silu(v);
            } else if (EPI == kBias) {
                v += bias[gn];
            }
            Cz[(long long)gm * ldc + gn] = v;
        }
    }
}

// out[i] = sum_z part[z n + i] in z order.
__global__ void k_sum_parts(const float* __restrict__ part, int nz, long long n, float* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < nz; ++z) s += part[(long long)z * n + i];
    out[i] = s;
}

// Column sums of D (rows x cols) over row chunks: part[z][c] = sum over rows of chunk z.
__global__ void k_colsum_parts(const float* __restrict__ D, int rows, int cols, int chunk, float* __restrict__ part) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const int r0 = blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
    float s = 0.0f;
    for (int r = r0; r < r1; ++r) s += D[(long long)r * cols + c];
    part[(long long)blockIdx.y * cols + c] = s;
}

// dZ = dY * silu'(Z) elementwise.
__global__ void k_silu_bwd(const float* dY, const float* __restrict__ Z, float* dZ, long long n) {  // dY may alias dZ
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = dY[i] * silu_grad(Z[i]);
}

// Masked MSE: dY = 2 m (y - t) / count; per-workgroup partial sums of m (y - t)^2.
__global__ __launch_bounds__(kThr) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                              const float* __restrict__ Mk, long long n, float inv_count,
                                              float* __restrict__ dY, float* __restrict__ part) {
    __shared__ float sw[kThr / 64];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float d = Mk[i] * (Y[i] - T[i]);
        l = d * d;
        if (dY) dY[i] = 2.0f * d * inv_count;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
#pragma unroll
        for (int w = 0; w < kThr / 64; ++w) s += sw[w];
        part[blockIdx.x] = s;
    }
}

// Adam (Kingma and Ba, 2015): m, v moments; bias-corrected step.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// rows of src selected by idx -> dst (gather of a minibatch), width w.
__global__ void k_gather(const float* __restrict__ src, const int* __restrict__ idx, int rows, int w,
                         float* __restrict__ dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * w) return;
    dst[i] = src[(long long)idx[i / w] * w + (i % w)];
}

inline unsigned nb(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    bool ensure(size_t b) {
        if (bytes >= b) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, std::max<size_t>(b, 16)) != hipSuccess) return false;
        bytes = std::max<size_t>(b, 16);
        return true;
    }
    float* f() const { return static_cast<float*>(p); }
    int* i() const { return static_cast<int*>(p); }
};

struct Net {
    int device = 0, d_in = 0, h = 0, d_out = 0;
    long long n_params = 0;
    long long off_w1, off_b1, off_w2, off_b2, off_w3, off_b3;
    Buf params, grads, adam_m, adam_v;
    long long adam_t = 0;
    // activations of the current batch
    Buf x, z1, a1, z2, a2, y, t, mk, d3, d2, d1, tmp, part, idx;
    int cap = 0;
    // resident training set
    Buf data_x, data_t, data_m;
    std::vector<float> mask_count;  // per sample
    int n_data = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    ~Net() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    bool reserve(int n) {
        if (n <= cap) return true;
        const size_t N = n;
        if (!x.ensure(sizeof(float) * N * d_in) || !z1.ensure(sizeof(float) * N * h) || !a1.ensure(sizeof(float) * N * h) ||
            !z2.ensure(sizeof(float) * N * h) || !a2.ensure(sizeof(float) * N * h) || !y.ensure(sizeof(float) * N * d_out) ||
            !t.ensure(sizeof(float) * N * d_out) || !mk.ensure(sizeof(float) * N * d_out) ||
            !d3.ensure(sizeof(float) * N * d_out) || !d2.ensure(sizeof(float) * N * h) || !d1.ensure(sizeof(float) * N * h) ||
            !idx.ensure(sizeof(int) * N))
            return false;
        cap = n;
        return true;
    }
};

std::mutex g_mu;
std::unordered_map<uint64_t, std::shared_ptr<Net>> g_nets;

std::shared_ptr<Net> find_net(uint64_t h) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_nets.find(h);
    return it == g_nets.end() ? nullptr : it->second;
}

bool dev_ok() {
    const char* dis = std::getenv("CALCHEM_HIP_DISABLE");
    if (dis && dis[0] == '1') return false;
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

template <bool TA, bool TB, int EPI>
hipError_t gemm(hipStream_t s, int M, int N, int K, int kchunk, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc, const float* bias = nullptr, float* Z = nullptr) {
    // kchunk: reduction length per partial (0: one partial over all of K)
    if (kchunk <= 0) kchunk = K;
    const int nz = (K + kchunk - 1) / kchunk;
    dim3 grid((N + kT - 1) / kT, (M + kT - 1) / kT, std::max(nz, 1));
    hipLaunchKernelGGL((k_sgemm<TA, TB, EPI>), grid, dim3(kThr), 0, s, M, N, K, kchunk, A, lda, B, ldb, C, ldc, bias, Z);
    return hipGetLastError();
}

// Forward pass of the n rows staged in net.x; output in net.y.
int forward(Net& W, int n) {
    const float* P = W.params.f();
    hipStream_t s = W.stream;
    if (gemm<false, true, kBiasSilu>(s, n, W.h, W.d_in, 0, W.x.f(), W.d_in, P + W.off_w1, W.d_in, W.a1.f(), W.h,
                                     P + W.off_b1, W.z1.f()) != hipSuccess ||
        gemm<false, true, kBiasSilu>(s, n, W.h, W.h, 0, W.a1.f(), W.h, P + W.off_w2, W.h, W.a2.f(), W.h, P + W.off_b2,
                                     W.z2.f()) != hipSuccess ||
        gemm<false, true, kBias>(s, n, W.d_out, W.h, 0, W.a2.f(), W.h, P + W.off_w3, W.h, W.y.f(), W.d_out,
                                 P + W.off_b3) != hipSuccess)
        return kLaunchFail;
    return kOk;
}

// dW (rows x cols) = dZ^T X over the batch, db = column sums of dZ: split over
// batch chunks of 1024 rows, partials summed in chunk order.
int weight_grad(Net& W, int n, const float* dZ, int rows, const float* X, int cols, float* dW, float* db) {
    hipStream_t s = W.stream;
    const int chunk = 1024;
    const int nz = (n + chunk - 1) / chunk;
    if (!W.tmp.ensure(sizeof(float) * (size_t)nz * rows * (cols + 1))) return kAllocFail;
    float* part = W.tmp.f();
    if (gemm<true, false, kStore>(s, rows, cols, n, chunk, dZ, rows, X, cols, part, cols) != hipSuccess) return kLaunchFail;
    hipLaunchKernelGGL(k_sum_parts, dim3(nb((long long)rows * cols)), dim3(256), 0, s, part, nz, (long long)rows * cols, dW);
    float* bpart = part + (size_t)nz * rows * cols;
    hipLaunchKernelGGL(k_colsum_parts, dim3(nb(rows), nz), dim3(256), 0, s, dZ, n, rows, chunk, bpart);
    hipLaunchKernelGGL(k_sum_parts, dim3(nb(rows)), dim3(256), 0, s, bpart, nz, (long long)rows, db);
    return hipGetLastError() == hipSuccess ? kOk : kLaunchFail;
}

// Loss and parameter gradient of the n rows staged in x / t / mk (count = sum of the mask).
int backward(Net& W, int n, float count, float* loss_host) {
    hipStream_t s = W.stream;
    const long long ny = (long long)n * W.d_out;
    const int nwg = (int)nb(ny, kThr);
    if (!W.part.ensure(sizeof(float) * (size_t)(nwg + 1))) return kAllocFail;
    int rc = forward(W, n);
    if (rc != kOk) return rc;
    hipLaunchKernelGGL(k_mse, dim3(nwg), dim3(kThr), 0, s, W.y.f(), W.t.f(), W.mk.f(), ny, 1.0f / count, W.d3.f(),
                       W.part.f());
    float* G = W.grads.f();
    const float* P = W.params.f();
    // layer 3: dW3 = d3^T a2, db3; d(a2) = d3 W3
    if ((rc = weight_grad(W, n, W.d3.f(), W.d_out, W.a2.f(), W.h, G + W.off_w3, G + W.off_b3)) != kOk) return rc;
    if (gemm<false, false, kStore>(s, n, W.h, W.d_out, 0, W.d3.f(), W.d_out, P + W.off_w3, W.h, W.d2.f(), W.h) != hipSuccess)
        return kLaunchFail;
    hipLaunchKernelGGL(k_silu_bwd, dim3(nb((long long)n * W.h)), dim3(256), 0, s, W.d2.f(), W.z2.f(), W.d2.f(),
                       (long long)n * W.h);
    if ((rc = weight_grad(W, n, W.d2.f(), W.h, W.a1.f(), W.h, G + W.off_w2, G + W.off_b2)) != kOk) return rc;
    if (gemm<false, false, kStore>(s, n, W.h, W.h, 0, W.d2.f(), W.h, P + W.off_w2, W.h, W.d1.f(), W.h) != hipSuccess)
        return kLaunchFail;
    hipLaunchKernelGGL(k_silu_bwd, dim3(nb((long long)n * W.h)), dim3(256), 0, s, W.d1.f(), W.z1.f(), W.d1.f(),
                       (long long)n * W.h);
    if ((rc = weight_grad(W, n, W.d1.f(), W.h, W.x.f(), W.d_in, G + W.off_w1, G + W.off_b1)) != kOk) return rc;
    if (loss_host) {
        // the loss: the workgroup partials summed on the host in order
        std::vector<float> parts(nwg);
        if (hipMemcpyAsync(parts.data(), W.part.p, sizeof(float) * nwg, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return kLaunchFail;
        double l = 0.0;
        for (float v : parts) l += v;
        *loss_host = (float)(l / count);
    }
    return kOk;
}

}  // namespace
}  // namespace mlp
}  // namespace chx

using namespace chx;
using namespace chx::mlp;

extern "C" {

int calchem_hip_mlp_create(int d_in, int d_hidden, int d_out, const float* params, uint64_t handle) {
    if (d_in <= 0 || d_hidden <= 0 || d_out <= 0 || !params || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_nets.count(handle)) return kMismatch;
    }
    auto net = std::make_shared<Net>();
    Net& W = *net;
    if (hipGetDevice(&W.device) != hipSuccess) return kNoDevice;
    W.d_in = d_in;
    W.h = d_hidden;
    W.d_out = d_out;
    W.off_w1 = 0;
    W.off_b1 = W.off_w1 + (long long)d_hidden * d_in;
    W.off_w2 = W.off_b1 + d_hidden;
    W.off_b2 = W.off_w2 + (long long)d_hidden * d_hidden;
    W.off_w3 = W.off_b2 + d_hidden;
    W.off_b3 = W.off_w3 + (long long)d_out * d_hidden;
    W.n_params = W.off_b3 + d_out;
    const size_t pb = sizeof(float) * W.n_params;
    if (hipStreamCreate(&W.stream) != hipSuccess) return kNoDevice;
    if (!W.params.ensure(pb) || !W.grads.ensure(pb) || !W.adam_m.ensure(pb) || !W.adam_v.ensure(pb)) return kAllocFail;
    std::vector<float> zeros(W.n_params, 0.0f);
    if (hipMemcpy(W.params.p, params, pb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(W.adam_m.p, zeros.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(W.adam_v.p, zeros.data(), pb, hipMemcpyHostToDevice) != hipSuccess)
        return kLaunchFail;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_nets.count(handle)) return kMismatch;
    g_nets[handle] = net;
    return kOk;
}

int calchem_hip_mlp_release(uint64_t handle) {
    std::shared_ptr<Net> p;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_nets.find(handle);
        if (it == g_nets.end()) return kOk;
        p = it->second;
        g_nets.erase(it);
    }
    std::lock_guard<std::mutex> lk(p->mu);
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    return kOk;
}

int calchem_hip_mlp_get_params(uint64_t handle, float* params) {
    if (!params || handle == 0) return kBadArgs;
    auto net = find_net(handle);
    if (!net) return kMismatch;
    std::lock_guard<std::mutex> lk(net->mu);
    hipSetDevice(net->device);
    if (hipStreamSynchronize(net->stream) != hipSuccess ||
        hipMemcpy(params, net->params.p, sizeof(float) * net->n_params, hipMemcpyDeviceToHost) != hipSuccess)
        return kLaunchFail;
    return kOk;
}

int calchem_hip_mlp_forward(uint64_t handle, int n, const float* x, float* y) {
    if (n < 0 || (n > 0 && (!x || !y)) || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto net = find_net(handle);
    if (!net) return kMismatch;
    Net& W = *net;
    std::lock_guard<std::mutex> lk(W.mu);
    hipSetDevice(W.device);
    if (n == 0) return kOk;
    if (!W.reserve(n)) return kAllocFail;
    if (hipMemcpyAsync(W.x.p, x, sizeof(float) * (size_t)n * W.d_in, hipMemcpyHostToDevice, W.stream) != hipSuccess)
        return kLaunchFail;
    const int rc = forward(W, n);
    if (rc != kOk) return rc;
    if (hipMemcpyAsync(y, W.y.p, sizeof(float) * (size_t)n * W.d_out, hipMemcpyDeviceToHost, W.stream) != hipSuccess ||
        hipStreamSynchronize(W.stream) != hipSuccess)
        return kLaunchFail;
    return kOk;
}

int calchem_hip_mlp_loss_grad(uint64_t handle, int n, const float* x, const float* t, const float* mask, float* loss,
                              float* grad) {
    if (n <= 0 || !x || !t || !mask || !loss || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto net = find_net(handle);
    if (!net) return kMismatch;
    Net& W = *net;
    std::lock_guard<std::mutex> lk(W.mu);
    hipSetDevice(W.device);
    if (!W.reserve(n)) return kAllocFail;
    const size_t no = (size_t)n * W.d_out;
    double count = 0.0;
    for (size_t i = 0; i < no; ++i) count += mask[i];
    if (!(count > 0.0)) return kBadArgs;
    if (hipMemcpyAsync(W.x.p, x, sizeof(float) * (size_t)n * W.d_in, hipMemcpyHostToDevice, W.stream) != hipSuccess ||
        hipMemcpyAsync(W.t.p, t, sizeof(float) * no, hipMemcpyHostToDevice, W.stream) != hipSuccess ||
        hipMemcpyAsync(W.mk.p, mask, sizeof(float) * no, hipMemcpyHostToDevice, W.stream) != hipSuccess)
        return kLaunchFail;
    const int rc = backward(W, n, (float)count, loss);
    if (rc != kOk) return rc;
    if (grad && (hipMemcpyAsync(grad, W.grads.p, sizeof(float) * W.n_params, hipMemcpyDeviceToHost, W.stream) != hipSuccess ||
                 hipStreamSynchronize(W.stream) != hipSuccess))
        return kLaunchFail;
    return kOk;
}

int calchem_hip_mlp_set_data(uint64_t handle, int n, const float* x, const float* t, const float* mask) {
    if (n <= 0 || !x || !t || !mask || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto net = find_net(handle);
    if (!net) return kMismatch;
    Net& W = *net;
    std::lock_guard<std::mutex> lk(W.mu);
    hipSetDevice(W.device);
    const size_t no = (size_t)n * W.d_out;
    if (!W.data_x.ensure(sizeof(float) * (size_t)n * W.d_in) || !W.data_t.ensure(sizeof(float) * no) ||
        !W.data_m.ensure(sizeof(float) * no))
        return kAllocFail;
    if (hipMemcpy(W.data_x.p, x, sizeof(float) * (size_t)n * W.d_in, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(W.data_t.p, t, sizeof(float) * no, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(W.data_m.p, mask, sizeof(float) * no, hipMemcpyHostToDevice) != hipSuccess)
        return kLaunchFail;
    W.mask_count.assign(n, 0.0f);
    for (int r = 0; r < n; ++r) {
        float c = 0.0f;
        for (int k = 0; k < W.d_out; ++k) c += mask[(size_t)r * W.d_out + k];
        W.mask_count[r] = c;
    }
    W.n_data = n;
    return kOk;
}

// Adam over the resident training set: `epochs` passes of shuffled minibatches
// of `batch` rows (order from a SplitMix64 stream seeded with `seed`), learning
// rate lr * 0.5 (1 + cos(pi step / steps)); loss_out[e] = mean minibatch loss of epoch e.
int calchem_hip_mlp_train(uint64_t handle, int epochs, int batch, float lr, uint64_t seed, float* loss_out) {
    if (epochs <= 0 || batch <= 0 || !(lr > 0.0f) || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto net = find_net(handle);
    if (!net) return kMismatch;
    Net& W = *net;
    std::lock_guard<std::mutex> lk(W.mu);
    hipSetDevice(W.device);
    if (W.n_data <= 0) return kMismatch;
    const int n = W.n_data, bsz = std::min(batch, n);
    if (!W.reserve(bsz)) return kAllocFail;
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    uint64_t st = seed;
    auto next = [&st]() {
        st += 0x9E3779B97F4A7C15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    const int steps_per_epoch = (n + bsz - 1) / bsz;
    const long long total = (long long)epochs * steps_per_epoch;
    long long step = 0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
        for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)(next() % (uint64_t)(i + 1))]);
        double eloss = 0.0;
        for (int b0 = 0; b0 < n; b0 += bsz) {
            const int m = std::min(bsz, n - b0);
            float count = 0.0f;
            for (int k = 0; k < m; ++k) count += W.mask_count[perm[b0 + k]];
            if (hipMemcpyAsync(W.idx.p, perm.data() + b0, sizeof(int) * m, hipMemcpyHostToDevice, W.stream) != hipSuccess)
                return kLaunchFail;
            hipLaunchKernelGGL(k_gather, dim3(nb((long long)m * W.d_in)), dim3(256), 0, W.stream, W.data_x.f(), W.idx.i(), m,
                               W.d_in, W.x.f());
            hipLaunchKernelGGL(k_gather, dim3(nb((long long)m * W.d_out)), dim3(256), 0, W.stream, W.data_t.f(), W.idx.i(),
                               m, W.d_out, W.t.f());
            hipLaunchKernelGGL(k_gather, dim3(nb((long long)m * W.d_out)), dim3(256), 0, W.stream, W.data_m.f(), W.idx.i(),
                               m, W.d_out, W.mk.f());
            if (!(count > 0.0f)) continue;
            float l = 0.0f;
            const int rc = backward(W, m, count, &l);  // synchronises (loss read back): perm stays valid
            if (rc != kOk) return rc;
            eloss += l;
            ++W.adam_t;
            const float lr_t = lr * 0.5f * (1.0f + cosf(3.14159265f * (float)step / (float)total));
            const float c1 = 1.0f - powf(b1, (float)W.adam_t), c2 = 1.0f - powf(b2, (float)W.adam_t);
            hipLaunchKernelGGL(k_adam, dim3(nb(W.n_params)), dim3(256), 0, W.stream, W.params.f(), W.grads.f(),
                               W.adam_m.f(), W.adam_v.f(), W.n_params, lr_t, b1, b2, eps, c1, c2);
            ++step;
        }
        if (loss_out) loss_out[e] = (float)(eloss / steps_per_epoch);
    }
    if (hipStreamSynchronize(W.stream) != hipSuccess) return kLaunchFail;
    return kOk;
}

}  // extern "C"
