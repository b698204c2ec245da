// vxc_kernels.hip — gfx950 kernels of the RKS Vxc / E_xc / N build.
//
// One build (crates/calchem-dft/src/vxc.rs:801-985 restated for CDNA4):
//   k_rho     per (block, 32 or 16 points): AO values of the block's AO list
//             evaluated straight into LDS, T = P_sub Phi on MFMA (P rows read
//             from L2), rho = sum_i Phi_i T_i and grad rho = 2 sum_i dPhi_i T_i,
//             then the functional in FP64: per-point (w v_rho/2, w u grad rho)
//             and per-workgroup partial E_xc and N.
//   k_vxc     per (group, 64x64 lower tile pair): Z = w v_rho/2 Phi + wu . dPhi
//             and Phi staged in LDS 16 points at a time, acc += Z_I Phi_J^T +
//             Phi_I Z_J^T on MFMA; the tile goes to a per-group scratch slot.
//   k_reduce_v   deterministic sum of the group tiles into the full symmetric V.
//   k_reduce_en  deterministic sum of the E_xc / N partials.
// T = double runs FP64 MFMA (v_mfma_f64_16x16x4f64); T = float runs FP32 MFMA
// (v_mfma_f32_16x16x4f32) with FP32 accumulation inside a tile and FP64 for
// the functional, the per-point reductions and every sum across tiles.
#include "ao_device.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "xc_device.hpp"

namespace chx {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mma;
template <>
struct Mma<double> {
    using acc = v4d;
    // Result row of element v in lanes 16 q .. 16 q + 15.
    __device__ static __forceinline__ int row(int q, int v) { return 4 * v + q; }
    __device__ static __forceinline__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
    __device__ static __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mma<float> {
    using acc = v4f;
    __device__ static __forceinline__ int row(int q, int v) { return 4 * q + v; }
    __device__ static __forceinline__ acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
    __device__ static __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// MFMA 16x16x4 operand layout (one value per lane): A[i][k] at lane i + 16 k,
// B[k][j] at lane j + 16 k.  Element v of lane j + 16 q of the 16x16 result
// holds D[Mma<T>::row(q, v)][j]: row 4 q + v for the f32 instruction, row
// 4 v + q for the f64 one (one output row group per register).

// ---------------------------------------------------------------------------
// rho stage
// ---------------------------------------------------------------------------
template <int PTS>
__host__ __device__ constexpr size_t rho_lds_fixed() {
    return sizeof(double4) * PTS + sizeof(double) * 16 * PTS + sizeof(double) * 2 * PTS;
}

template <typename T, int PTS, bool XC>
__global__ __launch_bounds__(kThreads) void k_rho(RhoArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const RhoTile tile = a.tiles[blockIdx.x];
    const RhoBlock blk = a.blocks[tile.block];
    const int m = blk.m, mp = blk.m_pad, np = tile.n_pts;
    double4* sPt = reinterpret_cast<double4*>(smem);
    double* sRed = reinterpret_cast<double*>(sPt + PTS);  // [4 waves][4 sums][PTS]
    double* sEN = sRed + 16 * PTS;                         // [2][PTS]
    int* sAo = reinterpret_cast<int*>(sEN + 2 * PTS);      // [mp]
    T* sPhi = reinterpret_cast<T*>(sAo + mp);              // [mp][PTS]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int r = tid; r < mp; r += kThreads) sAo[r] = r < m ? a.block_ao[blk.ao_begin + r] : -1;
    if (tid < PTS) sPt[tid] = tid < np ? a.pts[tile.pt_begin + tid] : make_double4(0.0, 0.0, 0.0, 0.0);
    __syncthreads();

    // AO values of every listed AO at the tile's points, straight into LDS.
    for (int idx = tid; idx < mp * PTS; idx += kThreads) {
        const int r = idx / PTS, p = idx % PTS;
        const int mu = sAo[r];
        double v = 0.0;
        if (mu >= 0 && p < np) {
            const double4 q = sPt[p];
            v = ao_eval(a.shells, a.prims, a.aos[mu], q.x, q.y, q.z).v;
        }
        sPhi[idx] = static_cast<T>(v);
    }
    __syncthreads();

    using M = Mma<T>;
    constexpr int NH = PTS / 16;
    const T* __restrict__ P = static_cast<const T*>(a.P);
    const int n = a.n_ao;
    const int col = lane & 15, kq = lane >> 4;
    double sr[NH], sx[NH], sy[NH], sz[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) sr[h] = sx[h] = sy[h] = sz[h] = 0.0;

    for (int it = wave; it < (mp >> 4); it += 4) {
        const int i0 = it << 4;
        const int arow = sAo[i0 + col];
        const T* __restrict__ Prow = P + static_cast<size_t>(arow < 0 ? 0 : arow) * n;
        typename M::acc acc[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) acc[h] = M::zero();
#pragma unroll 4
        for (int k = 0; k < mp; k += 4) {
            const int kk = k + kq;
            const int acol = sAo[kk];
            const T av = (arow >= 0 && acol >= 0) ? Prow[acol] : T(0);
#pragma unroll
            for (int h = 0; h < NH; ++h) acc[h] = M::op(av, sPhi[kk * PTS + h * 16 + col], acc[h]);
        }
        // Epilogue: this lane holds T[i0 + row(kq, v)][h*16 + col].
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int mu = sAo[i0 + M::row(kq, v)];
            if (mu < 0) continue;
            const DevAo ao = a.aos[mu];
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const int p = h * 16 + col;
                if (p >= np) continue;
                const double t = static_cast<double>(acc[h][v]);
                const double4 q = sPt[p];
                const AoVal f = ao_eval(a.shells, a.prims, ao, q.x, q.y, q.z);
                sr[h] += f.v * t;
                sx[h] += f.gx * t;
                sy[h] += f.gy * t;
                sz[h] += f.gz * t;
            }
        }
    }
    // Sum over the four row groups of the wave, then over the waves.
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        sr[h] += __shfl_xor(sr[h], 16);
        sr[h] += __shfl_xor(sr[h], 32);
        sx[h] += __shfl_xor(sx[h], 16);
        sx[h] += __shfl_xor(sx[h], 32);
        sy[h] += __shfl_xor(sy[h], 16);
        sy[h] += __shfl_xor(sy[h], 32);
        sz[h] += __shfl_xor(sz[h], 16);
        sz[h] += __shfl_xor(sz[h], 32);
        if (kq == 0) {
            const int p = h * 16 + col;
            sRed[(wave * 4 + 0) * PTS + p] = sr[h];
            sRed[(wave * 4 + 1) * PTS + p] = sx[h];
            sRed[(wave * 4 + 2) * PTS + p] = sy[h];
            sRed[(wave * 4 + 3) * PTS + p] = sz[h];
        }
    }
    __syncthreads();
    if (tid < PTS) {
        const int p = tid;
        double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            rho += sRed[(w * 4 + 0) * PTS + p];
            gx += sRed[(w * 4 + 1) * PTS + p];
            gy += sRed[(w * 4 + 2) * PTS + p];
            gz += sRed[(w * 4 + 3) * PTS + p];
        }
        gx *= 2.0;
        gy *= 2.0;
        gz *= 2.0;
        double e = 0.0, ne = 0.0;
        if (p < np) {
            const int g = tile.pt_begin + p;
            if constexpr (XC) {
                const double w = sPt[p].w;
                ne = w * rho;
                double4 out = make_double4(0.0, 0.0, 0.0, 0.0);
                if (rho >= 1e-18) {  // DENSITY_SCREEN, vxc.rs:23
                    const double rs = 0.5 * rho;
                    const double gss = 0.25 * (gx * gx + gy * gy + gz * gz);
                    const XcOut o = eval_xc(a.func, rs, rs, gss, gss, gss);
                    e = w * o.f;
                    const double us = w * (o.gaa + 0.5 * o.gab);
                    out = make_double4(0.5 * w * o.va, us * gx, us * gy, us * gz);
                }
                a.wv[g] = out;
            } else {
                a.rho_out[a.perm[g]] = make_double4(rho, gx, gy, gz);
            }
        }
        sEN[p] = e;
        sEN[PTS + p] = ne;
    }
    if constexpr (XC) {
        __syncthreads();
        if (tid == 0) {
            double e = 0.0, ne = 0.0;
            for (int p = 0; p < PTS; ++p) {
                e += sEN[p];
                ne += sEN[PTS + p];
            }
            a.e_part[blockIdx.x] = e;
            a.n_part[blockIdx.x] = ne;
        }
    }
}

// ---------------------------------------------------------------------------
// V stage
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_vxc(VArgs a) {
    constexpr int LD = kVPts + 1;  // padded row stride against LDS bank conflicts
    __shared__ T sZI[kVTile * LD], sPI[kVTile * LD], sZJ[kVTile * LD], sPJ[kVTile * LD];
    __shared__ int sAoI[kVTile], sAoJ[kVTile];
    __shared__ double4 sPt[kVPts], sWv[kVPts];

    const VPair pr = a.pairs[blockIdx.x];
    const VGroup G = a.groups[pr.group];
    const bool diag = pr.ti == pr.tj;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;

    if (tid < kVTile) {
        const int gi = pr.ti * kVTile + tid, gj = pr.tj * kVTile + tid;
        sAoI[tid] = gi < G.m ? a.group_ao[G.ao_begin + gi] : -1;
        sAoJ[tid] = gj < G.m ? a.group_ao[G.ao_begin + gj] : -1;
    }
    using M = Mma<T>;
    typename M::acc acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = M::zero();

    const T* ZJ = diag ? sZI : sZJ;
    const T* PJ = diag ? sPI : sPJ;
    const int nrows = diag ? kVTile : 2 * kVTile;

    for (int q0 = 0; q0 < G.n_pts; q0 += kVPts) {
        __syncthreads();
        if (tid < kVPts) {
            const int g = q0 + tid;
            if (g < G.n_pts) {
                const int sg = G.pt_begin + g;
                sPt[tid] = a.pts[sg];
                sWv[tid] = a.wv[a.wv_perm ? a.perm[sg] : sg];
            } else {
                sPt[tid] = make_double4(0.0, 0.0, 0.0, 0.0);
                sWv[tid] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
        }
        __syncthreads();
        for (int idx = tid; idx < nrows * kVPts; idx += kThreads) {
            const int side = idx / (kVTile * kVPts);
            const int r = (idx / kVPts) % kVTile, p = idx % kVPts;
            const int mu = side ? sAoJ[r] : sAoI[r];
            double v = 0.0, z = 0.0;
            if (mu >= 0 && q0 + p < G.n_pts) {
                const double4 q = sPt[p];
                const AoVal f = ao_eval(a.shells, a.prims, a.aos[mu], q.x, q.y, q.z);
                const double4 w = sWv[p];
                v = f.v;
                z = w.x * f.v + w.y * f.gx + w.z * f.gy + w.w * f.gz;
            }
            if (side) {
                sPJ[r * LD + p] = static_cast<T>(v);
                sZJ[r * LD + p] = static_cast<T>(z);
            } else {
                sPI[r * LD + p] = static_cast<T>(v);
                sZI[r * LD + p] = static_cast<T>(z);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kVPts; k += 4) {
            const int kk = k + kq;
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const int ri = (wr * 32 + x * 16 + col) * LD + kk;
                const T zi = sZI[ri], pi = sPI[ri];
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    const int rj = (wc * 32 + y * 16 + col) * LD + kk;
                    acc[x][y] = M::op(zi, PJ[rj], acc[x][y]);
                    acc[x][y] = M::op(pi, ZJ[rj], acc[x][y]);
                }
            }
        }
    }
    double* out = a.scratch + G.scratch_off +
                  static_cast<long long>(pr.ti * (pr.ti + 1) / 2 + pr.tj) * (kVTile * kVTile);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wr * 32 + x * 16 + M::row(kq, v);
                const int cj = wc * 32 + y * 16 + col;
                out[row * kVTile + cj] = static_cast<double>(acc[x][y][v]);
            }
}

__global__ __launch_bounds__(kThreads) void k_reduce_v(ReduceArgs a) {
    const int mu = blockIdx.y;
    const int nu = blockIdx.x * kThreads + threadIdx.x;
    if (nu > mu || nu >= a.n_ao) return;
    const int n = a.n_ao;
    double s = 0.0;
    for (int g = 0; g < a.n_groups; ++g) {
        const int li = a.loc[static_cast<long long>(g) * n + mu];
        if (li < 0) continue;
        const int lj = a.loc[static_cast<long long>(g) * n + nu];
        if (lj < 0) continue;
        const int ti = li >> 6, tj = lj >> 6;
        s += a.scratch[a.groups[g].scratch_off +
                       static_cast<long long>(ti * (ti + 1) / 2 + tj) * (kVTile * kVTile) + (li & 63) * kVTile +
                       (lj & 63)];
    }
    a.V[static_cast<long long>(mu) * n + nu] = s;
    a.V[static_cast<long long>(nu) * n + mu] = s;
}

__global__ __launch_bounds__(kThreads) void k_reduce_en(const double* __restrict__ e, const double* __restrict__ ne,
                                                        int n, double* __restrict__ out) {
    __shared__ double se[kThreads], sn[kThreads];
    double x = 0.0, y = 0.0;
    for (int i = threadIdx.x; i < n; i += kThreads) {
        x += e[i];
        y += ne[i];
    }
    se[threadIdx.x] = x;
    sn[threadIdx.x] = y;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            se[threadIdx.x] += se[threadIdx.x + s];
            sn[threadIdx.x] += sn[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = se[0];
        out[1] = sn[0];
    }
}

__global__ void k_to_float(const double* __restrict__ x, float* __restrict__ y, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = static_cast<float>(x[i]);
}

// Per (group-local candidate AO): does the AO survive the reference's value
// rule at any point of its block?  Setup only.
__global__ void k_significance(SigArgs a) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n_entries) return;
    const int mu = a.cand_ao[i];
    const int2 rng = a.cand_pts[i];  // sorted point range of the entry's block
    const DevAo ao = a.aos[mu];
    unsigned char keep = 0;
    for (int g = rng.x; g < rng.x + rng.y && !keep; ++g) {
        const double4 q = a.pts[g];
        const AoVal f = ao_eval(a.shells, a.prims, ao, q.x, q.y, q.z);
        keep = (f.v != 0.0 || f.gx != 0.0 || f.gy != 0.0 || f.gz != 0.0);
    }
    a.flag[i] = keep;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename T, int PTS, bool XC>
static hipError_t launch_rho_t(const RhoArgs& a, int n_tiles, int max_mpad, hipStream_t s) {
    const size_t lds = rho_lds_fixed<PTS>() + sizeof(int) * max_mpad + sizeof(T) * (size_t)max_mpad * PTS;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rho<T, PTS, XC>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    if (n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_rho<T, PTS, XC>), dim3(n_tiles), dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

size_t rho_lds_bytes(int prec_bytes, int pts, int max_mpad) {
    const size_t fixed = pts == 32 ? rho_lds_fixed<32>() : rho_lds_fixed<16>();
    return fixed + sizeof(int) * (size_t)max_mpad + (size_t)prec_bytes * max_mpad * pts;
}

hipError_t launch_rho(const RhoArgs& a, bool fp32, int pts, bool xc, int n_tiles, int max_mpad, hipStream_t s) {
    if (!fp32) {
        if (pts == 32) return xc ? launch_rho_t<double, 32, true>(a, n_tiles, max_mpad, s)
                                 : launch_rho_t<double, 32, false>(a, n_tiles, max_mpad, s);
        return xc ? launch_rho_t<double, 16, true>(a, n_tiles, max_mpad, s)
                  : launch_rho_t<double, 16, false>(a, n_tiles, max_mpad, s);
    }
    if (pts == 32) return xc ? launch_rho_t<float, 32, true>(a, n_tiles, max_mpad, s)
                             : launch_rho_t<float, 32, false>(a, n_tiles, max_mpad, s);
    return xc ? launch_rho_t<float, 16, true>(a, n_tiles, max_mpad, s)
              : launch_rho_t<float, 16, false>(a, n_tiles, max_mpad, s);
}

hipError_t launch_vxc(const VArgs& a, bool fp32, int n_pairs, hipStream_t s) {
    if (n_pairs <= 0) return hipSuccess;
    if (fp32)
        hipLaunchKernelGGL(k_vxc<float>, dim3(n_pairs), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_vxc<double>, dim3(n_pairs), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_v(const ReduceArgs& a, hipStream_t s) {
    if (a.n_ao <= 0) return hipSuccess;
    dim3 grid((a.n_ao + kThreads - 1) / kThreads, a.n_ao);
    hipLaunchKernelGGL(k_reduce_v, grid, dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_en(const double* e, const double* ne, int n, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_en, dim3(1), dim3(kThreads), 0, s, e, ne, n, out);
    return hipGetLastError();
}

hipError_t launch_to_float(const double* x, float* y, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_significance(const SigArgs& a, hipStream_t s) {
    if (a.n_entries <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_significance, dim3((unsigned)((a.n_entries + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace chx
