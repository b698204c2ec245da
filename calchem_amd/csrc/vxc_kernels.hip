// vxc_kernels.hip — gfx950 kernels of the RKS Vxc / E_xc / N build.
//
// One build (crates/calchem-dft/src/vxc.rs:801-985 restated for CDNA4):
//   k_rho     per (block, 16 points): the block's shells, primitives and
//             AO rows staged in LDS, radial sums per (shell, point), AO values
//             into LDS, T = P_sub Phi on MFMA (P_sub rows read from L2),
//             rho = sum_i Phi_i T_i and grad rho = 2 sum_i dPhi_i T_i.
//   k_xc      per point: the functional in FP64, (w v_rho/2, w u grad rho)
//             and per-workgroup partial E_xc and N.
//   k_vxc_group  per group (up to 272 AO rows): every lower 16x16 tile pair
//             of the group's V in registers for the whole pass over its
//             points; Z = w v_rho/2 Phi + wu . dPhi and Phi staged in LDS 16
//             points at a time, acc += Z_I Phi_J^T + Phi_I Z_J^T on MFMA
//             (FP64); k_vxc_group_pipe is its pipelined FP32 form for the
//             mixed build.  Both come in a 17-tile and an 8-tile class.
//   k_vxc     the same per 64x64 tile pair, for groups too wide for the above.
//   k_reduce_v   deterministic sum of the group tiles into the full symmetric V
//             (k_reduce_v_combine adds the partial rows of a split row).
//   k_reduce_en  deterministic sum of the E_xc / N partials.
// T = double runs FP64 MFMA (v_mfma_f64_16x16x4f64); T = float runs FP32 MFMA
// (v_mfma_f32_16x16x4f32) with FP32 accumulation inside a tile and FP64 for
// the functional, the per-point reductions and every sum across tiles.
#include <algorithm>

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

// BF16 operands of v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][8 (l / 16) + e]
// and B[8 (l / 16) + e][l % 16], e = 0..7, element e in half e % 2 of dword
// e / 2; the result D[4 (l / 16) + v][l % 16] lands as in the FP32 16x16x4 form.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma_bf16(const v4u& a, const v4u& b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// (x, y) rounded to BF16 (nearest, ties to even) in one dword, x in the low half.
__device__ __forceinline__ unsigned pack_bf16(float x, float y) {
    const bf16x2 h = {static_cast<__bf16>(x), static_cast<__bf16>(y)};
    return __builtin_bit_cast(unsigned, h);
}

// Three-term BF16 split of the pair (x, y): x = x0 + x1 + x2 + e with each
// term the nearest BF16 to what the terms before it leave (every remainder is
// exact in FP32), |e| <= 2^-24 |x| up to FP32 underflow; the same for y.  The
// products of two such splits that reach FP32 resolution are the six with
// term indices summing to at most 2.
struct Bf16Split3 {
    unsigned t[3];
};
__device__ __forceinline__ Bf16Split3 split3_bf16(float x, float y) {
    Bf16Split3 o;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o.t[k] = pack_bf16(x, y);
        if (k < 2) {
            x -= __builtin_bit_cast(float, o.t[k] << 16);
            y -= __builtin_bit_cast(float, o.t[k] & 0xffff0000u);
        }
    }
    return o;
}

template <typename T>
struct Vec4;
template <>
struct Vec4<double> {
    using type = v4d;
};
template <>
struct Vec4<float> {
    using type = v4f;
};

// MFMA 16x16x4 operand layout (one value per lane): A[i][k] at lane i + 16 k,
// B[k][j] at lane j + 16 k.  Element v of lane j + 16 q of the 16x16 result
// holds D[Mma<T>::row(q, v)][j]: row 4 q + v for the f32 instruction, row
// 4 v + q for the f64 one (one output row group per register).

// ---------------------------------------------------------------------------
// rho stage
// ---------------------------------------------------------------------------
// Dynamic LDS of one rho workgroup, carved in this order (every piece is a
// multiple of 16 bytes): points, per-wave sums, the exponential's table,
// radial table [shell][point],
// rows, shells, primitives (the block's packed records, copied as one run),
// FP64 leading-row values, row values (rho_lds_bytes).

// Workgroups are dispatched round robin over the XCDs (workgroup i to XCD
// i mod kXcds), each with its own L2.  Workgroup i = kXcds k + x takes tile k
// of segment x, so XCD x walks its segment in order: consecutive tiles of a
// block (which read the same density block) share an L2.  Workgroups past
// the end of their segment exit at once.

// Position in an f64 16x16x4 A tile that makes its result element v of lane
// (q, j) land on row 4 q + v, the row the f32 instruction gives that slot.
__device__ __forceinline__ int f64_row_as_f32(int i) { return 4 * (i & 3) + (i >> 2); }

// One workgroup (4 waves) per tile of 16 points of a rho block.  The block's
// rows, shells and primitives are staged in LDS; radial sums once per (shell,
// point); then wave w evaluates (phi, grad phi) for its 16-row tiles
// w, w + 4, ... in exactly the lanes and slots where the MFMA result T of
// those rows will land (lane (j, q), slot v: row Mma<TS>::row(q, v), point
// j); phi goes to LDS as the MFMA B operand, (phi, grad phi) of the wave's
// first two row tiles stay in registers for the epilogue (later row tiles,
// in blocks of more than 128 rows, recompute them).  T = P_sub Phi on MFMA (density rows from L2, loaded two
// chunks ahead), rho += phi T and grad rho += 2 grad phi T, summed over rows
// with cross-lane shuffles and then over the waves.

//
// NS = 2 (unrestricted builds): the same AO values serve both spin densities;
// each row tile is contracted with the alpha and then the beta density block
// and both sets of sums are formed (rho_out gets alpha, rho_out_b beta).
//
// AOC: kAocNone evaluates the AOs; kAocFill evaluates them and stores them
// in the per-geometry cache a.aoc (no contraction, no output); kAocLoad reads
// them from the cache instead of evaluating (mixed build only: the leading
// rows' values are FP64 there, the others FP32).
enum AocMode { kAocNone = 0, kAocFill = 1, kAocLoad = 2 };
template <typename TS, bool HYB, int NS, int AOC>
__global__ __launch_bounds__(kThreads, 2) void k_rho(RhoArgs a) {
    static_assert(AOC == kAocNone || (HYB && sizeof(TS) == 4), "the AO cache has the mixed layout");
    static_assert(AOC != kAocFill || NS == 1, "one fill pass per geometry");
    extern __shared__ __align__(16) unsigned char smem[];
    // row tiles per wave whose (phi, grad phi) stay in registers for the
    // epilogue: two when they are evaluated (later ones are evaluated again),
    // one when they come from the AO cache (later ones are read again, from
    // L2), which keeps that kernel at 120 VGPRs: four waves per SIMD
    constexpr int PTS = kRhoPts, LDP = kRhoPts, kRhoCached = AOC == kAocLoad ? 1 : 2;
    const int xcd = blockIdx.x % kXcds, t = a.seg[xcd] + blockIdx.x / kXcds;
    if (t >= a.seg[xcd + 1]) return;
    const int2 tile = a.tiles[t];  // (block, first point within it): up to kRhoSub x 16 points
    const RhoBlock blk = a.blocks[tile.x];
    const int m = blk.m, mp = blk.m_pad, nsh = blk.n_sh, T16 = blk.m_pad >> 4;
    const int m64 = HYB ? blk.m64 : 0;
    // (a launch that reads the AO cache stages no points, records or radial
    // sums: those pieces are empty)
    constexpr bool kEval = AOC != kAocLoad;
    unsigned char* q = smem;
    double4* sPtAll = reinterpret_cast<double4*>(q);  // [kRhoSub * PTS]
    q += kEval ? sizeof(double4) * kRhoSub * kRhoPts : 0;
    double* sRed = reinterpret_cast<double*>(q);  // [4 waves][4 NS sums][PTS]
    q += sizeof(double) * 16 * NS * PTS;
    double* sExp = reinterpret_cast<double*>(q);  // [64] the exponential's table
    q += kEval ? sizeof(double) * 64 : 0;
    double2* sRad = reinterpret_cast<double2*>(q);  // [nsh][PTS] (R0, R1)
    q += kEval ? sizeof(double2) * nsh * PTS : 0;
    RowL* sRow = reinterpret_cast<RowL*>(q);
    q += kEval ? sizeof(RowL) * mp : 0;
    DevShell* sSh = reinterpret_cast<DevShell*>(q);
    q += kEval ? sizeof(DevShell) * nsh : 0;
    double2* sPrim = reinterpret_cast<double2*>(q);
    q += kEval ? sizeof(double2) * blk.n_prim : 0;
    double* sPhi64 = reinterpret_cast<double*>(q);  // [m64][LDP]
    q += HYB ? sizeof(double) * m64 * LDP : 0;
    TS* sPhi = reinterpret_cast<TS*>(q);  // [mp][LDP]

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kq = lane >> 4;
    using M = Mma<TS>;
    const TS* __restrict__ PS = static_cast<const TS*>(a.PS);
    const TS* __restrict__ PSb = static_cast<const TS*>(a.PSb);
    using V4 = typename Vec4<TS>::type;
    // The first row tile's first density chunks are loaded now, so that their
    // latency passes during the staging and AO evaluation below (the loads
    // are the ones contract() makes for that tile; see there for the tail).
    V4 pre0{}, pre1{};
    double pre64[4] = {0.0, 0.0, 0.0, 0.0};
    auto preload = [&]() {
        if (AOC != kAocFill && wave < T16) {
            const TS* __restrict__ Arow = PS + blk.psub + static_cast<long long>((wave << 4) + col) * mp;
            pre0 = *reinterpret_cast<const V4*>(Arow + m64 + 4 * kq);
            pre1 = *reinterpret_cast<const V4*>(Arow + m64 + 16 + 4 * kq);
            if (HYB && m64 >= 16) {
                const double* __restrict__ A64 =
                    a.P64 + blk.psub + static_cast<long long>((wave << 4) + f64_row_as_f32(col)) * mp;
#pragma unroll
                for (int s = 0; s < 4; ++s) pre64[s] = A64[4 * s + kq];
            }
        }
    };
    preload();
    if constexpr (AOC != kAocLoad) {
        // rows, shells and primitives: one coalesced copy of the packed records
        const uint4* __restrict__ src = a.meta + blk.meta;
        uint4* dst = reinterpret_cast<uint4*>(sRow);
        const int n16 = mp + nsh * static_cast<int>(sizeof(DevShell) / sizeof(uint4)) + blk.n_prim;
        for (int i = tid; i < n16; i += kThreads) dst[i] = src[i];
        if (tid < 64) sExp[tid] = kExpNegTab[tid];
    }
    // The workgroup's points: kRhoSub sub-tiles of 16, taken one after the
    // other with the block's records staged once.
    const int t0 = tile.y, np_all = min(kRhoSub * PTS, blk.n_pts - t0);
    if (AOC != kAocLoad && tid < kRhoSub * PTS)
        sPtAll[tid] = tid < np_all ? a.pts[blk.pt_begin + t0 + tid] : make_double4(0.0, 0.0, 0.0, 0.0);
    __syncthreads();
    // This workgroup's part of the AO cache (see aoc_sub_bytes).
    const int L16 = m64 >> 4;
    unsigned char* const cache_base =
        AOC != kAocNone ? static_cast<unsigned char*>(a.aoc) + (a.aoc_off[t] << 4) : nullptr;
    const long long sub_bytes = aoc_sub_bytes(mp, m64);
    // (each sub-tile is a separate inlined copy of the body, not a loop:
    // a loop lets the compiler hoist per-tile addresses across both passes
    // and nearly doubles the registers)
    auto process = [&](const int sub) {
        const double4* sPt = sPtAll + sub * PTS;
        const int np = min(PTS, np_all - sub * PTS), tsub = t0 + sub * PTS;
        if (sub > 0) preload();

        // Slot v of row tile it of this sub-tile in the AO cache: its phi
        // and its gradient (see aoc_sub_bytes).
        unsigned char* const csub = AOC != kAocNone ? cache_base + sub * sub_bytes : nullptr;
        auto cphi = [&](int it, int v) {
            return csub + aoc_phi_off(it, L16) + (v * 64 + lane) * (it < L16 ? 8 : 4);
        };
        auto cgrad = [&](int it, int v) {
            return csub + aoc_grad_off(it, L16, T16) + (v * 64 + lane) * (it < L16 ? 24 : 12);
        };
        // (phi, grad phi) of the four slots of row tile it
        auto cload = [&](int it, AoVal (&f)[4]) {
            if (it < L16) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double* g = reinterpret_cast<const double*>(cgrad(it, v));
                    f[v] = AoVal{*reinterpret_cast<const double*>(cphi(it, v)), g[0], g[1], g[2]};
                }
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float* g = reinterpret_cast<const float*>(cgrad(it, v));
                    f[v] = AoVal{*reinterpret_cast<const float*>(cphi(it, v)), g[0], g[1], g[2]};
                }
            }
        };
        auto cload_phi = [&](int it, int v) {
            return it < L16 ? *reinterpret_cast<const double*>(cphi(it, v))
                            : static_cast<double>(*reinterpret_cast<const float*>(cphi(it, v)));
        };
        auto cstore = [&](int it, int v, const AoVal& f) {
            if (it < L16) {
                *reinterpret_cast<double*>(cphi(it, v)) = f.v;
                double* g = reinterpret_cast<double*>(cgrad(it, v));
                g[0] = f.gx;
                g[1] = f.gy;
                g[2] = f.gz;
            } else {
                *reinterpret_cast<float*>(cphi(it, v)) = static_cast<float>(f.v);
                float* g = reinterpret_cast<float*>(cgrad(it, v));
                g[0] = static_cast<float>(f.gx);
                g[1] = static_cast<float>(f.gy);
                g[2] = static_cast<float>(f.gz);
            }
        };

        // Radial sums once per (shell, point).
        for (int idx = tid; AOC != kAocLoad && idx < nsh * PTS; idx += kThreads) {
            const int sh = idx / PTS, p = idx % PTS;
            double2 rr = make_double2(0.0, 0.0);
            if (p < np) {
                const double4 x = sPt[p];
                rr = shell_radial_tab(sSh[sh], sPrim, x.x, x.y, x.z, sExp);
            }
            sRad[idx] = rr;
        }
        __syncthreads();

        // (phi, grad phi) of row Mma::row(kq, v) of row tile it at point col,
        // without a per-lane branch: padding rows (shell -1, scale 0) read
        // shell 0 and come out zero, and so do points past the tile's end
        // (zero radial sums).  A block without rows has no shell 0 and skips
        // this as a whole.
        auto eval = [&](int it, int v) {
            AoVal f{0.0, 0.0, 0.0, 0.0};
            if (nsh > 0) {
                const int r = (it << 4) + M::row(kq, v);
                const RowL row = sRow[r];
                const int sh = max(row.sh, 0);
                const double4 xp = sPt[col];
                f = ao_from_radial_row(sSh[sh], row, xp.x, xp.y, xp.z, sRad[sh * PTS + col]);
            }
            return f;
        };
        if constexpr (AOC == kAocFill) {
            for (int it = wave; it < T16; it += 4)
#pragma unroll
                for (int v = 0; v < 4; ++v) cstore(it, v, eval(it, v));
            return;
        }
        AoVal cache[kRhoCached][4];
        if constexpr (AOC == kAocLoad) {
#pragma unroll
            for (int j = 0; j < kRhoCached; ++j)
                if (wave + 4 * j < T16) cload(wave + 4 * j, cache[j]);
        }
#pragma unroll
        for (int j = 0; j < kRhoCached; ++j) {
            const int it = wave + 4 * j;
            if (it < T16) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (it << 4) + M::row(kq, v);
                    if constexpr (AOC != kAocLoad) cache[j][v] = eval(it, v);
                    sPhi[r * LDP + col] = static_cast<TS>(cache[j][v].v);
                    if (HYB && r < m64) sPhi64[r * LDP + col] = cache[j][v].v;
                }
            }
        }
        for (int it = wave + 4 * kRhoCached; it < T16; it += 4) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = (it << 4) + M::row(kq, v);
                const double fv = AOC == kAocLoad ? cload_phi(it, v) : eval(it, v).v;
                sPhi[r * LDP + col] = static_cast<TS>(fv);
                if (HYB && r < m64) sPhi64[r * LDP + col] = fv;
            }
        }
        __syncthreads();

        // T for row tile it: this lane ends up with T[row(kq, v)][col].
        auto contract = [&](int it, typename M::acc& acc, v4d& acc64, bool first, int spin) {
            const TS* __restrict__ PSs = (NS == 2 && spin == 1) ? PSb : PS;
            const double* __restrict__ P64s = (NS == 2 && spin == 1) ? a.P64b : a.P64;
            const int i0 = it << 4;
            acc = M::zero();
            // The first two density chunks of the rows in the precision TS are
            // requested before the FP64 leading part, so that they arrive
            // while it runs.
            const TS* __restrict__ Ap = PSs + blk.psub + static_cast<long long>(i0 + col) * mp + m64 + 4 * kq;
            auto load = [&](int c) { return *reinterpret_cast<const V4*>(Ap + 16 * c); };
            V4 r0 = first ? pre0 : load(0);
            V4 r1 = first ? pre1 : load(1);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (HYB) {
                acc64 = Mma<double>::zero();
                // Leading rows (large |phi|) in FP64 with the FP64 density
                // block, four k-steps per round with their loads issued together.
                const double* __restrict__ Arow =
                    P64s + blk.psub + static_cast<long long>(i0 + f64_row_as_f32(col)) * mp;
                int k = 0;
                for (; k < m64; k += 16) {
                    double av[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) av[s] = first && k == 0 ? pre64[s] : Arow[k + 4 * s + kq];
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        acc64 = Mma<double>::op(av[s], sPhi64[(k + 4 * s + kq) * LDP + col], acc64);
                }
            }
            // K in chunks of 16 whose slots are permuted so that lane (i, q)
            // reads four consecutive density entries per load: step s of a
            // chunk pairs A[i][c + 4 q + s] with Phi[c + 4 q + s].  The
            // density loads run two chunks ahead through a ring of three
            // registers, unrolled by three so that no register is copied
            // (a copy would wait for its load at the copy); reads past the
            // row end land in the next row or the buffer tail and are not
            // used.  m64 is a multiple of 16, so the chunks start there.
            const TS* __restrict__ Bp = sPhi + (m64 + 4 * kq) * LDP + col;
            const int nch = (mp - m64) >> 4;
            auto chunk = [&](const V4& av, int c) {
                const TS* b = Bp + c * 16 * LDP;
                const TS b0 = b[0], b1 = b[LDP], b2 = b[2 * LDP], b3 = b[3 * LDP];
                __builtin_amdgcn_sched_barrier(0);  // the four reads in flight together
                acc = M::op(av[0], b0, acc);
                acc = M::op(av[1], b1, acc);
                acc = M::op(av[2], b2, acc);
                acc = M::op(av[3], b3, acc);
            };
            V4 r2;
            int c = 0;
            // The scheduling barriers keep each load ahead of the chunk that
            // follows it (the scheduler would otherwise sink it to its use).
            for (; c + 3 <= nch; c += 3) {
                r2 = load(c + 2);
                __builtin_amdgcn_sched_barrier(0);
                chunk(r0, c);
                __builtin_amdgcn_sched_barrier(0);
                r0 = load(c + 3);
                __builtin_amdgcn_sched_barrier(0);
                chunk(r1, c + 1);
                __builtin_amdgcn_sched_barrier(0);
                r1 = load(c + 4);
                __builtin_amdgcn_sched_barrier(0);
                chunk(r2, c + 2);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (c < nch) chunk(r0, c);
            if (c + 1 < nch) chunk(r1, c + 1);
        };
        // sums[s] = (rho, grad rho / 2) of spin s at this lane's point
        double sums[NS][4];
#pragma unroll
        for (int sp = 0; sp < NS; ++sp)
#pragma unroll
            for (int k = 0; k < 4; ++k) sums[sp][k] = 0.0;
        auto accumulate = [&](int sp, const typename M::acc& acc, const v4d& acc64, int v, const AoVal& f) {
            double t = static_cast<double>(acc[v]);
            if constexpr (HYB) t += acc64[v];
            sums[sp][0] += f.v * t;
            sums[sp][1] += f.gx * t;
            sums[sp][2] += f.gy * t;
            sums[sp][3] += f.gz * t;
        };
#pragma unroll
        for (int j = 0; j < kRhoCached; ++j) {
            const int it = wave + 4 * j;
            if (it < T16) {
#pragma unroll
                for (int sp = 0; sp < NS; ++sp) {
                    typename M::acc acc;
                    v4d acc64;
                    contract(it, acc, acc64, j == 0 && sp == 0, sp);
#pragma unroll
                    for (int v = 0; v < 4; ++v) accumulate(sp, acc, acc64, v, cache[j][v]);
                }
            }
        }
        for (int it = wave + 4 * kRhoCached; it < T16; it += 4) {
            typename M::acc acc[NS];
            v4d acc64[NS];
            AoVal w[4];
            if constexpr (AOC == kAocLoad) cload(it, w);
#pragma unroll
            for (int sp = 0; sp < NS; ++sp) contract(it, acc[sp], acc64[sp], false, sp);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const AoVal f = AOC == kAocLoad ? w[v] : eval(it, v);
#pragma unroll
                for (int sp = 0; sp < NS; ++sp) accumulate(sp, acc[sp], acc64[sp], v, f);
            }
        }
        // Sum over the four row groups of the wave, then over the waves.
        // (all shuffles first, then the stores under one branch)
#pragma unroll
        for (int sp = 0; sp < NS; ++sp)
#pragma unroll
            for (int k = 0; k < 4; ++k) sums[sp][k] += __shfl_xor(sums[sp][k], 16);
#pragma unroll
        for (int sp = 0; sp < NS; ++sp)
#pragma unroll
            for (int k = 0; k < 4; ++k) sums[sp][k] += __shfl_xor(sums[sp][k], 32);
        if (kq == 0) {
#pragma unroll
            for (int sp = 0; sp < NS; ++sp)
#pragma unroll
                for (int k = 0; k < 4; ++k) sRed[(wave * 4 * NS + 4 * sp + k) * PTS + col] = sums[sp][k];
        }
        __syncthreads();
        double4* const outs[2] = {a.rho_out, a.rho_out_b};
        if (tid < np) {
            const int g = blk.pt_begin + tsub + tid;
            const int go = a.rho_perm ? a.perm[g] : g;
#pragma unroll
            for (int sp = 0; sp < NS; ++sp) {
                double r[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) r[k] = sRed[(4 * sp + k) * PTS + tid];
#pragma unroll
                for (int w = 1; w < 4; ++w)
#pragma unroll
                    for (int k = 0; k < 4; ++k) r[k] += sRed[(w * 4 * NS + 4 * sp + k) * PTS + tid];
                outs[sp][go] = make_double4(r[0], 2.0 * r[1], 2.0 * r[2], 2.0 * r[3]);
            }
        }
    };
    static_assert(kRhoSub == 2, "two sub-tiles per workgroup");
    process(0);
    if (np_all > PTS) {
        __syncthreads();  // the second sub-tile rewrites the radial table, phi and the sums
        process(1);
    }
}

// Mixed rho stage from the AO cache, NS spins (1, or 2 for unrestricted
// builds, whose two density blocks share one read of the cache), with the
// workgroup's two 16-point sub-tiles contracted together: every density
// chunk of T = P_sub Phi that a wave loads feeds the MFMAs of both sub-tiles
// (32 points), so the density blocks are read half as often as by one
// sub-tile at a time.  LDS holds the B operands of both sub-tiles: Phi of the
// FP64 leading rows (below m64) and of the FP32 rows (from m64 on).  Wave w
// takes the 16-row tiles w, w + 4, ...; after each tile's contraction (per
// spin) the epilogue reads the tile's gradient slots of one sub-tile, then
// the other, from the cache's gradient plane and phi from LDS, so the cache
// is read once.  (Requesting the gradient slots before the contraction was
// slower: 0.570 against 0.549 ms per 20-water stage; overlaying the sums on
// the operands, for less LDS per workgroup, made no difference.)  The sums of
// each spin are those of k_rho<float, true, 1, kAocLoad>, formed in the same
// order; with NS = 2 the unrestricted 20-water build took 2.54 against 3.01
// ms with k_rho<float, true, 2, kAocLoad>, bitwise equal.
template <int NS>
__global__ __launch_bounds__(kThreads, 2) void k_rho_aoc2(RhoArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int PTS = kRhoPts, PT2 = 2 * kRhoPts;
    const int xcd = blockIdx.x % kXcds, t = a.seg[xcd] + blockIdx.x / kXcds;
    if (t >= a.seg[xcd + 1]) return;
    const int2 tile = a.tiles[t];  // (block, first point within it): up to 32 points
    const RhoBlock blk = a.blocks[tile.x];
    const int mp = blk.m_pad, T16 = mp >> 4, m64 = blk.m64, L16 = m64 >> 4, mr = mp - m64;
    const int np_all = min(PT2, blk.n_pts - tile.y);
    const bool two = np_all > PTS;  // the second sub-tile exists (and has its cache slots)
    double* sRed = reinterpret_cast<double*>(smem);                   // [4 waves][NS][4 sums][PT2]
    double* sPhi64 = sRed + 16 * NS * PT2;                             // [2 sub-tiles][m64][PTS]
    float* sPhi = reinterpret_cast<float*>(sPhi64 + 2 * m64 * PTS);    // [2 sub-tiles][mr][PTS]

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kq = lane >> 4;
    // the block's density blocks (per spin): uniform bases, 32-bit per-lane offsets
    const float* __restrict__ PS0 = static_cast<const float*>(a.PS) + blk.psub;
    const double* __restrict__ P640 = a.P64 + blk.psub;
    const float* __restrict__ PS1 = NS == 2 ? static_cast<const float*>(a.PSb) + blk.psub : PS0;
    const double* __restrict__ P641 = NS == 2 ? a.P64b + blk.psub : P640;
    // Slot v of row tile it of sub-tile s in the AO cache: its phi and its
    // gradient (see aoc_sub_bytes).  A tile without a second sub-tile reads
    // the first one in its place: that sub-tile's sums are then never written.
    const unsigned char* const cbase = static_cast<const unsigned char*>(a.aoc) + (a.aoc_off[t] << 4);
    const long long sub1 = two ? aoc_sub_bytes(mp, m64) : 0;
    auto phi_at = [&](int it, int v, int s) {
        return cbase + (s ? sub1 : 0) + aoc_phi_off(it, L16) +
               static_cast<unsigned>((v * 64 + lane) * (it < L16 ? 8 : 4));
    };
    auto grad_at = [&](int it, int v, int s) {
        return cbase + (s ? sub1 : 0) + aoc_grad_off(it, L16, T16) +
               static_cast<unsigned>((v * 64 + lane) * (it < L16 ? 24 : 12));
    };
    // B operands: phi of both sub-tiles.
    for (int it = wave; it < T16; it += 4) {
        if (it < L16) {
            double f[2][4];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    f[s][v] = *reinterpret_cast<const double*>(phi_at(it, v, s));
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int v = 0; v < 4; ++v) sPhi64[(s * m64 + (it << 4) + 4 * kq + v) * PTS + col] = f[s][v];
        } else {
            float f[2][4];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    f[s][v] = *reinterpret_cast<const float*>(phi_at(it, v, s));
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    sPhi[(s * mr + (it << 4) - m64 + 4 * kq + v) * PTS + col] = f[s][v];
        }
    }
    __syncthreads();

    // sums[sp][s] = (rho, grad rho / 2) of spin sp at this lane's point of sub-tile s
    double sums[NS][2][4];
#pragma unroll
    for (int sp = 0; sp < NS; ++sp)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < 4; ++k) sums[sp][s][k] = 0.0;
    auto accumulate = [&](int sp, int s, const v4f& acc, const v4d& acc64, int v, const AoVal& f) {
        const double tv = static_cast<double>(acc[v]) + acc64[v];
        sums[sp][s][0] += f.v * tv;
        sums[sp][s][1] += f.gx * tv;
        sums[sp][s][2] += f.gy * tv;
        sums[sp][s][3] += f.gz * tv;
    };
    // T of row tile it for both sub-tiles: this lane ends up with
    // T[4 kq + v][col] of sub-tile s in acc[s][v] + acc64[s][v].
    auto contract = [&](int it, v4f (&acc)[2], v4d (&acc64)[2], int sp) {
        const float* __restrict__ PSb = sp ? PS1 : PS0;
        const double* __restrict__ P64b = sp ? P641 : P640;
        const int i0 = it << 4;
        acc[0] = acc[1] = Mma<float>::zero();
        acc64[0] = acc64[1] = Mma<double>::zero();
        const unsigned o32 = static_cast<unsigned>((i0 + col) * mp + m64 + 4 * kq);
        auto load = [&](int c) { return *reinterpret_cast<const v4f*>(PSb + (o32 + 16u * c)); };
        v4f r0 = load(0);
        v4f r1 = load(1);
        __builtin_amdgcn_sched_barrier(0);
        // leading columns in FP64 with the FP64 density block, in chunks of
        // 16 whose slots are permuted as in the FP32 part: step s of a chunk
        // pairs A[i][k + 4 q + s] with Phi[k + 4 q + s], so that lane (i, q)
        // reads its four density entries as one 32-byte load
        const unsigned o64 = static_cast<unsigned>((i0 + f64_row_as_f32(col)) * mp + 4 * kq);
        for (int k = 0; k < m64; k += 16) {
            const v4d av = *reinterpret_cast<const v4d*>(P64b + (o64 + static_cast<unsigned>(k)));
            const double* b = sPhi64 + (k + 4 * kq) * PTS + col;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                acc64[0] = Mma<double>::op(av[s], b[s * PTS], acc64[0]);
                acc64[1] = Mma<double>::op(av[s], b[(m64 + s) * PTS], acc64[1]);
            }
        }
        // the FP32 columns in chunks of 16, density loads two chunks ahead
        // through a ring of three registers (as in k_rho)
        const float* __restrict__ Bp = sPhi + (4 * kq) * PTS + col;
        const int nch = mr >> 4;
        auto chunk = [&](const v4f& av, int c) {
            const float* b = Bp + c * 16 * PTS;
            const float* e = b + mr * PTS;
            const float b0 = b[0], b1 = b[PTS], b2 = b[2 * PTS], b3 = b[3 * PTS];
            const float e0 = e[0], e1 = e[PTS], e2 = e[2 * PTS], e3 = e[3 * PTS];
            __builtin_amdgcn_sched_barrier(0);  // the eight reads in flight together
            acc[0] = Mma<float>::op(av[0], b0, acc[0]);
            acc[1] = Mma<float>::op(av[0], e0, acc[1]);
            acc[0] = Mma<float>::op(av[1], b1, acc[0]);
            acc[1] = Mma<float>::op(av[1], e1, acc[1]);
            acc[0] = Mma<float>::op(av[2], b2, acc[0]);
            acc[1] = Mma<float>::op(av[2], e2, acc[1]);
            acc[0] = Mma<float>::op(av[3], b3, acc[0]);
            acc[1] = Mma<float>::op(av[3], e3, acc[1]);
        };
        v4f r2;
        int c = 0;
        for (; c + 3 <= nch; c += 3) {
            r2 = load(c + 2);
            __builtin_amdgcn_sched_barrier(0);
            chunk(r0, c);
            __builtin_amdgcn_sched_barrier(0);
            r0 = load(c + 3);
            __builtin_amdgcn_sched_barrier(0);
            chunk(r1, c + 1);
            __builtin_amdgcn_sched_barrier(0);
            r1 = load(c + 4);
            __builtin_amdgcn_sched_barrier(0);
            chunk(r2, c + 2);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (c < nch) chunk(r0, c);
        if (c + 1 < nch) chunk(r1, c + 1);
    };
    for (int it = wave; it < T16; it += 4) {
        v4f acc[NS][2];
        v4d acc64[NS][2];
        if (it < L16) {
            // leading row tile: its 32-byte slots after the contraction
#pragma unroll
            for (int sp = 0; sp < NS; ++sp) contract(it, acc[sp], acc64[sp], sp);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                AoVal w[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double* g = reinterpret_cast<const double*>(grad_at(it, v, s));
                    w[v] = AoVal{sPhi64[(s * m64 + (it << 4) + 4 * kq + v) * PTS + col], g[0], g[1], g[2]};
                }
#pragma unroll
                for (int sp = 0; sp < NS; ++sp)
#pragma unroll
                    for (int v = 0; v < 4; ++v) accumulate(sp, s, acc[sp][s], acc64[sp][s], v, w[v]);
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int sp = 0; sp < NS; ++sp) contract(it, acc[sp], acc64[sp], sp);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                float4 w[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float* g = reinterpret_cast<const float*>(grad_at(it, v, s));
                    w[v] = make_float4(sPhi[(s * mr + (it << 4) - m64 + 4 * kq + v) * PTS + col], g[0], g[1], g[2]);
                }
#pragma unroll
                for (int sp = 0; sp < NS; ++sp)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        accumulate(sp, s, acc[sp][s], acc64[sp][s], v, AoVal{w[v].x, w[v].y, w[v].z, w[v].w});
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // Sum over the four row groups of the wave, then over the waves, in the
    // order k_rho uses.
#pragma unroll
    for (int sp = 0; sp < NS; ++sp)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < 4; ++k) sums[sp][s][k] += __shfl_xor(sums[sp][s][k], 16);
#pragma unroll
    for (int sp = 0; sp < NS; ++sp)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < 4; ++k) sums[sp][s][k] += __shfl_xor(sums[sp][s][k], 32);
    if (kq == 0) {
#pragma unroll
        for (int sp = 0; sp < NS; ++sp)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < 4; ++k) sRed[((wave * NS + sp) * 4 + k) * PT2 + s * PTS + col] = sums[sp][s][k];
    }
    __syncthreads();
    if (tid < np_all) {
        const int g = blk.pt_begin + tile.y + tid;
        const int go = a.rho_perm ? a.perm[g] : g;
        double4* const outs[2] = {a.rho_out, a.rho_out_b};
#pragma unroll
        for (int sp = 0; sp < NS; ++sp) {
            double r[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = sRed[(sp * 4 + k) * PT2 + tid];
#pragma unroll
            for (int w = 1; w < 4; ++w)
#pragma unroll
                for (int k = 0; k < 4; ++k) r[k] += sRed[((w * NS + sp) * 4 + k) * PT2 + tid];
            outs[sp][go] = make_double4(r[0], 2.0 * r[1], 2.0 * r[2], 2.0 * r[3]);
        }
    }
}

// Sums of e and n over the workgroup (kThreads threads) into *oe and *on, in
// a fixed order: within each wave by xor shuffles, then the waves in order.
__device__ __forceinline__ void wg_sum2(double e, double n, double* oe, double* on) {
    __shared__ double sw[2][kThreads / 64];
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
        for (int w = 0; w < kThreads / 64; ++w) {
            te += sw[0][w];
            tn += sw[1][w];
        }
        *oe = te;
        *on = tn;
    }
}

// Functional stage, one thread per sorted point: from (rho, grad rho) and the
// weight, wv = (w v_rho_a / 2, w (v_gamma_aa + v_gamma_ab / 2) grad rho) and
// the workgroup's sums of w f_xc and w rho (vxc.rs:892-920).
// One instance per functional (F = a.func), so that the functional's code is
// inlined and its registers are allocated for that functional alone.
template <int F>
__global__ __launch_bounds__(kThreads) void k_xc(XcArgs a) {
    const int2 wg = a.wg[blockIdx.x];
    const int g = wg.x + threadIdx.x;
    double e = 0.0, ne = 0.0;
    if (static_cast<int>(threadIdx.x) < wg.y) {
        const double4 r = a.rho[g];
        const double w = a.pts[g].w;
        ne = w * r.x;
        double4 out = make_double4(0.0, 0.0, 0.0, 0.0);
        if (r.x >= 1e-18) {  // DENSITY_SCREEN, vxc.rs:23
            const double rs = 0.5 * r.x;
            const double gss = 0.25 * (r.y * r.y + r.z * r.z + r.w * r.w);
            const XcOut o = eval_xc(F, rs, rs, gss, gss, gss);
            e = w * o.f;
            const double us = w * (o.gaa + 0.5 * o.gab);
            out = make_double4(0.5 * w * o.va, us * r.y, us * r.z, us * r.w);
        }
        a.wv[g] = out;
    }
    wg_sum2(e, ne, a.e_part + blockIdx.x, a.n_part + blockIdx.x);
}

// Unrestricted functional stage, one thread per sorted point
// (crates/calchem-dft/src/vxc.rs:1141-1348, accumulate_uks_points): spin
// densities clamped at zero, skipped when rho_a + rho_b < 1e-18; gamma_aa,
// gamma_ab, gamma_bb from the spin gradients; u_a = 2 v_gaa grad rho_a +
// v_gab grad rho_b (and a <-> b); wv_s = (w v_s / 2, w u_s); N sums the
// unclamped densities.
template <int F>
__global__ __launch_bounds__(kThreads) void k_xc_uks(XcUksArgs a) {
    const int2 wg = a.wg[blockIdx.x];
    const int g = wg.x + threadIdx.x;
    double e = 0.0, ne = 0.0;
    if (static_cast<int>(threadIdx.x) < wg.y) {
        const double4 ra = a.rho_a[g], rb = a.rho_b[g];
        const double w = a.pts[g].w;
        ne = w * (ra.x + rb.x);
        double4 oa = make_double4(0.0, 0.0, 0.0, 0.0), ob = oa;
        const double pa = ra.x > 0.0 ? ra.x : 0.0, pb = rb.x > 0.0 ? rb.x : 0.0;
        if (pa + pb >= 1e-18) {  // DENSITY_SCREEN, vxc.rs:23
            const double gaa = ra.y * ra.y + ra.z * ra.z + ra.w * ra.w;
            const double gab = ra.y * rb.y + ra.z * rb.z + ra.w * rb.w;
            const double gbb = rb.y * rb.y + rb.z * rb.z + rb.w * rb.w;
            const XcOut o = eval_xc(F, pa, pb, gaa, gab, gbb);
            e = w * o.f;
            oa = make_double4(0.5 * w * o.va, w * (2.0 * o.gaa * ra.y + o.gab * rb.y),
                              w * (2.0 * o.gaa * ra.z + o.gab * rb.z), w * (2.0 * o.gaa * ra.w + o.gab * rb.w));
            ob = make_double4(0.5 * w * o.vb, w * (2.0 * o.gbb * rb.y + o.gab * ra.y),
                              w * (2.0 * o.gbb * rb.z + o.gab * ra.z), w * (2.0 * o.gbb * rb.w + o.gab * ra.w));
        }
        a.wv_a[g] = oa;
        a.wv_b[g] = ob;
    }
    wg_sum2(e, ne, a.e_part + blockIdx.x, a.n_part + blockIdx.x);
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
        sAoI[tid] = gi < G.m ? a.group_rows[G.ao_begin + gi].x : -1;
        sAoJ[tid] = gj < G.m ? a.group_rows[G.ao_begin + gj].x : -1;
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

// Register-resident V kernel: one workgroup of NW wavefronts per group.
// Every lower pair of 16-row tiles of the group's V lives in the accumulators
// of one wavefront for the whole pass over the group's points; 16 points at a
// time, radial sums (per shell) then phi and Z = w v_rho/2 phi + wu . dphi
// (per row) are staged in LDS, each evaluated once, and every pair takes
// acc += Z_I Phi_J^T + Phi_I Z_J^T.  FP64: 8 waves, AO evaluation in FP64.
// (The mixed build uses k_vxc_group_pipe below: FP32 tiles and AO values.)
template <typename T, int MAXT>
struct VGroupCfg;
template <>
struct VGroupCfg<double, 17> {
    static constexpr int kWaves = 8;  // 20 FP64 pairs per wave: two waves per SIMD
    using Rad = double2;
};
template <>
struct VGroupCfg<double, 8> {
    static constexpr int kWaves = 8;  // 5 FP64 pairs per wave; small LDS, two workgroups per CU
    using Rad = double2;
};
template <>
struct VGroupCfg<double, 12> {
    static constexpr int kWaves = 16;  // 5 FP64 pairs per wave: four waves per SIMD
    using Rad = double2;
};

template <int MAXT>
__host__ __device__ constexpr int vgroup_max_pairs() {
    return MAXT * (MAXT + 1) / 2;
}

template <typename T, int MAXT>
__host__ __device__ constexpr int vgroup_pairs_per_wave() {
    return (vgroup_max_pairs<MAXT>() + VGroupCfg<T, MAXT>::kWaves - 1) / VGroupCfg<T, MAXT>::kWaves;
}

// Row stride (elements) of the staged phi and Z tiles.  Lane (col, kq) of an
// MFMA operand reads the 4 consecutive points 4 kq .. 4 kq + 3 of row col with
// one 16-byte (FP32) or two 16-byte (FP64) LDS reads; with these strides the
// 16 rows of a tile start on distinct 16-byte bank groups (80 B and 144 B
// rows), so those reads are free of bank conflicts.
template <typename T>
__host__ __device__ constexpr int vgroup_ld() {
    return sizeof(T) == 4 ? 20 : 18;
}

// Four consecutive elements from a 16-byte aligned LDS address: one 16-byte
// read for FP32, two for FP64 (whose 4-vector type would claim 32 bytes of
// alignment, which the rows do not keep).
typedef double v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v4f lds_load4(const float* p) { return *reinterpret_cast<const v4f*>(p); }
__device__ __forceinline__ v4d lds_load4(const double* p) {
    const v2d lo = reinterpret_cast<const v2d*>(p)[0], hi = reinterpret_cast<const v2d*>(p)[1];
    return v4d{lo.x, lo.y, hi.x, hi.y};
}

// Point / wv slots of the V kernels (a ring of steps of 16 points).
constexpr int kVSlots = 4;

// LDS before the per-shell tables; the pair table is rounded up to keep
// 16-byte alignment after it.
template <typename T, int MAXT>
__host__ __device__ constexpr size_t vgroup_lds_fixed() {
    return sizeof(double4) * 2 * kVSlots * kVPts + sizeof(double) * 64 +
           sizeof(int2) * ((vgroup_max_pairs<MAXT>() + 1) & ~1) +
           sizeof(RowL) * MAXT * 16 + sizeof(T) * 2 * MAXT * 16 * vgroup_ld<T>();
}

template <typename T, int MAXT>
__global__ __launch_bounds__((VGroupCfg<T, MAXT>::kWaves * 64)) void k_vxc_group(VArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NW = VGroupCfg<T, MAXT>::kWaves, NT = NW * 64, PPW = vgroup_pairs_per_wave<T, MAXT>();
    using Rad = typename VGroupCfg<T, MAXT>::Rad;
    using V4 = typename Vec4<T>::type;
    constexpr int LD = vgroup_ld<T>();
    static_assert((LD * sizeof(T)) % 16 == 0, "phi / Z rows must keep 16-byte alignment");
    const VGroup G = a.groups[a.group_list[blockIdx.x]];
    const int m = G.m, nsh = G.n_sh, npts = G.n_pts;
    const int T16 = (m + 15) >> 4, mp = T16 << 4;
    const int npairs = T16 * (T16 + 1) / 2;
    const int nsteps = (npts + kVPts - 1) / kVPts;
    unsigned char* q = smem;
    double4* sPt = reinterpret_cast<double4*>(q);  // [kVSlots][kVPts] points, then [kVSlots][kVPts] wv
    double4* sWv = sPt + kVSlots * kVPts;
    q += sizeof(double4) * 2 * kVSlots * kVPts;
    double* sExp = reinterpret_cast<double*>(q);  // [64] the exponential's table
    q += sizeof(double) * 64;
    int2* sPair = reinterpret_cast<int2*>(q);
    q += sizeof(int2) * ((vgroup_max_pairs<MAXT>() + 1) & ~1);
    RowL* sRow = reinterpret_cast<RowL*>(q);
    q += sizeof(RowL) * MAXT * 16;
    T* sP = reinterpret_cast<T*>(q);
    q += sizeof(T) * MAXT * 16 * LD;
    T* sZ = reinterpret_cast<T*>(q);
    q += sizeof(T) * MAXT * 16 * LD;
    Rad* sRad = reinterpret_cast<Rad*>(q);  // [nsh][kVPts]
    q += sizeof(Rad) * ((nsh * kVPts + 1) & ~1);
    DevShell* sSh = reinterpret_cast<DevShell*>(q);
    q += sizeof(DevShell) * nsh;
    double2* sPrim = reinterpret_cast<double2*>(q);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kq = lane >> 4;
    {
        // rows, then shells and primitives: two coalesced copies of the packed records
        const uint4* __restrict__ src = a.gmeta + G.meta;
        for (int i = tid; i < mp; i += NT) reinterpret_cast<uint4*>(sRow)[i] = src[i];
        const int nrest = nsh * static_cast<int>(sizeof(DevShell) / sizeof(uint4)) + G.n_prim;
        for (int i = tid; i < nrest; i += NT) reinterpret_cast<uint4*>(sSh)[i] = src[mp + i];
        if (tid < 64) sExp[tid] = kExpNegTab[tid];
    }

    // Pairs p = ti (ti + 1) / 2 + tj (ti >= tj) in this order; wave w takes
    // the contiguous range [w c, (w + 1) c), c = ceil(npairs / NW), so
    // consecutive pairs of a wave mostly share ti and its fragments are read
    // from LDS once.  The LDS table holds the LDS row offsets of tiles ti and
    // tj (in elements).
    const int per_wave = (npairs + NW - 1) / NW;
    for (int p = tid; p < npairs; p += NT) {
        int ti = 0;
        while ((ti + 1) * (ti + 2) / 2 <= p) ++ti;
        const int tj = p - ti * (ti + 1) / 2;
        sPair[p] = make_int2(ti * 16 * LD, tj * 16 * LD);
    }
    using M = Mma<T>;
    typename M::acc acc[PPW];
#pragma unroll
    for (int w = 0; w < PPW; ++w) acc[w] = M::zero();

    // Two barriers per step of 16 points: between them the MFMA updates of
    // step s run beside the radial sums of step s + 1 and the loads of the
    // points and wv of step s + 2 (slot (s + 2) % kVSlots); then the AO
    // values of step s + 1 replace those of step s.  Slots past the group's
    // last point get w = 0 and Z is zeroed there, so they add nothing.
    auto load_points = [&](int st) {
        if (tid < 2 * kVPts) {
            const int j = tid % kVPts, g = st * kVPts + j, slot = st % kVSlots;
            double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
            if (g < npts) {
                const int sg = G.pt_begin + g;
                v = tid < kVPts ? a.pts[sg] : a.wv[a.wv_perm ? a.perm[sg] : sg];
            }
            (tid < kVPts ? sPt : sWv)[slot * kVPts + j] = v;
        }
    };
    auto radial = [&](int st) {
        const double4* pt = sPt + (st % kVSlots) * kVPts;
        for (int idx = tid; idx < nsh * kVPts; idx += NT) {
            const int sh = idx / kVPts, p = idx % kVPts;
            const double4 x = pt[p];
            sRad[idx] = st * kVPts + p < npts ? shell_radial_tab(sSh[sh], sPrim, x.x, x.y, x.z, sExp)
                                              : make_double2(0.0, 0.0);
        }
    };
    auto evaluate = [&](int st) {
        const double4* pt = sPt + (st % kVSlots) * kVPts;
        const double4* wv = sWv + (st % kVSlots) * kVPts;
        // a thread takes one row at four points (p0, p0 + 4, p0 + 8, p0 + 12):
        // the row's record, its decoding and its shell centre serve all four
        for (int idx = tid; idx < mp * 4; idx += NT) {
            const int r = idx >> 2, p0 = idx & 3;
            // padding rows (shell -1, scale 0) read shell 0 and come out zero
            const RowL row = sRow[r];
            const int sh = max(row.sh, 0);
            const DevShell shl = sSh[sh];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int p = p0 + 4 * k;
                const double4 x = pt[p];
                const double4 w = wv[p];
                const AoVal f = ao_from_radial_row(shl, row, x.x, x.y, x.z, sRad[sh * kVPts + p]);
                const T v = f.v;
                const T z = w.x * f.v + w.y * f.gx + w.z * f.gy + w.w * f.gz;
                sP[r * LD + p] = v;
                sZ[r * LD + p] = st * kVPts + p < npts ? z : T(0);
            }
        }
    };
    // MFMA step k takes point 4 kq + k in lane (col, kq): A and B use the
    // same point order, so the sum over the 16 points is unchanged.
    auto update = [&]() {
        int cur_i = -1;
        V4 zi, pi;
#pragma unroll
        for (int w = 0; w < PPW; ++w) {
            const int p = wave * per_wave + w;
            if (w >= per_wave || p >= npairs) continue;
            const int2 pr = sPair[p];
            if (w == 0 || pr.x != cur_i) {
                cur_i = pr.x;
                const int ri = pr.x + col * LD + 4 * kq;
                zi = lds_load4(sZ + ri);
                pi = lds_load4(sP + ri);
            }
            const int rj = pr.y + col * LD + 4 * kq;
            const V4 pj = lds_load4(sP + rj);
            const V4 zj = lds_load4(sZ + rj);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                acc[w] = M::op(zi[k], pj[k], acc[w]);
                acc[w] = M::op(pi[k], zj[k], acc[w]);
            }
        }
    };

    load_points(0);
    if (nsteps > 1) load_points(1);
    __syncthreads();
    radial(0);
    __syncthreads();
    evaluate(0);
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
        update();
        if (st + 1 < nsteps) radial(st + 1);
        if (st + 2 < nsteps) load_points(st + 2);
        __syncthreads();
        if (st + 1 < nsteps) {
            evaluate(st + 1);
            __syncthreads();
        }
    }
#pragma unroll
    for (int w = 0; w < PPW; ++w) {
        const int p = wave * per_wave + w;
        if (w >= per_wave || p >= npairs) continue;
        double* out = a.scratch + G.scratch_off + static_cast<long long>(p) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v) out[M::row(kq, v) * 16 + col] = static_cast<double>(acc[w][v]);
    }
}

// Pipelined form of k_vxc_group for the mixed build (FP32 AO values, 16
// waves): one barrier per step of 16 points.  Between two barriers the
// workgroup does four independent things: the MFMA updates of step s
// (operand buffer s % 2), the AO values of step s + 1 into the other buffer,
// the radial sums of step s + 2 (radial buffer s % 2) and the loads of the
// points and wv of step s + 3 (slot (s + 3) % 4 of four).  Each buffer is
// written in one interval and read in a later one, so the barrier between
// them is the only ordering needed; waves that finish their MFMA share early
// go on with AO evaluation, so MFMA and VALU work overlap on the SIMDs.
//
// Two operand forms (template parameter BF), chosen per launch.  FP32: P
// and Z stored as floats (row stride vgroup_ld), acc += Z_I P_J^T + P_I Z_J^T
// on v_mfma_f32_16x16x4f32.  BF16 three-term: the tile-pair update
// Z_I P_J^T + P_I Z_J^T of a step is one 16 x 16 x 32 product over
// k = 2 q + s (slot q of the step's 16 points, s = 0, 1): A[i][2q] = Z_I[i][.],
// A[i][2q + 1] = P_I[i][.], B[2q][j] = P_J[j][.], B[2q + 1][j] = Z_J[j][.] at
// the point of slot q; slot 4 c + e holds point c + 4 e (any order of the
// points is the same sum while A and B share it).
// P and Z are formed in FP32 and split in three BF16 terms each
// (split3_bf16), and the six products of terms whose indices sum to at most
// 2 run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation: FP32 operand
// resolution at a sixth of the matrix-core cycles of the FP32 16x16x4 form
// (6 x 16 against 8 x 32 cycles per pair and step).  LDS holds per buffer and
// term one dword (P, Z) per (row, point): 16 dwords per row, the 16-byte chunk
// of slots 4c .. 4c + 3 stored at chunk c ^ ((row / 4) % 4), so that the 16
// rows of an operand read fall on distinct bank groups.  A reads the same
// dwords with the halves swapped.  The BF16 form saves matrix-core time and
// costs the split (VALU) and 1.5x the operand LDS.  The 9-17-tile class
// always runs it (20 waters: 152 -> 127 us per build; 64 clusters: 6.1 -> 5.1
// ms per set); for the small-group class it is a plan option (the
// autotuner's choice): faster on the 20-water cluster (0.593 -> 0.538 ms per
// V stage), slower on the 64-cluster set (14.5 -> 19.5 ms), where the FP32
// form, the default, is kept.

// Point slots of the pipelined kernel: in the interval of step s it reads the
// points of steps s + 1 and s + 2 and loads those of step s + 3, so three
// slots suffice; wv is kept as the FP32 values the kernel uses.
constexpr int kVPipeSlots = 3;

template <int MAXT, bool BF>
__host__ __device__ constexpr size_t vgroup_pipe_lds_fixed() {
    return (sizeof(double4) + sizeof(float4)) * kVPipeSlots * kVPts +
           sizeof(int2) * ((vgroup_max_pairs<MAXT>() + 1) & ~1) +
           sizeof(RowL) * MAXT * 16 +
           (BF ? sizeof(unsigned) * 2 * 3 * MAXT * 16 * kVPts : sizeof(float) * 4 * MAXT * 16 * vgroup_ld<float>());
}

// Waves of the pipelined kernel: 16 for groups of up to 12 or 17 tiles, 8
// for the small-group class (up to 8 tiles), whose LDS lets two workgroups
// share a CU.
template <int MAXT>
__host__ __device__ constexpr int vgroup_pipe_waves() {
    return MAXT <= 8 ? 8 : 16;
}

template <int MAXT, bool BF>
__global__ __launch_bounds__(vgroup_pipe_waves<MAXT>() * 64) void k_vxc_group_pipe(VArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    using T = float;
    constexpr int NW = vgroup_pipe_waves<MAXT>(), NT = NW * 64, PPW = (vgroup_max_pairs<MAXT>() + NW - 1) / NW;
    static_assert(kVPts == 16, "one 16-dword row of (P, Z) per 16 points");
    constexpr int LD = vgroup_ld<T>();  // FP32 form: row stride of P and Z
    const VGroup G = a.groups[a.group_list[blockIdx.x]];
    const int m = G.m, nsh = G.n_sh, npts = G.n_pts;
    const int T16 = (m + 15) >> 4, mp = T16 << 4;
    const int npairs = T16 * (T16 + 1) / 2;
    const int nsteps = (npts + kVPts - 1) / kVPts;
    unsigned char* q = smem;
    double4* sPt = reinterpret_cast<double4*>(q);                 // [kVPipeSlots][kVPts] points
    float4* sWv = reinterpret_cast<float4*>(sPt + kVPipeSlots * kVPts);  // [kVPipeSlots][kVPts] wv
    q += (sizeof(double4) + sizeof(float4)) * kVPipeSlots * kVPts;
    int2* sPair = reinterpret_cast<int2*>(q);
    q += sizeof(int2) * ((vgroup_max_pairs<MAXT>() + 1) & ~1);
    RowL* sRow = reinterpret_cast<RowL*>(q);
    q += sizeof(RowL) * MAXT * 16;
    // BF16 form: buffer b, term t at sW + (3 b + t) S, [row][16 points];
    // FP32 form: buffer b: phi at sPZ + 2 b SF, Z at sPZ + (2 b + 1) SF
    unsigned* sW = reinterpret_cast<unsigned*>(q);
    T* sPZ = reinterpret_cast<T*>(q);
    constexpr int S = MAXT * 16 * kVPts, SF = MAXT * 16 * LD;
    q += BF ? sizeof(unsigned) * 6 * S : sizeof(T) * 4 * SF;
    float2* sRad = reinterpret_cast<float2*>(q);  // buffer b at sRad + b nsh kVPts
    q += sizeof(float2) * 2 * nsh * kVPts;
    DevShell* sSh = reinterpret_cast<DevShell*>(q);
    q += sizeof(DevShell) * nsh;
    double2* sPrim = reinterpret_cast<double2*>(q);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kq = lane >> 4;
    {
        // rows, then shells and primitives: two coalesced copies of the packed records
        const uint4* __restrict__ src = a.gmeta + G.meta;
        for (int i = tid; i < mp; i += NT) reinterpret_cast<uint4*>(sRow)[i] = src[i];
        const int nrest = nsh * static_cast<int>(sizeof(DevShell) / sizeof(uint4)) + G.n_prim;
        for (int i = tid; i < nrest; i += NT) reinterpret_cast<uint4*>(sSh)[i] = src[mp + i];
    }
    const int per_wave = (npairs + NW - 1) / NW;
    for (int p = tid; p < npairs; p += NT) {
        int ti = 0;
        while ((ti + 1) * (ti + 2) / 2 <= p) ++ti;
        const int tj = p - ti * (ti + 1) / 2;
        sPair[p] = BF ? make_int2(ti * 16 * kVPts, tj * 16 * kVPts) : make_int2(ti * 16 * LD, tj * 16 * LD);
    }

    // points and wv of step st into its slot (zero past the group's points)
    auto load_points = [&](int st) {
        if (tid < 2 * kVPts) {
            const int j = tid % kVPts, g = st * kVPts + j, slot = st % kVPipeSlots;
            double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
            if (g < npts) {
                const int sg = G.pt_begin + g;
                v = tid < kVPts ? a.pts[sg] : a.wv[a.wv_perm ? a.perm[sg] : sg];
            }
            if (tid < kVPts)
                sPt[slot * kVPts + j] = v;
            else
                sWv[slot * kVPts + j] = make_float4(static_cast<float>(v.x), static_cast<float>(v.y),
                                                    static_cast<float>(v.z), static_cast<float>(v.w));
        }
    };
    // (R0, A) per (shell, point) of step st, A = w0 R0 + R1 (w . d): the
    // points and wv of step st arrived two intervals earlier
    auto radial = [&](int st) {
        const double4* pt = sPt + (st % kVPipeSlots) * kVPts;
        const float4* wv = sWv + (st % kVPipeSlots) * kVPts;
        float2* rad = sRad + (st & 1) * nsh * kVPts;
        for (int idx = tid; idx < nsh * kVPts; idx += NT) {
            const int sh = idx / kVPts, p = idx % kVPts;
            const double4 x = pt[p];
            const float4 w = wv[p];
            rad[idx] = st * kVPts + p < npts ? shell_radial_za_f(sSh[sh], sPrim, x.x, x.y, x.z, w)
                                             : make_float2(0.0f, 0.0f);
        }
    };
    auto evaluate = [&](int st) {
        const double4* pt = sPt + (st % kVPipeSlots) * kVPts;
        const float4* wv = sWv + (st % kVPipeSlots) * kVPts;
        const float2* rad = sRad + (st & 1) * nsh * kVPts;
        if constexpr (!BF) {
            T* P = sPZ + 2 * (st & 1) * SF;
            T* Z = P + SF;
            // a thread takes one row at four points (p0, p0 + 4, p0 + 8, p0 + 12):
            // the row's record, its decoding and its shell centre serve all four
            for (int idx = tid; idx < mp * 4; idx += NT) {
                const int r = idx >> 2, p0 = idx & 3;
                // padding rows (shell -1, scale 0) read shell 0 and come out zero,
                // and so do points past the group's end (zero (R0, A))
                const RowL row = sRow[r];
                const int sh = max(row.sh, 0);
                const DevShell shl = sSh[sh];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int p = p0 + 4 * k;
                    const double4 x = pt[p];
                    const float2 pz = pz_from_radial_row_f(shl, row, x.x, x.y, x.z, rad[sh * kVPts + p], wv[p]);
                    P[r * LD + p] = pz.x;
                    Z[r * LD + p] = pz.y;
                }
            }
            return;
        }
        unsigned* W = sW + 3 * (st & 1) * S;
        // a thread takes one row at the four points c, c + 4, c + 8, c + 12
        // (slots 4 c .. 4 c + 3; the threads of a row read four consecutive
        // points at a time, free of LDS bank conflicts): the row's record, its
        // decoding and its shell centre serve all four, and each term's four
        // dwords go to LDS as one 16-byte store
        for (int idx = tid; idx < mp * 4; idx += NT) {
            const int r = idx >> 2, c = idx & 3;
            // padding rows (shell -1, scale 0) read shell 0 and come out zero,
            // and so do points past the group's end (zero (R0, A))
            const RowL row = sRow[r];
            const int sh = max(row.sh, 0);
            const DevShell shl = sSh[sh];
            v4u t[3];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int p = c + 4 * k;
                const double4 x = pt[p];
                const float4 w = wv[p];
                const float2 pz = pz_from_radial_row_f(shl, row, x.x, x.y, x.z, rad[sh * kVPts + p], w);
                const Bf16Split3 sp = split3_bf16(pz.x, pz.y);
#pragma unroll
                for (int j = 0; j < 3; ++j) t[j][k] = sp.t[j];
            }
            const int at = r * kVPts + 4 * (c ^ ((r >> 2) & 3));
#pragma unroll
            for (int j = 0; j < 3; ++j) *reinterpret_cast<v4u*>(W + j * S + at) = t[j];
        }
    };
    using M = Mma<T>;
    typename M::acc acc[PPW];
#pragma unroll
    for (int w = 0; w < PPW; ++w) acc[w] = M::zero();
    // this lane's 16-byte chunk of a tile row: slots 4 kq .. 4 kq + 3 of row col
    const int lane_at = col * kVPts + 4 * (kq ^ ((col >> 2) & 3));
    auto update = [&](int st) {
        if constexpr (!BF) {
            using V4 = typename Vec4<T>::type;
            const T* P = sPZ + 2 * (st & 1) * SF;
            const T* Z = P + SF;
            int cur_i = -1;
            V4 zi, pi;
#pragma unroll
            for (int w = 0; w < PPW; ++w) {
                const int p = wave * per_wave + w;
                if (w >= per_wave || p >= npairs) continue;
                const int2 pr = sPair[p];
                if (w == 0 || pr.x != cur_i) {
                    cur_i = pr.x;
                    const int ri = pr.x + col * LD + 4 * kq;
                    zi = lds_load4(Z + ri);
                    pi = lds_load4(P + ri);
                }
                const int rj = pr.y + col * LD + 4 * kq;
                const V4 pj = lds_load4(P + rj);
                const V4 zj = lds_load4(Z + rj);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    acc[w] = M::op(zi[k], pj[k], acc[w]);
                    acc[w] = M::op(pi[k], zj[k], acc[w]);
                }
            }
            return;
        }
        const unsigned* W = sW + 3 * (st & 1) * S;
        int cur_i = -1;
        v4u a[3];
#pragma unroll
        for (int w = 0; w < PPW; ++w) {
            const int p = wave * per_wave + w;
            if (w >= per_wave || p >= npairs) continue;
            const int2 pr = sPair[p];
            if (w == 0 || pr.x != cur_i) {
                cur_i = pr.x;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const v4u v = *reinterpret_cast<const v4u*>(W + j * S + pr.x + lane_at);
#pragma unroll
                    for (int e = 0; e < 4; ++e) a[j][e] = __builtin_amdgcn_alignbit(v[e], v[e], 16);  // (Z, P)
                }
            }
            v4u b[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) b[j] = *reinterpret_cast<const v4u*>(W + j * S + pr.y + lane_at);
            acc[w] = mfma_bf16(a[1], b[1], acc[w]);
            acc[w] = mfma_bf16(a[2], b[0], acc[w]);
            acc[w] = mfma_bf16(a[0], b[2], acc[w]);
            acc[w] = mfma_bf16(a[1], b[0], acc[w]);
            acc[w] = mfma_bf16(a[0], b[1], acc[w]);
            acc[w] = mfma_bf16(a[0], b[0], acc[w]);
        }
    };

    // prologue: points of steps 0-2, radial sums of steps 0-1, AO values of step 0
    for (int st = 0; st < kVPipeSlots && st < nsteps; ++st) load_points(st);
    __syncthreads();
    radial(0);
    if (nsteps > 1) radial(1);
    __syncthreads();
    evaluate(0);
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
        update(st);
        if (st + 1 < nsteps) evaluate(st + 1);
        if (st + 2 < nsteps) radial(st + 2);
        if (st + 3 < nsteps) load_points(st + 3);
        __syncthreads();
    }
#pragma unroll
    for (int w = 0; w < PPW; ++w) {
        const int p = wave * per_wave + w;
        if (w >= per_wave || p >= npairs) continue;
        double* out = a.scratch + G.scratch_off + static_cast<long long>(p) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v) out[M::row(kq, v) * 16 + col] = static_cast<double>(acc[w][v]);
    }
}

// One workgroup per AO row mu: V[mu][nu] (nu <= mu) = sum over the groups
// holding mu of row li(mu) of that group's lower tiles, scattered through the
// group's row list.  Wave w of the workgroup sums the w-th of W contiguous
// runs of the entries (ascending group order) into its own LDS copy of the
// row (a wave's LDS operations execute in order, so its lanes need no barrier
// between groups);
// the W copies are then added in wave order, the row written to V and
// mirrored (V is exactly symmetric).  Group rows are in ascending AO order, so
// nu <= mu is lj <= li.  Deterministic for a given plan.
__global__ __launch_bounds__(1024) void k_reduce_v(ReduceArgs a) {
    extern __shared__ double srow[];  // [W][local mu + 1]
    const int mu = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const MolInfo mol = a.mols[a.ao_mol[mu]];
    const int base = mol.ao_base, lmu = mu - base;
    const int nw = blockDim.x >> 6, len = lmu + 1;
    for (int j = tid; j < nw * len; j += blockDim.x) srow[j] = 0.0;
    __syncthreads();
    double* __restrict__ my = srow + w * len;
    // Wave w of segment seg takes run seg nw + w of the n_seg nw contiguous
    // runs of mu's entries, B at a time: the row indices and tile values of
    // the B are requested together (one memory round trip), then added in
    // entry order.
    const int seg = blockIdx.y, runs = a.n_seg * nw;
    const int g0 = a.ao_gbegin[mu], e1 = a.ao_gbegin[mu + 1];
    const int per = (e1 - g0 + runs - 1) / runs;
    const int eb = min(g0 + (seg * nw + w) * per, e1), ee = min(eb + per, e1);
    constexpr int B = 4;
    for (int e0 = eb; e0 < ee; e0 += B) {
        int li[B], ts[B], rows_at[B];
        long long tbase[B];
        int maxli = -1;
#pragma unroll
        for (int u = 0; u < B; ++u) {
            li[u] = -1;
            ts[u] = 0;
            rows_at[u] = 0;
            tbase[u] = 0;
            if (e0 + u < ee) {
                const RedEntry r = a.entries[e0 + u];
                li[u] = r.li_ts & 0xffff;
                ts[u] = r.li_ts >> 16;
                rows_at[u] = r.rows;
                tbase[u] = r.base;
                maxli = max(maxli, li[u]);
            }
        }
        for (int lj = lane; lj <= maxli; lj += 64) {
            int c[B];
            double v[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                c[u] = -1;
                v[u] = 0.0;
                if (lj <= li[u]) {
                    c[u] = a.group_ao[rows_at[u] + lj];
                    v[u] = a.scratch[tbase[u] + (static_cast<long long>(lj >> ts[u]) << (2 * ts[u])) +
                                     (lj & ((1 << ts[u]) - 1))];
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (c[u] >= 0) my[c[u] - base] += v[u];
        }
    }
    __syncthreads();
    const long long n = mol.n_ao;
    double* __restrict__ V = a.V + mol.p_off;
    double* __restrict__ part = a.part + (static_cast<long long>(seg) * a.n_ao + mu) * a.row_stride;
    for (int j = tid; j < len; j += blockDim.x) {
        double v = 0.0;
        for (int c = 0; c < nw; ++c) v += srow[c * len + j];
        if (a.n_seg > 1) {
            part[j] = v;
        } else {
            V[lmu * n + j] = v;
            V[j * n + lmu] = v;
        }
    }
}

// Second pass of a segmented reduction: row mu of V is the sum of its
// segments' partial rows in segment order, written and mirrored.
__global__ __launch_bounds__(256) void k_reduce_v_combine(ReduceArgs a) {
    const int mu = blockIdx.x;
    const MolInfo mol = a.mols[a.ao_mol[mu]];
    const int lmu = mu - mol.ao_base, len = lmu + 1;
    const long long n = mol.n_ao;
    double* __restrict__ V = a.V + mol.p_off;
    for (int j = threadIdx.x; j < len; j += blockDim.x) {
        double v = 0.0;
        for (int sg = 0; sg < a.n_seg; ++sg) v += a.part[(static_cast<long long>(sg) * a.n_ao + mu) * a.row_stride + j];
        V[lmu * n + j] = v;
        V[j * n + lmu] = v;
    }
}

// One workgroup per molecule: its functional-stage partials, summed in a
// fixed order.
__global__ __launch_bounds__(1024) void k_reduce_en(const double* __restrict__ e_all, const double* __restrict__ ne_all,
                                                    const MolInfo* __restrict__ mols, double* __restrict__ out) {
    __shared__ double se[1024], sn[1024];
    const MolInfo mol = mols[blockIdx.x];
    const double* __restrict__ e = e_all + mol.xc_wg_begin;
    const double* __restrict__ ne = ne_all + mol.xc_wg_begin;
    const int n = mol.n_xc_wg;
    double x[4] = {0.0, 0.0, 0.0, 0.0}, y[4] = {0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + 3 * 1024 < n; i += 4 * 1024) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u] += e[i + u * 1024];
            y[u] += ne[i + u * 1024];
        }
    }
    for (; i < n; i += 1024) {
        x[0] += e[i];
        y[0] += ne[i];
    }
    se[threadIdx.x] = (x[0] + x[1]) + (x[2] + x[3]);
    sn[threadIdx.x] = (y[0] + y[1]) + (y[2] + y[3]);
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            se[threadIdx.x] += se[threadIdx.x + s];
            sn[threadIdx.x] += sn[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = se[0];
        out[2 * blockIdx.x + 1] = sn[0];
    }
}

__global__ void k_to_float(const double* __restrict__ x, float* __restrict__ y, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = static_cast<float>(x[i]);
}

// One workgroup per block: its m_pad x m_pad density sub-block, four rows
// per wave at a time (wave w takes rows 4 w .. 4 w + 3, then 16 further, and
// so on; lanes run along the rows, so the stores are contiguous).  Each
// thread requests its four rows x four column chunks (up to 256 columns) of
// loads together, before any store: the loads read P through clamped
// indices (padding rows and columns read P[0][0]) and padding is zeroed
// after, so no load sits behind a branch.  The block's AO indices are staged
// in LDS first (dynamic LDS of 4 x the largest m_pad of the launch).
__global__ __launch_bounds__(kThreads) void k_gather_p(GatherArgs a) {
    extern __shared__ int sIdx[];
    const RhoBlock B = a.blocks[a.block_list ? a.block_list[blockIdx.x] : static_cast<int>(blockIdx.x)];
    const MolInfo mol = a.mols[B.mol];
    const double* __restrict__ P = a.P + mol.p_off;
    const int mp = B.m_pad, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int r = threadIdx.x; r < mp; r += kThreads) {
        const int ao = a.block_rows[B.ao_begin + r].x;
        sIdx[r] = ao >= 0 ? ao - mol.ao_base : -1;
    }
    __syncthreads();
    constexpr int R = 4, C = 4;
    for (int i0 = wave * R; i0 < mp; i0 += R * (kThreads / 64)) {  // mp is a multiple of 16
        int ri[R];
#pragma unroll
        for (int u = 0; u < R; ++u) ri[u] = sIdx[i0 + u];
        for (int j0 = lane; j0 < mp; j0 += 64 * C) {
            int rj[C];
#pragma unroll
            for (int c = 0; c < C; ++c) rj[c] = j0 + 64 * c < mp ? sIdx[j0 + 64 * c] : -1;
            double v[R][C];
#pragma unroll
            for (int u = 0; u < R; ++u)
#pragma unroll
                for (int c = 0; c < C; ++c)
                    v[u][c] = P[static_cast<long long>(max(ri[u], 0)) * mol.n_ao + max(rj[c], 0)];
#pragma unroll
            for (int u = 0; u < R; ++u)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int j = j0 + 64 * c;
                    if (j >= mp) continue;
                    const double x = (ri[u] >= 0 && rj[c] >= 0) ? v[u][c] : 0.0;
                    const long long o = B.psub + static_cast<long long>(i0 + u) * mp + j;
                    if (a.out64 && (!a.out64_leading_only || j < B.m64)) a.out64[o] = x;
                    if (a.out32) a.out32[o] = static_cast<float>(x);
                }
        }
    }
}

__global__ void k_fill_zero(double* __restrict__ x, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// Per (block, candidate AO): does the AO survive the reference's value rule
// at any point of the block, and its largest |phi| there.  Setup only.
__global__ void k_significance(SigArgs a) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n_entries) return;
    const int mu = a.cand_ao[i];
    const int2 rng = a.cand_pts[i];  // sorted point range of the entry's block
    const DevAo ao = a.aos[mu];
    unsigned char keep = 0;
    double mx = 0.0;
    for (int g = rng.x; g < rng.x + rng.y; ++g) {
        const double4 q = a.pts[g];
        const AoVal f = ao_eval(a.shells, a.prims, ao, q.x, q.y, q.z);
        keep |= (f.v != 0.0 || f.gx != 0.0 || f.gy != 0.0 || f.gz != 0.0);
        mx = fmax(mx, fabs(f.v));
    }
    a.flag[i] = keep;
    a.maxabs[i] = mx;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename TS, bool HYB, int NS, int AOC = kAocNone>
static hipError_t launch_rho_t(const RhoArgs& a, const RhoClass& c, hipStream_t s) {
    const size_t lds = rho_lds_bytes(HYB ? kRhoMixed : kRhoFp64, c, NS, AOC == kAocLoad);
    static int attr_set = 0;
    if (!attr_set) {
        int dev = 0, lim = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rho<TS, HYB, NS, AOC>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lim);
        if (e != hipSuccess) return e;
        attr_set = 1;
    }
    if (c.n_tiles <= 0) return hipSuccess;
    RhoArgs b = a;
    b.tiles = a.tiles + c.tile_begin;
    if (AOC != kAocNone) b.aoc_off = a.aoc_off + c.tile_begin;
    int longest = 0;
    for (int x = 0; x < kXcds; ++x) {
        b.seg[x] = c.seg[x];
        longest = std::max(longest, c.seg[x + 1] - c.seg[x]);
    }
    b.seg[kXcds] = c.seg[kXcds];
    hipLaunchKernelGGL((k_rho<TS, HYB, NS, AOC>), dim3(kXcds * longest), dim3(kThreads), lds, s, b);
    return hipGetLastError();
}

// LDS of k_rho_aoc2<NS>: the sums, then the B operands of both sub-tiles.
static size_t rho_aoc2_lds_bytes(const RhoClass& c, int ns) {
    return sizeof(double) * 16 * (size_t)ns * 2 * kRhoPts +
           2 * kRhoPts * (sizeof(double) * (size_t)c.max_m64 + sizeof(float) * (size_t)c.max_mpad);
}

// The mixed rho stage from the AO cache: k_rho_aoc2<NS> where its LDS fits,
// else k_rho<float, true, NS, kAocLoad> (one sub-tile at a time, less LDS).
template <int NS>
static hipError_t launch_rho_aoc2(const RhoArgs& a, const RhoClass& c, hipStream_t s) {
    const size_t lds = rho_aoc2_lds_bytes(c, NS);
    static int lim = 0;
    if (lim == 0) {
        int dev = 0, l = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&l, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rho_aoc2<NS>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, l);
        if (e != hipSuccess) return e;
        lim = l;
    }
    if (lds > static_cast<size_t>(lim)) return launch_rho_t<float, true, NS, kAocLoad>(a, c, s);
    if (c.n_tiles <= 0) return hipSuccess;
    RhoArgs b = a;
    b.tiles = a.tiles + c.tile_begin;
    b.aoc_off = a.aoc_off + c.tile_begin;
    int longest = 0;
    for (int x = 0; x < kXcds; ++x) {
        b.seg[x] = c.seg[x];
        longest = std::max(longest, c.seg[x + 1] - c.seg[x]);
    }
    b.seg[kXcds] = c.seg[kXcds];
    hipLaunchKernelGGL(k_rho_aoc2<NS>, dim3(kXcds * longest), dim3(kThreads), lds, s, b);
    return hipGetLastError();
}

size_t rho_lds_bytes(int mode, const RhoClass& c, int ns, bool cached) {
    const size_t ts = mode == kRhoMixed ? sizeof(float) : sizeof(double);
    const size_t staged = cached ? 0
                                 : sizeof(double4) * kRhoSub * kRhoPts + sizeof(double) * 64 +
                                       sizeof(double2) * (size_t)c.max_nsh * kRhoPts +
                                       sizeof(DevShell) * (size_t)c.max_nsh + sizeof(double2) * (size_t)c.max_prim +
                                       sizeof(RowL) * (size_t)c.max_mpad;
    return sizeof(double) * 16 * (size_t)ns * kRhoPts + staged +
           (mode == kRhoMixed ? sizeof(double) * (size_t)c.max_m64 * kRhoPts : 0) +
           ts * (size_t)c.max_mpad * kRhoPts;
}

hipError_t launch_rho(const RhoArgs& a, int mode, const RhoClass& c, hipStream_t s, int ns) {
    if (mode == kRhoFp64)
        return ns == 2 ? launch_rho_t<double, false, 2>(a, c, s) : launch_rho_t<double, false, 1>(a, c, s);
    if (a.aoc)
        return ns == 2 ? launch_rho_aoc2<2>(a, c, s) : launch_rho_aoc2<1>(a, c, s);
    return ns == 2 ? launch_rho_t<float, true, 2>(a, c, s) : launch_rho_t<float, true, 1>(a, c, s);
}

hipError_t launch_rho_fill_cache(const RhoArgs& a, const RhoClass& c, hipStream_t s) {
    if (!a.aoc || !a.aoc_off) return hipErrorInvalidValue;
    return launch_rho_t<float, true, 1, kAocFill>(a, c, s);
}

// The kernel instance of functional a.func (an XcId); unknown ids are
// rejected by the caller before a build.
template <typename Args, typename Fn>
static hipError_t launch_by_functional(const Args& a, hipStream_t s, Fn pick) {
    if (a.n_wg <= 0) return hipSuccess;
    const void* k = pick(a.func);
    if (!k) return hipErrorInvalidValue;
    void* args[] = {const_cast<Args*>(&a)};
    return hipLaunchKernel(k, dim3(a.n_wg), dim3(kThreads), args, 0, s);
}

hipError_t launch_xc(const XcArgs& a, hipStream_t s) {
    return launch_by_functional(a, s, [](int f) -> const void* {
        switch (f) {
        case XC_SVWN5: return reinterpret_cast<const void*>(&k_xc<XC_SVWN5>);
        case XC_BLYP: return reinterpret_cast<const void*>(&k_xc<XC_BLYP>);
        case XC_PBE: return reinterpret_cast<const void*>(&k_xc<XC_PBE>);
        case XC_B3LYP: return reinterpret_cast<const void*>(&k_xc<XC_B3LYP>);
        case XC_B3LYP5: return reinterpret_cast<const void*>(&k_xc<XC_B3LYP5>);
        case XC_PBE0: return reinterpret_cast<const void*>(&k_xc<XC_PBE0>);
        default: return nullptr;
        }
    });
}

hipError_t launch_xc_uks(const XcUksArgs& a, hipStream_t s) {
    return launch_by_functional(a, s, [](int f) -> const void* {
        switch (f) {
        case XC_SVWN5: return reinterpret_cast<const void*>(&k_xc_uks<XC_SVWN5>);
        case XC_BLYP: return reinterpret_cast<const void*>(&k_xc_uks<XC_BLYP>);
        case XC_PBE: return reinterpret_cast<const void*>(&k_xc_uks<XC_PBE>);
        case XC_B3LYP: return reinterpret_cast<const void*>(&k_xc_uks<XC_B3LYP>);
        case XC_B3LYP5: return reinterpret_cast<const void*>(&k_xc_uks<XC_B3LYP5>);
        case XC_PBE0: return reinterpret_cast<const void*>(&k_xc_uks<XC_PBE0>);
        default: return nullptr;
        }
    });
}

hipError_t launch_vxc(const VArgs& a, bool fp32, int n_pairs, hipStream_t s) {
    if (n_pairs <= 0) return hipSuccess;
    if (fp32)
        hipLaunchKernelGGL(k_vxc<float>, dim3(n_pairs), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_vxc<double>, dim3(n_pairs), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

template <typename T, int MAXT>
static size_t vgroup_lds_bytes_t(int max_nsh, int max_prim, bool bf16) {
    if constexpr (sizeof(T) == 4) {  // k_vxc_group_pipe
        return (bf16 ? vgroup_pipe_lds_fixed<MAXT, true>() : vgroup_pipe_lds_fixed<MAXT, false>()) + sizeof(float2) * 2 * (size_t)max_nsh * kVPts +
               sizeof(DevShell) * (size_t)max_nsh + sizeof(double2) * (size_t)max_prim;
    } else {
        const size_t rad = sizeof(typename VGroupCfg<T, MAXT>::Rad) * (size_t)((max_nsh * kVPts + 1) & ~1);
        return vgroup_lds_fixed<T, MAXT>() + rad + sizeof(DevShell) * (size_t)max_nsh +
               sizeof(double2) * (size_t)max_prim;
    }
}

size_t vgroup_lds_bytes(bool fp32, int max_tiles, int max_nsh, int max_prim, bool bf16) {
    if (max_tiles == 8)
        return fp32 ? vgroup_lds_bytes_t<float, 8>(max_nsh, max_prim, bf16)
                    : vgroup_lds_bytes_t<double, 8>(max_nsh, max_prim, false);
    if (max_tiles == 12)
        return fp32 ? vgroup_lds_bytes_t<float, 12>(max_nsh, max_prim, bf16)
                    : vgroup_lds_bytes_t<double, 12>(max_nsh, max_prim, false);
    return fp32 ? vgroup_lds_bytes_t<float, 17>(max_nsh, max_prim, bf16)
                : vgroup_lds_bytes_t<double, 17>(max_nsh, max_prim, false);
}

template <typename T, int MAXT, bool BF = false>
static hipError_t launch_vxc_group_t(const VArgs& a, int n_groups, int max_nsh, int max_prim, hipStream_t s) {
    static int attr_set = 0;
    if (!attr_set) {
        int dev = 0, lim = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        const void* fn;
        if constexpr (sizeof(T) == 4)
            fn = reinterpret_cast<const void*>(&k_vxc_group_pipe<MAXT, BF>);
        else
            fn = reinterpret_cast<const void*>(&k_vxc_group<T, MAXT>);
        if (e == hipSuccess) e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
        if (e != hipSuccess) return e;
        attr_set = 1;
    }
    if (n_groups <= 0) return hipSuccess;
    const size_t lds = vgroup_lds_bytes_t<T, MAXT>(max_nsh, max_prim, BF);
    if constexpr (sizeof(T) == 4) {
        hipLaunchKernelGGL((k_vxc_group_pipe<MAXT, BF>), dim3(n_groups), dim3(vgroup_pipe_waves<MAXT>() * 64), lds, s,
                           a);
    } else {
        constexpr int threads = VGroupCfg<T, MAXT>::kWaves * 64;
        hipLaunchKernelGGL((k_vxc_group<T, MAXT>), dim3(n_groups), dim3(threads), lds, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_vxc_group(const VArgs& a, bool fp32, int max_tiles, int n_groups, int max_nsh, int max_prim,
                            hipStream_t s, bool bf16) {
    if (max_tiles == 8) {
        if (!fp32) return launch_vxc_group_t<double, 8>(a, n_groups, max_nsh, max_prim, s);
        return bf16 ? launch_vxc_group_t<float, 8, true>(a, n_groups, max_nsh, max_prim, s)
                    : launch_vxc_group_t<float, 8, false>(a, n_groups, max_nsh, max_prim, s);
    }
    if (max_tiles == 12)
        return fp32 ? launch_vxc_group_t<float, 12, true>(a, n_groups, max_nsh, max_prim, s)
                    : launch_vxc_group_t<double, 12>(a, n_groups, max_nsh, max_prim, s);
    return fp32 ? launch_vxc_group_t<float, 17, true>(a, n_groups, max_nsh, max_prim, s)
                : launch_vxc_group_t<double, 17>(a, n_groups, max_nsh, max_prim, s);
}

// Waves (row copies) per reduction workgroup that fit in LDS: up to 16.
int reduce_v_waves(int max_mol_ao, size_t lds_limit) {
    const size_t row = sizeof(double) * (size_t)std::max(max_mol_ao, 1);
    return (int)std::min<size_t>(16, lds_limit / row);
}

hipError_t launch_reduce_v(const ReduceArgs& a, int waves, int max_mol_ao, hipStream_t s) {
    if (a.n_ao <= 0) return hipSuccess;
    static int attr_set = 0;
    if (!attr_set) {
        int dev = 0, lim = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_reduce_v),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lim);
        if (e != hipSuccess) return e;
        attr_set = 1;
    }
    // Row mu needs (local mu + 1) doubles per wave; size for the largest molecule.
    hipLaunchKernelGGL(k_reduce_v, dim3(a.n_ao, std::max(a.n_seg, 1)), dim3(64 * waves),
                       sizeof(double) * (size_t)max_mol_ao * waves, s, a);
    if (a.n_seg > 1) hipLaunchKernelGGL(k_reduce_v_combine, dim3(a.n_ao), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_en(const double* e, const double* ne, const MolInfo* mols, int n_mol, double* out,
                            hipStream_t s) {
    if (n_mol <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_en, dim3(n_mol), dim3(1024), 0, s, e, ne, mols, out);
    return hipGetLastError();
}

hipError_t launch_to_float(const double* x, float* y, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_fill_zero(double* x, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n);
    return hipGetLastError();
}

hipError_t launch_gather_p(const GatherArgs& a, int n_blocks, int max_mpad, hipStream_t s) {
    if (n_blocks <= 0) return hipSuccess;
    if (max_mpad > 16384) return hipErrorInvalidValue;  // 64 KiB of row indices
    hipLaunchKernelGGL(k_gather_p, dim3(n_blocks), dim3(kThreads), sizeof(int) * (size_t)std::max(max_mpad, 1), s, a);
    return hipGetLastError();
}

hipError_t launch_significance(const SigArgs& a, hipStream_t s) {
    if (a.n_entries <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_significance, dim3((unsigned)((a.n_entries + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace chx
