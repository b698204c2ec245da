// grad_kernels.hip — gfx950 kernels of the XC nuclear gradient dE_xc/dR_A.
//
// Restates crates/calchem-dft/src/vxc.rs:1769-1899 (accumulate_xc_grad_points)
// and crates/calchem-dft/src/grid.rs:423-584 (becke_weight_nuclear_gradient)
// for spin densities Pa, Pb at fixed density matrix:
//   k_xc_grad     per point: functional -> (w v_rho_s, w u_s) per spin and f_xc.
//   k_grad_ao     per (rho block, 16 points): S = P_s,sub Phi and
//                 T = P_s,sub Y with Y_j = u_s . grad phi_j on FP64 MFMA, then
//                 F_iq = -2 [(w v_rho_s d_q phi_i + w u_s . H_i,q) S_i + d_q phi_i T_i]
//                 onto the atom of AO i, and minus its sum over AOs onto the
//                 host atom of the point (the point moves with its host).
//   k_grad_becke  one wavefront per run of points, lanes over atoms: the Becke
//                 weight derivative raw_w f_xc d(P_host / Z)/dR_A.
//   k_grad_sum    deterministic two-level sum of the per-workgroup partials.
// AO values here follow eval_aos_hess (ao_grid.rs:234-299): no per-AO drop rule.
#include <algorithm>

#include "ao_device.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "xc_device.hpp"

namespace chx {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Functional stage of the gradient, one thread per sorted point.  Restricted
// (rho_b null): the spin densities are half the total.  Writes (w v_rho_s,
// w u_s) for both spins and f_xc (zero where the density is screened).
__global__ __launch_bounds__(kThreads) void k_xc_grad(XcGradArgs a) {
    const int g = blockIdx.x * kThreads + threadIdx.x;
    if (g >= a.n_pts) return;
    double4 ra = a.rho_a[g], rb;
    if (a.rho_b) {
        rb = a.rho_b[g];
    } else {
        ra = make_double4(0.5 * ra.x, 0.5 * ra.y, 0.5 * ra.z, 0.5 * ra.w);
        rb = ra;
    }
    const double w = a.pts[g].w;
    double4 oa = make_double4(0.0, 0.0, 0.0, 0.0), ob = oa;
    double f = 0.0;
    const double pa = ra.x > 0.0 ? ra.x : 0.0, pb = rb.x > 0.0 ? rb.x : 0.0;
    if (pa + pb >= 1e-18) {  // DENSITY_SCREEN, vxc.rs:23
        const double gaa = ra.y * ra.y + ra.z * ra.z + ra.w * ra.w;
        const double gab = ra.y * rb.y + ra.z * rb.z + ra.w * rb.w;
        const double gbb = rb.y * rb.y + rb.z * rb.z + rb.w * rb.w;
        const XcOut o = eval_xc(a.func, pa, pb, gaa, gab, gbb);
        f = o.f;
        if (fabs(w) > 1e-24) {
            oa = make_double4(w * o.va, w * (2.0 * o.gaa * ra.y + o.gab * rb.y), w * (2.0 * o.gaa * ra.z + o.gab * rb.z),
                              w * (2.0 * o.gaa * ra.w + o.gab * rb.w));
            ob = make_double4(w * o.vb, w * (2.0 * o.gbb * rb.y + o.gab * ra.y), w * (2.0 * o.gbb * rb.z + o.gab * ra.z),
                              w * (2.0 * o.gbb * rb.w + o.gab * ra.w));
        }
    }
    a.wx_a[g] = oa;
    a.wx_b[g] = ob;
    a.fxc[g] = f;
}

// AO-force stage for one spin (k_rho's tiling: one workgroup per 16 points of
// a block).  LDS: points, (w v, w u) per point, host per point, radial
// (R0, R1, R2) per (shell, point), shells, primitives, rows, shell atoms, Phi
// and Y [row][point], per-row force sums, the exponential's table, per-wave
// per-point force sums.  The staging pass forms phi and u . grad phi from
// values and gradients alone; the epilogue evaluates the Hessians of the
// wave's rows.
__global__ __launch_bounds__(kThreads) void k_grad_ao(GradAoArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int PTS = kRhoPts;
    const int2 tile = a.tiles[blockIdx.x];
    const RhoBlock blk = a.blocks[tile.x];
    const int m = blk.m, mp = blk.m_pad, nsh = blk.n_sh, T16 = mp >> 4;
    const int t0 = tile.y, np = min(PTS, blk.n_pts - t0);
    unsigned char* q = smem;
    double4* sPt = reinterpret_cast<double4*>(q);
    q += sizeof(double4) * PTS;
    double4* sW = reinterpret_cast<double4*>(q);
    q += sizeof(double4) * PTS;
    double* sPtF = reinterpret_cast<double*>(q);  // [4 waves][PTS][3]
    q += sizeof(double) * 4 * PTS * 3;
    int* sHost = reinterpret_cast<int*>(q);
    q += sizeof(int) * PTS;
    Rad3* sRad = reinterpret_cast<Rad3*>(q);  // [nsh][PTS]
    q += sizeof(Rad3) * ((nsh * PTS + 1) & ~1);
    DevShell* sSh = reinterpret_cast<DevShell*>(q);
    q += sizeof(DevShell) * nsh;
    double2* sPrim = reinterpret_cast<double2*>(q);
    q += sizeof(double2) * blk.n_prim;
    RowL* sRow = reinterpret_cast<RowL*>(q);
    q += sizeof(RowL) * mp;
    int* sShAtom = reinterpret_cast<int*>(q);
    q += sizeof(int) * ((nsh + 3) & ~3);
    double* sPhi = reinterpret_cast<double*>(q);  // [mp][PTS]
    q += sizeof(double) * mp * PTS;
    double* sY = reinterpret_cast<double*>(q);  // [mp][PTS]
    q += sizeof(double) * mp * PTS;
    double* sRowF = reinterpret_cast<double*>(q);  // [mp][3]
    double* sExp = sRowF + 3 * mp;                 // [64] the exponential's table

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, kq = lane >> 4;
    if (tid < 64) sExp[tid] = kExpNegTab[tid];
    stage_meta(a.block_rows + blk.ao_begin, m, mp, a.block_sh + blk.sh_begin, nsh, a.shells, a.prims, a.aos, sRow,
               sSh, sPrim, tid, kThreads);
    for (int s = tid; s < nsh; s += kThreads) sShAtom[s] = a.shell_atom[a.block_sh[blk.sh_begin + s].x];
    if (tid < PTS) {
        const bool live = tid < np;
        const int g = blk.pt_begin + t0 + tid;
        sPt[tid] = live ? a.pts[g] : make_double4(0.0, 0.0, 0.0, 0.0);
        sW[tid] = live ? a.wx[g] : make_double4(0.0, 0.0, 0.0, 0.0);
        sHost[tid] = live ? a.host[g] : -1;
    }
    for (int i = tid; i < mp * 3; i += kThreads) sRowF[i] = 0.0;
    __syncthreads();

    for (int idx = tid; idx < nsh * PTS; idx += kThreads) {
        const int sh = idx / PTS, p = idx % PTS;
        Rad3 rr{0.0, 0.0, 0.0};
        if (p < np) {
            const double4 x = sPt[p];
            rr = shell_radial3(sSh[sh], sPrim, x.x, x.y, x.z, sExp);
        }
        sRad[idx] = rr;
    }
    __syncthreads();
    for (int idx = tid; idx < mp * PTS; idx += kThreads) {
        const int r = idx / PTS, p = idx % PTS;
        const RowL row = sRow[r];
        double v = 0.0, y = 0.0;
        if (row.sh >= 0 && p < np) {
            const double4 x = sPt[p];
            const AoVal f = ao_grad_row(sSh[row.sh], row, x.x, x.y, x.z, sRad[row.sh * PTS + p]);
            const double4 w = sW[p];
            v = f.v;
            y = w.y * f.gx + w.z * f.gy + w.w * f.gz;
        }
        sPhi[r * PTS + p] = v;
        sY[r * PTS + p] = y;
    }
    __syncthreads();

    double pf[3] = {0.0, 0.0, 0.0};  // this lane's point (col): sum over its rows
    for (int it = wave; it < T16; it += 4) {
        const int i0 = it << 4;
        v4d accS = {0.0, 0.0, 0.0, 0.0}, accT = accS;
        const double* __restrict__ Arow = a.PS + blk.psub + static_cast<long long>(i0 + col) * mp;
        // four k-steps per round, their density loads issued together (m_pad is a multiple of 16)
        for (int k = 0; k < mp; k += 16) {
            double av[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) av[s] = Arow[k + 4 * s + kq];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int kk = k + 4 * s + kq;
                accS = mfma64(av[s], sPhi[kk * PTS + col], accS);
                accT = mfma64(av[s], sY[kk * PTS + col], accT);
            }
        }
        // Element v of this lane: row i0 + 4 v + kq (the f64 result layout), point col.
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = i0 + 4 * v + kq;
            const RowL row = sRow[r];
            double fr[3] = {0.0, 0.0, 0.0};
            if (row.sh >= 0 && col < np) {
                const double4 x = sPt[col];
                const AoHess f = ao_hess_row(sSh[row.sh], row, x.x, x.y, x.z, sRad[row.sh * PTS + col]);
                const double4 w = sW[col];
                const double S = accS[v], T = accT[v];
                const int hq[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double hu = w.y * f.h[hq[c][0]] + w.z * f.h[hq[c][1]] + w.w * f.h[hq[c][2]];
                    fr[c] = a.factor * ((w.x * f.g[c] + hu) * S + f.g[c] * T);
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pf[c] += fr[c];
                // sum over the 16 points of the tile (lanes of the same kq)
                double t = fr[c];
                t += __shfl_xor(t, 1);
                t += __shfl_xor(t, 2);
                t += __shfl_xor(t, 4);
                t += __shfl_xor(t, 8);
                if (col == 0) sRowF[r * 3 + c] = t;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double t = pf[c];
        t += __shfl_xor(t, 16);
        t += __shfl_xor(t, 32);
        if (kq == 0) sPtF[(wave * PTS + col) * 3 + c] = t;
    }
    __syncthreads();
    // This tile's partial gradient: rows onto their atoms (row order), minus
    // each point's total onto its host (point order).
    for (int t = tid; t < 3 * a.n_atoms; t += kThreads) {
        const int atom = t / 3, c = t % 3;
        double s = 0.0;
        for (int r = 0; r < m; ++r)
            if (sShAtom[sRow[r].sh] == atom) s += sRowF[r * 3 + c];
        for (int p = 0; p < np; ++p)
            if (sHost[p] == atom)
                s -= sPtF[(0 * PTS + p) * 3 + c] + sPtF[(1 * PTS + p) * 3 + c] + sPtF[(2 * PTS + p) * 3 + c] +
                     sPtF[(3 * PTS + p) * 3 + c];
        a.partial[static_cast<long long>(blockIdx.x) * 3 * a.n_atoms + t] = s;
    }
}

// Becke weight derivative (becke_weight_nuclear_gradient, grid.rs:423-584).
// One wavefront per run of consecutive points; lane a handles atoms
// a, a + 64, ...; per-atom sums stay in registers over the run and go to
// this wave's partial row at the end.  The host atom gets minus the sum over
// the other atoms' terms of each point.
//
// Per point, after the cell functions: an atom c is active when its largest
// leave-one-out product (the product of all its factors s_cd but the
// smallest) exceeds kBeckeScreen Z, Z the sum of the cell functions.  A pair
// (a, b) enters the derivative through (P_a / s_ab - P_b / s_ba) ds_ab/dR_a,
// and each leave-one-out product of an inactive atom is at most
// kBeckeScreen Z: what is dropped per pair and point is at most
// 2 kBeckeScreen raw |f_xc| |ds_ab/dR_a| <= 2e-20 x 6.75 / R_ab raw |f_xc|
// (|ds/dnu| <= 27/16, |1 - 2 a mu| <= 2).  Pairs of two inactive atoms are
// skipped; pairs with an active second atom b run in full with lanes over
// the first atom a; pairs of an active first atom and an inactive second
// one keep the P_a / s_ab term, with lanes over b, summed across the lanes
// in a fixed order.  About 10 of 60 atoms are active at a point of the
// 20-water grid.
constexpr int kBeckeAtomsPerLane = 4;  // up to 256 atoms (APL = 1, 2 or 4 atoms per lane)
constexpr double kBeckeScreen = 1e-20;

__device__ __forceinline__ void becke_f3_deriv(double nu, double& f3, double& ds) {
    const double x = fmin(fmax(nu, -1.0), 1.0);
    const double f1 = 0.5 * x * (3.0 - x * x), d1 = 1.5 * (1.0 - x * x);
    const double f2 = 0.5 * f1 * (3.0 - f1 * f1), d2 = 1.5 * (1.0 - f1 * f1);
    f3 = 0.5 * f2 * (3.0 - f2 * f2);
    const double d3 = 1.5 * (1.0 - f2 * f2);
    ds = -0.5 * d3 * d2 * d1;
}

__device__ __forceinline__ void becke_step_deriv(double nu, double& s, double& ds) {
    double f3;
    becke_f3_deriv(nu, f3, ds);
    s = 0.5 * (1.0 - f3);
}

// 1 / x to about an ulp: the hardware reciprocal and two Newton steps.
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// Points whose cell functions share one pass through the pair table (one
// read of inv_r and asym per pair for both; 20 waters: gradient 21.4-21.8 ->
// 20.9-21.2 ms, bitwise equal; four points per pass took 23.8 ms).
constexpr int kBeckePointsPerPass = 2;

size_t grad_becke_lds_bytes(int n_atoms) {
    return ((sizeof(double) * 5 + sizeof(int) * 2) * kBeckePointsPerPass + sizeof(int) * 2) * (size_t)n_atoms;
}

template <int APL>
__global__ __launch_bounds__(64) void k_grad_becke(GradBeckeArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = a.n_atoms, lane = threadIdx.x;
    constexpr int Q = kBeckePointsPerPass;
    double* sDistQ = reinterpret_cast<double*>(smem);  // [Q][n]
    double* sCellQ = sDistQ + Q * n;                  // [Q][n] P_c = prod_{d != c} s_cd
    double* sBigQ = sCellQ + Q * n;                   // [Q][n] product of the factors s_cd > 1e-12
    double* sTinyQ = sBigQ + Q * n;                   // [Q][n] product of the nonzero factors s_cd <= 1e-12
    double* sMaxQ = sTinyQ + Q * n;                   // [Q][n] largest leave-one-out product
    int* sNTinyQ = reinterpret_cast<int*>(sMaxQ + Q * n);  // [Q][n] how many factors are <= 1e-12
    int* sNZeroQ = sNTinyQ + Q * n;                         // [Q][n] how many factors are exactly 0
    int* sAct = sNZeroQ + Q * n;                            // [n] the active atoms, ascending
    int* sIsAct = sAct + n;                                 // [n] 1 for an active atom
    double acc[APL][3];
#pragma unroll
    for (int k = 0; k < APL; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    const long long g0 = static_cast<long long>(blockIdx.x) * a.pts_per_wave;
    const long long g1 = min(g0 + a.pts_per_wave, static_cast<long long>(a.n_pts));
    long long g = g0;
    while (true) {
        // the next Q points that carry a weight and a functional value
        // (uniform over the wave), in order
        long long gq[Q];
        int nq = 0;
        for (; g < g1 && nq < Q; ++g) {
            if (fabs(a.fxc[g]) < 1e-18 || a.raw_w[g] == 0.0) continue;
            gq[nq++] = g;
        }
        if (nq == 0) break;
        for (int q = 0; q < nq; ++q) {
            const double4 p = a.pts[gq[q]];
            for (int c = lane; c < n; c += 64) {
                const double dx = p.x - a.xyz[3 * c], dy = p.y - a.xyz[3 * c + 1], dz = p.z - a.xyz[3 * c + 2];
                sDistQ[q * n + c] = sqrt(dx * dx + dy * dy + dz * dz);
            }
        }
        __syncthreads();
        // cell functions of the pass's points, one read of the pair table for
        // all of them: P_c = prod_{d != c} s_cd, the same product split into
        // its factors above and at or below 1e-12 (for the products that
        // leave one factor out), and the product of all factors but the
        // smallest (the largest of those products)
        for (int c = lane; c < n; c += 64) {
            double pc[Q], big[Q], tiny[Q], smin[Q], pmax[Q];
            int nt[Q], nz[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                pc[q] = big[q] = tiny[q] = smin[q] = pmax[q] = 1.0;
                nt[q] = nz[q] = 0;
            }
            for (int d = 0; d < n; ++d) {
                if (d == c) continue;
                // inv_r is symmetric and asym antisymmetric: the transposed
                // reads [d][c] are contiguous over the lanes
                const double inv = a.inv_r[d * n + c];
                if (inv == 0.0) continue;
                const double asy = a.asym[d * n + c];
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    if (q >= nq) continue;
                    const double mu = fmin(fmax((sDistQ[q * n + c] - sDistQ[q * n + d]) * inv, -1.0), 1.0);
                    const double nu = mu - asy * (1.0 - mu * mu);
                    double sv, ds;
                    becke_step_deriv(nu, sv, ds);
                    pc[q] *= sv;
                    if (sv < smin[q]) {
                        pmax[q] *= smin[q];
                        smin[q] = sv;
                    } else {
                        pmax[q] *= sv;
                    }
                    if (sv > 1e-12) {
                        big[q] *= sv;
                    } else {
                        ++nt[q];
                        if (sv > 0.0)
                            tiny[q] *= sv;
                        else
                            ++nz[q];
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (q >= nq) continue;
                sCellQ[q * n + c] = pc[q];
                sBigQ[q * n + c] = big[q];
                sTinyQ[q * n + c] = tiny[q];
                sMaxQ[q * n + c] = pmax[q];
                sNTinyQ[q * n + c] = nt[q];
                sNZeroQ[q * n + c] = nz[q];
            }
        }
        __syncthreads();
        for (int q = 0; q < nq; ++q) {
        const long long gg = gq[q];
        const double fxc = a.fxc[gg], raw = a.raw_w[gg];
        const double4 p = a.pts[gg];
        const int host = a.host[gg];
        const double* sDist = sDistQ + q * n;
        const double* sCell = sCellQ + q * n;
        const double* sBig = sBigQ + q * n;
        const double* sTiny = sTinyQ + q * n;
        const double* sMax = sMaxQ + q * n;
        const int* sNTiny = sNTinyQ + q * n;
        const int* sNZero = sNZeroQ + q * n;
        double zs = 0.0;
        for (int c = 0; c < n; ++c) zs += sCell[c];
        if (zs > 1e-30) {
            // the active atoms, ascending (ballot prefix per 64 atoms)
            const double thr = kBeckeScreen * zs;
            int n_act = 0;
#pragma unroll
            for (int k = 0; k < APL; ++k) {
                const int c = lane + 64 * k;
                const bool act = c < n && sMax[c] > thr;
                const unsigned long long m = __ballot(act);
                if (act) sAct[n_act + __popcll(m & ((1ull << lane) - 1ull))] = c;
                if (c < n) sIsAct[c] = act ? 1 : 0;
                n_act += __popcll(m);
            }
            __syncthreads();
            const double wn = sCell[host] / zs, pref = raw * fxc / zs;
            // products over the other pairs of c: P / s when s is above 1e-12;
            // else from the split products, in which s is one of the factors at
            // or below 1e-12 (and one of the zeros if 0)
            auto except = [&](int c, double sv) {
                if (sv > 1e-12) return sCell[c] * recip(sv);
                if (sNTiny[c] == 1) return sBig[c];
                if (sv > 0.0) return sNZero[c] > 0 ? 0.0 : sBig[c] * (sTiny[c] / sv);
                return sNZero[c] > 1 ? 0.0 : sBig[c] * sTiny[c];
            };
            // the terms of pair (at, b) for atom at (ra = |r - R_at| >= 1e-14,
            // dra = d ra / dR_at) into dz (dZ / dR_at) and dp (dP_host / dR_at)
            auto pair = [&](int at, double ra, const double (&dra)[3], int b, double (&dz)[3], double (&dp)[3]) {
                const double inv = a.inv_r[b * n + at];
                if (inv == 0.0) return;
                const double mu = fmin(fmax((ra - sDist[b]) * inv, -1.0), 1.0);
                const double asy = -a.asym[b * n + at];
                const double nu = mu + asy * (1.0 - mu * mu);
                double f3, dsdnu;
                becke_f3_deriv(nu, f3, dsdnu);
                const double s_ab = 0.5 * (1.0 - f3);
                const double slope = dsdnu * (1.0 - 2.0 * asy * mu);
                if (fabs(slope) < 1e-30) return;
                const double rv[3] = {a.xyz[3 * at] - a.xyz[3 * b], a.xyz[3 * at + 1] - a.xyz[3 * b + 1],
                                      a.xyz[3 * at + 2] - a.xyz[3 * b + 2]};
                // s_ba as the cell loop formed it for atom b (so that it is one
                // of the factors counted there): its nu is exactly -nu, f3 is
                // odd, so s_ba = (1 + f3) / 2 bit for bit
                const double s_ba = 0.5 * (1.0 + f3);
                double p_ab, p_ba;
                if (s_ab > 1e-12 && s_ba > 1e-12) {
                    // both leave-one-out products by division, from one reciprocal
                    const double r = recip(s_ab * s_ba);
                    p_ab = sCell[at] * (s_ba * r);
                    p_ba = sCell[b] * (s_ab * r);
                } else {
                    p_ab = except(at, s_ab);
                    p_ba = except(b, s_ba);
                }
                // dsab_c = slope (dra_c - mu rv_c / R) / R, with t = slope / R and
                // u = mu / R formed once per pair
                const double t = slope * inv, u = mu * inv, dpp = p_ab - p_ba;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double dsab = t * fma(-u, rv[c], dra[c]);
                    dz[c] += dpp * dsab;
                    if (b == host) dp[c] -= p_ba * dsab;
                }
            };
            // the same for a second atom b that is inactive: its leave-one-out
            // product p_ba is at most kBeckeScreen Z, so only p_ab ds_ab/dR_at
            // is kept (and nothing for P_host when b is the host)
            auto pair_ab = [&](int at, double ra, const double (&dra)[3], int b, double (&dz)[3]) {
                const double inv = a.inv_r[b * n + at];
                if (inv == 0.0) return;
                const double mu = fmin(fmax((ra - sDist[b]) * inv, -1.0), 1.0);
                const double asy = -a.asym[b * n + at];
                const double nu = mu + asy * (1.0 - mu * mu);
                double f3, dsdnu;
                becke_f3_deriv(nu, f3, dsdnu);
                const double slope = dsdnu * (1.0 - 2.0 * asy * mu);
                if (fabs(slope) < 1e-30) return;
                const double p_ab = except(at, 0.5 * (1.0 - f3));
                const double t = slope * inv * p_ab, u = mu * inv;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    dz[c] += t * fma(-u, a.xyz[3 * at + c] - a.xyz[3 * b + c], dra[c]);
            };
            auto geometry = [&](int at, double& ra, double (&dra)[3]) {
                ra = sDist[at];
                dra[0] = -(p.x - a.xyz[3 * at]) / ra;
                dra[1] = -(p.y - a.xyz[3 * at + 1]) / ra;
                dra[2] = -(p.z - a.xyz[3 * at + 2]) / ra;
            };
            double dzs[APL][3], dph[APL][3], ra[APL], dra[APL][3];
            bool ok[APL];
#pragma unroll
            for (int k = 0; k < APL; ++k) {
                const int at = lane + 64 * k;
                ok[k] = at < n && at != host && sDist[at] >= 1e-14;
                dzs[k][0] = dzs[k][1] = dzs[k][2] = 0.0;
                dph[k][0] = dph[k][1] = dph[k][2] = 0.0;
                ra[k] = 1.0;
                dra[k][0] = dra[k][1] = dra[k][2] = 0.0;
                if (ok[k]) geometry(at, ra[k], dra[k]);
            }
            // pairs whose second atom is active: lanes over the first
            for (int j = 0; j < n_act; ++j) {
                const int b = sAct[j];
#pragma unroll
                for (int k = 0; k < APL; ++k)
                    if (ok[k] && b != lane + 64 * k) pair(lane + 64 * k, ra[k], dra[k], b, dzs[k], dph[k]);
            }
            // pairs of an active first atom and an inactive second one: lanes
            // over the second, summed across the lanes (xor order), then onto
            // the first atom's lane
            for (int j = 0; j < n_act; ++j) {
                const int at = __builtin_amdgcn_readfirstlane(sAct[j]);
                if (at == host || sDist[at] < 1e-14) continue;  // uniform
                double rat, drat[3];
                geometry(at, rat, drat);
                double pz[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int k = 0; k < APL; ++k) {
                    const int b = lane + 64 * k;
                    if (b < n && b != at && !sIsAct[b]) pair_ab(at, rat, drat, b, pz);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    for (int o = 32; o > 0; o >>= 1) pz[c] += __shfl_xor(pz[c], o);
#pragma unroll
                for (int k = 0; k < APL; ++k)
                    if (lane + 64 * k == at)
#pragma unroll
                        for (int c = 0; c < 3; ++c) dzs[k][c] += pz[c];
            }
            double hsum[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < APL; ++k) {
                if (!ok[k]) continue;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double dwf = pref * (dph[k][c] - wn * dzs[k][c]);
                    acc[k][c] += dwf;
                    hsum[c] += dwf;
                }
            }
            // the host atom gets minus the others' terms
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double t = hsum[c];
                for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
#pragma unroll
                for (int k = 0; k < APL; ++k)
                    if (lane + 64 * k == host) acc[k][c] -= t;
            }
        }
        __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < APL; ++k) {
        const int at = lane + 64 * k;
        if (at < n)
            for (int c = 0; c < 3; ++c) a.partial[static_cast<long long>(blockIdx.x) * 3 * n + 3 * at + c] = acc[k][c];
    }
}

// out[t] (+)= sum over rows r of part[r * width + t], rows in order; first
// pass: one workgroup per chunk of rows, second pass over the chunk sums.
__global__ __launch_bounds__(kThreads) void k_grad_sum(const double* __restrict__ part, long long rows, int width,
                                                     long long rows_per_chunk, double* __restrict__ out) {
    const long long r0 = static_cast<long long>(blockIdx.y) * rows_per_chunk;
    const long long r1 = min(r0 + rows_per_chunk, rows);
    for (int t = blockIdx.x * kThreads + threadIdx.x; t < width; t += gridDim.x * kThreads) {
        double s = 0.0;
        for (long long r = r0; r < r1; ++r) s += part[r * width + t];
        out[static_cast<long long>(blockIdx.y) * width + t] = s;
    }
}

size_t grad_ao_lds_bytes(int max_mpad, int max_nsh, int max_prim) {
    return sizeof(double4) * 2 * kRhoPts + sizeof(double) * 4 * kRhoPts * 3 + sizeof(int) * kRhoPts +
           sizeof(Rad3) * (size_t)((max_nsh * kRhoPts + 1) & ~1) + sizeof(DevShell) * (size_t)max_nsh +
           sizeof(double2) * (size_t)max_prim + sizeof(RowL) * (size_t)max_mpad +
           sizeof(int) * (size_t)((max_nsh + 3) & ~3) + sizeof(double) * 2 * (size_t)max_mpad * kRhoPts +
           sizeof(double) * 3 * (size_t)max_mpad + sizeof(double) * 64;
}

hipError_t launch_xc_grad(const XcGradArgs& a, hipStream_t s) {
    if (a.n_pts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xc_grad, dim3((a.n_pts + kThreads - 1) / kThreads), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_ao(const GradAoArgs& a, int n_tiles, size_t lds, hipStream_t s) {
    static int attr_set = 0;
    if (!attr_set) {
        int dev = 0, lim = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_grad_ao),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lim);
        if (e != hipSuccess) return e;
        attr_set = 1;
    }
    if (n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grad_ao, dim3(n_tiles), dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_becke(const GradBeckeArgs& a, int n_waves, hipStream_t s) {
    if (n_waves <= 0) return hipSuccess;
    const size_t lds = grad_becke_lds_bytes(a.n_atoms);
    if (a.n_atoms <= 64)
        hipLaunchKernelGGL(k_grad_becke<1>, dim3(n_waves), dim3(64), lds, s, a);
    else if (a.n_atoms <= 128)
        hipLaunchKernelGGL(k_grad_becke<2>, dim3(n_waves), dim3(64), lds, s, a);
    else
        hipLaunchKernelGGL(k_grad_becke<kBeckeAtomsPerLane>, dim3(n_waves), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_sum(const double* part, long long rows, int width, long long rows_per_chunk, double* out,
                           hipStream_t s) {
    const long long chunks = rows_per_chunk > 0 ? (rows + rows_per_chunk - 1) / rows_per_chunk : 0;
    if (chunks <= 0 || width <= 0) return hipSuccess;
    const unsigned gx = static_cast<unsigned>(std::min<long long>((width + kThreads - 1) / kThreads, 64));
    hipLaunchKernelGGL(k_grad_sum, dim3(gx, static_cast<unsigned>(chunks)), dim3(kThreads), 0, s, part, rows, width,
                       rows_per_chunk, out);
    return hipGetLastError();
}

}  // namespace chx
