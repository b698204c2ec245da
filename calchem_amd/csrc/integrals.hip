// integrals.hip — host runtime and C ABI of the Gaussian integral engine that
// the SCF row (crates/calchem-scf/src/rks.rs:131-615) needs beside the Vxc
// build: one-electron integrals S, T, V_ne (one_electron.rs:115-168), the
// electron-repulsion integrals and J / K.
//
// The repulsion integrals are used in Cholesky-decomposed form,
// (mu nu|lam sig) ~ sum_Q L^Q_{mu nu} L^Q_{lam sig}, built on the device by
// pivoted column generation (Beebe and Linderberg, Int. J. Quantum Chem. 12,
// 683 (1977); Koch, Sanchez de Meras and Pedersen, J. Chem. Phys. 118, 9481
// (2003)): the residual diagonal picks shell pairs, the engine computes their
// ERI columns (k_eri), the columns are orthogonalised against the vectors so
// far with one GEMM, and pivots are taken in the block until the residual
// falls below the threshold.  The vectors stay resident in HBM as
// L[Q][mu][nu]; per SCF iteration
//   J = sum_Q L^Q (L^Q . P)                   (two GEMVs)
//   K = sum_Q (L^Q C)(L^Q C)^T                (one GEMM and one SYRK)
// for P = C+ C+^T - C- C-^T given by its factors.  The reference contracts
// exact integrals directly (direct_fock.rs:1028); the decomposition error is
// bounded by the threshold (1e-10 by default).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

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
#include "int_device.hpp"

namespace chx {
namespace ints {

hipError_t launch_eri_bra0(int ket_class, bool diag, const EriArgs& a, hipStream_t s);
hipError_t launch_eri_bra1(int ket_class, bool diag, const EriArgs& a, hipStream_t s);
hipError_t launch_eri_bra2(int ket_class, bool diag, const EriArgs& a, hipStream_t s);
hipError_t launch_eri_bra3(int ket_class, bool diag, const EriArgs& a, hipStream_t s);
hipError_t launch_eri_bra4(int ket_class, bool diag, const EriArgs& a, hipStream_t s);
hipError_t launch_eri_bra5(int ket_class, bool diag, const EriArgs& a, hipStream_t s);

namespace {

hipError_t launch_eri(int bra_class, int ket_class, bool diag, const EriArgs& a, hipStream_t s) {
    switch (bra_class) {
    case 0: return launch_eri_bra0(ket_class, diag, a, s);
    case 1: return launch_eri_bra1(ket_class, diag, a, s);
    case 2: return launch_eri_bra2(ket_class, diag, a, s);
    case 3: return launch_eri_bra3(ket_class, diag, a, s);
    case 4: return launch_eri_bra4(ket_class, diag, a, s);
    case 5: return launch_eri_bra5(ket_class, diag, a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_one(int cls, const OneArgs& a, hipStream_t s) {
    static const int na[6] = {1, 3, 3, 6, 6, 6}, nb[6] = {1, 1, 3, 1, 3, 6};
    const long long n = (long long)a.n_list * na[cls] * nb[cls];
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 127) / 128);
    switch (cls) {
    case 0: hipLaunchKernelGGL((k_one_electron<0, 0>), dim3(blocks), dim3(128), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_one_electron<1, 0>), dim3(blocks), dim3(128), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_one_electron<1, 1>), dim3(blocks), dim3(128), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_one_electron<2, 0>), dim3(blocks), dim3(128), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_one_electron<2, 1>), dim3(blocks), dim3(128), 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_one_electron<2, 2>), dim3(blocks), dim3(128), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Small kernels of the decomposition and of J / K.
// ---------------------------------------------------------------------------
// out[i + t * n] = G[i + cols[t] * n]: the pivot columns of a block.
__global__ void k_gather_cols(const double* __restrict__ G, long long n, const int* __restrict__ cols, int m,
                              double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * m) return;
    out[i] = G[(i % n) + (long long)cols[i / n] * n];
}
// D[i] -= sum_t V[i + t n]^2 over the m new vectors.
__global__ void k_diag_update_block(double* __restrict__ D, const double* __restrict__ V, long long n, int m) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int t = 0; t < m; ++t) s += V[i + (long long)t * n] * V[i + (long long)t * n];
    D[i] -= s;
}
// out[j + k * ldo] = src[rows[j] + k * lds] for j < nrow, k < ncol (row gather).
__global__ void k_gather_rows(const double* __restrict__ src, long long lds, const int* __restrict__ rows, int nrow,
                              int ncol, double* __restrict__ out, long long ldo) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nrow * ncol) return;
    const int j = (int)(i % nrow), k = (int)(i / nrow);
    out[j + (long long)k * ldo] = src[rows[j] + (long long)k * lds];
}
// Lf[(Q n + mu) n + nu] = Lp[Q n_pair + pair(max, min)]: the packed vectors as full symmetric matrices.
__global__ void k_expand(const double* __restrict__ Lp, long long n_pair, int n, int r, double* __restrict__ Lf) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nn = (long long)n * n;
    if (i >= nn * r) return;
    const long long q = i / nn, rem = i % nn;
    const int mu = (int)(rem / n), nu = (int)(rem % n);
    const long long hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
    Lf[i] = Lp[q * n_pair + hi * (hi + 1) / 2 + lo];
}
__global__ void k_zero(double* __restrict__ x, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}
// Copy the lower triangle of a column-major n x n matrix to its upper triangle.
__global__ void k_symmetrize_lower(double* __restrict__ A, int n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * n) return;
    const int r = (int)(i % n), c = (int)(i / n);
    if (r < c) A[i] = A[(long long)r * n + c];
}

inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    ~Buf() { reset(); }
    void reset() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    bool alloc(size_t b) {
        reset();
        if (b == 0) b = 16;
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = b;
        return true;
    }
    bool ensure(size_t b) { return bytes >= b ? true : alloc(b); }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct PairHost {
    int a, b;  // shells, la >= lb
    int la, lb, ao_a, ao_b, na, nb;
};

struct IntPlan {
    int device = 0;
    int n_ao = 0, n_shells = 0, n_atoms = 0;
    long long n_pair = 0;  // packed AO pairs n (n + 1) / 2
    std::vector<PairHost> pairs;       // id = i (i + 1) / 2 + j for shells i >= j
    std::vector<ShellPairDev> pairs_dev_h;
    std::vector<int> class_list[kPairClasses];
    Buf pairs_dev, prims, beta, ao_scale, atoms, class_lists, diag;
    int class_off[kPairClasses] = {};
    // decomposition
    int rank = 0;
    double tau = 0.0;
    Buf Lf;       // [rank][n][n]
    Buf work_y, work_p, work_d, work_c, work_k;
    rocblas_handle blas = nullptr;
    hipStream_t stream = nullptr;
    std::mutex mu;
    ~IntPlan() {
        if (blas) rocblas_destroy_handle(blas);
        if (stream) hipStreamDestroy(stream);
    }
};

std::mutex g_int_mu;
std::unordered_map<uint64_t, std::shared_ptr<IntPlan>> g_int;

std::shared_ptr<IntPlan> find_int(uint64_t h) {
    std::lock_guard<std::mutex> lk(g_int_mu);
    auto it = g_int.find(h);
    return it == g_int.end() ? nullptr : it->second;
}

bool dev_ok() {
    const char* dis = std::getenv("CALCHEM_HIP_DISABLE");
    if (dis && dis[0] == '1') return false;
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

inline long long pidx(int mu, int nu) {
    const long long hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
    return hi * (hi + 1) / 2 + lo;
}

// ERI columns of the ket pairs `kets` (pair ids) for all packed bra pairs,
// written from column `col0` of out (leading dimension n_pair); returns the
// column count or a negative status.
int eri_columns(IntPlan& P, const std::vector<int>& kets, double* out, double schwarz, std::vector<int>* col_of) {
    std::vector<int> by_class[kPairClasses];
    for (int k : kets) by_class[pair_class(P.pairs[k].la, P.pairs[k].lb)].push_back(k);
    std::vector<int> ket_ids, ket_cols;
    int col = 0;
    if (col_of) col_of->assign(kets.size(), 0);
    std::vector<int> cls_begin(kPairClasses + 1, 0);
    for (int c = 0; c < kPairClasses; ++c) {
        cls_begin[c] = (int)ket_ids.size();
        for (int k : by_class[c]) {
            ket_ids.push_back(k);
            ket_cols.push_back(col);
            col += P.pairs[k].na * P.pairs[k].nb;
        }
    }
    cls_begin[kPairClasses] = (int)ket_ids.size();
    if (col_of)
        for (size_t i = 0; i < kets.size(); ++i)
            for (size_t j = 0; j < ket_ids.size(); ++j)
                if (ket_ids[j] == kets[i]) (*col_of)[i] = ket_cols[j];
    if (!P.work_c.ensure(sizeof(int) * 2 * (ket_ids.size() + 1))) return kAllocFail;
    int* dk = P.work_c.as<int>();
    if (hipMemcpyAsync(dk, ket_ids.data(), sizeof(int) * ket_ids.size(), hipMemcpyHostToDevice, P.stream) != hipSuccess ||
        hipMemcpyAsync(dk + ket_ids.size(), ket_cols.data(), sizeof(int) * ket_ids.size(), hipMemcpyHostToDevice,
                       P.stream) != hipSuccess)
        return kLaunchFail;
    for (int cb = 0; cb < kPairClasses; ++cb) {
        if (P.class_list[cb].empty()) continue;
        for (int ck = 0; ck < kPairClasses; ++ck) {
            const int nk = cls_begin[ck + 1] - cls_begin[ck];
            if (nk == 0) continue;
            EriArgs a;
            a.pairs = P.pairs_dev.as<ShellPairDev>();
            a.bra = P.class_lists.as<int>() + P.class_off[cb];
            a.n_bra = (int)P.class_list[cb].size();
            a.ket = dk + cls_begin[ck];
            a.ket_col = dk + ket_ids.size() + cls_begin[ck];
            a.n_ket = nk;
            a.prims = P.prims.as<PrimPair>();
            a.ao_scale = P.ao_scale.as<double>();
            a.out = out;
            a.ld = P.n_pair;
            a.schwarz = schwarz;
            if (launch_eri(cb, ck, false, a, P.stream) != hipSuccess) return kLaunchFail;
        }
    }
    // the device copies of the ket lists must outlive the launches
    if (hipStreamSynchronize(P.stream) != hipSuccess) return kLaunchFail;
    return col;
}

int compute_diag(IntPlan& P) {
    if (!P.diag.alloc(sizeof(double) * P.n_pair)) return kAllocFail;
    for (int c = 0; c < kPairClasses; ++c) {
        if (P.class_list[c].empty()) continue;
        EriArgs a;
        std::memset(&a, 0, sizeof a);
        a.pairs = P.pairs_dev.as<ShellPairDev>();
        a.bra = P.class_lists.as<int>() + P.class_off[c];
        a.n_bra = (int)P.class_list[c].size();
        a.prims = P.prims.as<PrimPair>();
        a.ao_scale = P.ao_scale.as<double>();
        a.out = P.diag.as<double>();
        a.ld = P.n_pair;
        if (launch_eri(c, c, true, a, P.stream) != hipSuccess) return kLaunchFail;
    }
    std::vector<double> d(P.n_pair);
    if (hipMemcpyAsync(d.data(), P.diag.p, sizeof(double) * P.n_pair, hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
        hipStreamSynchronize(P.stream) != hipSuccess)
        return kLaunchFail;
    // Schwarz factors per shell pair
    for (size_t k = 0; k < P.pairs.size(); ++k) {
        const PairHost& h = P.pairs[k];
        double m = 0.0;
        for (int i = 0; i < h.na; ++i)
            for (int j = 0; j < h.nb; ++j) m = std::max(m, d[pidx(h.ao_a + i, h.ao_b + j)]);
        P.pairs_dev_h[k].q = std::sqrt(std::max(m, 0.0));
    }
    if (hipMemcpyAsync(P.pairs_dev.p, P.pairs_dev_h.data(), sizeof(ShellPairDev) * P.pairs_dev_h.size(),
                       hipMemcpyHostToDevice, P.stream) != hipSuccess ||
        hipStreamSynchronize(P.stream) != hipSuccess)
        return kLaunchFail;
    return kOk;
}

// Quartets whose Schwarz bound q_ab q_cd falls below this are skipped (zeros).
constexpr double kSchwarz = 1e-15;

int build_cholesky(IntPlan& P, double tau, int max_rank, int* rank_out) {
    const long long np = P.n_pair;
    const int n = P.n_ao;
    int cap = max_rank > 0 ? max_rank : (int)std::min<long long>(np, 25LL * n);
    cap = (int)std::min<long long>(cap, np);
    Buf Lp, G, Dd, coef, rows, sub;
    if (!Lp.alloc(sizeof(double) * (size_t)np * cap)) return kAllocFail;
    if (!Dd.alloc(sizeof(double) * np)) return kAllocFail;
    if (hipMemcpyAsync(Dd.p, P.diag.p, sizeof(double) * np, hipMemcpyDeviceToDevice, P.stream) != hipSuccess)
        return kLaunchFail;
    std::vector<double> Dh(np);
    if (hipMemcpyAsync(Dh.data(), P.diag.p, sizeof(double) * np, hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
        hipStreamSynchronize(P.stream) != hipSuccess)
        return kLaunchFail;
    constexpr int kMaxCols = 512;
    if (!G.alloc(sizeof(double) * (size_t)np * kMaxCols) || !coef.alloc(sizeof(double) * kMaxCols * kMaxCols) ||
        !rows.alloc(sizeof(int) * kMaxCols) || !sub.alloc(sizeof(double) * (size_t)kMaxCols * std::max(cap, kMaxCols)))
        return kAllocFail;
    double* Lpd = Lp.as<double>();
    double* Gd = G.as<double>();
    int r = 0;
    const double sigma = 1e-2;
    std::vector<std::pair<double, int>> order;
    order.reserve(P.pairs.size());
    while (true) {
        double dmax = 0.0;
        for (long long i = 0; i < np; ++i) dmax = std::max(dmax, Dh[i]);
        if (dmax < tau || r >= cap) break;
        const double thr = std::max(tau, sigma * dmax);
        // candidate shell pairs: largest residual diagonal first, up to kMaxCols columns
        order.clear();
        for (size_t k = 0; k < P.pairs.size(); ++k) {
            const PairHost& h = P.pairs[k];
            double m = 0.0;
            for (int i = 0; i < h.na; ++i)
                for (int j = 0; j < h.nb; ++j) m = std::max(m, Dh[pidx(h.ao_a + i, h.ao_b + j)]);
            if (m >= thr) order.push_back({-m, (int)k});
        }
        std::sort(order.begin(), order.end());
        std::vector<int> kets;
        int ncol = 0;
        for (auto& o : order) {
            const int c = P.pairs[o.second].na * P.pairs[o.second].nb;
            if (ncol + c > kMaxCols && !kets.empty()) break;
            kets.push_back(o.second);
            ncol += c;
        }
        std::vector<int> col_of;
        const int got = eri_columns(P, kets, Gd, kSchwarz, &col_of);
        if (got < 0) return got;
        // the function pair of each column; columns repeating a pair of one shell are skipped
        std::vector<int> cfp(ncol, -1);
        for (size_t i = 0; i < kets.size(); ++i) {
            const PairHost& h = P.pairs[kets[i]];
            for (int a = 0; a < h.na; ++a)
                for (int b = 0; b < h.nb; ++b) {
                    if (h.a == h.b && a < b) continue;
                    cfp[col_of[i] + a * h.nb + b] = (int)pidx(h.ao_a + a, h.ao_b + b);
                }
        }
        std::vector<int> rows_h(ncol);
        for (int j = 0; j < ncol; ++j) rows_h[j] = cfp[j] >= 0 ? cfp[j] : 0;
        if (hipMemcpyAsync(rows.p, rows_h.data(), sizeof(int) * ncol, hipMemcpyHostToDevice, P.stream) != hipSuccess)
            return kLaunchFail;
        if (r > 0) {
            // G -= Lp[:, :r] * Lp[rows, :r]^T
            double* Lc = sub.as<double>();  // ncol x r, leading dimension ncol
            hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long long)ncol * r)), dim3(256), 0, P.stream, Lpd, np,
                               rows.as<int>(), ncol, r, Lc, (long long)ncol);
            const double m1 = -1.0, one = 1.0;
            if (rocblas_dgemm(P.blas, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)np, ncol, r, &m1,
                              Lpd, (rocblas_int)np, Lc, ncol, &one, Gd, (rocblas_int)np) != rocblas_status_success)
                return kLaunchFail;
        }
        // the block's residual submatrix on the host
        std::vector<double> Gs((size_t)ncol * ncol);
        {
            double* Sd = sub.as<double>();
            hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long long)ncol * ncol)), dim3(256), 0, P.stream, Gd, np,
                               rows.as<int>(), ncol, ncol, Sd, (long long)ncol);
            if (hipMemcpyAsync(Gs.data(), Sd, sizeof(double) * ncol * ncol, hipMemcpyDeviceToHost, P.stream) !=
                    hipSuccess ||
                hipStreamSynchronize(P.stream) != hipSuccess)
                return kLaunchFail;
        }
        // Gs[j * ncol + i] = residual (fp_i | col j).  Pivoted Cholesky of this
        // block on the host: pivots p_t and the factor's rows at the pivots,
        // Ls[t][s] = l_s[p_t]; the block's vectors are then V = G[:, p] Ls^{-T}.
        std::vector<int> piv;
        std::vector<double> lcols;  // ncol x m, column t = l_t
        std::vector<char> used(ncol, 0);
        std::vector<double> dres(ncol);  // residual diagonal of the block
        for (int j = 0; j < ncol; ++j) dres[j] = Gs[(size_t)j * ncol + j];
        while (r + (int)piv.size() < cap) {
            int jb = -1;
            double best = -1.0;
            for (int j = 0; j < ncol; ++j)
                if (!used[j] && cfp[j] >= 0 && dres[j] > best) {
                    best = dres[j];
                    jb = j;
                }
            if (jb < 0 || best < thr) break;
            const double inv = 1.0 / std::sqrt(best);
            const size_t t = piv.size();
            lcols.resize((t + 1) * ncol);
            double* lt = lcols.data() + t * ncol;
            for (int i = 0; i < ncol; ++i) {
                double v = Gs[(size_t)jb * ncol + i];
                for (size_t s2 = 0; s2 < t; ++s2) v -= lcols[s2 * ncol + i] * lcols[s2 * ncol + jb];
                lt[i] = v * inv;
            }
            for (int i = 0; i < ncol; ++i) dres[i] -= lt[i] * lt[i];
            used[jb] = 1;
            piv.push_back(jb);
        }
        const int m = (int)piv.size();
        if (m == 0) break;  // no candidate above the threshold: the residual is below it to rounding
        {
            std::vector<double> Ls((size_t)m * m, 0.0);  // column-major, lower
            std::vector<int> pcol(m);
            for (int t = 0; t < m; ++t) {
                pcol[t] = piv[t];
                for (int s2 = 0; s2 <= t; ++s2) Ls[t + (size_t)s2 * m] = lcols[(size_t)s2 * ncol + piv[t]];
            }
            double* Lsd = coef.as<double>();
            int* pcd = rows.as<int>();
            if (hipMemcpyAsync(Lsd, Ls.data(), sizeof(double) * m * m, hipMemcpyHostToDevice, P.stream) != hipSuccess ||
                hipMemcpyAsync(pcd, pcol.data(), sizeof(int) * m, hipMemcpyHostToDevice, P.stream) != hipSuccess)
                return kLaunchFail;
            double* V = Lpd + (size_t)r * np;
            hipLaunchKernelGGL(k_gather_cols, dim3(nblk(np * m)), dim3(256), 0, P.stream, Gd, np, pcd, m, V);
            const double one = 1.0;
            if (rocblas_dtrsm(P.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                              rocblas_diagonal_non_unit, (rocblas_int)np, m, &one, Lsd, m, V, (rocblas_int)np) !=
                rocblas_status_success)
                return kLaunchFail;
            hipLaunchKernelGGL(k_diag_update_block, dim3(nblk(np)), dim3(256), 0, P.stream, Dd.as<double>(), V, np, m);
            // the host buffers above must outlive the copies
            if (hipStreamSynchronize(P.stream) != hipSuccess) return kLaunchFail;
            r += m;
            for (int t = 0; t < m; ++t) used[piv[t]] = 1;
        }
        if (hipMemcpyAsync(Dh.data(), Dd.p, sizeof(double) * np, hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
            hipStreamSynchronize(P.stream) != hipSuccess)
            return kLaunchFail;
        for (int j = 0; j < ncol; ++j)
            if (used[j]) Dh[cfp[j]] = 0.0;  // pivots are exhausted
        if (hipMemcpyAsync(Dd.p, Dh.data(), sizeof(double) * np, hipMemcpyHostToDevice, P.stream) != hipSuccess)
            return kLaunchFail;
    }
    // the vectors as full symmetric matrices, L[Q][mu][nu]
    const long long nn = (long long)n * n;
    if (!P.Lf.alloc(sizeof(double) * (size_t)nn * std::max(r, 1))) return kAllocFail;
    if (r > 0)
        hipLaunchKernelGGL(k_expand, dim3(nblk(nn * r)), dim3(256), 0, P.stream, Lpd, np, n, r, P.Lf.as<double>());
    if (hipStreamSynchronize(P.stream) != hipSuccess) return kLaunchFail;
    P.rank = r;
    P.tau = tau;
    if (rank_out) *rank_out = r;
    return kOk;
}

// K += sign * sum_Q (L^Q C)(L^Q C)^T for C (n x k, row-major on the device:
// column-major k x n).  Y^Q = L^Q C for a block of Q by one GEMM; the sum over
// (Q, i) of Y Y^T has a tiny n x n output and a long reduction, so it runs as
// a strided-batched GEMM over chunks of Q (enough workgroups to fill the
// device) whose partial matrices are then added in chunk order.
__global__ void k_sum_partials(const double* __restrict__ part, int n_part, long long nn, double sign,
                               double* __restrict__ K, int accumulate) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double s = 0.0;
    for (int b = 0; b < n_part; ++b) s += part[(long long)b * nn + i];
    K[i] = accumulate ? K[i] + sign * s : sign * s;
}

int add_exchange(IntPlan& P, const double* dC, int k, double sign, double* dK, bool first) {
    const int n = P.n_ao, r = P.rank;
    const long long nn = (long long)n * n, per_q = (long long)n * k;
    const long long qb = std::max<long long>(1, std::min<long long>(r, (256LL << 20) / std::max<long long>(per_q, 1)));
    // Q per batch entry: about 4096 columns of Z each
    const int cq = (int)std::max<long long>(1, std::min<long long>(qb, 4096 / std::max(k, 1)));
    const int max_parts = (int)((qb + cq - 1) / cq);
    if (!P.work_y.ensure(sizeof(double) * (size_t)(per_q * qb)) ||
        !P.work_k.ensure(sizeof(double) * (size_t)(nn * max_parts)))
        return kAllocFail;
    double* Y = P.work_y.as<double>();
    double* part = P.work_k.as<double>();
    bool acc = !first;
    for (long long q0 = 0; q0 < r; q0 += qb) {
        const int nq = (int)std::min<long long>(qb, r - q0);
        const double one = 1.0, zero = 0.0;
        // Y (n nq x k, column-major) = M1^T C with M1 = L block viewed as n x (n nq)
        if (rocblas_dgemm(P.blas, rocblas_operation_transpose, rocblas_operation_transpose, n * nq, k, n, &one,
                          P.Lf.as<double>() + q0 * nn, n, dC, k, &zero, Y, n * nq) != rocblas_status_success)
            return kLaunchFail;
        // Z = Y viewed as n x (nq k); chunk b = columns [b cq k, (b + 1) cq k), the last one shorter
        const int full = nq / cq, rest = nq - full * cq;
        if (full > 0 &&
            rocblas_dgemm_strided_batched(P.blas, rocblas_operation_none, rocblas_operation_transpose, n, n, cq * k,
                                          &one, Y, n, (rocblas_stride)cq * k * n, Y, n, (rocblas_stride)cq * k * n,
                                          &zero, part, n, (rocblas_stride)nn, full) != rocblas_status_success)
            return kLaunchFail;
        if (rest > 0 &&
            rocblas_dgemm(P.blas, rocblas_operation_none, rocblas_operation_transpose, n, n, rest * k, &one,
                          Y + (long long)full * cq * k * n, n, Y + (long long)full * cq * k * n, n, &zero,
                          part + (long long)full * nn, n) != rocblas_status_success)
            return kLaunchFail;
        hipLaunchKernelGGL(k_sum_partials, dim3(nblk(nn)), dim3(256), 0, P.stream, part, full + (rest > 0 ? 1 : 0),
                           nn, sign, dK, acc ? 1 : 0);
        acc = true;
    }
    return kOk;
}

}  // namespace
}  // namespace ints
}  // namespace chx

using namespace chx;
using namespace chx::ints;

extern "C" {

int calchem_hip_int_create(const CalchemAoShell* shells, int n_shells, const CalchemAoPrim* prims, int n_prims,
                           const CalchemAoComp* comps, int n_comps, int n_ao, int n_atoms, const double* atoms_xyzz,
                           uint64_t handle) {
    if (!shells || !prims || !comps || n_shells <= 0 || n_prims <= 0 || n_comps <= 0 || n_ao <= 0 || n_atoms < 0 ||
        (n_atoms > 0 && !atoms_xyzz) || handle == 0)
        return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    {
        std::lock_guard<std::mutex> lk(g_int_mu);
        if (g_int.count(handle)) return kMismatch;
    }
    std::vector<int> sh_l(n_shells);
    std::vector<double> scale(n_ao, 0.0);
    int ao_count = 0;
    for (int s = 0; s < n_shells; ++s) {
        const CalchemAoShell& x = shells[s];
        int l = -1;
        for (int q = 0; q <= kMaxL; ++q)
            if (ncart(q) == x.n_funcs) l = q;
        if (l < 0) return kUnsupported;  // f shells and above: not in the bases of this engine
        if (x.prim_start < 0 || x.n_prims <= 0 || x.prim_start + x.n_prims > n_prims || x.comp_start < 0 ||
            x.comp_start + x.n_funcs > n_comps || x.ao_offset < 0 || x.ao_offset + x.n_funcs > n_ao)
            return kMismatch;
        for (int c = 0; c < x.n_funcs; ++c) {
            const CalchemAoComp& q = comps[x.comp_start + c];
            if (q.lx != cart_x(l, c) || q.ly != cart_y(l, c) || q.lz != cart_z(l, c)) return kMismatch;
            scale[x.ao_offset + c] = q.scale;
        }
        sh_l[s] = l;
        ao_count += x.n_funcs;
    }
    if (ao_count != n_ao) return kMismatch;
    auto plan = std::make_shared<IntPlan>();
    IntPlan& P = *plan;
    if (hipGetDevice(&P.device) != hipSuccess) return kNoDevice;
    P.n_ao = n_ao;
    P.n_shells = n_shells;
    P.n_atoms = n_atoms;
    P.n_pair = (long long)n_ao * (n_ao + 1) / 2;
    const size_t npairs = (size_t)n_shells * (n_shells + 1) / 2;
    P.pairs.resize(npairs);
    P.pairs_dev_h.resize(npairs);
    std::vector<PrimPair> pp;
    std::vector<double> beta;
    for (int i = 0; i < n_shells; ++i)
        for (int j = 0; j <= i; ++j) {
            const size_t id = (size_t)i * (i + 1) / 2 + j;
            int a = i, b = j;
            if (sh_l[a] < sh_l[b]) std::swap(a, b);
            const CalchemAoShell &A = shells[a], &B = shells[b];
            PairHost& h = P.pairs[id];
            h.a = a;
            h.b = b;
            h.la = sh_l[a];
            h.lb = sh_l[b];
            h.ao_a = A.ao_offset;
            h.ao_b = B.ao_offset;
            h.na = A.n_funcs;
            h.nb = B.n_funcs;
            ShellPairDev& d = P.pairs_dev_h[id];
            d.prim_begin = (int)pp.size();
            d.ao_a = h.ao_a;
            d.ao_b = h.ao_b;
            d.la = h.la;
            d.lb = h.lb;
            d.q = 0.0;
            double AB2 = 0.0;
            for (int q = 0; q < 3; ++q) AB2 += (A.center[q] - B.center[q]) * (A.center[q] - B.center[q]);
            for (int pa = 0; pa < A.n_prims; ++pa)
                for (int pb = 0; pb < B.n_prims; ++pb) {
                    const double al = prims[A.prim_start + pa].exponent, be = prims[B.prim_start + pb].exponent;
                    const double p = al + be;
                    const double k = prims[A.prim_start + pa].weighted_norm * prims[B.prim_start + pb].weighted_norm *
                                     std::exp(-al * be / p * AB2);
                    if (std::fabs(k) < 1e-20) continue;  // below any contribution that survives FP64 sums
                    PrimPair r;
                    r.p = p;
                    r.inv2p = 0.5 / p;
                    r.px = (al * A.center[0] + be * B.center[0]) / p;
                    r.py = (al * A.center[1] + be * B.center[1]) / p;
                    r.pz = (al * A.center[2] + be * B.center[2]) / p;
                    r.pax = r.px - A.center[0];
                    r.pay = r.py - A.center[1];
                    r.paz = r.pz - A.center[2];
                    r.pbx = r.px - B.center[0];
                    r.pby = r.py - B.center[1];
                    r.pbz = r.pz - B.center[2];
                    r.k = k;
                    pp.push_back(r);
                    beta.push_back(be);
                }
            d.n_prim = (int)pp.size() - d.prim_begin;
            P.class_list[pair_class(h.la, h.lb)].push_back((int)id);
        }
    // within a class, pairs with equal primitive counts side by side (one wave, one trip count)
    std::vector<int> lists;
    for (int c = 0; c < kPairClasses; ++c) {
        std::stable_sort(P.class_list[c].begin(), P.class_list[c].end(),
                         [&](int x, int y) { return P.pairs_dev_h[x].n_prim > P.pairs_dev_h[y].n_prim; });
        P.class_off[c] = (int)lists.size();
        lists.insert(lists.end(), P.class_list[c].begin(), P.class_list[c].end());
    }
    std::vector<double4> at(std::max(n_atoms, 1));
    for (int c = 0; c < n_atoms; ++c)
        at[c] = make_double4(atoms_xyzz[4 * c], atoms_xyzz[4 * c + 1], atoms_xyzz[4 * c + 2], atoms_xyzz[4 * c + 3]);
    if (hipStreamCreate(&P.stream) != hipSuccess) return kNoDevice;
    if (rocblas_create_handle(&P.blas) != rocblas_status_success) return kNoDevice;
    rocblas_set_stream(P.blas, P.stream);
    if (!P.pairs_dev.alloc(sizeof(ShellPairDev) * npairs) || !P.prims.alloc(sizeof(PrimPair) * std::max<size_t>(pp.size(), 1)) ||
        !P.beta.alloc(sizeof(double) * std::max<size_t>(beta.size(), 1)) || !P.ao_scale.alloc(sizeof(double) * n_ao) ||
        !P.atoms.alloc(sizeof(double4) * at.size()) || !P.class_lists.alloc(sizeof(int) * lists.size()))
        return kAllocFail;
    if (hipMemcpy(P.pairs_dev.p, P.pairs_dev_h.data(), sizeof(ShellPairDev) * npairs, hipMemcpyHostToDevice) != hipSuccess ||
        (!pp.empty() && hipMemcpy(P.prims.p, pp.data(), sizeof(PrimPair) * pp.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        (!beta.empty() && hipMemcpy(P.beta.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(P.ao_scale.p, scale.data(), sizeof(double) * n_ao, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P.atoms.p, at.data(), sizeof(double4) * at.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P.class_lists.p, lists.data(), sizeof(int) * lists.size(), hipMemcpyHostToDevice) != hipSuccess)
        return kLaunchFail;
    const int rc = compute_diag(P);
    if (rc != kOk) return rc;
    std::lock_guard<std::mutex> lk(g_int_mu);
    if (g_int.count(handle)) return kMismatch;
    g_int[handle] = plan;
    return kOk;
}

int calchem_hip_int_release(uint64_t handle) {
    std::shared_ptr<IntPlan> p;
    {
        std::lock_guard<std::mutex> lk(g_int_mu);
        auto it = g_int.find(handle);
        if (it == g_int.end()) return kOk;
        p = it->second;
        g_int.erase(it);
    }
    std::lock_guard<std::mutex> lk(p->mu);
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    return kOk;
}

int calchem_hip_int_one_electron(uint64_t handle, double* S, double* T, double* V) {
    if (!S || !T || !V || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto plan = find_int(handle);
    if (!plan) return kMismatch;
    IntPlan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    const size_t nn = (size_t)P.n_ao * P.n_ao;
    Buf d;
    if (!d.alloc(sizeof(double) * 3 * nn)) return kAllocFail;
    double* base = d.as<double>();
    for (int c = 0; c < kPairClasses; ++c) {
        if (P.class_list[c].empty()) continue;
        OneArgs a;
        a.pairs = P.pairs_dev.as<ShellPairDev>();
        a.list = P.class_lists.as<int>() + P.class_off[c];
        a.n_list = (int)P.class_list[c].size();
        a.prims = P.prims.as<PrimPair>();
        a.beta = P.beta.as<double>();
        a.ao_scale = P.ao_scale.as<double>();
        a.atoms = P.atoms.as<double4>();
        a.n_atoms = P.n_atoms;
        a.n_ao = P.n_ao;
        a.S = base;
        a.T = base + nn;
        a.V = base + 2 * nn;
        if (launch_one(c, a, P.stream) != hipSuccess) return kLaunchFail;
    }
    if (hipMemcpyAsync(S, base, sizeof(double) * nn, hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
        hipMemcpyAsync(T, base + nn, sizeof(double) * nn, hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
        hipMemcpyAsync(V, base + 2 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
        hipStreamSynchronize(P.stream) != hipSuccess)
        return kLaunchFail;
    return kOk;
}

int calchem_hip_int_eri_diag(uint64_t handle, double* out) {
    if (!out || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto plan = find_int(handle);
    if (!plan) return kMismatch;
    IntPlan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    if (hipMemcpyAsync(out, P.diag.p, sizeof(double) * P.n_pair, hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
        hipStreamSynchronize(P.stream) != hipSuccess)
        return kLaunchFail;
    return kOk;
}

int calchem_hip_int_eri_columns(uint64_t handle, int n_kets, const int* shell_a, const int* shell_b, double* out) {
    if (n_kets <= 0 || !shell_a || !shell_b || !out || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto plan = find_int(handle);
    if (!plan) return kMismatch;
    IntPlan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    std::vector<int> kets(n_kets);
    int ncol = 0;
    for (int k = 0; k < n_kets; ++k) {
        const int i = std::max(shell_a[k], shell_b[k]), j = std::min(shell_a[k], shell_b[k]);
        if (j < 0 || i >= P.n_shells) return kBadArgs;
        kets[k] = i * (i + 1) / 2 + j;
        ncol += P.pairs[kets[k]].na * P.pairs[kets[k]].nb;
    }
    Buf d;
    if (!d.alloc(sizeof(double) * (size_t)P.n_pair * ncol)) return kAllocFail;
    std::vector<int> col_of;
    const int got = eri_columns(P, kets, d.as<double>(), 0.0, &col_of);
    if (got < 0) return got;
    std::vector<double> tmp((size_t)P.n_pair * ncol);
    if (hipMemcpy(tmp.data(), d.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost) != hipSuccess) return kLaunchFail;
    // columns in the caller's order and orientation: pair k's (i_a, i_b) components, i_a of shell_a[k]
    int oc = 0;
    for (int k = 0; k < n_kets; ++k) {
        const PairHost& h = P.pairs[kets[k]];
        const bool swapped = h.a != shell_a[k];
        const int n1 = swapped ? h.nb : h.na, n2 = swapped ? h.na : h.nb;
        for (int x = 0; x < n1; ++x)
            for (int y = 0; y < n2; ++y) {
                const int ca = swapped ? y : x, cb = swapped ? x : y;
                std::memcpy(out + (size_t)oc * P.n_pair, tmp.data() + (size_t)(col_of[k] + ca * h.nb + cb) * P.n_pair,
                            sizeof(double) * P.n_pair);
                ++oc;
            }
    }
    return kOk;
}

int calchem_hip_int_cholesky(uint64_t handle, double tau, int max_rank, int* rank_out) {
    if (!(tau > 0.0) || max_rank < 0 || handle == 0) return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto plan = find_int(handle);
    if (!plan) return kMismatch;
    IntPlan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    return build_cholesky(P, tau, max_rank, rank_out);
}

int calchem_hip_int_jk(uint64_t handle, int n_pos, const double* Cpos, int n_neg, const double* Cneg, double* J,
                       double* K) {
    if (handle == 0 || n_pos < 0 || n_neg < 0 || (n_pos > 0 && !Cpos) || (n_neg > 0 && !Cneg) || (!J && !K))
        return kBadArgs;
    if (!dev_ok()) return kNoDevice;
    auto plan = find_int(handle);
    if (!plan) return kMismatch;
    IntPlan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    if (P.rank <= 0) return kMismatch;  // calchem_hip_int_cholesky first
    const int n = P.n_ao, r = P.rank;
    const size_t nn = (size_t)n * n;
    const int kmax = std::max(std::max(n_pos, n_neg), 1);
    if (!P.work_p.ensure(sizeof(double) * (4 * nn + (size_t)r)) ||
        !P.work_d.ensure(sizeof(double) * (size_t)n * kmax * 2))
        return kAllocFail;
    double* dP = P.work_p.as<double>();
    double* dJ = dP + nn;
    double* dK = dP + 2 * nn;
    double* dd = dP + 4 * nn;
    double* dCp = P.work_d.as<double>();
    double* dCn = dCp + (size_t)n * kmax;
    if ((n_pos > 0 && hipMemcpyAsync(dCp, Cpos, sizeof(double) * n * n_pos, hipMemcpyHostToDevice, P.stream) != hipSuccess) ||
        (n_neg > 0 && hipMemcpyAsync(dCn, Cneg, sizeof(double) * n * n_neg, hipMemcpyHostToDevice, P.stream) != hipSuccess))
        return kLaunchFail;
    const double one = 1.0, m1 = -1.0, zero = 0.0;
    if (J) {
        // P = C+ C+^T - C- C-^T (lower triangle by SYRK, then mirrored); C row-major n x k = column-major k x n
        bool first = true;
        if (n_pos > 0) {
            if (rocblas_dsyrk(P.blas, rocblas_fill_lower, rocblas_operation_transpose, n, n_pos, &one, dCp, n_pos, &zero,
                              dP, n) != rocblas_status_success)
                return kLaunchFail;
            first = false;
        }
        if (n_neg > 0) {
            const double b = first ? 0.0 : 1.0;
            if (rocblas_dsyrk(P.blas, rocblas_fill_lower, rocblas_operation_transpose, n, n_neg, &m1, dCn, n_neg, &b, dP,
                              n) != rocblas_status_success)
                return kLaunchFail;
            first = false;
        }
        if (first) {
            hipLaunchKernelGGL(k_zero, dim3(nblk((long long)nn)), dim3(256), 0, P.stream, dP, (long long)nn);
        } else {
            hipLaunchKernelGGL(k_symmetrize_lower, dim3(nblk((long long)nn)), dim3(256), 0, P.stream, dP, n);
        }
        // d = M2^T p, J = M2 d, M2 = L viewed as (n^2) x r
        if (rocblas_dgemv(P.blas, rocblas_operation_transpose, (rocblas_int)nn, r, &one, P.Lf.as<double>(),
                          (rocblas_int)nn, dP, 1, &zero, dd, 1) != rocblas_status_success ||
            rocblas_dgemv(P.blas, rocblas_operation_none, (rocblas_int)nn, r, &one, P.Lf.as<double>(), (rocblas_int)nn,
                          dd, 1, &zero, dJ, 1) != rocblas_status_success)
            return kLaunchFail;
        if (hipMemcpyAsync(J, dJ, sizeof(double) * nn, hipMemcpyDeviceToHost, P.stream) != hipSuccess) return kLaunchFail;
    }
    if (K) {
        bool first = true;
        int rc = kOk;
        if (n_pos > 0) {
            rc = add_exchange(P, dCp, n_pos, 1.0, dK, first);
            first = false;
        }
        if (rc == kOk && n_neg > 0) {
            rc = add_exchange(P, dCn, n_neg, -1.0, dK, first);
            first = false;
        }
        if (rc != kOk) return rc;
        if (first)
            hipLaunchKernelGGL(k_zero, dim3(nblk((long long)nn)), dim3(256), 0, P.stream, dK, (long long)nn);
        else  // exact symmetry (the GEMM's two triangles may differ in the last bit)
            hipLaunchKernelGGL(k_symmetrize_lower, dim3(nblk((long long)nn)), dim3(256), 0, P.stream, dK, n);
        if (hipMemcpyAsync(K, dK, sizeof(double) * nn, hipMemcpyDeviceToHost, P.stream) != hipSuccess) return kLaunchFail;
    }
    if (hipStreamSynchronize(P.stream) != hipSuccess) return kLaunchFail;
    return kOk;
}

}  // extern "C"
