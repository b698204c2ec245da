// api.hip — host runtime and C ABI of the calchem_hip Vxc engine.
//
// A plan holds everything that is fixed per geometry (shells, sorted points,
// blocks, AO lists, scratch) resident in HBM; a build runs the per-iteration
// kernels against it.  Plans live in a table keyed by a caller-chosen 64-bit
// id, like the reference GPU backend's grid table
// (crates/calchem-cuda/src/cuda_kernels.cu:6261, 7354-7562).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/calchem_hip.h"
#include "common.hpp"
#include "kernels.hpp"

namespace chx {
namespace {

std::atomic<size_t> g_allocated{0};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) {
            hipFree(p);
            g_allocated -= bytes;
        }
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
        g_allocated += b;
        return true;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct Profile {
    bool on = false;
    std::vector<hipEvent_t> ev;  // 5 events per build: start, after rho, after vxc, after reduce, end
    double ms[4] = {0, 0, 0, 0};
    long builds = 0;
};

struct Plan {
    int device = 0;
    int n_ao = 0, n_shells = 0;
    int n_in = 0;       // points handed in by the caller
    int n_pts = 0;      // points kept (|w| >= 1e-24), sorted into blocks
    bool has_weights = true;
    int n_blocks = 0, n_groups = 0, n_pairs = 0;
    int n_tiles[2] = {0, 0};  // rho tiles of 32 / 16 points
    int max_mpad_rho = 0, max_m_v = 0;
    int pts_fp64 = 32, pts_fp32 = 32;
    double flops_rho = 0, flops_v = 0, flops_rho_pad[2] = {0, 0}, flops_v_pad = 0;
    size_t lds_limit = 0;
    DevBuf shells, prims, aos, pts, perm, blocks, block_ao, tiles32, tiles16, groups, group_ao, pairs, loc,
        scratch, wv, wv_caller, epart, npart, en, P64, P32, V, rho;
    hipStream_t stream = nullptr;
    std::mutex mu;
    Profile prof;
    ~Plan() {
        for (hipEvent_t e : prof.ev) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
    size_t device_bytes() const {
        const DevBuf* b[] = {&shells, &prims, &aos,     &pts,   &perm,  &blocks,    &block_ao, &tiles32,
                             &tiles16, &groups, &group_ao, &pairs, &loc,   &scratch,   &wv,       &wv_caller,
                             &epart,  &npart,  &en,      &P64,   &P32,   &V,         &rho};
        size_t s = 0;
        for (const DevBuf* x : b) s += x->bytes;
        return s;
    }
};

std::mutex g_table_mu;
std::unordered_map<uint64_t, std::shared_ptr<Plan>> g_table;

int g_avail = -1;

bool device_available() {
    if (const char* e = std::getenv("CALCHEM_HIP_DISABLE")) {
        if (std::strcmp(e, "1") == 0) return false;
    }
    if (g_avail < 0) {
        int n = 0;
        g_avail = (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
    }
    return g_avail == 1;
}

std::shared_ptr<Plan> find_plan(uint64_t id) {
    std::lock_guard<std::mutex> lk(g_table_mu);
    auto it = g_table.find(id);
    return it == g_table.end() ? nullptr : it->second;
}

// Recursive bisection of points along the widest axis; split positions are
// multiples of 32 so rho tiles are full except at the very end.
void kd_blocks(std::vector<int>& idx, int b, int e, const std::vector<double4>& p, int nb,
               std::vector<std::pair<int, int>>& leaves) {
    if (e - b <= nb) {
        leaves.push_back({b, e});
        return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = b; i < e; ++i) {
        const double4& q = p[idx[i]];
        const double c[3] = {q.x, q.y, q.z};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], c[k]);
            hi[k] = std::max(hi[k], c[k]);
        }
    }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
    int mid = b + ((e - b) / 2 + 31) / 32 * 32;
    if (mid >= e) mid = b + (e - b) / 2;
    auto key = [&](int i) { return ax == 0 ? p[i].x : (ax == 1 ? p[i].y : p[i].z); };
    std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                     [&](int x, int y) { return key(x) < key(y) || (key(x) == key(y) && x < y); });
    kd_blocks(idx, b, mid, p, nb, leaves);
    kd_blocks(idx, mid, e, p, nb, leaves);
}

template <typename T>
bool upload(DevBuf& d, const std::vector<T>& h) {
    if (!d.alloc(sizeof(T) * h.size())) return false;
    if (h.empty()) return true;
    return hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) == hipSuccess;
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

int create_plan(const CalchemAoShell* sh, int n_shells, const CalchemAoPrim* pr, int n_prims,
                const CalchemAoComp* cp, int n_comps, const double* xyzw, int stride, int n_pts, int n_ao,
                bool has_w, const CalchemVxcOptions* opt, uint64_t id) {
    if (!sh || !pr || !cp || !xyzw || n_shells <= 0 || n_prims <= 0 || n_comps <= 0 || n_pts <= 0 || n_ao <= 0 ||
        id == 0)
        return kBadArgs;
    if (!device_available()) return kNoDevice;
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        auto it = g_table.find(id);
        if (it != g_table.end())
            return (it->second->n_in == n_pts && it->second->n_ao == n_ao) ? kOk : kMismatch;
    }
    // Validate shell records against the arrays they index.
    int ao_count = 0;
    for (int s = 0; s < n_shells; ++s) {
        const CalchemAoShell& x = sh[s];
        const int nf = (x.n_funcs);
        if (x.prim_start < 0 || x.n_prims <= 0 || x.prim_start + x.n_prims > n_prims || x.comp_start < 0 ||
            nf <= 0 || x.comp_start + nf > n_comps || x.ao_offset < 0 || x.ao_offset + nf > n_ao)
            return kMismatch;
        for (int c = 0; c < nf; ++c) {
            const CalchemAoComp& q = cp[x.comp_start + c];
            if (q.lx < 0 || q.ly < 0 || q.lz < 0 || q.lx > 3 || q.ly > 3 || q.lz > 3) return kUnsupported;
        }
        ao_count += nf;
    }
    if (ao_count != n_ao) return kMismatch;

    auto plan = std::make_shared<Plan>();
    Plan& P = *plan;
    if (hipGetDevice(&P.device) != hipSuccess) return kNoDevice;
    P.n_ao = n_ao;
    P.n_shells = n_shells;
    P.n_in = n_pts;
    P.has_weights = has_w;
    const int nb_rho = (opt && opt->block_points > 0) ? round_up(opt->block_points, 32) : 256;
    const int nb_v = (opt && opt->group_points > 0) ? opt->group_points : 1024;
    const size_t mem_limit = opt ? opt->memory_limit_bytes : 0;

    // Device shell / AO tables.
    std::vector<DevShell> hs(n_shells);
    std::vector<DevAo> hao(n_ao);
    std::vector<double2> hp(n_prims);
    for (int i = 0; i < n_prims; ++i) hp[i] = make_double2(pr[i].exponent, pr[i].weighted_norm);
    int lmax = 0;
    for (int s = 0; s < n_shells; ++s) {
        const CalchemAoShell& x = sh[s];
        int L = 0;
        for (int c = 0; c < x.n_funcs; ++c) {
            const CalchemAoComp& q = cp[x.comp_start + c];
            L = std::max(L, q.lx + q.ly + q.lz);
            DevAo& a = hao[x.ao_offset + c];
            a.scale = q.scale;
            a.shell = s;
            a.lx = q.lx;
            a.ly = q.ly;
            a.lz = q.lz;
        }
        lmax = std::max(lmax, L);
        hs[s] = DevShell{x.center[0], x.center[1], x.center[2], x.r2_cutoff, x.prim_start, x.n_prims, L, x.ao_offset};
    }

    // Points: keep |w| >= 1e-24 (vxc.rs:454,539), sort into compact blocks.
    std::vector<double4> all;
    std::vector<int> keep;
    all.reserve(n_pts);
    for (int i = 0; i < n_pts; ++i) {
        const double* q = xyzw + (size_t)stride * i;
        const double w = has_w ? q[3] : 1.0;
        if (!std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2]) || !std::isfinite(w)) return kBadArgs;
        all.push_back(make_double4(q[0], q[1], q[2], w));
        if (std::fabs(w) >= 1e-24) keep.push_back(i);
    }
    P.n_pts = static_cast<int>(keep.size());
    std::vector<std::pair<int, int>> leaves;
    if (P.n_pts > 0) kd_blocks(keep, 0, P.n_pts, all, nb_rho, leaves);
    std::vector<double4> spts(P.n_pts);
    for (int i = 0; i < P.n_pts; ++i) spts[i] = all[keep[i]];
    P.n_blocks = static_cast<int>(leaves.size());

    // Candidate shells per block: |R_shell - c| <= r_box + r_cut with c, r_box
    // the block's bounding-box centre and half diagonal (vxc.rs:446-493).
    std::vector<double> rcut(n_shells);
    for (int s = 0; s < n_shells; ++s) rcut[s] = std::sqrt(sh[s].r2_cutoff);
    std::vector<int> cand_ao;
    std::vector<int2> cand_rng;
    std::vector<int> cand_begin(P.n_blocks + 1, 0);
    for (int b = 0; b < P.n_blocks; ++b) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = leaves[b].first; i < leaves[b].second; ++i) {
            const double c[3] = {spts[i].x, spts[i].y, spts[i].z};
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        }
        const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
        const double rb = std::sqrt((hi[0] - cx) * (hi[0] - cx) + (hi[1] - cy) * (hi[1] - cy) +
                                    (hi[2] - cz) * (hi[2] - cz));
        cand_begin[b] = static_cast<int>(cand_ao.size());
        for (int s = 0; s < n_shells; ++s) {
            const double dx = hs[s].cx - cx, dy = hs[s].cy - cy, dz = hs[s].cz - cz;
            const double th = rb + rcut[s];
            if (dx * dx + dy * dy + dz * dz <= th * th) {
                for (int c = 0; c < sh[s].n_funcs; ++c) {
                    cand_ao.push_back(sh[s].ao_offset + c);
                    cand_rng.push_back(make_int2(leaves[b].first, leaves[b].second - leaves[b].first));
                }
            }
        }
    }
    cand_begin[P.n_blocks] = static_cast<int>(cand_ao.size());

    // AO significance on the GPU (exact pruning: an AO that the value rule
    // drops at every point of a block contributes nothing there).
    bool ok = upload(P.shells, hs) && upload(P.prims, hp) && upload(P.aos, hao) && upload(P.pts, spts);
    if (!ok) return kAllocFail;
    std::vector<unsigned char> flag(cand_ao.size(), 0);
    {
        DevBuf dca, dcr, dfl;
        if (!upload(dca, cand_ao) || !upload(dcr, cand_rng) || !dfl.alloc(flag.size() + 1)) return kAllocFail;
        SigArgs sa{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
                   dca.as<int>(), dcr.as<int2>(), dfl.as<unsigned char>(), (long long)cand_ao.size()};
        if (launch_significance(sa, 0) != hipSuccess) return kLaunchFail;
        if (!flag.empty() && hipMemcpy(flag.data(), dfl.p, flag.size(), hipMemcpyDeviceToHost) != hipSuccess)
            return kLaunchFail;
    }

    // rho blocks, tiles and their AO lists.
    std::vector<RhoBlock> hb(P.n_blocks);
    std::vector<int> hbao;
    std::vector<std::vector<int>> blist(P.n_blocks);
    std::vector<RhoTile> t32, t16;
    for (int b = 0; b < P.n_blocks; ++b) {
        for (int k = cand_begin[b]; k < cand_begin[b + 1]; ++k)
            if (flag[k]) blist[b].push_back(cand_ao[k]);
        std::sort(blist[b].begin(), blist[b].end());
        const int m = static_cast<int>(blist[b].size());
        const int mp = round_up(std::max(m, 1), 32);
        hb[b] = RhoBlock{static_cast<int>(hbao.size()), m, mp, 0};
        hbao.insert(hbao.end(), blist[b].begin(), blist[b].end());
        hbao.resize(hbao.size() + (mp - m), -1);
        P.max_mpad_rho = std::max(P.max_mpad_rho, mp);
        const int nb = leaves[b].second - leaves[b].first;
        P.flops_rho += 2.0 * m * (double)m * nb;
        if (m == 0) continue;
        for (int o = 0; o < nb; o += 32) {
            t32.push_back(RhoTile{b, leaves[b].first + o, std::min(32, nb - o), 0});
            P.flops_rho_pad[0] += 2.0 * mp * (double)mp * 32;
        }
        for (int o = 0; o < nb; o += 16) {
            t16.push_back(RhoTile{b, leaves[b].first + o, std::min(16, nb - o), 0});
            P.flops_rho_pad[1] += 2.0 * mp * (double)mp * 16;
        }
    }
    P.n_tiles[0] = static_cast<int>(t32.size());
    P.n_tiles[1] = static_cast<int>(t16.size());

    // V groups: consecutive blocks (k-d order) merged up to nb_v points; the
    // group's AO list is the union of its blocks' lists.
    std::vector<VGroup> hg;
    std::vector<int> hgao;
    std::vector<VPair> hpairs;
    std::vector<int> hloc;
    long long scratch = 0;
    for (int b = 0; b < P.n_blocks;) {
        int e = b, npt = 0;
        while (e < P.n_blocks && (npt == 0 || npt + (leaves[e].second - leaves[e].first) <= nb_v)) {
            npt += leaves[e].second - leaves[e].first;
            ++e;
        }
        std::vector<int> u;
        for (int k = b; k < e; ++k) u.insert(u.end(), blist[k].begin(), blist[k].end());
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        const int m = static_cast<int>(u.size());
        if (m > 0) {
            VGroup G;
            G.ao_begin = static_cast<int>(hgao.size());
            G.m = m;
            G.n_tiles64 = (m + kVTile - 1) / kVTile;
            G.m_pad = G.n_tiles64 * kVTile;
            G.pt_begin = leaves[b].first;
            G.n_pts = npt;
            G.scratch_off = scratch;
            const int gidx = static_cast<int>(hg.size());
            for (int i = 0; i < G.n_tiles64; ++i)
                for (int j = 0; j <= i; ++j) hpairs.push_back(VPair{gidx, i, j, 0});
            scratch += (long long)G.n_tiles64 * (G.n_tiles64 + 1) / 2 * kVTile * kVTile;
            hgao.insert(hgao.end(), u.begin(), u.end());
            const size_t lo = hloc.size();
            hloc.resize(lo + n_ao, -1);
            for (int i = 0; i < m; ++i) hloc[lo + u[i]] = i;
            P.flops_v += 2.0 * m * (double)m * npt;
            P.flops_v_pad += 2.0 * G.n_tiles64 * (G.n_tiles64 + 1) * (double)(kVTile * kVTile) *
                             (double)round_up(npt, kVPts);
            P.max_m_v = std::max(P.max_m_v, m);
            hg.push_back(G);
        }
        b = e;
    }
    P.n_groups = static_cast<int>(hg.size());
    P.n_pairs = static_cast<int>(hpairs.size());

    // LDS budget of the rho kernel decides 32- or 16-point tiles per precision.
    int lds_max = 0;
    hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, P.device);
    P.lds_limit = static_cast<size_t>(std::max(lds_max, 65536));
    auto pick = [&](int bytes) {
        if (rho_lds_bytes(bytes, 32, P.max_mpad_rho) <= P.lds_limit) return 32;
        if (rho_lds_bytes(bytes, 16, P.max_mpad_rho) <= P.lds_limit) return 16;
        return 0;
    };
    P.pts_fp64 = pick(8);
    P.pts_fp32 = pick(4);

    const size_t need = sizeof(double) * (size_t)scratch + sizeof(int) * hloc.size() +
                        sizeof(double) * (size_t)n_ao * n_ao * 2 + sizeof(float) * (size_t)n_ao * n_ao +
                        sizeof(double4) * (size_t)(2 * P.n_pts + n_pts);
    if (mem_limit > 0 && need > mem_limit) return kLimit;

    const int ntile_max = std::max(P.n_tiles[0], P.n_tiles[1]);
    std::vector<int> hperm(P.n_pts);
    for (int i = 0; i < P.n_pts; ++i) hperm[i] = keep[i];
    ok = upload(P.perm, hperm) && upload(P.blocks, hb) && upload(P.block_ao, hbao) && upload(P.tiles32, t32) &&
         upload(P.tiles16, t16) && upload(P.groups, hg) && upload(P.group_ao, hgao) && upload(P.pairs, hpairs) &&
         upload(P.loc, hloc) && P.scratch.alloc(sizeof(double) * (size_t)std::max<long long>(scratch, 1)) &&
         P.wv.alloc(sizeof(double4) * (size_t)std::max(P.n_pts, 1)) &&
         P.wv_caller.alloc(sizeof(double4) * (size_t)n_pts) && P.epart.alloc(sizeof(double) * (ntile_max + 1)) &&
         P.npart.alloc(sizeof(double) * (ntile_max + 1)) && P.en.alloc(sizeof(double) * 2) &&
         P.P64.alloc(sizeof(double) * (size_t)n_ao * n_ao) && P.P32.alloc(sizeof(float) * (size_t)n_ao * n_ao) &&
         P.V.alloc(sizeof(double) * (size_t)n_ao * n_ao) && P.rho.alloc(sizeof(double4) * (size_t)n_pts);
    if (!ok) return kAllocFail;
    if (hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking) != hipSuccess) return kNoDevice;
    if (hipDeviceSynchronize() != hipSuccess) return kLaunchFail;
    std::lock_guard<std::mutex> lk(g_table_mu);
    g_table[id] = plan;
    return kOk;
}

// Record a profiling event if profiling is on.
void mark(Plan& P, hipStream_t s) {
    if (!P.prof.on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    hipEventRecord(e, s);
    P.prof.ev.push_back(e);
}

void resolve_profile(Plan& P) {
    auto& ev = P.prof.ev;
    for (size_t i = 0; i + 5 <= ev.size(); i += 5) {
        hipEventSynchronize(ev[i + 4]);
        for (int k = 0; k < 4; ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[i + k], ev[i + k + 1]) == hipSuccess) P.prof.ms[k] += ms;
        }
        P.prof.builds++;
    }
    for (hipEvent_t e : ev) hipEventDestroy(e);
    ev.clear();
}

// Fused RKS build on device pointers, asynchronous on stream s.
int run_build(Plan& P, int func, int prec, const double* dP, double* dV, double* den, hipStream_t s) {
    if (func < 0 || func > 5) return kUnsupported;
    if (prec != CALCHEM_PREC_FP64 && prec != CALCHEM_PREC_FP32) return kUnsupported;
    if (!P.has_weights) return kBadArgs;
    const bool f32 = prec == CALCHEM_PREC_FP32;
    const int pts = f32 ? P.pts_fp32 : P.pts_fp64;
    if (pts == 0) return kLimit;
    const long long nn = (long long)P.n_ao * P.n_ao;
    mark(P, s);
    const void* Pk = dP;
    if (f32) {
        if (launch_to_float(dP, P.P32.as<float>(), nn, s) != hipSuccess) return kLaunchFail;
        Pk = P.P32.p;
    }
    const int ti = pts == 32 ? 0 : 1;
    RhoArgs ra{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
               ti == 0 ? P.tiles32.as<RhoTile>() : P.tiles16.as<RhoTile>(), P.blocks.as<RhoBlock>(),
               P.block_ao.as<int>(), Pk, P.n_ao, func, P.wv.as<double4>(), nullptr, P.perm.as<int>(),
               P.epart.as<double>(), P.npart.as<double>()};
    if (launch_rho(ra, f32, pts, true, P.n_tiles[ti], P.max_mpad_rho, s) != hipSuccess) return kLaunchFail;
    mark(P, s);
    VArgs va{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
             P.groups.as<VGroup>(), P.pairs.as<VPair>(), P.group_ao.as<int>(), P.wv.as<double4>(),
             P.perm.as<int>(), 0, P.scratch.as<double>()};
    if (launch_vxc(va, f32, P.n_pairs, s) != hipSuccess) return kLaunchFail;
    mark(P, s);
    ReduceArgs rd{P.loc.as<int>(), P.groups.as<VGroup>(), P.n_groups, P.n_ao, P.scratch.as<double>(), dV};
    if (P.n_groups == 0) {
        if (hipMemsetAsync(dV, 0, sizeof(double) * nn, s) != hipSuccess) return kLaunchFail;
    } else if (launch_reduce_v(rd, s) != hipSuccess) {
        return kLaunchFail;
    }
    mark(P, s);
    if (P.n_tiles[ti] == 0) {
        if (hipMemsetAsync(den, 0, sizeof(double) * 2, s) != hipSuccess) return kLaunchFail;
    } else if (launch_reduce_en(P.epart.as<double>(), P.npart.as<double>(), P.n_tiles[ti], den, s) != hipSuccess) {
        return kLaunchFail;
    }
    mark(P, s);
    return kOk;
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" {

int calchem_hip_is_available(void) { return device_available() ? 1 : 0; }

int calchem_hip_get_memory_info(size_t* free_b, size_t* total_b) {
    if (!free_b || !total_b) return kBadArgs;
    if (!device_available()) return kNoDevice;
    return hipMemGetInfo(free_b, total_b) == hipSuccess ? kOk : kNoDevice;
}

size_t calchem_hip_total_allocated_bytes(void) { return g_allocated.load(); }

int calchem_hip_vxc_upload_grid(const CalchemAoShell* shells, int n_shells, const CalchemAoPrim* prims, int n_prims,
                                const CalchemAoComp* comps, int n_comps, const CalchemGridPt* pts, int n_pts,
                                const int* cand_offsets, int n_cand_offsets, const int* cand_shells,
                                int n_cand_shells, int n_ao, uint64_t grid_id, size_t memory_limit_bytes) {
    if (!shells || !prims || !comps || !pts || !cand_offsets || !cand_shells || n_shells <= 0 || n_prims <= 0 ||
        n_comps <= 0 || n_pts <= 0 || n_cand_offsets < 2 || n_cand_shells <= 0 || n_ao <= 0 || grid_id == 0)
        return kBadArgs;
    // The caller's batch CSR must be well formed; the library re-blocks the
    // points spatially (the per-point cutoff rule makes the result independent
    // of the batch partition).
    if (cand_offsets[0] != 0 || cand_offsets[n_cand_offsets - 1] != n_cand_shells) return kMismatch;
    for (int i = 0; i + 1 < n_cand_offsets; ++i)
        if (cand_offsets[i + 1] < cand_offsets[i]) return kMismatch;
    for (int i = 0; i < n_cand_shells; ++i)
        if (cand_shells[i] < 0 || cand_shells[i] >= n_shells) return kMismatch;
    for (int i = 0; i < n_pts; ++i)
        if (pts[i].batch_id < 0 || pts[i].batch_id >= n_cand_offsets - 1) return kMismatch;
    std::vector<double> xyz((size_t)n_pts * 4);
    for (int i = 0; i < n_pts; ++i) {
        xyz[4 * (size_t)i] = pts[i].x;
        xyz[4 * (size_t)i + 1] = pts[i].y;
        xyz[4 * (size_t)i + 2] = pts[i].z;
        xyz[4 * (size_t)i + 3] = 1.0;
    }
    CalchemVxcOptions o;
    std::memset(&o, 0, sizeof o);
    o.memory_limit_bytes = memory_limit_bytes;
    return create_plan(shells, n_shells, prims, n_prims, comps, n_comps, xyz.data(), 4, n_pts, n_ao, false, &o,
                       grid_id);
}

int calchem_hip_vxc_eval_rho(uint64_t grid_id, const double* P_mat, double* rho_grad, int n_pts, int n_ao) {
    if (!P_mat || !rho_grad || n_pts <= 0 || n_ao <= 0 || grid_id == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(grid_id);
    if (!plan || plan->n_in != n_pts || plan->n_ao != n_ao) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    const size_t nn = (size_t)n_ao * n_ao;
    hipStream_t s = P.stream;
    if (hipMemcpyAsync(P.P64.p, P_mat, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess) return kLaunchFail;
    if (hipMemsetAsync(P.rho.p, 0, sizeof(double4) * (size_t)n_pts, s) != hipSuccess) return kLaunchFail;
    const int pts = P.pts_fp64;
    if (pts == 0) return kLimit;
    const int ti = pts == 32 ? 0 : 1;
    RhoArgs ra{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
               ti == 0 ? P.tiles32.as<RhoTile>() : P.tiles16.as<RhoTile>(), P.blocks.as<RhoBlock>(),
               P.block_ao.as<int>(), P.P64.p, n_ao, -1, nullptr, P.rho.as<double4>(), P.perm.as<int>(),
               P.epart.as<double>(), P.npart.as<double>()};
    if (launch_rho(ra, false, pts, false, P.n_tiles[ti], P.max_mpad_rho, s) != hipSuccess) return kLaunchFail;
    if (hipMemcpyAsync(rho_grad, P.rho.p, sizeof(double4) * (size_t)n_pts, hipMemcpyDeviceToHost, s) != hipSuccess)
        return kLaunchFail;
    return hipStreamSynchronize(s) == hipSuccess ? kOk : kLaunchFail;
}

int calchem_hip_vxc_accum(uint64_t grid_id, const double* wv_in, double* vxc_out, int n_pts, int n_ao) {
    if (!wv_in || !vxc_out || n_pts <= 0 || n_ao <= 0 || grid_id == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(grid_id);
    if (!plan || plan->n_in != n_pts || plan->n_ao != n_ao) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    hipStream_t s = P.stream;
    const size_t nn = (size_t)n_ao * n_ao;
    if (hipMemcpyAsync(P.wv_caller.p, wv_in, sizeof(double4) * (size_t)n_pts, hipMemcpyHostToDevice, s) != hipSuccess)
        return kLaunchFail;
    VArgs va{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
             P.groups.as<VGroup>(), P.pairs.as<VPair>(), P.group_ao.as<int>(), P.wv_caller.as<double4>(),
             P.perm.as<int>(), 1, P.scratch.as<double>()};
    if (launch_vxc(va, false, P.n_pairs, s) != hipSuccess) return kLaunchFail;
    ReduceArgs rd{P.loc.as<int>(), P.groups.as<VGroup>(), P.n_groups, n_ao, P.scratch.as<double>(),
                  P.V.as<double>()};
    if (P.n_groups == 0) {
        if (hipMemsetAsync(P.V.p, 0, sizeof(double) * nn, s) != hipSuccess) return kLaunchFail;
    } else if (launch_reduce_v(rd, s) != hipSuccess) {
        return kLaunchFail;
    }
    if (hipMemcpyAsync(vxc_out, P.V.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s) != hipSuccess)
        return kLaunchFail;
    return hipStreamSynchronize(s) == hipSuccess ? kOk : kLaunchFail;
}

int calchem_hip_vxc_release(uint64_t grid_id) {
    if (grid_id == 0) return kOk;
    std::shared_ptr<Plan> p;
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        auto it = g_table.find(grid_id);
        if (it == g_table.end()) return kOk;
        p = it->second;
        g_table.erase(it);
    }
    std::lock_guard<std::mutex> lk(p->mu);
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    return kOk;
}

int calchem_hip_vxc_create(const CalchemAoShell* shells, int n_shells, const CalchemAoPrim* prims, int n_prims,
                           const CalchemAoComp* comps, int n_comps, const double* xyzw, int n_pts, int n_ao,
                           const CalchemVxcOptions* opts, uint64_t handle) {
    return create_plan(shells, n_shells, prims, n_prims, comps, n_comps, xyzw, 4, n_pts, n_ao, true, opts, handle);
}

int calchem_hip_vxc_build(uint64_t handle, int xc, int precision, const double* P_mat, double* V_out, double* exc,
                          double* nelec) {
    if (!P_mat || !V_out || !exc || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    hipStream_t s = P.stream;
    const size_t nn = (size_t)P.n_ao * P.n_ao;
    if (hipMemcpyAsync(P.P64.p, P_mat, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess) return kLaunchFail;
    const int rc = run_build(P, xc, precision, P.P64.as<double>(), P.V.as<double>(), P.en.as<double>(), s);
    if (rc != kOk) return rc;
    double en[2] = {0.0, 0.0};
    if (hipMemcpyAsync(V_out, P.V.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(en, P.en.p, sizeof(double) * 2, hipMemcpyDeviceToHost, s) != hipSuccess)
        return kLaunchFail;
    if (hipStreamSynchronize(s) != hipSuccess) return kLaunchFail;
    *exc = en[0];
    if (nelec) *nelec = en[1];
    return kOk;
}

int calchem_hip_vxc_build_device(uint64_t handle, int xc, int precision, const double* d_P, double* d_V,
                                 double* d_exc_nelec, void* stream) {
    if (!d_P || !d_V || !d_exc_nelec || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    return run_build(P, xc, precision, d_P, d_V, d_exc_nelec, static_cast<hipStream_t>(stream));
}

int calchem_hip_vxc_set_profiling(uint64_t handle, int on) {
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    std::lock_guard<std::mutex> lk(plan->mu);
    hipSetDevice(plan->device);
    resolve_profile(*plan);
    plan->prof.on = on != 0;
    for (double& x : plan->prof.ms) x = 0.0;
    plan->prof.builds = 0;
    return kOk;
}

int calchem_hip_vxc_stats(uint64_t handle, CalchemVxcStats* out) {
    if (!out) return kBadArgs;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    resolve_profile(P);
    std::memset(out, 0, sizeof *out);
    out->n_points_in = P.n_in;
    out->n_points = P.n_pts;
    out->n_blocks = P.n_blocks;
    out->n_groups = P.n_groups;
    out->n_rho_tiles = P.n_tiles[0];
    out->n_v_pairs = P.n_pairs;
    out->flops_rho = P.flops_rho;
    out->flops_v = P.flops_v;
    out->flops_rho_executed_fp64 = P.flops_rho_pad[P.pts_fp64 == 16 ? 1 : 0];
    out->flops_rho_executed_fp32 = P.flops_rho_pad[P.pts_fp32 == 16 ? 1 : 0];
    out->flops_v_executed = P.flops_v_pad;
    out->max_m_rho = P.max_mpad_rho;
    out->max_m_v = P.max_m_v;
    out->rho_points_fp64 = P.pts_fp64;
    out->rho_points_fp32 = P.pts_fp32;
    out->device_bytes = P.device_bytes();
    out->lds_bytes_per_block = P.lds_limit;
    out->profiled_builds = P.prof.builds;
    out->ms_convert_and_rho = P.prof.ms[0];
    out->ms_vxc = P.prof.ms[1];
    out->ms_reduce_v = P.prof.ms[2];
    out->ms_reduce_en = P.prof.ms[3];
    return kOk;
}

long calchem_hip_grid_build(int n_atoms, const int* z, const double* xyz_bohr, int level, int n_rad, int n_ang,
                            int radial_min, double* out, long max_out) {
    if (!device_available()) return kNoDevice;
    int st = kOk;
    const long cnt = build_grid_host(n_atoms, z, xyz_bohr, level, n_rad, n_ang, radial_min, out, max_out, &st);
    if (st != kOk) return st;
    if (!out) return cnt;
    return cnt > max_out ? kLimit : cnt;
}

}  // extern "C"
