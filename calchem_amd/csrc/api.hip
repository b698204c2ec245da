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

// Radial shells per streaming task (STREAMING_RADIAL_BATCH_SHELLS, grid.rs:54).
constexpr int kStreamingRadialShells = 8;

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
    static constexpr int kMarks = 7;  // start, after gather, rho, xc, vxc, reduce_v, reduce_en
    std::vector<hipEvent_t> ev;
    double ms[kMarks - 1] = {0, 0, 0, 0, 0, 0};
    long builds = 0;
};

struct Plan {
    int device = 0;
    int n_ao = 0, n_shells = 0;
    int n_in = 0;       // points handed in by the caller
    int n_pts = 0;      // points kept (|w| >= 1e-24), sorted into blocks
    bool has_weights = true;
    int n_blocks = 0, n_groups = 0, n_pairs = 0;
    long n_rho_tiles = 0;     // 16-point rho tiles
    int max_mpad_rho = 0, max_m_v = 0;
    // Rho launch classes and, per mode (FP64 / mixed), whether each fits LDS.
    std::vector<RhoClass> classes;
    std::vector<int> class_order;  // classes by decreasing cost: the order of the rho stage's launches
    std::vector<char> class_fits[2];
    bool v_fits_fp32 = true;  // the mixed build's V kernel fits the LDS limit
    double flops_rho = 0, flops_v = 0, flops_rho_pad = 0, flops_v_pad = 0, flops_rho_fp64_rows = 0;
    double mixed_threshold = 0.1;
    size_t lds_limit = 0;
    int n_group_list = 0, reduce_waves = 1;
    int reduce_segments = 1;  // workgroups per AO row in the V reduction (fewer rows than CUs x 8: more)
    // The register-resident group list is three launches: groups of more than
    // kSmallGroupTiles tiles first (group_max_tiles kernel), then the rest
    // (kSmallGroupTiles kernel) in two lists, those whose shells and
    // primitives keep the mixed kernel's LDS above half a CU's (one workgroup
    // per CU) and then the others (two per CU), each costliest first.
    int n_group_big = 0, max_small_nsh = 0, max_small_nprim = 0, max_big_nsh = 0, max_big_nprim = 0;
    int n_group_heavy = 0, max_heavy_nsh = 0, max_heavy_nprim = 0;
    bool v_small_bf16 = false;  // mixed small-group V kernel in the BF16 three-term form
    int group_max_tiles = kGroupMaxTiles;
    int n_mol = 1, max_mol_ao = 0, n_xc_wg = 0;
    long long p_elems = 0;  // sum over molecules of n_ao^2 (packed P / V)
    std::vector<MolInfo> mol_host;
    DevBuf shells, prims, aos, pts, perm, blocks, block_rows, block_sh, block_meta, rho_tiles, grad_tiles, groups,
        group_rows,
        group_meta, red_part, group_ao,
        group_sh, group_list, pairs, ao_gbegin, ao_groups, scratch, wv, wv_caller, epart, npart, en, P64, Pb_host,
        V, Vb_host, rho, rho_s, psub64, psub32, psub64b, psub32b, class_blocks, rho_sb, wv_b, mols, ao_mol, xc_wg,
        shell_atom,
        atom_xyz, inv_r, asym, raw_s, host_s, gwx_a, gwx_b, gfxc, gpart, gsum,
        aoc, aoc_off,  // the mixed build's AO cache (empty: AOs evaluated per build)
        scratch_b, red_part_b;  // the second spin's V-stage tiles and reduction rows (unrestricted builds)
    int n_atoms = 0;  // atoms set by calchem_hip_vxc_set_atoms (0 = not set)
    bool uks_ready = false;  // rho_sb / wv_b / psub64b / psub32b / scratch_b / red_part_b allocated (first unrestricted build)
    long long psub_elems = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    Profile prof;
    std::vector<hipGraphExec_t> graphs;  // captured builds; id = index + 1, null once destroyed
    // Side streams and events that let the independent launches of the rho
    // stage (its LDS classes) run side by side: the caller's stream forks to
    // them after the gather and joins them before the functional.
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kSide] = {};
    ~Plan() {
        for (hipGraphExec_t g : graphs)
            if (g) hipGraphExecDestroy(g);
        for (hipEvent_t e : prof.ev) hipEventDestroy(e);
        for (int i = 0; i < kSide; ++i) {
            if (join_ev[i]) hipEventDestroy(join_ev[i]);
            if (side[i]) hipStreamDestroy(side[i]);
        }
        if (fork_ev) hipEventDestroy(fork_ev);
        if (stream) hipStreamDestroy(stream);
    }
    size_t device_bytes() const {
        const DevBuf* b[] = {&shells,   &prims,     &aos,      &pts,   &perm,      &blocks,   &block_rows,
                             &block_sh, &block_meta, &group_meta, &red_part, &rho_tiles, &grad_tiles, &groups, &group_rows, &group_ao, &group_sh, &group_list,
                             &pairs,    &ao_gbegin, &ao_groups, &scratch, &wv,     &wv_caller,
                             &epart,    &npart,     &en,       &P64,   &Pb_host,   &V,        &rho, &rho_s, &psub64, &psub32,
                             &psub64b, &psub32b, &class_blocks, &rho_sb, &wv_b, &Vb_host, &mols, &ao_mol, &xc_wg, &shell_atom, &atom_xyz,
                             &scratch_b, &red_part_b,
                             &inv_r, &asym, &raw_s, &host_s, &gwx_a, &gwx_b, &gfxc, &gpart, &gsum, &aoc, &aoc_off};
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

// Appends n records of a type whose size is a multiple of 16 bytes.
template <typename T>
void push_records(std::vector<uint4>& out, const T* rec, int n) {
    static_assert(sizeof(T) % sizeof(uint4) == 0, "records are whole 16-byte units");
    const size_t at = out.size();
    out.resize(at + sizeof(T) / sizeof(uint4) * (size_t)n);
    if (n > 0) std::memcpy(out.data() + at, rec, sizeof(T) * (size_t)n);
}

// Block- or group-local shell list of an ordered row list: every shell of
// the rows once, those with more primitives first (ties in order of first
// use), so that the (shell, point) lanes of one wave's radial sums loop over
// equally many primitives.  Appends (global shell, local primitive offset) to
// shl and (AO, local shell) to out_rows; returns the list's primitive count.
// `local` is n_shells entries of -1 on entry and on return.
int append_shell_list(const std::vector<int>& rows, const std::vector<DevAo>& hao, const std::vector<DevShell>& hs,
                      std::vector<int>& local, std::vector<int2>& shl, std::vector<int2>& out_rows) {
    std::vector<int> order;
    for (int mu : rows) {
        const int s = hao[mu].shell;
        if (local[s] < 0) {
            local[s] = static_cast<int>(order.size());
            order.push_back(s);
        }
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return hs[x].n_prims > hs[y].n_prims; });
    int n_prim = 0;
    for (int k = 0; k < static_cast<int>(order.size()); ++k) {
        local[order[k]] = k;
        shl.push_back(make_int2(order[k], n_prim));
        n_prim += hs[order[k]].n_prims;
    }
    for (int mu : rows) out_rows.push_back(make_int2(mu, local[hao[mu].shell]));
    for (int x : order) local[x] = -1;
    return n_prim;
}

// One molecule of a plan as the caller hands it in: its shells
// [shell_begin, shell_end) (ao_offset local to the molecule), its points
// [pt_begin, pt_end) of the caller's point array, its AO count.
struct MolSpec {
    int shell_begin, shell_end, pt_begin, pt_end, n_ao;
};

int create_plan(const CalchemAoShell* sh, int n_shells, const CalchemAoPrim* pr, int n_prims,
                const CalchemAoComp* cp, int n_comps, const double* xyzw, int stride, int n_pts,
                const std::vector<MolSpec>& mspec, bool has_w, const CalchemVxcOptions* opt, uint64_t id,
                bool reuse_ok) {
    if (!sh || !pr || !cp || !xyzw || n_shells <= 0 || n_prims <= 0 || n_comps <= 0 || n_pts <= 0 ||
        mspec.empty() || id == 0)
        return kBadArgs;
    if (!device_available()) return kNoDevice;
    const int n_mol = static_cast<int>(mspec.size());
    long long n_ao_ll = 0;
    for (const MolSpec& m : mspec) {
        if (m.n_ao <= 0 || m.shell_begin < 0 || m.shell_end > n_shells || m.shell_begin >= m.shell_end ||
            m.pt_begin < 0 || m.pt_end > n_pts || m.pt_begin > m.pt_end)
            return kBadArgs;
        n_ao_ll += m.n_ao;
    }
    for (int k = 0; k + 1 < n_mol; ++k)
        if (mspec[k].shell_end != mspec[k + 1].shell_begin || mspec[k].pt_end != mspec[k + 1].pt_begin)
            return kMismatch;
    if (mspec[0].shell_begin != 0 || mspec[n_mol - 1].shell_end != n_shells || mspec[0].pt_begin != 0 ||
        mspec[n_mol - 1].pt_end != n_pts || n_ao_ll > (1 << 30))
        return kMismatch;
    const int n_ao = static_cast<int>(n_ao_ll);
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        auto it = g_table.find(id);
        // upload_grid keeps the reference's reuse rule (same id and sizes: ok,
        // cuda_kernels.cu:7382-7386); create / create_batched are per-geometry
        // set-up, so an id in use must be released first.
        if (it != g_table.end())
            return (reuse_ok && it->second->n_in == n_pts && it->second->n_ao == n_ao && it->second->n_mol == n_mol)
                       ? kOk
                       : kMismatch;
    }
    // Validate shell records against the arrays they index; AO offsets are
    // local to each molecule.
    std::vector<int> shell_mol(n_shells), ao_base(n_mol);
    for (int k = 0, base = 0; k < n_mol; ++k) {
        ao_base[k] = base;
        int ao_count = 0;
        for (int s = mspec[k].shell_begin; s < mspec[k].shell_end; ++s) {
            const CalchemAoShell& x = sh[s];
            const int nf = (x.n_funcs);
            if (x.prim_start < 0 || x.n_prims <= 0 || x.prim_start + x.n_prims > n_prims || x.comp_start < 0 ||
                nf <= 0 || x.comp_start + nf > n_comps || x.ao_offset < 0 || x.ao_offset + nf > mspec[k].n_ao)
                return kMismatch;
            for (int c = 0; c < nf; ++c) {
                const CalchemAoComp& q = cp[x.comp_start + c];
                if (q.lx < 0 || q.ly < 0 || q.lz < 0 || q.lx > 3 || q.ly > 3 || q.lz > 3) return kUnsupported;
            }
            ao_count += nf;
            shell_mol[s] = k;
        }
        if (ao_count != mspec[k].n_ao) return kMismatch;
        base += mspec[k].n_ao;
    }

    auto plan = std::make_shared<Plan>();
    Plan& P = *plan;
    if (hipGetDevice(&P.device) != hipSuccess) return kNoDevice;
    P.n_ao = n_ao;
    P.n_shells = n_shells;
    P.n_in = n_pts;
    P.has_weights = has_w;
    P.n_mol = n_mol;
    const int nb_rho = (opt && opt->block_points > 0) ? round_up(opt->block_points, 32) : kDefaultBlockPoints;
    const int nb_v = (opt && opt->group_points > 0) ? opt->group_points : 384;
    const size_t mem_limit = opt ? opt->memory_limit_bytes : 0;
    if (opt && opt->mixed_exact_threshold > 0.0) P.mixed_threshold = opt->mixed_exact_threshold;
    const int want_pts = opt ? opt->rho_tile_points : 0;
    if (want_pts != 0 && want_pts != 16 && want_pts != 32) return kBadArgs;
    if (want_pts == 32) return kUnsupported;  // the rho stage walks blocks 16 points at a time
    const int want_rows = opt ? opt->group_max_rows : 0;
    if (want_rows != 0 && want_rows != 192 && want_rows != 272) return kBadArgs;
    const int aoc_mode = opt ? opt->ao_cache : 0;
    if (aoc_mode < 0 || aoc_mode > 2) return kBadArgs;
    const int vform = opt ? opt->v_small_bf16 : 0;
    if (vform < 0 || vform > 2) return kBadArgs;
    P.v_small_bf16 = vform == 2;
    P.group_max_tiles = want_rows == 192 ? 12 : kGroupMaxTiles;

    // Device shell / AO tables.
    std::vector<DevShell> hs(n_shells);
    std::vector<DevAo> hao(n_ao);
    std::vector<double2> hp(n_prims);
    for (int i = 0; i < n_prims; ++i) hp[i] = make_double2(pr[i].exponent, pr[i].weighted_norm);
    int lmax = 0;
    for (int s = 0; s < n_shells; ++s) {
        const CalchemAoShell& x = sh[s];
        int L = 0;
        const int gao = ao_base[shell_mol[s]] + x.ao_offset;
        for (int c = 0; c < x.n_funcs; ++c) {
            const CalchemAoComp& q = cp[x.comp_start + c];
            L = std::max(L, q.lx + q.ly + q.lz);
            DevAo& a = hao[gao + c];
            a.scale = q.scale;
            a.shell = s;
            a.lx = q.lx;
            a.ly = q.ly;
            a.lz = q.lz;
        }
        lmax = std::max(lmax, L);
        hs[s] = DevShell{x.center[0], x.center[1], x.center[2], x.r2_cutoff, x.prim_start, x.n_prims, L, gao};
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
    // Blocks never mix molecules: each molecule's kept points are blocked on
    // their own (keep is in caller order, so a molecule's points are a range).
    std::vector<std::pair<int, int>> leaves;
    std::vector<int> block_mol;
    P.mol_host.assign(n_mol, MolInfo{});
    for (int k = 0, kb = 0; k < n_mol; ++k) {
        int ke = kb;
        while (ke < P.n_pts && keep[ke] < mspec[k].pt_end) ++ke;
        const size_t l0 = leaves.size();
        if (ke > kb) kd_blocks(keep, kb, ke, all, nb_rho, leaves);
        block_mol.resize(leaves.size(), k);
        MolInfo& mi = P.mol_host[k];
        mi.ao_base = ao_base[k];
        mi.n_ao = mspec[k].n_ao;
        mi.p_off = P.p_elems;
        P.p_elems += (long long)mi.n_ao * mi.n_ao;
        mi.pt_begin = kb;
        mi.n_pts = ke - kb;
        P.max_mol_ao = std::max(P.max_mol_ao, mi.n_ao);
        (void)l0;
        kb = ke;
    }
    std::vector<double4> spts(P.n_pts);
    for (int i = 0; i < P.n_pts; ++i) spts[i] = all[keep[i]];
    P.n_blocks = static_cast<int>(leaves.size());
    // Functional-stage workgroups: 256 points of one molecule each.
    std::vector<int2> hxc;
    std::vector<int> hao_mol(n_ao);
    for (int k = 0; k < n_mol; ++k) {
        MolInfo& mi = P.mol_host[k];
        mi.xc_wg_begin = static_cast<int>(hxc.size());
        for (int o = 0; o < mi.n_pts; o += kThreads)
            hxc.push_back(make_int2(mi.pt_begin + o, std::min(kThreads, mi.n_pts - o)));
        mi.n_xc_wg = static_cast<int>(hxc.size()) - mi.xc_wg_begin;
        for (int i = 0; i < mi.n_ao; ++i) hao_mol[mi.ao_base + i] = k;
    }
    P.n_xc_wg = static_cast<int>(hxc.size());

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
        const MolSpec& ms = mspec[block_mol[b]];
        for (int s = ms.shell_begin; s < ms.shell_end; ++s) {
            const double dx = hs[s].cx - cx, dy = hs[s].cy - cy, dz = hs[s].cz - cz;
            const double th = rb + rcut[s];
            if (dx * dx + dy * dy + dz * dz <= th * th) {
                for (int c = 0; c < sh[s].n_funcs; ++c) {
                    cand_ao.push_back(hs[s].ao_offset + c);
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
    std::vector<double> maxabs(cand_ao.size(), 0.0);
    {
        DevBuf dca, dcr, dfl, dmx;
        if (!upload(dca, cand_ao) || !upload(dcr, cand_rng) || !dfl.alloc(flag.size() + 1) ||
            !dmx.alloc(sizeof(double) * (maxabs.size() + 1)))
            return kAllocFail;
        SigArgs sa{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
                   dca.as<int>(), dcr.as<int2>(), dfl.as<unsigned char>(), dmx.as<double>(),
                   (long long)cand_ao.size()};
        if (launch_significance(sa, 0) != hipSuccess) return kLaunchFail;
        if (!flag.empty() &&
            (hipMemcpy(flag.data(), dfl.p, flag.size(), hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(maxabs.data(), dmx.p, sizeof(double) * maxabs.size(), hipMemcpyDeviceToHost) != hipSuccess))
            return kLaunchFail;
    }

    // rho blocks: rows are the AOs that survive somewhere in the block, those
    // with max |phi| above the mixed threshold first (each part ascending);
    // each row names its shell in the block's shell list.
    std::vector<RhoBlock> hb(P.n_blocks);
    std::vector<uint4> hmeta;
    std::vector<int2> hrows;
    std::vector<int2> hsh;  // (global shell, block-local primitive offset)
    std::vector<std::vector<int>> blist(P.n_blocks);
    std::vector<int> local_sh(n_shells, -1);
    for (int b = 0; b < P.n_blocks; ++b) {
        std::vector<int> big, small;
        for (int k = cand_begin[b]; k < cand_begin[b + 1]; ++k) {
            if (!flag[k]) continue;
            blist[b].push_back(cand_ao[k]);
            (maxabs[k] > P.mixed_threshold ? big : small).push_back(cand_ao[k]);
        }
        std::sort(blist[b].begin(), blist[b].end());
        const int m = static_cast<int>(blist[b].size());
        const int mp = round_up(std::max(m, 1), 16);
        std::vector<int> rows = big;
        rows.insert(rows.end(), small.begin(), small.end());
        RhoBlock B{};
        B.mol = block_mol[b];
        B.ao_begin = static_cast<int>(hrows.size());
        B.m = m;
        B.m_pad = mp;
        B.m64 = std::min(round_up(static_cast<int>(big.size()), 16), mp);  // whole 16-column chunks
        B.sh_begin = static_cast<int>(hsh.size());
        B.n_prim = append_shell_list(rows, hao, hs, local_sh, hsh, hrows);
        B.n_sh = static_cast<int>(hsh.size()) - B.sh_begin;
        hrows.resize(hrows.size() + (mp - m), make_int2(-1, 0));
        // Packed records, the layout the rho kernel copies into LDS as is.
        B.meta = static_cast<long long>(hmeta.size());
        for (int r = 0; r < mp; ++r) {
            RowL o{0.0, 0, -1};
            if (r < m) {
                const int2 e = hrows[B.ao_begin + r];
                const DevAo& ao = hao[e.x];
                o.scale = ao.scale;
                o.lmn = ao.lx | (ao.ly << 2) | (ao.lz << 4);
                o.sh = e.y;
            }
            push_records(hmeta, &o, 1);
        }
        for (int k = 0; k < B.n_sh; ++k) {
            DevShell d = hs[hsh[B.sh_begin + k].x];
            d.prim_start = hsh[B.sh_begin + k].y;
            push_records(hmeta, &d, 1);
        }
        for (int k = 0; k < B.n_sh; ++k) {
            const DevShell& d = hs[hsh[B.sh_begin + k].x];
            push_records(hmeta, hp.data() + d.prim_start, d.n_prims);
        }
        B.psub = P.psub_elems;
        P.psub_elems += (long long)mp * mp;
        hb[b] = B;
        P.max_mpad_rho = std::max(P.max_mpad_rho, mp);
        const int nb = leaves[b].second - leaves[b].first;
        P.flops_rho += 2.0 * m * (double)m * nb;
        P.flops_rho_fp64_rows += 2.0 * m * (double)big.size() * nb;
    }
    // Tiles of 16 points grouped into launch classes by the block's m_pad so
    // each launch sizes its LDS for similar blocks.  Within a class the blocks
    // are dealt to kXcds segments of about equal cost (largest first, each to
    // the segment with the least cost so far), each segment ordered costliest
    // block first; a block's tiles stay consecutive.  Cost of a tile: m_pad^2
    // plus a fixed part for its staging and AO work (that of m_pad = 128).
    std::vector<int2> htiles, hgtiles;  // rho workgroups; 16-point tiles of the gradient kernel
    std::vector<int> hclsblk;           // blocks in class order (each class gathers its own)
    {
        const int ncls = std::max((P.max_mpad_rho + 63) / 64, 1);
        std::vector<std::vector<int>> bycls(ncls);
        std::vector<RhoClass> cls(ncls, RhoClass{});
        for (int b = 0; b < P.n_blocks; ++b) {
            const int nb = leaves[b].second - leaves[b].first;
            hb[b].pt_begin = leaves[b].first;
            hb[b].n_pts = nb;
            if (hb[b].m == 0) continue;
            const int c = (hb[b].m_pad - 1) / 64;
            bycls[c].push_back(b);
            P.n_rho_tiles += (nb + kRhoPts - 1) / kRhoPts;
            P.flops_rho_pad += 2.0 * hb[b].m_pad * (double)hb[b].m_pad * round_up(nb, kRhoPts);
            cls[c].max_mpad = std::max(cls[c].max_mpad, hb[b].m_pad);
            cls[c].max_nsh = std::max(cls[c].max_nsh, hb[b].n_sh);
            cls[c].max_m64 = std::max(cls[c].max_m64, hb[b].m64);
            cls[c].max_prim = std::max(cls[c].max_prim, hb[b].n_prim);
        }
        auto block_cost = [&](int b) {
            return (double)((hb[b].n_pts + kRhoPts - 1) / kRhoPts) * ((double)hb[b].m_pad * hb[b].m_pad + 16384.0);
        };
        const int step = kRhoSub * kRhoPts;  // points per rho workgroup
        for (int c = 0; c < ncls; ++c) {
            if (bycls[c].empty()) continue;
            std::vector<int>& bl = bycls[c];
            std::stable_sort(bl.begin(), bl.end(), [&](int x, int y) { return block_cost(x) > block_cost(y); });
            std::vector<std::vector<int>> seg(kXcds);
            double load[kXcds] = {};
            for (int b : bl) {
                const int x = static_cast<int>(std::min_element(load, load + kXcds) - load);
                seg[x].push_back(b);
                load[x] += block_cost(b);
                cls[c].cost += block_cost(b);
            }
            cls[c].block_begin = static_cast<int>(hclsblk.size());
            cls[c].n_blocks = static_cast<int>(bl.size());
            hclsblk.insert(hclsblk.end(), bl.begin(), bl.end());
            cls[c].tile_begin = static_cast<int>(htiles.size());
            for (int x = 0; x < kXcds; ++x) {
                cls[c].seg[x] = static_cast<int>(htiles.size()) - cls[c].tile_begin;
                for (int b : seg[x])
                    for (int o = 0; o < hb[b].n_pts; o += step) htiles.push_back(make_int2(b, o));
            }
            // the XC gradient's AO kernel keeps one workgroup per 16 points
            cls[c].grad_tile_begin = static_cast<int>(hgtiles.size());
            for (int b : bl)
                for (int o = 0; o < hb[b].n_pts; o += kRhoPts) hgtiles.push_back(make_int2(b, o));
            cls[c].grad_n_tiles = static_cast<int>(hgtiles.size()) - cls[c].grad_tile_begin;
            cls[c].seg[kXcds] = static_cast<int>(htiles.size()) - cls[c].tile_begin;
            cls[c].n_tiles = cls[c].seg[kXcds];
            P.classes.push_back(cls[c]);
        }
        // the rho stage starts the costliest class first, on the build's own
        // stream, and the others beside it in decreasing cost
        P.class_order.resize(P.classes.size());
        for (size_t c = 0; c < P.classes.size(); ++c) P.class_order[c] = static_cast<int>(c);
        std::stable_sort(P.class_order.begin(), P.class_order.end(),
                         [&](int x, int y) { return P.classes[x].cost > P.classes[y].cost; });
    }

    // V groups: consecutive blocks (k-d order) merged up to nb_v points and,
    // for the register-resident kernel, up to 272 rows; the group's rows are
    // the union of its blocks' AOs, ascending.  A single block wider than that
    // goes to the tile-pair kernel.
    std::vector<VGroup> hg;
    std::vector<int2> hgrows, hgsh;  // hgsh: (global shell, group-local primitive offset)
    std::vector<uint4> hgmeta;
    std::vector<int> hglist;
    std::vector<VPair> hpairs;
    long long scratch = 0;
    const int max_rows = P.group_max_tiles * 16;
    auto union_rows = [&](int b, int e) {
        std::vector<int> u;
        for (int k = b; k < e; ++k) u.insert(u.end(), blist[k].begin(), blist[k].end());
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        return u;
    };
    for (int b = 0; b < P.n_blocks;) {
        int e = b + 1, npt = leaves[b].second - leaves[b].first;
        std::vector<int> u = union_rows(b, e);
        while (e < P.n_blocks && block_mol[e] == block_mol[b] && npt + (leaves[e].second - leaves[e].first) <= nb_v) {
            std::vector<int> u2 = union_rows(b, e + 1);
            if ((int)u2.size() > max_rows && (int)u.size() <= max_rows) break;
            u.swap(u2);
            npt += leaves[e].second - leaves[e].first;
            ++e;
        }
        const int m = static_cast<int>(u.size());
        if (m > 0) {
            VGroup G{};
            const int gidx = static_cast<int>(hg.size());
            G.ao_begin = static_cast<int>(hgrows.size());
            G.m = m;
            G.pt_begin = leaves[b].first;
            G.n_pts = npt;
            G.scratch_off = scratch;
            G.sh_begin = static_cast<int>(hgsh.size());
            G.n_prim = append_shell_list(u, hao, hs, local_sh, hgsh, hgrows);
            G.n_sh = static_cast<int>(hgsh.size()) - G.sh_begin;
            G.meta = -1;
            if (m <= max_rows) {
                G.tile_shift = 4;
                G.n_tiles = (m + 15) / 16;
                // packed records, copied into LDS as is by the group kernels
                G.meta = static_cast<long long>(hgmeta.size());
                for (int r = 0; r < G.n_tiles * 16; ++r) {
                    RowL o{0.0, 0, -1};
                    if (r < m) {
                        const int2 e = hgrows[G.ao_begin + r];
                        const DevAo& ao = hao[e.x];
                        o.scale = ao.scale;
                        o.lmn = ao.lx | (ao.ly << 2) | (ao.lz << 4);
                        o.sh = e.y;
                    }
                    push_records(hgmeta, &o, 1);
                }
                for (int k = 0; k < G.n_sh; ++k) {
                    DevShell d = hs[hgsh[G.sh_begin + k].x];
                    d.prim_start = hgsh[G.sh_begin + k].y;
                    push_records(hgmeta, &d, 1);
                }
                for (int k = 0; k < G.n_sh; ++k) {
                    const DevShell& d = hs[hgsh[G.sh_begin + k].x];
                    push_records(hgmeta, hp.data() + d.prim_start, d.n_prims);
                }
                hglist.push_back(gidx);
                P.flops_v_pad += 2.0 * G.n_tiles * (G.n_tiles + 1) * 256.0 * (double)round_up(npt, kVPts);
            } else {
                G.tile_shift = 6;
                G.n_tiles = (m + kVTile - 1) / kVTile;
                for (int i = 0; i < G.n_tiles; ++i)
                    for (int j = 0; j <= i; ++j) hpairs.push_back(VPair{gidx, i, j, 0});
                P.flops_v_pad += 2.0 * G.n_tiles * (G.n_tiles + 1) * (double)(kVTile * kVTile) *
                                 (double)round_up(npt, kVPts);
            }
            scratch += (long long)G.n_tiles * (G.n_tiles + 1) / 2 << (2 * G.tile_shift);
            P.flops_v += 2.0 * m * (double)m * npt;
            P.max_m_v = std::max(P.max_m_v, m);
            hg.push_back(G);
        }
        b = e;
    }
    P.n_groups = static_cast<int>(hg.size());
    std::vector<int> hshatom(n_shells);
    for (int s = 0; s < n_shells; ++s) hshatom[s] = sh[s].atom_index;
    std::vector<int> hgao(hgrows.size());
    for (size_t i = 0; i < hgrows.size(); ++i) hgao[i] = hgrows[i].x;
    P.n_pairs = static_cast<int>(hpairs.size());
    P.n_group_list = static_cast<int>(hglist.size());
    // Launch the register-resident groups costliest first (tile pairs x
    // points), so that the last workgroups to start are short ones.
    std::stable_sort(hglist.begin(), hglist.end(), [&](int x, int y) {
        const auto cost = [&](int g) {
            return (double)hg[g].n_tiles * (hg[g].n_tiles + 1) * (double)round_up(hg[g].n_pts, kVPts);
        };
        return cost(x) > cost(y);
    });
    std::stable_partition(hglist.begin(), hglist.end(), [&](int g) { return hg[g].n_tiles > kSmallGroupTiles; });
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, P.device) != hipSuccess)
        lds_max = 65536;
    P.lds_limit = static_cast<size_t>(std::max(lds_max, 65536));
    const auto heavy = [&](int g) {
        return hg[g].n_tiles <= kSmallGroupTiles &&
               vgroup_lds_bytes(true, kSmallGroupTiles, hg[g].n_sh, hg[g].n_prim, P.v_small_bf16) > P.lds_limit / 2;
    };
    std::stable_partition(hglist.begin(), hglist.end(),
                          [&](int g) { return hg[g].n_tiles > kSmallGroupTiles || heavy(g); });
    for (int g : hglist) {
        const bool big = hg[g].n_tiles > kSmallGroupTiles, hv = heavy(g);
        P.n_group_big += big;
        P.n_group_heavy += hv;
        int& nsh = big ? P.max_big_nsh : (hv ? P.max_heavy_nsh : P.max_small_nsh);
        int& npr = big ? P.max_big_nprim : (hv ? P.max_heavy_nprim : P.max_small_nprim);
        nsh = std::max(nsh, hg[g].n_sh);
        npr = std::max(npr, hg[g].n_prim);
    }
    // CSR: for each AO the groups holding it, ascending, with its row there.
    std::vector<int> hgb(n_ao + 1, 0);
    std::vector<RedEntry> hgg;
    for (int g = 0; g < P.n_groups; ++g)
        for (int i = 0; i < hg[g].m; ++i) hgb[hgrows[hg[g].ao_begin + i].x + 1]++;
    for (int mu = 0; mu < n_ao; ++mu) hgb[mu + 1] += hgb[mu];
    hgg.resize(hgb[n_ao]);
    {
        std::vector<int> fill(hgb.begin(), hgb.end() - 1);
        for (int g = 0; g < P.n_groups; ++g)
            for (int i = 0; i < hg[g].m; ++i) {
                const VGroup& G = hg[g];
                const int ts = G.tile_shift, ti = i >> ts;
                RedEntry r;
                r.base = G.scratch_off + ((long long)(ti * (ti + 1) / 2) << (2 * ts)) + ((i & ((1 << ts) - 1)) << ts);
                r.rows = G.ao_begin;
                r.li_ts = i | (ts << 16);
                hgg[fill[hgrows[G.ao_begin + i].x]++] = r;
            }
    }

    // LDS budget per class and mode.
    for (int mode = 0; mode < 2; ++mode) {
        P.class_fits[mode].assign(P.classes.size(), 0);
        for (size_t c = 0; c < P.classes.size(); ++c)
            P.class_fits[mode][c] = rho_lds_bytes(mode, P.classes[c]) <= P.lds_limit;
    }

    const int n_small = P.n_group_list - P.n_group_big - P.n_group_heavy;
    if ((P.n_group_big > 0 &&
         vgroup_lds_bytes(false, P.group_max_tiles, P.max_big_nsh, P.max_big_nprim) > P.lds_limit) ||
        (P.n_group_heavy > 0 &&
         vgroup_lds_bytes(false, kSmallGroupTiles, P.max_heavy_nsh, P.max_heavy_nprim) > P.lds_limit) ||
        (n_small > 0 && vgroup_lds_bytes(false, kSmallGroupTiles, P.max_small_nsh, P.max_small_nprim) > P.lds_limit))
        return kLimit;
    P.v_fits_fp32 =
        (P.n_group_big == 0 || vgroup_lds_bytes(true, P.group_max_tiles, P.max_big_nsh, P.max_big_nprim) <= P.lds_limit) &&
        (P.n_group_heavy == 0 || vgroup_lds_bytes(true, kSmallGroupTiles, P.max_heavy_nsh, P.max_heavy_nprim,
                                                  P.v_small_bf16) <= P.lds_limit) &&
        (n_small == 0 || vgroup_lds_bytes(true, kSmallGroupTiles, P.max_small_nsh, P.max_small_nprim,
                                          P.v_small_bf16) <= P.lds_limit);
    if (P.max_mpad_rho > 16384) return kLimit;  // the gather stages a block's AO indices in LDS
    P.reduce_waves = reduce_v_waves(P.max_mol_ao, P.lds_limit);
    if (P.reduce_waves < 1) return kLimit;
    P.reduce_segments = std::max(1, std::min(4, 2048 / std::max(n_ao, 1)));
    const size_t need = sizeof(double) * (size_t)scratch + sizeof(RedEntry) * hgg.size() +
                        sizeof(double) * (size_t)P.p_elems * 2 +
                        sizeof(double4) * (size_t)(2 * P.n_pts + n_pts) + 12 * (size_t)P.psub_elems;
    if (mem_limit > 0 && need > mem_limit) return kLimit;

    // The mixed build's AO cache: per rho workgroup tile, its sub-tiles' slots
    // in the kernel's lane order (aoc_sub_bytes).  Offsets in 16-byte units.
    std::vector<long long> haoc_off(htiles.size());
    size_t aoc_bytes = 0;
    for (size_t t = 0; t < htiles.size(); ++t) {
        const RhoBlock& B = hb[htiles[t].x];
        const int n_sub = std::min(kRhoSub * kRhoPts, B.n_pts - htiles[t].y) > kRhoPts ? 2 : 1;
        haoc_off[t] = static_cast<long long>(aoc_bytes >> 4);
        aoc_bytes += static_cast<size_t>(n_sub * aoc_sub_bytes(B.m_pad, B.m64));
    }
    bool want_aoc = aoc_mode != 1 && aoc_bytes > 0;
    if (want_aoc && aoc_mode == 0) {
        size_t free_b = 0, total_b = 0;
        want_aoc = hipMemGetInfo(&free_b, &total_b) == hipSuccess && aoc_bytes <= free_b / 4 &&
                   (mem_limit == 0 || need + aoc_bytes <= mem_limit);
    }
    if (want_aoc && aoc_mode == 2 && mem_limit > 0 && need + aoc_bytes > mem_limit) return kLimit;


    std::vector<int> hperm(P.n_pts);
    for (int i = 0; i < P.n_pts; ++i) hperm[i] = keep[i];
    ok = upload(P.perm, hperm) && upload(P.blocks, hb) && upload(P.block_rows, hrows) && upload(P.block_sh, hsh) &&
         upload(P.block_meta, hmeta) &&
         upload(P.rho_tiles, htiles) && upload(P.grad_tiles, hgtiles) && upload(P.class_blocks, hclsblk) && upload(P.group_ao, hgao) && upload(P.shell_atom, hshatom) &&
         upload(P.groups, hg) && upload(P.group_rows, hgrows) &&
         upload(P.group_sh, hgsh) && upload(P.group_meta, hgmeta) &&
         (P.reduce_segments == 1 ||
          P.red_part.alloc(sizeof(double) * (size_t)P.reduce_segments * n_ao * std::max(P.max_mol_ao, 1))) && upload(P.group_list, hglist) && upload(P.pairs, hpairs) &&
         upload(P.ao_gbegin, hgb) && upload(P.ao_groups, hgg) && P.scratch.alloc(sizeof(double) * (size_t)std::max<long long>(scratch, 1)) &&
         P.wv.alloc(sizeof(double4) * (size_t)std::max(P.n_pts, 1)) &&
         P.wv_caller.alloc(sizeof(double4) * (size_t)n_pts) && P.epart.alloc(sizeof(double) * (P.n_xc_wg + 1)) &&
         P.npart.alloc(sizeof(double) * (P.n_xc_wg + 1)) && P.en.alloc(sizeof(double) * 2 * n_mol) && upload(P.mols, P.mol_host) && upload(P.ao_mol, hao_mol) &&
         upload(P.xc_wg, hxc) &&
         P.P64.alloc(sizeof(double) * (size_t)P.p_elems) &&
         P.V.alloc(sizeof(double) * (size_t)P.p_elems) && P.rho.alloc(sizeof(double4) * (size_t)n_pts) &&
         P.rho_s.alloc(sizeof(double4) * (size_t)std::max(P.n_pts, 1)) &&
         P.psub64.alloc(sizeof(double) * (size_t)(P.psub_elems + kPsubTail)) &&
         P.psub32.alloc(sizeof(float) * (size_t)(P.psub_elems + kPsubTail));
    if (!ok) return kAllocFail;
    // Points of blocks with no significant AO are never written by the rho
    // kernel: their sorted rho stays zero for the life of the plan.
    if (hipMemset(P.rho_s.p, 0, P.rho_s.bytes) != hipSuccess) return kLaunchFail;
    if (hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking) != hipSuccess) return kNoDevice;
    for (int i = 0; i < Plan::kSide; ++i)
        if (hipStreamCreateWithFlags(&P.side[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&P.join_ev[i], hipEventDisableTiming) != hipSuccess)
            return kNoDevice;
    if (hipEventCreateWithFlags(&P.fork_ev, hipEventDisableTiming) != hipSuccess) return kNoDevice;
    if (want_aoc) {
        if (!upload(P.aoc_off, haoc_off) || !P.aoc.alloc(aoc_bytes)) {
            (void)hipGetLastError();  // clear the failed allocation's error
            P.aoc_off.reset();
            P.aoc.reset();
            if (aoc_mode == 2) return kAllocFail;
        } else {
            // one evaluation per geometry, by the mixed rho kernel's own AO code
            RhoArgs fa{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
                       P.rho_tiles.as<int2>(), P.blocks.as<RhoBlock>(), P.block_rows.as<int2>(),
                       P.block_sh.as<int2>(), P.block_meta.as<uint4>(), nullptr, nullptr, nullptr, nullptr, 0};
            fa.aoc = P.aoc.p;
            fa.aoc_off = P.aoc_off.as<long long>();
            for (size_t c = 0; c < P.classes.size(); ++c)
                if (P.class_fits[kRhoMixed][c] && launch_rho_fill_cache(fa, P.classes[c], P.stream) != hipSuccess)
                    return kLaunchFail;
        }
    }
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
    constexpr int K = Profile::kMarks;
    for (size_t i = 0; i + K <= ev.size(); i += K) {
        hipEventSynchronize(ev[i + K - 1]);
        for (int k = 0; k < K - 1; ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[i + k], ev[i + k + 1]) == hipSuccess) P.prof.ms[k] += ms;
        }
        P.prof.builds++;
    }
    for (hipEvent_t e : ev) hipEventDestroy(e);
    ev.clear();
}

// Rho stage: gather the density blocks, then every launch class with its own
// LDS budget.  rho goes to `sorted` per sorted point, or, when sorted is
// null, to P.rho per caller point (points dropped at plan time keep what
// P.rho held).  With gather_mark the profiling mark after the gather is set.
// Runs fn(i, stream) for i in [0, n): i = 0 on s, the others on the plan's
// side streams, forked from s after what s has queued so far; s then waits
// for all of them.  The launches must be independent of each other.
template <class F>
int fan_out(Plan& P, hipStream_t s, int n, F fn) {
    if (n <= 1) return n == 1 ? fn(0, s) : kOk;
    if (hipEventRecord(P.fork_ev, s) != hipSuccess) return kLaunchFail;
    const int used = std::min(n - 1, Plan::kSide);
    for (int j = 0; j < used; ++j)
        if (hipStreamWaitEvent(P.side[j], P.fork_ev, 0) != hipSuccess) return kLaunchFail;
    for (int i = 0; i < n; ++i) {
        const int rc = fn(i, i == 0 ? s : P.side[(i - 1) % Plan::kSide]);
        if (rc != kOk) return rc;
    }
    for (int j = 0; j < used; ++j)
        if (hipEventRecord(P.join_ev[j], P.side[j]) != hipSuccess || hipStreamWaitEvent(s, P.join_ev[j], 0) != hipSuccess)
            return kLaunchFail;
    return kOk;
}

int run_rho(Plan& P, int mode, double4* sorted, const double* dP, hipStream_t s, bool gather_mark,
            double4* sorted_b = nullptr, const double* dPb = nullptr) {
    const bool two = dPb != nullptr;
    for (size_t c = 0; c < P.classes.size(); ++c)
        if (!P.class_fits[mode][c]) return kLimit;
    // The density blocks are gathered class by class, each on the stream of
    // its class's rho launch, so one class's gather (memory-bound) runs beside
    // another class's contraction.  With profiling marks on they are gathered
    // in one launch first instead, so that the marks time the gather and the
    // rho kernels apart.
    auto gather = [&](const int* list, int n_blk, int max_mpad, hipStream_t st) {
        GatherArgs ga{P.blocks.as<RhoBlock>(), P.block_rows.as<int2>(), P.mols.as<MolInfo>(), dP,
                      P.psub64.as<double>(), mode == kRhoMixed ? P.psub32.as<float>() : nullptr,
                      mode == kRhoMixed ? 1 : 0, list};
        if (launch_gather_p(ga, n_blk, max_mpad, st) != hipSuccess) return false;
        if (!two) return true;
        GatherArgs gb = ga;
        gb.P = dPb;
        gb.out64 = P.psub64b.as<double>();
        gb.out32 = mode == kRhoMixed ? P.psub32b.as<float>() : nullptr;
        return launch_gather_p(gb, n_blk, max_mpad, st) == hipSuccess;
    };
    const bool per_class = !P.prof.on && P.classes.size() > 1;
    if (!per_class && !gather(nullptr, P.n_blocks, P.max_mpad_rho, s)) return kLaunchFail;
    if (gather_mark) mark(P, s);
    const double* dP64 = P.psub64.as<double>();
    const void* dPS = mode == kRhoMixed ? static_cast<const void*>(P.psub32.p) : P.psub64.p;
    // the LDS classes cover disjoint tiles: side by side (their tails overlap)
    return fan_out(P, s, static_cast<int>(P.classes.size()), [&](int i, hipStream_t st) {
        const RhoClass& cl = P.classes[P.class_order[i]];
        if (per_class && !gather(P.class_blocks.as<int>() + cl.block_begin, cl.n_blocks, cl.max_mpad, st))
            return kLaunchFail;
        RhoArgs ra{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
                   P.rho_tiles.as<int2>(), P.blocks.as<RhoBlock>(), P.block_rows.as<int2>(), P.block_sh.as<int2>(),
                   P.block_meta.as<uint4>(), dP64, dPS, sorted ? sorted : P.rho.as<double4>(), P.perm.as<int>(),
                   sorted ? 0 : 1};
        if (mode == kRhoMixed && P.aoc.p) {
            ra.aoc = P.aoc.p;
            ra.aoc_off = P.aoc_off.as<long long>();
        }
        if (!two) return launch_rho(ra, mode, cl, st) == hipSuccess ? kOk : kLaunchFail;
        // both spins from one AO evaluation when the class's LDS allows,
        // else one pass per spin
        if (rho_lds_bytes(mode, cl, 2) <= P.lds_limit) {
            ra.P64b = P.psub64b.as<double>();
            ra.PSb = mode == kRhoMixed ? static_cast<const void*>(P.psub32b.p) : P.psub64b.p;
            ra.rho_out_b = sorted_b;
            return launch_rho(ra, mode, cl, st, 2) == hipSuccess ? kOk : kLaunchFail;
        }
        if (launch_rho(ra, mode, cl, st) != hipSuccess) return kLaunchFail;
        RhoArgs rb = ra;
        rb.P64 = P.psub64b.as<double>();
        rb.PS = mode == kRhoMixed ? static_cast<const void*>(P.psub32b.p) : P.psub64b.p;
        rb.rho_out = sorted_b;
        return launch_rho(rb, mode, cl, st) == hipSuccess ? kOk : kLaunchFail;
    });
}

// V stage: register-resident kernel over its groups, tile-pair kernel over
// the rest, then the deterministic reduction into the full symmetric V.
// With fused set this is part of a fused build and records its profiling mark.
int run_v(Plan& P, bool f32, const double4* wv, int wv_perm, double* dV, hipStream_t s, bool fused,
          double* scratch = nullptr, double* red_part = nullptr) {
    if (f32 && !P.v_fits_fp32) return kLimit;
    if (!scratch) scratch = P.scratch.as<double>();
    if (!red_part) red_part = P.red_part.as<double>();
    VArgs va{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
             P.groups.as<VGroup>(), P.pairs.as<VPair>(), P.group_list.as<int>(), P.group_rows.as<int2>(),
             P.group_sh.as<int2>(), P.group_meta.as<uint4>(), wv, P.perm.as<int>(), wv_perm, scratch};
    if (launch_vxc_group(va, f32, P.group_max_tiles, P.n_group_big, P.max_big_nsh, P.max_big_nprim, s) != hipSuccess)
        return kLaunchFail;
    VArgs vs = va;
    vs.group_list = va.group_list + P.n_group_big;
    if (launch_vxc_group(vs, f32, kSmallGroupTiles, P.n_group_heavy, P.max_heavy_nsh, P.max_heavy_nprim, s,
                         P.v_small_bf16) != hipSuccess)
        return kLaunchFail;
    vs.group_list = va.group_list + P.n_group_big + P.n_group_heavy;
    if (launch_vxc_group(vs, f32, kSmallGroupTiles, P.n_group_list - P.n_group_big - P.n_group_heavy,
                         P.max_small_nsh, P.max_small_nprim, s, P.v_small_bf16) != hipSuccess)
        return kLaunchFail;
    if (launch_vxc(va, f32, P.n_pairs, s) != hipSuccess) return kLaunchFail;
    if (fused) mark(P, s);
    ReduceArgs rd{P.ao_gbegin.as<int>(), P.ao_groups.as<RedEntry>(), P.group_ao.as<int>(), P.ao_mol.as<int>(),
                  P.mols.as<MolInfo>(), P.n_ao, scratch, dV, P.reduce_segments, std::max(P.max_mol_ao, 1), red_part};
    return launch_reduce_v(rd, P.reduce_waves, P.max_mol_ao, s) == hipSuccess ? kOk : kLaunchFail;
}

// Fused RKS build on device pointers, asynchronous on stream s.
int run_build(Plan& P, int func, int prec, const double* dP, double* dV, double* den, hipStream_t s) {
    if (func < 0 || func > 5) return kUnsupported;
    if (prec != CALCHEM_PREC_FP64 && prec != CALCHEM_PREC_FP32) return kUnsupported;
    if (!P.has_weights) return kBadArgs;
    const bool f32 = prec == CALCHEM_PREC_FP32;
    mark(P, s);
    const int rc = run_rho(P, f32 ? kRhoMixed : kRhoFp64, P.rho_s.as<double4>(), dP, s, true);
    if (rc != kOk) return rc;
    mark(P, s);
    XcArgs xa{P.rho_s.as<double4>(), P.pts.as<double4>(), P.xc_wg.as<int2>(), P.n_xc_wg, func, P.wv.as<double4>(),
              P.epart.as<double>(),
              P.npart.as<double>()};
    if (launch_xc(xa, s) != hipSuccess) return kLaunchFail;
    mark(P, s);
    const int rv = run_v(P, f32, P.wv.as<double4>(), 0, dV, s, true);
    if (rv != kOk) return rv;
    mark(P, s);
    if (launch_reduce_en(P.epart.as<double>(), P.npart.as<double>(), P.mols.as<MolInfo>(), P.n_mol, den, s) !=
        hipSuccess)
        return kLaunchFail;
    mark(P, s);
    return kOk;
}

// Unrestricted build (UksVxcEngine::evaluate, vxc.rs:1662-1766): both spin
// densities through the rho stage, the unrestricted functional, then the V
// stage once per spin.  Profiling marks: start, after the first gather, after
// both rho passes, after the functional, after both V passes (with their
// reductions), an empty reduce interval, after the E/N reduction.
int run_build_uks(Plan& P, int func, int prec, const double* dPa, const double* dPb, double* dVa, double* dVb,
                  double* den, hipStream_t s) {
    if (func < 0 || func > 5) return kUnsupported;
    if (prec != CALCHEM_PREC_FP64 && prec != CALCHEM_PREC_FP32) return kUnsupported;
    if (!P.has_weights) return kBadArgs;
    if (!P.uks_ready) {
        const size_t b = sizeof(double4) * (size_t)std::max(P.n_pts, 1);
        if (!P.rho_sb.alloc(b) || !P.wv_b.alloc(b) ||
            (!P.psub64b.p && !P.psub64b.alloc(sizeof(double) * (size_t)(P.psub_elems + kPsubTail))) ||
            !P.psub32b.alloc(sizeof(float) * (size_t)(P.psub_elems + kPsubTail)) ||
            !P.scratch_b.alloc(P.scratch.bytes) || !P.red_part_b.alloc(P.red_part.bytes))
            return kAllocFail;
        // Points of blocks without significant AOs are never written.
        if (hipMemset(P.rho_sb.p, 0, b) != hipSuccess) return kLaunchFail;
        P.uks_ready = true;
    }
    const bool f32 = prec == CALCHEM_PREC_FP32;
    const int mode = f32 ? kRhoMixed : kRhoFp64;
    mark(P, s);
    int rc = run_rho(P, mode, P.rho_s.as<double4>(), dPa, s, true, P.rho_sb.as<double4>(), dPb);
    if (rc != kOk) return rc;
    mark(P, s);
    XcUksArgs xa{P.rho_s.as<double4>(), P.rho_sb.as<double4>(), P.pts.as<double4>(), P.xc_wg.as<int2>(), P.n_xc_wg, func,
                 P.wv.as<double4>(), P.wv_b.as<double4>(), P.epart.as<double>(), P.npart.as<double>()};
    if (launch_xc_uks(xa, s) != hipSuccess) return kLaunchFail;
    mark(P, s);
    // the two spins' V stages side by side (each with its own tiles and
    // reduction rows), so that one's tail overlaps the other's launches
    rc = fan_out(P, s, 2, [&](int i, hipStream_t st) {
        return i == 0 ? run_v(P, f32, P.wv.as<double4>(), 0, dVa, st, false)
                      : run_v(P, f32, P.wv_b.as<double4>(), 0, dVb, st, false, P.scratch_b.as<double>(),
                              P.red_part_b.as<double>());
    });
    if (rc != kOk) return rc;
    mark(P, s);
    mark(P, s);
    if (launch_reduce_en(P.epart.as<double>(), P.npart.as<double>(), P.mols.as<MolInfo>(), P.n_mol, den, s) !=
        hipSuccess)
        return kLaunchFail;
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
    const std::vector<MolSpec> one{MolSpec{0, n_shells, 0, n_pts, n_ao}};
    return create_plan(shells, n_shells, prims, n_prims, comps, n_comps, xyz.data(), 4, n_pts, one, false, &o,
                       grid_id, true);
}

int calchem_hip_vxc_eval_rho(uint64_t grid_id, const double* P_mat, double* rho_grad, int n_pts, int n_ao) {
    if (!P_mat || !rho_grad || n_pts <= 0 || n_ao <= 0 || grid_id == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(grid_id);
    if (!plan || plan->n_in != n_pts || plan->n_ao != n_ao) return kMismatch;
    Plan& P = *plan;
    if (P.n_mol != 1) return kMismatch;  // batched plans: calchem_hip_vxc_build_batched*
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    const size_t nn = (size_t)n_ao * n_ao;
    hipStream_t s = P.stream;
    if (hipMemcpyAsync(P.P64.p, P_mat, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess) return kLaunchFail;
    if (launch_fill_zero(P.rho.as<double>(), 4LL * n_pts, s) != hipSuccess) return kLaunchFail;
    const int rc = run_rho(P, kRhoFp64, nullptr, P.P64.as<double>(), s, false);
    if (rc != kOk) return rc;
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
    if (P.n_mol != 1) return kMismatch;  // batched plans: calchem_hip_vxc_build_batched*
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    hipStream_t s = P.stream;
    const size_t nn = (size_t)n_ao * n_ao;
    if (hipMemcpyAsync(P.wv_caller.p, wv_in, sizeof(double4) * (size_t)n_pts, hipMemcpyHostToDevice, s) != hipSuccess)
        return kLaunchFail;
    const int rv = run_v(P, false, P.wv_caller.as<double4>(), 1, P.V.as<double>(), s, false);
    if (rv != kOk) return rv;
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
    if (n_shells <= 0 || n_pts <= 0 || n_ao <= 0) return kBadArgs;
    const std::vector<MolSpec> one{MolSpec{0, n_shells, 0, n_pts, n_ao}};
    return create_plan(shells, n_shells, prims, n_prims, comps, n_comps, xyzw, 4, n_pts, one, true, opts, handle,
                       false);
}

int calchem_hip_vxc_create_batched(int n_mol, const int* mol_shell_begin, const int* mol_point_begin,
                                   const int* mol_n_ao, const CalchemAoShell* shells, const CalchemAoPrim* prims,
                                   int n_prims, const CalchemAoComp* comps, int n_comps, const double* xyzw,
                                   const CalchemVxcOptions* opts, uint64_t handle) {
    if (n_mol <= 0 || !mol_shell_begin || !mol_point_begin || !mol_n_ao) return kBadArgs;
    std::vector<MolSpec> ms(n_mol);
    for (int k = 0; k < n_mol; ++k)
        ms[k] = MolSpec{mol_shell_begin[k], mol_shell_begin[k + 1], mol_point_begin[k], mol_point_begin[k + 1],
                        mol_n_ao[k]};
    return create_plan(shells, mol_shell_begin[n_mol], prims, n_prims, comps, n_comps, xyzw, 4,
                       mol_point_begin[n_mol], ms, true, opts, handle, false);
}

int calchem_hip_vxc_build(uint64_t handle, int xc, int precision, const double* P_mat, double* V_out, double* exc,
                          double* nelec) {
    if (!P_mat || !V_out || !exc || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    if (P.n_mol != 1) return kMismatch;  // batched plans: calchem_hip_vxc_build_batched*
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
    if (P.n_mol != 1) return kMismatch;  // batched plans: calchem_hip_vxc_build_batched*
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    return run_build(P, xc, precision, d_P, d_V, d_exc_nelec, static_cast<hipStream_t>(stream));
}

int calchem_hip_vxc_build_uks(uint64_t handle, int xc, int precision, const double* Pa, const double* Pb,
                              double* Va_out, double* Vb_out, double* exc, double* nelec) {
    if (!Pa || !Pb || !Va_out || !Vb_out || !exc || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    if (P.n_mol != 1) return kMismatch;  // batched plans: calchem_hip_vxc_build_batched*
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    hipStream_t s = P.stream;
    const size_t nn = (size_t)P.n_ao * P.n_ao;
    // The host-buffer form stages the beta density and V in buffers of its
    // own, allocated on its first call.
    if (!P.Pb_host.p && (!P.Pb_host.alloc(sizeof(double) * nn) || !P.Vb_host.alloc(sizeof(double) * nn)))
        return kAllocFail;
    double* dPb = P.Pb_host.as<double>();
    if (hipMemcpyAsync(P.P64.p, Pa, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dPb, Pb, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess)
        return kLaunchFail;
    const int rc = run_build_uks(P, xc, precision, P.P64.as<double>(), dPb, P.V.as<double>(),
                                 P.Vb_host.as<double>(), P.en.as<double>(), s);
    if (rc != kOk) return rc;
    double en[2] = {0.0, 0.0};
    if (hipMemcpyAsync(Va_out, P.V.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(Vb_out, P.Vb_host.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(en, P.en.p, sizeof(double) * 2, hipMemcpyDeviceToHost, s) != hipSuccess)
        return kLaunchFail;
    if (hipStreamSynchronize(s) != hipSuccess) return kLaunchFail;
    *exc = en[0];
    if (nelec) *nelec = en[1];
    return kOk;
}

int calchem_hip_vxc_build_uks_device(uint64_t handle, int xc, int precision, const double* d_Pa, const double* d_Pb,
                                     double* d_Va, double* d_Vb, double* d_exc_nelec, void* stream) {
    if (!d_Pa || !d_Pb || !d_Va || !d_Vb || !d_exc_nelec || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    if (P.n_mol != 1) return kMismatch;  // unrestricted builds: single-molecule plans only
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    return run_build_uks(P, xc, precision, d_Pa, d_Pb, d_Va, d_Vb, d_exc_nelec, static_cast<hipStream_t>(stream));
}

int calchem_hip_vxc_build_batched(uint64_t handle, int xc, int precision, const double* P_packed,
                                  double* V_packed, double* exc, double* nelec) {
    if (!P_packed || !V_packed || !exc || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    hipStream_t s = P.stream;
    const size_t ne = (size_t)P.p_elems;
    if (hipMemcpyAsync(P.P64.p, P_packed, sizeof(double) * ne, hipMemcpyHostToDevice, s) != hipSuccess)
        return kLaunchFail;
    const int rc = run_build(P, xc, precision, P.P64.as<double>(), P.V.as<double>(), P.en.as<double>(), s);
    if (rc != kOk) return rc;
    std::vector<double> en(2 * (size_t)P.n_mol);
    if (hipMemcpyAsync(V_packed, P.V.p, sizeof(double) * ne, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(en.data(), P.en.p, sizeof(double) * en.size(), hipMemcpyDeviceToHost, s) != hipSuccess)
        return kLaunchFail;
    if (hipStreamSynchronize(s) != hipSuccess) return kLaunchFail;
    for (int k = 0; k < P.n_mol; ++k) {
        exc[k] = en[2 * k];
        if (nelec) nelec[k] = en[2 * k + 1];
    }
    return kOk;
}

int calchem_hip_vxc_build_batched_device(uint64_t handle, int xc, int precision, const double* d_P_packed,
                                         double* d_V_packed, double* d_exc_nelec, void* stream) {
    if (!d_P_packed || !d_V_packed || !d_exc_nelec || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    return run_build(P, xc, precision, d_P_packed, d_V_packed, d_exc_nelec, static_cast<hipStream_t>(stream));
}

int calchem_hip_vxc_graph_create(uint64_t handle, int xc, int precision, const double* d_P, double* d_V,
                                 double* d_exc_nelec, uint64_t* graph_out) {
    if (!d_P || !d_V || !d_exc_nelec || !graph_out || handle == 0) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    hipStream_t s = P.stream;
    // One eager build first: launch attributes are set and every buffer the
    // build touches exists before capture.
    int rc = run_build(P, xc, precision, d_P, d_V, d_exc_nelec, s);
    if (rc != kOk) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return kLaunchFail;
    const bool prof = P.prof.on;
    P.prof.on = false;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        P.prof.on = prof;
        return kLaunchFail;
    }
    rc = run_build(P, xc, precision, d_P, d_V, d_exc_nelec, s);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    P.prof.on = prof;
    if (rc != kOk || ec != hipSuccess || !g) {
        if (g) hipGraphDestroy(g);
        return rc != kOk ? rc : kLaunchFail;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess || !ex) return kLaunchFail;
    P.graphs.push_back(ex);
    *graph_out = P.graphs.size();
    return kOk;
}

int calchem_hip_vxc_graph_launch(uint64_t handle, uint64_t graph, void* stream) {
    if (handle == 0 || graph == 0) return kBadArgs;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    if (graph > P.graphs.size() || !P.graphs[graph - 1]) return kMismatch;
    hipSetDevice(P.device);
    return hipGraphLaunch(P.graphs[graph - 1], static_cast<hipStream_t>(stream)) == hipSuccess ? kOk : kLaunchFail;
}

int calchem_hip_vxc_graph_destroy(uint64_t handle, uint64_t graph) {
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    std::lock_guard<std::mutex> lk(P.mu);
    if (graph == 0 || graph > P.graphs.size()) return kMismatch;
    hipSetDevice(P.device);
    if (P.graphs[graph - 1]) {
        hipDeviceSynchronize();
        hipGraphExecDestroy(P.graphs[graph - 1]);
        P.graphs[graph - 1] = nullptr;
    }
    return kOk;
}

int calchem_hip_vxc_set_atoms(uint64_t handle, int n_atoms, const int* z, const double* xyz_bohr,
                              const double* raw_weight, const int* host_atom) {
    if (handle == 0 || n_atoms <= 0 || !z || !xyz_bohr || !raw_weight || !host_atom) return kBadArgs;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    if (P.n_mol != 1) return kUnsupported;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    if (n_atoms > 64 * 4) return kLimit;  // Becke derivative: up to 4 atoms per lane
    std::vector<int> sa(P.n_shells);
    if (hipMemcpy(sa.data(), P.shell_atom.p, sizeof(int) * sa.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return kLaunchFail;
    for (int a : sa)
        if (a < 0 || a >= n_atoms) return kMismatch;
    std::vector<int> perm(P.n_pts), host(P.n_pts);
    std::vector<double> raw(P.n_pts);
    if (P.n_pts > 0 &&
        hipMemcpy(perm.data(), P.perm.p, sizeof(int) * perm.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return kLaunchFail;
    for (int i = 0; i < P.n_pts; ++i) {
        const int c = perm[i];
        if (host_atom[c] < 0 || host_atom[c] >= n_atoms || !std::isfinite(raw_weight[c])) return kBadArgs;
        host[i] = host_atom[c];
        raw[i] = raw_weight[c];
    }
    std::vector<double> inv((size_t)n_atoms * n_atoms), asy((size_t)n_atoms * n_atoms);
    becke_pair_tables(n_atoms, z, xyz_bohr, inv.data(), asy.data());
    std::vector<double> xyz(xyz_bohr, xyz_bohr + 3 * (size_t)n_atoms);
    if (!upload(P.atom_xyz, xyz) || !upload(P.inv_r, inv) || !upload(P.asym, asy) || !upload(P.raw_s, raw) ||
        !upload(P.host_s, host))
        return kAllocFail;
    P.n_atoms = n_atoms;
    return kOk;
}

int calchem_hip_vxc_xc_gradient(uint64_t handle, int xc, const double* Pa, const double* Pb, double* grad) {
    if (handle == 0 || !Pa || !grad) return kBadArgs;
    if (xc < 0 || xc > 5) return kUnsupported;
    if (!device_available()) return kNoDevice;
    auto plan = find_plan(handle);
    if (!plan) return kMismatch;
    Plan& P = *plan;
    if (P.n_mol != 1) return kUnsupported;
    std::lock_guard<std::mutex> lk(P.mu);
    hipSetDevice(P.device);
    if (P.n_atoms <= 0 || !P.has_weights) return kMismatch;
    const int na = P.n_atoms, width = 3 * na;
    for (size_t c = 0; c < P.classes.size(); ++c)
        if (grad_ao_lds_bytes(P.classes[c].max_mpad, P.classes[c].max_nsh, P.classes[c].max_prim) > P.lds_limit)
            return kLimit;
    const bool uks = Pb != nullptr;
    const size_t nn = (size_t)P.n_ao * P.n_ao;
    const long long n_tiles = (long long)P.n_rho_tiles, passes = uks ? 2 : 1;
    // Becke-derivative points per wave: enough waves to fill the GPU several
    // times over (up to 65536 of them, which bounds the partial rows)
    int pts_per_wave = 16;
    while (pts_per_wave < 256 && (P.n_pts + pts_per_wave - 1) / pts_per_wave > 65536) pts_per_wave *= 2;
    const long long n_waves = (P.n_pts + pts_per_wave - 1) / pts_per_wave;
    const long long rows = n_tiles * passes + n_waves;
    const long long per_chunk = 512, chunks = (rows + per_chunk - 1) / per_chunk;
    hipStream_t s = P.stream;
    const size_t pb = sizeof(double4) * (size_t)std::max(P.n_pts, 1);
    if (!P.gwx_a.p && (!P.gwx_a.alloc(pb) || !P.gwx_b.alloc(pb) ||
                       !P.gfxc.alloc(sizeof(double) * (size_t)std::max(P.n_pts, 1))))
        return kAllocFail;
    if (!P.rho_sb.p) {
        if (!P.rho_sb.alloc(pb) || hipMemset(P.rho_sb.p, 0, pb) != hipSuccess) return kAllocFail;
    }
    if (P.gpart.bytes < sizeof(double) * (size_t)rows * width && !P.gpart.alloc(sizeof(double) * (size_t)rows * width))
        return kAllocFail;
    if (P.gsum.bytes < sizeof(double) * (size_t)(chunks + 1) * width &&
        !P.gsum.alloc(sizeof(double) * (size_t)(chunks + 1) * width))
        return kAllocFail;
    if (uks && !P.Pb_host.p && (!P.Pb_host.alloc(sizeof(double) * nn) || !P.Vb_host.alloc(sizeof(double) * nn)))
        return kAllocFail;
    if (hipMemcpyAsync(P.P64.p, Pa, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess) return kLaunchFail;
    if (uks && hipMemcpyAsync(P.Pb_host.p, Pb, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess)
        return kLaunchFail;
    const double* dPa = P.P64.as<double>();
    const double* dPb = uks ? P.Pb_host.as<double>() : nullptr;
    // Densities at the points (both spins from one AO evaluation), then the functional.
    if (uks && !P.psub64b.p && !P.psub64b.alloc(sizeof(double) * (size_t)(P.psub_elems + kPsubTail)))
        return kAllocFail;
    int rc = run_rho(P, kRhoFp64, P.rho_s.as<double4>(), dPa, s, false, uks ? P.rho_sb.as<double4>() : nullptr, dPb);
    if (rc != kOk) return rc;
    XcGradArgs xa{P.rho_s.as<double4>(), uks ? P.rho_sb.as<double4>() : nullptr, P.pts.as<double4>(), P.n_pts, xc,
                  P.gwx_a.as<double4>(), P.gwx_b.as<double4>(), P.gfxc.as<double>()};
    if (launch_xc_grad(xa, s) != hipSuccess) return kLaunchFail;
    // AO terms, one pass per spin (restricted: the total density once).
    for (int pass = 0; pass < passes; ++pass) {
        const double* dP = pass == 0 ? dPa : dPb;
        GatherArgs ga{P.blocks.as<RhoBlock>(), P.block_rows.as<int2>(), P.mols.as<MolInfo>(), dP,
                      P.psub64.as<double>(), nullptr, 0};
        if (launch_gather_p(ga, P.n_blocks, P.max_mpad_rho, s) != hipSuccess) return kLaunchFail;
        for (size_t c = 0; c < P.classes.size(); ++c) {
            const RhoClass& cl = P.classes[c];
            GradAoArgs g{P.shells.as<DevShell>(), P.prims.as<double2>(), P.aos.as<DevAo>(), P.pts.as<double4>(),
                         P.grad_tiles.as<int2>() + cl.grad_tile_begin, P.blocks.as<RhoBlock>(), P.block_rows.as<int2>(),
                         P.block_sh.as<int2>(), P.shell_atom.as<int>(), P.psub64.as<double>(),
                         pass == 0 ? P.gwx_a.as<double4>() : P.gwx_b.as<double4>(), P.host_s.as<int>(), na, -2.0,
                         P.gpart.as<double>() + (size_t)(pass * n_tiles + cl.grad_tile_begin) * width};
            if (launch_grad_ao(g, cl.grad_n_tiles, grad_ao_lds_bytes(cl.max_mpad, cl.max_nsh, cl.max_prim), s) !=
                hipSuccess)
                return kLaunchFail;
        }
    }
    // Becke weight derivatives.
    GradBeckeArgs gb{P.pts.as<double4>(), P.raw_s.as<double>(), P.host_s.as<int>(), P.gfxc.as<double>(),
                     P.atom_xyz.as<double>(), P.inv_r.as<double>(), P.asym.as<double>(), na, P.n_pts, pts_per_wave,
                     P.gpart.as<double>() + (size_t)(passes * n_tiles) * width};
    if (launch_grad_becke(gb, (int)n_waves, s) != hipSuccess) return kLaunchFail;
    // Fixed-order sum of all partial rows.
    if (launch_grad_sum(P.gpart.as<double>(), rows, width, per_chunk, P.gsum.as<double>(), s) != hipSuccess ||
        launch_grad_sum(P.gsum.as<double>(), chunks, width, chunks, P.gsum.as<double>() + (size_t)chunks * width, s) !=
            hipSuccess)
        return kLaunchFail;
    if (hipMemcpyAsync(grad, P.gsum.as<double>() + (size_t)chunks * width, sizeof(double) * width,
                       hipMemcpyDeviceToHost, s) != hipSuccess)
        return kLaunchFail;
    return hipStreamSynchronize(s) == hipSuccess ? kOk : kLaunchFail;
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
    out->n_rho_tiles = P.n_rho_tiles;
    out->n_v_pairs = P.n_pairs;
    out->flops_rho = P.flops_rho;
    out->flops_v = P.flops_v;
    out->flops_rho_executed_fp64 = P.flops_rho_pad;
    out->flops_rho_executed_fp32 = P.flops_rho_pad;
    out->flops_v_executed = P.flops_v_pad;
    out->max_m_rho = P.max_mpad_rho;
    out->max_m_v = P.max_m_v;
    out->rho_points_fp64 = kRhoPts;
    out->rho_points_fp32 = kRhoPts;
    out->flops_rho_fp64_rows = P.flops_rho_fp64_rows;
    out->device_bytes = P.device_bytes();
    out->ao_cache_bytes = P.aoc.bytes;
    out->lds_bytes_per_block = P.lds_limit;
    out->profiled_builds = P.prof.builds;
    out->ms_gather = P.prof.ms[0];
    out->ms_rho = P.prof.ms[1];
    out->ms_xc = P.prof.ms[2];
    out->ms_vxc = P.prof.ms[3];
    out->ms_reduce_v = P.prof.ms[4];
    out->ms_reduce_en = P.prof.ms[5];
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

long calchem_hip_grid_streaming_tasks(int n_atoms, const int* z, int level, int n_rad, int radial_min, int* tasks,
                                      long max_tasks) {
    if (n_atoms <= 0 || !z || level < 0 || level > 4 || (level == 4 && n_rad <= 0)) return kBadArgs;
    long cnt = 0;
    for (int a = 0; a < n_atoms; ++a) {
        const int nr = atom_radial_count(z[a], level, n_rad, radial_min);
        for (int s0 = 1; s0 <= nr; s0 += kStreamingRadialShells) {
            if (tasks && cnt < max_tasks) {
                tasks[3 * cnt] = a;
                tasks[3 * cnt + 1] = s0;
                tasks[3 * cnt + 2] = std::min(s0 + kStreamingRadialShells - 1, nr);
            }
            ++cnt;
        }
    }
    if (tasks && cnt > max_tasks) return kLimit;
    return cnt;
}

long calchem_hip_grid_task_points(int n_atoms, const int* z, const double* xyz_bohr, int level, int n_rad, int n_ang,
                                  int radial_min, int atom, int rad_start, int rad_end, double* out, long max_out) {
    if (atom < 0 || atom >= n_atoms || rad_start < 1 || rad_end < rad_start) return kBadArgs;
    if (!device_available()) return kNoDevice;
    int st = kOk;
    const long cnt = build_grid_task(n_atoms, z, xyz_bohr, level, n_rad, n_ang, radial_min, atom, rad_start, rad_end,
                                     out, max_out, &st);
    if (st != kOk) return st;
    if (!out) return cnt;
    return cnt > max_out ? kLimit : cnt;
}

}  // extern "C"
