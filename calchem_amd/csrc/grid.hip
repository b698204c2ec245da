// grid.hip — molecular DFT grid: Murray-Handy-Laming radial shells x
// Lebedev-Laikov spheres per atom, Becke (k = 3) partition weights on the GPU.
//
// Semantics follow crates/calchem-dft/src/grid.rs:
//   Bragg-Slater radii (Angstrom, H = 0.35)          grid.rs:58-101
//   heteronuclear size adjustment a_ij in [-0.5,0.5] grid.rs:104-115
//   radial shell counts and angular pruning          grid.rs:138-220
//   pair table and screened fast partition           grid.rs:268-419
//   point generation order (atom, radial, angular),
//   r = R t^2/(m-t)^2, w = 2 R^3 m t^5/(m-t)^7, keep w > 1e-18   grid.rs:732-782
// and the Lebedev-Laikov orbit generators of crates/calchem-dft/src/lebedev.rs:69-250
// (parameters from Lebedev & Laikov, Doklady Math. 59, 477 (1999)).
#include <cmath>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace chx {

namespace {

constexpr double kAngToBohr = 1.8897261245650618;

// Orbit types: 1 = (1,0,0) x6, 2 = (s,s,0) x12, 3 = (s,s,s) x8,
// 4 = (a,a,b) x24, 5 = (a,b,0) x24, 6 = (a,b,c) x48.
struct Orbit {
    int kind;
    double a, b, v;
};

const Orbit kLeb26[] = {{1, 0, 0, 0.4761904761904762e-1}, {2, 0, 0, 0.3809523809523810e-1},
                        {3, 0, 0, 0.3214285714285714e-1}};
const Orbit kLeb50[] = {{1, 0, 0, 0.1269841269841270e-1}, {2, 0, 0, 0.2257495590828924e-1},
                        {3, 0, 0, 0.2109375000000000e-1}, {4, 0.3015113445777636, 0, 0.2017333553791887e-1}};
const Orbit kLeb110[] = {{1, 0, 0, 0.3828270494937162e-2},          {3, 0, 0, 0.9793737512487512e-2},
                         {4, 0.1851156353447362, 0, 0.8211737283191111e-2}, {4, 0.6904210483822922, 0, 0.9942814891178103e-2},
                         {4, 0.3956894730559419, 0, 0.9595471336070963e-2}, {5, 0.4783690288121502, 0, 0.9694996361663028e-2}};
const Orbit kLeb194[] = {{1, 0, 0, 0.1782340447244611e-2},
                         {2, 0, 0, 0.5716905949977102e-2},
                         {3, 0, 0, 0.5573383178848738e-2},
                         {4, 0.6712973442695226, 0, 0.5608704082587997e-2},
                         {4, 0.2892465627575439, 0, 0.5158237711805383e-2},
                         {4, 0.4446933178717437, 0, 0.5518771467273614e-2},
                         {4, 0.1299335447650067, 0, 0.4106777028169394e-2},
                         {5, 0.3457702197611283, 0, 0.5051846064614808e-2},
                         {6, 0.1590417105383530, 0.8360360154824589, 0.5530248916233094e-2}};
const Orbit kLeb302[] = {{1, 0, 0, 0.8545911725128148e-3},
                         {3, 0, 0, 0.3599119285025571e-2},
                         {4, 0.3515640345570105, 0, 0.3449788424305883e-2},
                         {4, 0.6566329410219612, 0, 0.3604822601419882e-2},
                         {4, 0.4729054132581005, 0, 0.3576729661743367e-2},
                         {4, 0.9618308522614784e-1, 0, 0.2352101413689164e-2},
                         {4, 0.2219645236294178, 0, 0.3108953122413675e-2},
                         {4, 0.7011766416089545, 0, 0.3650045807677255e-2},
                         {5, 0.2644152887060663, 0, 0.2982344963171804e-2},
                         {5, 0.5718955891878961, 0, 0.3600820932216460e-2},
                         {6, 0.2510034751770465, 0.8000727494073952, 0.3571540554273387e-2},
                         {6, 0.1233548532583327, 0.4127724083168531, 0.3392312205006170e-2}};
const Orbit kLeb590[] = {{1, 0, 0, 0.3095121295306187e-3},
                         {3, 0, 0, 0.1852379698597489e-2},
                         {4, 0.7040954938227469, 0, 0.1871790639277744e-2},
                         {4, 0.6807744066455243, 0, 0.1858812585438317e-2},
                         {4, 0.6372546939258752, 0, 0.1852028828296213e-2},
                         {4, 0.5044419707800358, 0, 0.1846715956151242e-2},
                         {4, 0.4215761784010967, 0, 0.1818471778162769e-2},
                         {4, 0.3317920736472123, 0, 0.1749564657281154e-2},
                         {4, 0.2384736701421887, 0, 0.1617210647254411e-2},
                         {4, 0.1459036449157763, 0, 0.1384737234851692e-2},
                         {4, 0.06095034115507196, 0, 0.9764331165051050e-3},
                         {5, 0.6116843442009876, 0, 0.1857161196774078e-2},
                         {5, 0.3964755348199858, 0, 0.1705153996395864e-2},
                         {5, 0.1724782009907724, 0, 0.1300321685886048e-2},
                         {6, 0.5610263808622060, 0.3518280927733519, 0.1842866472905286e-2},
                         {6, 0.4742392842551980, 0.2634716655937950, 0.1802658934377451e-2},
                         {6, 0.5984126497885380, 0.1816640840360209, 0.1849830560443660e-2},
                         {6, 0.3791035407695563, 0.1720795225656878, 0.1713904507106709e-2},
                         {6, 0.2778673190586244, 0.08213021581932511, 0.1555213603396808e-2},
                         {6, 0.5033564271075117, 0.08999205842074875, 0.1802239128008525e-2}};

struct Dir {
    double x, y, z, w;
};

void sign_cube(std::vector<Dir>& o, double x, double y, double z, double w) {
    for (int m = 0; m < 8; ++m) {  // sign bits (x, y, z) from most to least significant, minus first
        const double sx = (m & 4) ? 1.0 : -1.0, sy = (m & 2) ? 1.0 : -1.0, sz = (m & 1) ? 1.0 : -1.0;
        o.push_back({sx * x, sy * y, sz * z, w});
    }
}

void expand(std::vector<Dir>& o, const Orbit& ob) {
    const double w = 4.0 * M_PI * ob.v;
    switch (ob.kind) {
    case 1:
        for (double s : {-1.0, 1.0}) {
            o.push_back({s, 0.0, 0.0, w});
            o.push_back({0.0, s, 0.0, w});
            o.push_back({0.0, 0.0, s, w});
        }
        break;
    case 2: {
        const double h = std::sqrt(0.5);
        for (double s1 : {-h, h})
            for (double s2 : {-h, h}) {
                o.push_back({s1, s2, 0.0, w});
                o.push_back({s1, 0.0, s2, w});
                o.push_back({0.0, s1, s2, w});
            }
        break;
    }
    case 3: {
        const double t = std::sqrt(1.0 / 3.0);
        sign_cube(o, t, t, t, w);
        break;
    }
    case 4: {
        const double m = std::sqrt(1.0 - 2.0 * ob.a * ob.a);
        sign_cube(o, ob.a, ob.a, m, w);
        sign_cube(o, ob.a, m, ob.a, w);
        sign_cube(o, m, ob.a, ob.a, w);
        break;
    }
    case 5: {
        const double q = std::sqrt(1.0 - ob.a * ob.a);
        const double pr[2][2] = {{ob.a, q}, {q, ob.a}};
        for (const auto& ut : pr)
            for (double su : {-1.0, 1.0})
                for (double st : {-1.0, 1.0}) {
                    const double u = su * ut[0], t = st * ut[1];
                    o.push_back({u, t, 0.0, w});
                    o.push_back({u, 0.0, t, w});
                    o.push_back({0.0, u, t, w});
                }
        break;
    }
    default: {
        const double a = ob.a, b = ob.b, c = std::sqrt(1.0 - a * a - b * b);
        const double perms[6][3] = {{a, b, c}, {b, a, c}, {c, a, b}, {a, c, b}, {c, b, a}, {b, c, a}};
        for (const auto& p : perms) sign_cube(o, p[0], p[1], p[2], w);
        break;
    }
    }
}

bool lebedev_rule(int n, std::vector<Dir>& out) {
    out.clear();
    const Orbit* t = nullptr;
    int k = 0;
    switch (n) {
    case 26: t = kLeb26; k = 3; break;
    case 50: t = kLeb50; k = 4; break;
    case 110: t = kLeb110; k = 6; break;
    case 194: t = kLeb194; k = 9; break;
    case 302: t = kLeb302; k = 12; break;
    case 590: t = kLeb590; k = 20; break;
    default: return false;
    }
    for (int i = 0; i < k; ++i) expand(out, t[i]);
    return static_cast<int>(out.size()) == n;
}

double bragg_slater_angstrom(int z) {
    static const double tab[57] = {1.75, 0.35, 0.35, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.45, 1.80, 1.50,
                                   1.25, 1.10, 1.00, 1.00, 1.00, 1.00, 2.20, 1.80, 1.40, 1.40, 1.40, 1.40, 1.40,
                                   1.40, 1.40, 1.40, 1.40, 1.40, 1.30, 1.25, 1.15, 1.15, 1.15, 1.10, 2.35, 2.00,
                                   1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.55, 1.45, 1.45,
                                   1.40, 1.40, 1.35, 2.60, 2.15};
    return (z >= 1 && z <= 56) ? tab[z] : 1.75;
}

double becke_a(int zi, int zj) {
    if (zi == zj) return 0.0;
    const double chi = bragg_slater_angstrom(zi) / bragg_slater_angstrom(zj);
    const double u = (chi - 1.0) / (chi + 1.0);
    const double a = u / (u * u - 1.0);
    return std::fmin(0.5, std::fmax(-0.5, a));
}

int radial_shells(int z, int level) {
    const int row = z <= 2 ? 0 : (z <= 18 ? 1 : 2);
    static const int tab[3][3] = {{25, 35, 45}, {35, 50, 65}, {55, 75, 90}};
    return level <= 2 ? tab[level][row] : 200;
}

int pruned_rule(double r, double rbs, int level) {
    const double x = r / std::fmax(rbs, 0.3);
    switch (level) {
    case 0: return x < 0.25 ? 26 : x < 0.75 ? 50 : x < 4.5 ? 110 : 50;
    case 1: return x < 0.20 ? 26 : x < 0.55 ? 50 : x < 1.10 ? 110 : x < 4.50 ? 194 : 110;
    case 2: return x < 0.18 ? 50 : x < 0.50 ? 110 : x < 0.95 ? 194 : x < 5.00 ? 302 : 194;
    default: return 590;
    }
}

double ipow(double x, int n) {
    double r = 1.0;
    while (true) {
        if (n & 1) r *= x;
        n >>= 1;
        if (!n) break;
        x *= x;
    }
    return r;
}

// The partition is evaluated without fused multiply-add so that its
// threshold tests (1e-20 products, 1e-17 activity, w > 1e-18) see the same
// roundings as the reference's separately rounded IEEE operations.
#pragma clang fp contract(off)

__device__ __forceinline__ double step3(double nu) {
    if (nu <= -1.0) return 1.0;
    if (nu >= 1.0) return 0.0;
    double f = 0.5 * nu * (3.0 - nu * nu);
    f = 0.5 * f * (3.0 - f * f);
    f = 0.5 * f * (3.0 - f * f);
    return 0.5 * (1.0 - f);
}

__device__ __forceinline__ double dist_to(const double* __restrict__ xyz, int a, double x, double y, double z) {
    const double dx = x - xyz[3 * a], dy = y - xyz[3 * a + 1], dz = z - xyz[3 * a + 2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Becke cell weight of host atom h at point (x, y, z) with the reference's
// screening: host product exits at 1e-20, only partners with 1 - s > 1e-17
// enter the normalisation, partner products exit at 1e-20.
__device__ double becke_cell_weight(const double* __restrict__ xyz, const double* __restrict__ inv_r,
                                    const double* __restrict__ aij, int n, int h, double x, double y, double z) {
    if (n <= 1) return 1.0;
    const double rh = dist_to(xyz, h, x, y, z);
    double ph = 1.0;
    bool any = false;
    for (int b = 0; b < n; ++b) {
        if (b == h) continue;
        const double inv = inv_r[h * n + b];
        if (inv == 0.0) continue;
        const double mu = fmin(1.0, fmax(-1.0, (rh - dist_to(xyz, b, x, y, z)) * inv));
        const double s = step3(mu + aij[h * n + b] * (1.0 - mu * mu));
        ph *= s;
        if (ph < 1e-20) return 0.0;
        any |= (1.0 - s > 1e-17);
    }
    if (!any) return 1.0;
    double sum = ph;
    for (int a = 0; a < n; ++a) {
        if (a == h) continue;
        const double inv_ha = inv_r[h * n + a];
        if (inv_ha == 0.0) continue;
        const double ra = dist_to(xyz, a, x, y, z);
        const double mu_h = fmin(1.0, fmax(-1.0, (rh - ra) * inv_ha));
        const double s_ha = step3(mu_h + aij[h * n + a] * (1.0 - mu_h * mu_h));
        if (!(1.0 - s_ha > 1e-17)) continue;
        double pa = 1.0 - s_ha;
        if (pa < 1e-20) continue;
        for (int b = 0; b < n; ++b) {
            if (b == h || b == a) continue;
            const double inv = inv_r[a * n + b];
            if (inv == 0.0) continue;
            const double mu = fmin(1.0, fmax(-1.0, (ra - dist_to(xyz, b, x, y, z)) * inv));
            pa *= step3(mu + aij[a * n + b] * (1.0 - mu * mu));
            if (pa < 1e-20) break;
        }
        sum += pa;
    }
    return sum <= 0.0 ? 0.0 : ph / sum;
}

// One radial shell of one atom: its centre, radius, radial weight and its
// Lebedev rule (offset and size in the concatenated direction table); its
// raw points start at raw_off, its kept points at keep_off.
struct GridShell {
    double cx, cy, cz, r, w_rad;
    int atom, dir_off, n_dir;
    long long raw_off, keep_off;
};

constexpr int kGridThreads = 256;

// One workgroup per radial shell: points r u + centre, raw weight w_rad w_u,
// Becke weight; raw rows (x, y, z, w, raw_w, host) and the shell's count of
// points with w > 1e-18 (grid.rs:732-782).
__global__ __launch_bounds__(kGridThreads) void k_grid_shell(const GridShell* __restrict__ shells,
                                                             const double4* __restrict__ dirs,
                                                             const double* __restrict__ xyz,
                                                             const double* __restrict__ inv_r,
                                                             const double* __restrict__ aij, int n_atoms,
                                                             double* __restrict__ raw, int* __restrict__ keep) {
    __shared__ int cnt;
    const GridShell s = shells[blockIdx.x];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int local = 0;
    for (int d = threadIdx.x; d < s.n_dir; d += kGridThreads) {
        const double4 u = dirs[s.dir_off + d];
        const double x = s.cx + s.r * u.x, y = s.cy + s.r * u.y, z = s.cz + s.r * u.z;
        const double raw_w = s.w_rad * u.w;
        const double w = raw_w * becke_cell_weight(xyz, inv_r, aij, n_atoms, s.atom, x, y, z);
        double* o = raw + 6 * (s.raw_off + d);
        o[0] = x;
        o[1] = y;
        o[2] = z;
        o[3] = w;
        o[4] = raw_w;
        o[5] = (double)s.atom;
        local += w > 1e-18 ? 1 : 0;
    }
    atomicAdd(&cnt, local);
    __syncthreads();
    if (threadIdx.x == 0) keep[blockIdx.x] = cnt;
}

// One workgroup per radial shell: the kept rows, in direction order, to
// out[keep_off ..]: per 256-direction chunk an in-order prefix of the keep
// flags (wave ballots, then the waves in order).
__global__ __launch_bounds__(kGridThreads) void k_grid_compact(const GridShell* __restrict__ shells,
                                                               const double* __restrict__ raw,
                                                               double* __restrict__ out) {
    __shared__ int wave_cnt[kGridThreads / 64];
    const GridShell s = shells[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    long long base = s.keep_off;
    for (int d0 = 0; d0 < s.n_dir; d0 += kGridThreads) {
        const int d = d0 + threadIdx.x;
        const double* r = raw + 6 * (s.raw_off + d);
        const bool keep = d < s.n_dir && r[3] > 1e-18;
        const unsigned long long m = __ballot(keep);
        const int before = __popcll(m & ((1ULL << lane) - 1ULL));
        if (lane == 0) wave_cnt[wave] = __popcll(m);
        __syncthreads();
        int off = 0, tot = 0;
        for (int w = 0; w < kGridThreads / 64; ++w) {
            if (w < wave) off += wave_cnt[w];
            tot += wave_cnt[w];
        }
        if (keep) {
            double* o = out + 6 * (base + off + before);
            for (int k = 0; k < 6; ++k) o[k] = r[k];
        }
        base += tot;
        __syncthreads();
    }
}

}  // namespace

// Radial shell count of atom z at a level, with the basis-aware minimum and
// Fine's 99-shell floor (grid.rs:745-752); explicit grids use n_rad.
int atom_radial_count(int z, int level, int n_rad, int radial_min) {
    if (level == 4) return n_rad;
    int nr = radial_shells(z, level);
    if (radial_min > nr) nr = radial_min;
    if (level == 2 && nr < 99) nr = 99;
    return nr;
}

long build_grid_host(int n_atoms, const int* z, const double* xyz, int level, int n_rad, int n_ang,
                     int radial_min, double* out, long max_out, int* status) {
    return build_grid_task(n_atoms, z, xyz, level, n_rad, n_ang, radial_min, -1, 0, 0, out, max_out, status);
}

// Points of atom `task_atom`'s radial shells rad_start..rad_end (1-based,
// inclusive; every atom and shell when task_atom < 0), generated, weighted
// and filtered on the device in the reference's order (atom, radial,
// angular).  The host only lists the radial shells.
long build_grid_task(int n_atoms, const int* z, const double* xyz, int level, int n_rad, int n_ang, int radial_min,
                     int task_atom, int rad_start, int rad_end, double* out, long max_out, int* status) {
    *status = kOk;
    if (n_atoms <= 0 || !z || !xyz || level < 0 || level > 4 || (level == 4 && n_rad <= 0) ||
        (task_atom >= n_atoms)) {
        *status = kBadArgs;
        return 0;
    }
    const int sizes[6] = {26, 50, 110, 194, 302, 590};
    std::vector<double4> dirs;
    int dir_off[6];
    for (int k = 0; k < 6; ++k) {
        std::vector<Dir> rule;
        lebedev_rule(sizes[k], rule);
        dir_off[k] = (int)dirs.size();
        for (const Dir& d : rule) dirs.push_back(make_double4(d.x, d.y, d.z, d.w));
    }
    std::vector<GridShell> shells;
    long long n_raw = 0;
    for (int A = 0; A < n_atoms; ++A) {
        if (task_atom >= 0 && A != task_atom) continue;
        const double R = bragg_slater_angstrom(z[A]) * kAngToBohr;
        const int nr = atom_radial_count(z[A], level, n_rad, radial_min);
        const int i0 = task_atom >= 0 ? std::max(rad_start, 1) : 1;
        const int i1 = task_atom >= 0 ? std::min(rad_end, nr) : nr;
        const double m = nr + 1.0;
        for (int i = i0; i <= i1; ++i) {
            const double t = i, mt = m - t;
            GridShell g;
            g.r = R * (t * t) / (mt * mt);
            g.w_rad = 2.0 * ipow(R, 3) * m * ipow(t, 5) / ipow(mt, 7);
            const int na = level == 4 ? n_ang : pruned_rule(g.r, R, level);
            int k = -1;
            for (int q = 0; q < 6; ++q)
                if (sizes[q] == na) k = q;
            if (k < 0) {
                *status = kBadArgs;
                return 0;
            }
            g.cx = xyz[3 * A];
            g.cy = xyz[3 * A + 1];
            g.cz = xyz[3 * A + 2];
            g.atom = A;
            g.dir_off = dir_off[k];
            g.n_dir = sizes[k];
            g.raw_off = n_raw;
            g.keep_off = 0;
            n_raw += sizes[k];
            shells.push_back(g);
        }
    }
    if (shells.empty()) return 0;
    std::vector<double> inv_r((size_t)n_atoms * n_atoms, 0.0), aij((size_t)n_atoms * n_atoms, 0.0);
    becke_pair_tables(n_atoms, z, xyz, inv_r.data(), aij.data());
    const int ns = (int)shells.size();
    double *d_xyz = nullptr, *d_inv = nullptr, *d_a = nullptr, *d_raw = nullptr, *d_out = nullptr;
    double4* d_dirs = nullptr;
    GridShell* d_sh = nullptr;
    int* d_keep = nullptr;
    std::vector<int> keep(ns);
    long cnt = 0;
    bool ok = hip_ok(hipMalloc(&d_xyz, sizeof(double) * 3 * n_atoms), "hipMalloc") &&
              hip_ok(hipMalloc(&d_inv, sizeof(double) * inv_r.size()), "hipMalloc") &&
              hip_ok(hipMalloc(&d_a, sizeof(double) * aij.size()), "hipMalloc") &&
              hip_ok(hipMalloc(&d_raw, sizeof(double) * 6 * n_raw), "hipMalloc") &&
              hip_ok(hipMalloc(&d_dirs, sizeof(double4) * dirs.size()), "hipMalloc") &&
              hip_ok(hipMalloc(&d_sh, sizeof(GridShell) * ns), "hipMalloc") &&
              hip_ok(hipMalloc(&d_keep, sizeof(int) * ns), "hipMalloc");
    if (!ok) *status = kAllocFail;
    if (ok) {
        ok = hip_ok(hipMemcpy(d_xyz, xyz, sizeof(double) * 3 * n_atoms, hipMemcpyHostToDevice), "copy") &&
             hip_ok(hipMemcpy(d_inv, inv_r.data(), sizeof(double) * inv_r.size(), hipMemcpyHostToDevice), "copy") &&
             hip_ok(hipMemcpy(d_a, aij.data(), sizeof(double) * aij.size(), hipMemcpyHostToDevice), "copy") &&
             hip_ok(hipMemcpy(d_dirs, dirs.data(), sizeof(double4) * dirs.size(), hipMemcpyHostToDevice), "copy") &&
             hip_ok(hipMemcpy(d_sh, shells.data(), sizeof(GridShell) * ns, hipMemcpyHostToDevice), "copy");
        if (ok) {
            hipLaunchKernelGGL(k_grid_shell, dim3(ns), dim3(kGridThreads), 0, 0, d_sh, d_dirs, d_xyz, d_inv, d_a, n_atoms,
                               d_raw, d_keep);
            ok = hip_ok(hipGetLastError(), "k_grid_shell launch") &&
                 hip_ok(hipMemcpy(keep.data(), d_keep, sizeof(int) * ns, hipMemcpyDeviceToHost), "copy");
        }
        if (ok) {
            for (int i = 0; i < ns; ++i) {
                shells[i].keep_off = cnt;
                cnt += keep[i];
            }
        }
        if (ok && out && cnt <= max_out && cnt > 0) {
            ok = hip_ok(hipMalloc(&d_out, sizeof(double) * 6 * cnt), "hipMalloc") &&
                 hip_ok(hipMemcpy(d_sh, shells.data(), sizeof(GridShell) * ns, hipMemcpyHostToDevice), "copy");
            if (ok) {
                hipLaunchKernelGGL(k_grid_compact, dim3(ns), dim3(kGridThreads), 0, 0, d_sh, d_raw, d_out);
                ok = hip_ok(hipGetLastError(), "k_grid_compact launch") &&
                     hip_ok(hipMemcpy(out, d_out, sizeof(double) * 6 * cnt, hipMemcpyDeviceToHost), "copy");
            }
        }
        if (!ok && *status == kOk) *status = kLaunchFail;
    }
    (void)hipFree(d_xyz);
    (void)hipFree(d_inv);
    (void)hipFree(d_a);
    (void)hipFree(d_raw);
    (void)hipFree(d_dirs);
    (void)hipFree(d_sh);
    (void)hipFree(d_keep);
    if (d_out) (void)hipFree(d_out);
    return ok ? cnt : 0;
}

void becke_pair_tables(int n_atoms, const int* z, const double* xyz, double* inv_r, double* aij) {
    for (int i = 0; i < n_atoms; ++i) {
        inv_r[(size_t)i * n_atoms + i] = 0.0;
        aij[(size_t)i * n_atoms + i] = 0.0;
        for (int j = i + 1; j < n_atoms; ++j) {
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                         dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double rij = std::sqrt(dx * dx + dy * dy + dz * dz);
            const double inv = rij >= 1e-14 ? 1.0 / rij : 0.0;
            const double a = becke_a(z[i], z[j]);
            inv_r[(size_t)i * n_atoms + j] = inv_r[(size_t)j * n_atoms + i] = inv;
            aij[(size_t)i * n_atoms + j] = a;
            aij[(size_t)j * n_atoms + i] = -a;
        }
    }
}

}  // namespace chx
