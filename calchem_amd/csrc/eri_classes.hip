// eri_classes.hip — template instances of the repulsion kernel k_eri for one
// bra class (ERI_BRA_CLASS, set on the compiler line) against every ket
// class.  The build compiles this file once per bra class, so the six
// objects compile side by side.
#include "int_device.hpp"

#ifndef ERI_BRA_CLASS
#error "compile with -DERI_BRA_CLASS=0..5"
#endif

namespace chx {
namespace ints {
namespace {

constexpr int kBraLa = ERI_BRA_CLASS == 0 ? 0 : (ERI_BRA_CLASS <= 2 ? 1 : 2);
constexpr int kBraLb = ERI_BRA_CLASS - pair_class(kBraLa, 0);

template <int LC, int LD>
hipError_t launch_one(const EriArgs& a, bool diag, hipStream_t s) {
    constexpr int NCD = ncart(LC) * ncart(LD);
    const long long n = diag ? (long long)a.n_bra * NCD : (long long)a.n_bra * a.n_ket * NCD;
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 127) / 128);
    if (diag) {
        if constexpr (LC == kBraLa && LD == kBraLb)
            hipLaunchKernelGGL((k_eri<kBraLa, kBraLb, LC, LD, true>), dim3(blocks), dim3(128), 0, s, a);
        else
            return hipErrorInvalidValue;
    } else {
        hipLaunchKernelGGL((k_eri<kBraLa, kBraLb, LC, LD, false>), dim3(blocks), dim3(128), 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace

#define CHX_CAT2(a, b) a##b
#define CHX_CAT(a, b) CHX_CAT2(a, b)

hipError_t CHX_CAT(launch_eri_bra, ERI_BRA_CLASS)(int ket_class, bool diag, const EriArgs& a, hipStream_t s) {
    switch (ket_class) {
    case 0: return launch_one<0, 0>(a, diag, s);
    case 1: return launch_one<1, 0>(a, diag, s);
    case 2: return launch_one<1, 1>(a, diag, s);
    case 3: return launch_one<2, 0>(a, diag, s);
    case 4: return launch_one<2, 1>(a, diag, s);
    case 5: return launch_one<2, 2>(a, diag, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ints
}  // namespace chx
