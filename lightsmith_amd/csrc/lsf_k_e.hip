// Instantiations of the bulk E kernel for one field type (LSF_E_CPLX), material mode (LSF_E_MAT) and
// periodicity (LSF_E_PER): pole registers x tile shape x persistent or one tile per block. LSF_E_ENTRY names
// the launcher that the build defines.

#include <type_traits>

#include "lsf_launch.h"

#if !defined(LSF_E_CPLX) || !defined(LSF_E_MAT) || !defined(LSF_E_PER) || !defined(LSF_E_ENTRY)
#error "build with -DLSF_E_CPLX=0|1 -DLSF_E_MAT=0|1 -DLSF_E_PER=0|1 -DLSF_E_ENTRY=<launcher name>"
#endif

namespace lsf {
namespace {

using TE = std::conditional_t<LSF_E_CPLX != 0, c32, float>;
constexpr bool PER = LSF_E_PER != 0;

template <int RY, int W, int XW, int NP>
void go(const KLaunch& L, hipStream_t s) {
  if (L.grid < L.sc.n_pad)
    hipLaunchKernelGGL((k_update_e<TE, LSF_E_MAT, RY, W, XW, PER, NP, true>), dim3(L.grid), dim3(64 * W), 0, s,
                       L.args, L.sc);
  else
    hipLaunchKernelGGL((k_update_e<TE, LSF_E_MAT, RY, W, XW, PER, NP, false>), dim3(L.sc.n_pad), dim3(64 * W), 0, s,
                       L.args, L.sc);
}

template <int NP>
void by_shape(const KLaunch& L, hipStream_t s) {
#define LSF_GO(RY, W, XW) \
  if (L.ry == RY && L.waves == W && L.xw == XW) return go<RY, W, XW, NP>(L, s);
  LSF_FOR_SHAPES(LSF_GO)
#undef LSF_GO
}

}  // namespace

void LSF_E_ENTRY(const KLaunch& L, hipStream_t s) {
  switch (L.np) {
    case 0: return by_shape<0>(L, s);
    case 1: return by_shape<1>(L, s);
    case 2: return by_shape<2>(L, s);
    case 3: return by_shape<3>(L, s);
    default: return by_shape<4>(L, s);
  }
}

}  // namespace lsf
