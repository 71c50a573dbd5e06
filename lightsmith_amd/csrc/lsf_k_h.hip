// Instantiations of the bulk H kernel for one field type (LSF_H_CPLX): periodic axes x tile shape x
// persistent or one tile per block. LSF_H_ENTRY names the launcher that the build defines.

#include <type_traits>

#include "lsf_launch.h"

#if !defined(LSF_H_CPLX) || !defined(LSF_H_ENTRY)
#error "build with -DLSF_H_CPLX=0|1 -DLSF_H_ENTRY=<launcher name>"
#endif

namespace lsf {
namespace {

using TH = std::conditional_t<LSF_H_CPLX != 0, c32, float>;

template <int RY, int W, int XW, bool PER>
void go(const KLaunch& L, hipStream_t s) {
  if (L.grid < L.sc.n_pad)
    hipLaunchKernelGGL((k_update_h<TH, RY, W, XW, PER, true>), dim3(L.grid), dim3(64 * W), 0, s, L.args, L.sc);
  else
    hipLaunchKernelGGL((k_update_h<TH, RY, W, XW, PER, false>), dim3(L.sc.n_pad), dim3(64 * W), 0, s, L.args, L.sc);
}

template <bool PER>
void by_shape(const KLaunch& L, hipStream_t s) {
#define LSF_GO(RY, W, XW) \
  if (L.ry == RY && L.waves == W && L.xw == XW) return go<RY, W, XW, PER>(L, s);
  LSF_FOR_SHAPES(LSF_GO)
#undef LSF_GO
}

}  // namespace

void LSF_H_ENTRY(const KLaunch& L, hipStream_t s) {
  if (L.per) by_shape<true>(L, s);
  else by_shape<false>(L, s);
}

}  // namespace lsf
