// Instantiations of the single-pass step kernel (real fields, no periodic axes, one row per wavefront, 64 x cells):
// material mode x pole registers x wavefronts per block (4 or 8).

#include "lsf_fused.h"
#include "lsf_launch.h"

namespace lsf {
namespace {

template <int MAT, int W, int NP>
void go(const FusedLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL((k_step_fused<MAT, W, NP>), dim3(L.sc.n_pad), dim3(64 * W), 0, s, L.args, L.sc);
}

template <int MAT, int NP>
void by_waves(const FusedLaunch& L, hipStream_t s) {
  if (L.waves == 8) go<MAT, 8, NP>(L, s);
  else go<MAT, 4, NP>(L, s);
}

template <int MAT>
void by_poles(const FusedLaunch& L, hipStream_t s) {
  switch (L.np) {
    case 0: return by_waves<MAT, 0>(L, s);
    case 2: return by_waves<MAT, 2>(L, s);
    case 3: return by_waves<MAT, 3>(L, s);
    default: return by_waves<MAT, 4>(L, s);
  }
}

}  // namespace

void launch_fused(const FusedLaunch& L, hipStream_t s) {
  if (L.mat) by_poles<1>(L, s);
  else by_poles<0>(L, s);
}

}  // namespace lsf
