// Instantiations of the single-pass step kernel (real fields, no periodic axes, one row per wavefront, 64 x cells):
// material mode x poles per D-form node x wavefronts per block (4 or 8) x CPML class (none, x only, any).

#include "lsf_fused.h"
#include "lsf_launch.h"

namespace lsf {
namespace {

template <int MAT, int W, int NP>
void go(const FusedLaunch& L, hipStream_t s) {
  const int pad[3] = {(L.count[0] + 7) & ~7, (L.count[1] + 7) & ~7, (L.count[2] + 7) & ~7};
  if (L.count[2])
    hipLaunchKernelGGL((k_step_fused<MAT, W, NP, 7>), dim3(pad[2]), dim3(64 * W), 0, s, L.args,
                       Sched{nullptr, nullptr, L.count[2], pad[2]}, L.tiles[2]);
  if (L.count[1])
    hipLaunchKernelGGL((k_step_fused<MAT, W, NP, 1>), dim3(pad[1]), dim3(64 * W), 0, s, L.args,
                       Sched{nullptr, nullptr, L.count[1], pad[1]}, L.tiles[1]);
  if (L.count[0])
    hipLaunchKernelGGL((k_step_fused<MAT, W, NP, 0>), dim3(pad[0]), dim3(64 * W), 0, s, L.args,
                       Sched{nullptr, nullptr, L.count[0], pad[0]}, L.tiles[0]);
}

template <int MAT, int NP>
void by_waves(const FusedLaunch& L, hipStream_t s) {
  if (L.waves == 8) go<MAT, 8, NP>(L, s);
  else go<MAT, 4, NP>(L, s);
}

template <int MAT>
void by_poles(const FusedLaunch& L, hipStream_t s) {
  switch (L.np) {  // the pole count is the slot stride: exact
    case 0: return by_waves<MAT, 0>(L, s);
    case 1: return by_waves<MAT, 1>(L, s);
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
