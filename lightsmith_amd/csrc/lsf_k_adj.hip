// Instantiations of the adjoint bulk kernels: tile shape (rows per lane 1 or 2, wavefronts 4 or 8, 64 x cells
// per row) x material mode (E kernel).

#include "lsf_adjoint.h"
#include "lsf_launch.h"

namespace lsf {
namespace {

template <int RY, int W>
void go_h(const KLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL((k_adjoint_h<RY, W>), dim3(L.sc.n_pad), dim3(64 * W), 0, s, L.args, L.sc);
}

template <int MAT, int RY, int W>
void go_e(const KLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL((k_adjoint_e<MAT, RY, W>), dim3(L.sc.n_pad), dim3(64 * W), 0, s, L.args, L.sc);
}

template <int MAT>
void e_by_shape(const KLaunch& L, hipStream_t s) {
  if (L.ry == 2) {
    if (L.waves == 8) go_e<MAT, 2, 8>(L, s);
    else go_e<MAT, 2, 4>(L, s);
  } else {
    if (L.waves == 8) go_e<MAT, 1, 8>(L, s);
    else go_e<MAT, 1, 4>(L, s);
  }
}

}  // namespace

void launch_adjoint_h(const KLaunch& L, hipStream_t s) {
  if (L.ry == 2) {
    if (L.waves == 8) go_h<2, 8>(L, s);
    else go_h<2, 4>(L, s);
  } else {
    if (L.waves == 8) go_h<1, 8>(L, s);
    else go_h<1, 4>(L, s);
  }
}

void launch_adjoint_e(const KLaunch& L, hipStream_t s) {
  if (L.mat) e_by_shape<1>(L, s);
  else e_by_shape<0>(L, s);
}

}  // namespace lsf
