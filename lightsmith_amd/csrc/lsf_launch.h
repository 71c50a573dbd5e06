// Host-side launchers of the bulk H and E kernels. The kernels are instantiated in their own translation
// units (lsf_k_h.hip, and lsf_k_e.hip built once per field type and material mode) so that they compile in
// parallel; the host code selects an instantiation from runtime parameters.
#pragma once

#include <hip/hip_runtime.h>

#include "lsf_step.h"

namespace lsf {

struct KLaunch {
  const StepArgs* args;  // device array, one entry per member
  Sched sc;              // the launch's tiles
  int grid;              // blocks launched: sc.n_pad (one tile per block), or fewer (a multiple of 8): persistent
  int ry, waves, xw;     // tile shape: rows per lane, wavefronts per block, x cells per row
  int cplx;              // complex64 fields
  int per;               // some axis periodic
  int mat;               // E: 1 = packed material ids, 0 = dense coefficients
  int np;                // E: pole registers (0 or the engine's max_poles, 1 to 4)
};

// The tile shapes (rows per lane, wavefronts per block, x cells per row) with compiled kernels.
#define LSF_FOR_SHAPES(X) \
  X(1, 4, 64) X(2, 4, 64) X(1, 8, 64) X(2, 8, 64) X(1, 4, 32) X(2, 4, 32) X(1, 8, 32) X(2, 8, 32)

inline bool tile_shape_ok(int ry, int waves, int xw) {
#define LSF_SHAPE_OK(RY, W, XW) if (ry == RY && waves == W && xw == XW) return true;
  LSF_FOR_SHAPES(LSF_SHAPE_OK)
#undef LSF_SHAPE_OK
  return false;
}

// one launcher per translation unit: H by field type; E by field type, material mode and periodicity
void launch_update_h_f(const KLaunch& L, hipStream_t s);
void launch_update_h_c(const KLaunch& L, hipStream_t s);
#define LSF_E_LAUNCHERS(X) X(f, 0, 0) X(f, 0, 1) X(f, 1, 0) X(f, 1, 1) X(c, 0, 0) X(c, 0, 1) X(c, 1, 0) X(c, 1, 1)
#define LSF_DECLARE_E(T, M, P) void launch_update_e_##T##M##P(const KLaunch& L, hipStream_t s);
LSF_E_LAUNCHERS(LSF_DECLARE_E)
#undef LSF_DECLARE_E

// the single-pass step (real fields, no periodic axes, one row per wavefront, 64 x cells): lsf_k_fused.hip
struct FusedArgs;
// The tiles are split three ways by the CPML axes they touch (their halos included): none, x only, any other;
// each class runs its own kernel (the psi paths of absent axes compiled out) over its tile list.
struct FusedLaunch {
  const FusedArgs* args;  // device, one entry
  const int* tiles[3];    // device tile lists: no CPML, x CPML only, other
  int count[3];
  int waves;
  int mat, np;
};
void launch_fused(const FusedLaunch& L, hipStream_t s);

// adjoint kernels (real fields, no periodic axes, no D-form nodes; one tile per block): lsf_k_adj.hip
void launch_adjoint_h(const KLaunch& L, hipStream_t s);
void launch_adjoint_e(const KLaunch& L, hipStream_t s);

inline void launch_update_h(const KLaunch& L, hipStream_t s) {
  if (L.cplx) launch_update_h_c(L, s);
  else launch_update_h_f(L, s);
}

inline void launch_update_e(const KLaunch& L, hipStream_t s) {
#define LSF_PICK_E(T, M, P) \
  if ((L.cplx != 0) == (#T[0] == 'c') && L.mat == M && (L.per != 0) == (P != 0)) return launch_update_e_##T##M##P(L, s);
  LSF_E_LAUNCHERS(LSF_PICK_E)
#undef LSF_PICK_E
}

}  // namespace lsf
