// C ABI of the FDTD time-stepping library: engine handles, per-step launch sequence, HIP graphs.
// Declarations and conventions: include/lsf.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "lsf.h"
#include "lsf_fused.h"
#include "lsf_launch.h"
#include "lsf_sparse.h"

using lsf::c32;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return LSF_OK;
  return fail(LSF_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

enum OpKind {
  OP_GATHER, OP_MODAL, OP_TABLE, OP_VIEW, OP_LINE_H, OP_LINE_E, OP_SUB_H, OP_SUB_E, OP_MIXED,
  OP_TAPE, OP_ADJ_GRAD, OP_ADJ_SRC, OP_TENSOR, OP_ADJ_ADE_PRE, OP_ADJ_ADE_POST
};

struct Op {
  int kind = 0, phase = 0, field = 0;
  long long n = 0;
  int width = 0;
  const long long* target = nullptr;
  const void* src = nullptr;
  const long long* sidx = nullptr;
  const void* w = nullptr;
  const float* scale = nullptr;
  const float* table = nullptr;
  long long table_len = 0;
  const int* row_ptr = nullptr;
  int m = 1;
  long long cmax = 0;
  const long long* base = nullptr;
  const long long* i0 = nullptr;
  const float* u0 = nullptr;
  const float* coef = nullptr;
  float* out = nullptr;
  // 1D line
  float *e1 = nullptr, *h1 = nullptr;
  const float *ca = nullptr, *cb = nullptr, *da = nullptr, *db = nullptr;
  int m_src = 0;
  // sub-engine
  lsf_engine* sub = nullptr;
  // mixed group
  int S = 0, maxp = 0, has_h = 0, has_a = 0;
  const float *f = nullptr, *wmix = nullptr, *h0 = nullptr, *inv_a = nullptr, *sm = nullptr;
  void* state = nullptr;
  // adjoint: tape rows, forward step count, DFT stride, gradient accumulator, second H target (adjoint source:
  // the contributing point of each CSR entry), H side, monitor points
  long long n_slots = 0, n_steps = 0, m_pts = 0;
  int dft_stride = 1, h_side = 0;
  double* gacc = nullptr;
  const long long* t2 = nullptr;
  const int* ade_medium = nullptr;
  const double* omega = nullptr;
  double dt = 0.0;
  // conformal vertex tensor
  lsf::TensorJob<float> tjf{};
  lsf::TensorJob<c32> tjc{};
  // E-side table launched with the D update of the tensor op at this index of the engine's ops (-1: alone); the
  // tensor op then skips its own D launch (lsf_fuse_table_tensor)
  int fuse_tensor = -1;
  int d_fused = 0;
};

struct Monitor {
  int nf = 0;
  long long np = 0;
  const double* omega = nullptr;
  const long long *e = nullptr, *h1 = nullptr, *h2 = nullptr;
  void *acc_e = nullptr, *acc_h = nullptr;
  const void *esrc = nullptr, *hsrc = nullptr;
  int src_complex = 0;
  int group = -1;
  int nblocks = 0;
  int poff = 0;
};

struct TimeMon {
  long long np = 0, n_max = 0;
  const long long* e = nullptr;
  void* buf = nullptr;
  long long* pos = nullptr;
};

}  // namespace

struct lsf_engine {
  lsf::Grid g{};
  bool cplx = false;
  double dt = 0.0;
  void* E = nullptr;
  void* H = nullptr;
  long long* counter = nullptr;
  int tick = 1;
  const float* cbd = nullptr;
  const uint32_t* mid = nullptr;
  const float* mtab = nullptr;
  lsf::PmlAx pml[2][3]{};  // [field E=0, H=1][axis]
  const int* chunk_base = nullptr;
  void* P = nullptr;
  long long n_ade = 0;
  int max_poles = 0;
  const float* media = nullptr;
  int n_media = 0;
  std::vector<Op> ops;
  std::vector<Monitor> mons;
  std::vector<TimeMon> tmons;
  float* energy = nullptr;
  float* partial = nullptr;  // library scratch: DFT energy partials
  int* mon_tab = nullptr;    // library scratch: [n_mon + 1] offsets, [n_mon] groups
  int n_partial = 0;
  int ry = 2, waves = 4, xw = 64, zc = 16;
  int rev_e = 1;  // the E launch walks each XCD's tiles backwards (lsf_set_tile_order)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  lsf::StepArgs* dev_args = nullptr;  // library scratch: [0] E kernel, [1] H kernel arguments
  const lsf::StepArgs* args_override = nullptr;  // during lsf_step_noalloc: the arguments staged in `scratch`
  lsf::StepArgs host_args[2];          // lsf_step_noalloc: host source of the staged arguments (outlives the copy)
  int nb[2] = {0, 0}, nbp[2] = {0, 0};  // tiles of the E [0] and H [1] launches (and padded to a multiple of 8)
  bool dirty = true;                   // dev_args out of date
  bool launched = false;               // kernels may be reading dev_args
  int held = 0;                        // batches holding this engine (its configuration is then frozen)
  int adjoint = 0;                     // steps the adjoint (transposed, time-reversed) update
  // single-pass stepping: both field buffers and H-side psi sets, the parity of the current ones
  void* Ebuf[2] = {nullptr, nullptr};
  void* Hbuf[2] = {nullptr, nullptr};
  void* psiH[2][3][3] = {};            // [parity][axis][component]
  int parity = 0;
  bool fused_set = false;              // second buffers given (lsf_set_fused)
  bool use_fused = false;              // the single-pass kernel steps this engine (decided at sync_args)
  lsf::FusedArgs* dev_fused = nullptr; // library scratch: [parity]
  int* dev_tiles = nullptr;            // single-pass tile lists by CPML class, back to back
  int fused_count[3] = {0, 0, 0};
  hipGraph_t graph2 = nullptr;         // the chunk graph starting at the other parity (odd strides)
  hipGraphExec_t exec2 = nullptr;
  int graph_parity = 0, graph_stride = 0;
};

namespace {

using lsf::StepArgs;

struct Tiling {
  int ry, waves, xw, zc;
};

Tiling own_tiling(const lsf_engine* e) { return Tiling{e->ry, e->waves, e->xw, e->zc}; }

// pole registers of the E kernel instantiation for this engine: its max_poles (the slot stride of the pole
// arrays); the single-pass kernel rounds 1 up to 2
int pole_regs(const lsf_engine* e) { return e->n_ade > 0 ? e->max_poles : 0; }

// Kernel arguments of engine e for field 0 (E kernel) or 1 (H kernel) under tiling t.
StepArgs make_args(const lsf_engine* e, int field, const Tiling& t) {
  StepArgs a{};
  a.g = e->g;
  a.E = e->E;
  a.H = e->H;
  a.cbd = e->cbd;
  a.mid = e->mid;
  a.mtab = e->mtab;
  for (int ax = 0; ax < 3; ++ax) {
    a.pml[ax] = e->pml[field][ax];
    a.xpml[ax] = e->pml[1 - field][ax];
  }
  a.chunk_base = e->chunk_base;
  a.P = e->P;
  a.n_ade = (int)e->n_ade;
  a.nchunk = (e->g.n[0] + 31) / 32;
  a.max_poles = e->max_poles;
  a.media_stride = 2 + 3 * e->max_poles;
  a.media = e->media;
  a.n_media_floats = e->n_ade > 0 ? e->n_media * (2 + 3 * e->max_poles) : 0;
  a.counter = e->counter;
  a.tick = e->tick;
  const int rows = t.ry * t.waves * (64 / t.xw);
  a.tiles_x = (e->g.n[0] + t.xw - 1) / t.xw;
  a.tiles_y = (e->g.n[1] + rows - 1) / rows;
  // the E kernel with pole state marches twice the z chunk (320^3, 10 % 3-pole nodes: E kernel 451 us at 8 planes
  // per tile, 442 us at 16; H kernel 269 us at 8, 276 us at 16)
  a.zc = field == 0 && !e->adjoint && pole_regs(e) > 0 ? 2 * t.zc : t.zc;
  const int tiles_z = (e->g.n[2] + a.zc - 1) / a.zc;
  a.nb = a.tiles_x * a.tiles_y * tiles_z;
  a.nbp = (a.nb + 7) / 8 * 8;
  return a;
}

bool any_periodic(const lsf_engine* e) { return e->g.per[0] || e->g.per[1] || e->g.per[2]; }


void set_parity(lsf_engine* e, int p) {
  e->parity = p;
  e->E = e->Ebuf[p];
  e->H = e->Hbuf[p];
  for (int ax = 0; ax < 3; ++ax)
    for (int c = 0; c < 3; ++c) e->pml[1][ax].psi[c] = e->psiH[p][ax][c];
}

bool has_h_phase_ops(const lsf_engine* e) {
  for (const Op& o : e->ops)
    if (o.phase <= LSF_PH_ADVANCE_H || o.phase == LSF_PH_BEFORE_H) return true;
  return false;
}

// elements from the first to the second psi array of (field, axis) when they are contiguous, else -1
long long psi_gap(const lsf_engine* e, int field, int ax, int parity_set) {
  const int c1 = ax == 0 ? 1 : 0, c2 = ax == 2 ? 1 : 2;
  void* const* ps = field == 0 ? e->pml[0][ax].psi : e->psiH[parity_set][ax];
  const long long ns = e->pml[field][ax].ns;
  long long size = ns;
  for (int a = 0; a < 3; ++a)
    if (a != ax) size *= e->g.n[a];
  if (!ps[c1] || !ps[c2]) return -1;
  return ((const char*)ps[c2] - (const char*)ps[c1]) == size * (long long)sizeof(float) ? size : -1;
}

bool fused_eligible(const lsf_engine* e) {
  if (!e->fused_set || e->cplx || any_periodic(e) || e->adjoint || e->ry != 1 || e->xw != 64 || has_h_phase_ops(e))
    return false;
  // the kernel assumes the stepped boxes of a grid without periodic axes or mirror planes
  for (int c = 0; c < 3; ++c)
    for (int ax = 0; ax < 3; ++ax) {
      const int n = e->g.n[ax];
      if (e->g.rE[c][ax][0] != (ax == c ? 0 : 1) || e->g.rE[c][ax][1] != n - 1) return false;
      if (e->g.rH[c][ax][0] != 0 || e->g.rH[c][ax][1] != (ax == c ? n : n - 1)) return false;
    }
  const long long lim = 1LL << 32;  // the kernel addresses fields and psi blocks with 32-bit byte offsets
  if (3 * e->g.N * 4 >= lim) return false;
  for (int f = 0; f < 2; ++f)
    for (int ax = 0; ax < 3; ++ax)
      if (e->pml[f][ax].prof)
        for (int p = 0; p < (f == 1 ? 2 : 1); ++p) {
          const long long gap = psi_gap(e, f, ax, p);
          if (gap < 0 || 2 * gap * 4 >= lim) return false;
        }
  return true;
}

lsf::FusedArgs fused_args(const lsf_engine* e, int p) {
  lsf::FusedArgs a{};
  a.g = e->g;
  a.E_in = (const float*)e->Ebuf[p];
  a.H_in = (const float*)e->Hbuf[p];
  a.E_out = (float*)e->Ebuf[p ^ 1];
  a.H_out = (float*)e->Hbuf[p ^ 1];
  a.cbd = e->cbd;
  a.mid = e->mid;
  a.mtab = e->mtab;
  for (int ax = 0; ax < 3; ++ax) {
    const int c1 = ax == 0 ? 1 : 0;
    a.prof[0][ax] = e->pml[0][ax].prof;
    a.prof[1][ax] = e->pml[1][ax].prof;
    a.psiE[ax] = (float*)e->pml[0][ax].psi[c1];
    a.psiHi[ax] = (const float*)e->psiH[p][ax][c1];
    a.psiHo[ax] = (float*)e->psiH[p ^ 1][ax][c1];
    for (int f = 0; f < 2; ++f) {
      a.psz[f][ax] = e->pml[f][ax].prof ? psi_gap(e, f, ax, 0) : 0;
      for (int c = 0; c < 3; ++c) a.pso[f][ax][c] = c == ax ? 0u : (unsigned)(lsf::psi_slot(c, ax) * a.psz[f][ax] * 4);
      a.ns[f][ax] = e->pml[f][ax].ns;
      a.has_k[f][ax] = e->pml[f][ax].has_k;
    }
  }
  a.chunk_base = e->chunk_base;
  a.P = e->P;
  a.nchunk = (e->g.n[0] + 31) / 32;
  a.max_poles = e->max_poles;
  a.media_stride = 2 + 3 * e->max_poles;
  a.n_media_floats = e->n_ade > 0 ? e->n_media * (2 + 3 * e->max_poles) : 0;
  a.media = e->media;
  a.counter = e->counter;
  a.tick = e->tick;
  a.tiles_x = (e->g.n[0] + 63) / 64;
  a.tiles_y = (e->g.n[1] + e->waves - 1) / e->waves;  // one row per wavefront
  a.zc = e->zc;
  return a;
}

// The single-pass tiles in three lists by the CPML axes they touch: a tile reads psi on the x span [x0 - 1, x0 + 64),
// the rows [j0 - 1, j0 + waves) and the planes [z0 - 1, z1) (the halo cells it recomputes included). Lists: no
// CPML, x CPML only, any other.
int fused_tiles(lsf_engine* e, const lsf::FusedArgs& a) {
  const int n[3] = {e->g.n[0], e->g.n[1], e->g.n[2]};
  std::vector<char> slab[3];
  for (int ax = 0; ax < 3; ++ax) {
    slab[ax].assign(n[ax], 0);
    for (int f = 0; f < 2; ++f) {
      if (!e->pml[f][ax].prof) continue;
      std::vector<float4> pr(n[ax]);
      const int rc = hip_check(hipMemcpy(pr.data(), e->pml[f][ax].prof, n[ax] * sizeof(float4), hipMemcpyDeviceToHost),
                               "CPML profile read");
      if (rc) return rc;
      for (int q = 0; q < n[ax]; ++q)
        if ((int)pr[q].w >= 0) slab[ax][q] = 1;
    }
  }
  auto touches = [&](int ax, int lo, int hi) {  // any slab node in [lo, hi) (clamped to the grid)
    for (int q = std::max(lo, 0); q < std::min(hi, n[ax]); ++q)
      if (slab[ax][q]) return true;
    return false;
  };
  std::vector<int> lists[3];
  const int tz_n = (n[2] + a.zc - 1) / a.zc;
  for (int tz = 0; tz < tz_n; ++tz)
    for (int ty = 0; ty < a.tiles_y; ++ty)
      for (int tx = 0; tx < a.tiles_x; ++tx) {
        const int z0 = tz * a.zc, z1 = std::min(z0 + a.zc, n[2]), j0 = ty * e->waves, x0 = tx * 64;
        const int m = (touches(0, x0 - 1, x0 + 64) ? 1 : 0) | (touches(1, j0 - 1, j0 + e->waves) ? 2 : 0) |
                      (touches(2, z0 - 1, z1) ? 4 : 0);
        lists[m == 0 ? 0 : (m == 1 ? 1 : 2)].push_back((tz * a.tiles_y + ty) * a.tiles_x + tx);
      }
  std::vector<int> all;
  for (int c = 0; c < 3; ++c) {
    e->fused_count[c] = (int)lists[c].size();
    all.insert(all.end(), lists[c].begin(), lists[c].end());
  }
  if (e->dev_tiles) (void)hipFree(e->dev_tiles);
  e->dev_tiles = nullptr;
  int rc = hip_check(hipMalloc(&e->dev_tiles, std::max<size_t>(1, all.size()) * sizeof(int)), "hipMalloc tile lists");
  if (!rc) rc = hip_check(hipMemcpy(e->dev_tiles, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice),
                          "tile list upload");
  return rc;
}

// Upload the engine's kernel arguments (and its sub-engines') when its configuration changed. Not called
// during graph capture.
int sync_args(lsf_engine* e) {
  for (const Op& o : e->ops)
    if (o.kind == OP_SUB_H) {
      const int rc = sync_args(o.sub);
      if (rc) return rc;
    }
  if (!e->dirty) return LSF_OK;
  int rc = LSF_OK;
  if (!e->dev_args) rc = hip_check(hipMalloc(&e->dev_args, 2 * sizeof(StepArgs)), "hipMalloc kernel arguments");
  else if (e->launched) rc = hip_check(hipDeviceSynchronize(), "synchronize before an argument update");
  if (rc) return rc;
  const Tiling t = own_tiling(e);
  const StepArgs h[2] = {make_args(e, 0, t), make_args(e, 1, t)};
  rc = hip_check(hipMemcpy(e->dev_args, h, sizeof h, hipMemcpyHostToDevice), "kernel argument upload");
  if (rc) return rc;
  for (int f = 0; f < 2; ++f) {
    e->nb[f] = h[f].nb;
    e->nbp[f] = h[f].nbp;
  }
  e->use_fused = fused_eligible(e);
  if (e->use_fused) {
    if (!e->dev_fused) rc = hip_check(hipMalloc(&e->dev_fused, 2 * sizeof(lsf::FusedArgs)), "hipMalloc fused arguments");
    if (rc) return rc;
    const lsf::FusedArgs f[2] = {fused_args(e, 0), fused_args(e, 1)};
    rc = hip_check(hipMemcpy(e->dev_fused, f, sizeof f, hipMemcpyHostToDevice), "fused argument upload");
    if (rc) return rc;
    rc = fused_tiles(e, f[0]);
    if (rc) return rc;
  }
  e->dirty = false;
  return LSF_OK;
}

void launch_fused_step(lsf_engine* e, hipStream_t s) {
  lsf::FusedLaunch L{};
  L.args = e->dev_fused + e->parity;
  L.tiles[0] = e->dev_tiles;
  L.tiles[1] = e->dev_tiles + e->fused_count[0];
  L.tiles[2] = e->dev_tiles + e->fused_count[0] + e->fused_count[1];
  for (int c = 0; c < 3; ++c) L.count[c] = e->fused_count[c];
  L.waves = e->waves;
  L.mat = e->mid ? 1 : 0;
  L.np = pole_regs(e);
  lsf::launch_fused(L, s);
  e->launched = true;
  set_parity(e, e->parity ^ 1);
}

lsf::KLaunch own_launch(lsf_engine* e, int field) {
  lsf::KLaunch L{};
  L.args = (e->args_override ? e->args_override : e->dev_args) + field;
  L.sc = lsf::Sched{nullptr, nullptr, e->nb[field], e->nbp[field]};
  L.grid = e->nbp[field];
  L.ry = e->ry;
  L.waves = e->waves;
  L.xw = e->xw;
  L.cplx = e->cplx ? 1 : 0;
  L.per = any_periodic(e) ? 1 : 0;
  L.mat = e->mid ? 1 : 0;
  L.np = pole_regs(e);
  return L;
}

void launch_h(lsf_engine* e, hipStream_t s) {
  if (e->adjoint) lsf::launch_adjoint_h(own_launch(e, 1), s);
  else lsf::launch_update_h(own_launch(e, 1), s);
  e->launched = true;
}

void launch_e(lsf_engine* e, hipStream_t s) {
  lsf::KLaunch L = own_launch(e, 0);
  L.sc.rev = e->rev_e;
  if (e->adjoint) lsf::launch_adjoint_e(L, s);
  else lsf::launch_update_e(L, s);
  e->launched = true;
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void run_phase(lsf_engine* e, int phase, hipStream_t s);

template <class T>
lsf::GatherJob<T> gather_job(lsf_engine* e, const Op& o) {
  T* F = (T*)(o.field == 0 ? e->E : e->H);
  const T* src = o.src ? (const T*)o.src : (const T*)e->H;  // null: the engine's own (current) H
  return lsf::GatherJob<T>{o.n, o.width, o.target, F, src, o.sidx, (const T*)o.w, o.scale};
}

template <class T>
lsf::ModalJob<T> modal_job(lsf_engine* e, const Op& o) {
  T* F = (T*)(o.field == 0 ? e->E : e->H);
  const int slot = o.phase <= LSF_PH_ADVANCE_H ? 0 : 1;
  return lsf::ModalJob<T>{o.n, o.width, o.target, F, (const T*)o.w, o.table, o.table_len,
                          (const long long*)e->counter, slot, o.scale};
}

template <class T>
void run_op(lsf_engine* e, const Op& o, hipStream_t s) {
  const int slot = o.phase <= LSF_PH_ADVANCE_H ? 0 : 1;
  T* F = (T*)(o.field == 0 ? e->E : e->H);
  switch (o.kind) {
    case OP_GATHER:
      if (o.n > 0)
        hipLaunchKernelGGL(lsf::k_gather<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, gather_job<T>(e, o));
      break;
    case OP_MODAL:
      if (o.n > 0)
        hipLaunchKernelGGL(lsf::k_modal<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, modal_job<T>(e, o));
      break;
    case OP_TABLE:
      if (o.fuse_tensor >= 0) {
        const Op& t = e->ops[o.fuse_tensor];
        lsf::TensorJob<T> J;
        if constexpr (sizeof(T) == sizeof(float)) J = t.tjf;
        else J = t.tjc;
        J.E = (T*)e->E;
        const int tb = (int)nblk(o.n, 256);
        hipLaunchKernelGGL(lsf::k_table_tensor_d<T>, dim3(tb + nblk(J.nN, 256)), dim3(256), 0, s, o.n, tb, o.target,
                           o.row_ptr, F, o.table, o.m, o.cmax, o.base, o.i0, o.u0, o.coef,
                           (const long long*)e->counter, slot, o.scale, J);
      } else if (o.n > 0) {
        hipLaunchKernelGGL(lsf::k_table<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, o.n, o.target, o.row_ptr, F,
                           o.table, o.m, o.cmax, o.base, o.i0, o.u0, o.coef, (const long long*)e->counter, slot,
                           o.scale);
      }
      break;
    case OP_VIEW:
      hipLaunchKernelGGL(lsf::k_table_view, dim3(1), dim3(64), 0, s, (int)o.n, o.out, o.table, o.m, o.cmax, o.base,
                         o.i0, o.u0, (const long long*)e->counter, slot);
      break;
    case OP_LINE_H:
      hipLaunchKernelGGL(lsf::k_line_h, dim3(1), dim3(1024), 0, s, (int)o.n, (const float*)o.e1, o.h1, o.da, o.db);
      break;
    case OP_LINE_E:
      hipLaunchKernelGGL(lsf::k_line_e, dim3(1), dim3(1024), 0, s, (int)o.n, o.e1, (const float*)o.h1, o.ca, o.cb,
                         o.m_src, o.table, o.table_len, (const long long*)e->counter, slot);
      break;
    case OP_SUB_H:
      run_phase(o.sub, LSF_PH_BEFORE_H, s);
      launch_h(o.sub, s);
      run_phase(o.sub, LSF_PH_CORRECT_H, s);
      run_phase(o.sub, LSF_PH_ADVANCE_H, s);
      break;
    case OP_SUB_E:
      launch_e(o.sub, s);
      for (int ph = LSF_PH_OFFDIAG; ph <= LSF_PH_ADVANCE_E; ++ph) run_phase(o.sub, ph, s);
      break;
    case OP_TAPE: {
      lsf::TapeJob J{o.n, o.target, (const float*)e->H, e->g.n[0], e->g.n[1], e->g.N,
                     {e->g.s[0], e->g.s[1], e->g.s[2]}, o.out, o.n_slots, o.base, (const long long*)e->counter};
      if (o.n > 0) hipLaunchKernelGGL(lsf::k_tape, dim3(nblk(o.n, 256)), dim3(256), 0, s, J);
      break;
    }
    case OP_ADJ_GRAD: {
      lsf::GradJob J{o.n, o.target, (const float*)e->E, o.coef, o.table, o.n_slots, o.base,
                     (const long long*)e->counter, o.n_steps, o.gacc};
      if (o.n > 0) hipLaunchKernelGGL(lsf::k_adj_grad, dim3(nblk(o.n, 256)), dim3(256), 0, s, J);
      break;
    }
    case OP_ADJ_ADE_PRE:
    case OP_ADJ_ADE_POST: {
      lsf::AdjAdeJob J{o.n, o.target, o.ade_medium, o.coef, 2 + 3 * o.maxp, o.maxp, (float*)o.state, o.out, (float*)e->E};
      if (o.n > 0) {
        if (o.kind == OP_ADJ_ADE_PRE) hipLaunchKernelGGL(lsf::k_adj_ade_pre, dim3(nblk(o.n, 256)), dim3(256), 0, s, J);
        else hipLaunchKernelGGL(lsf::k_adj_ade_post, dim3(nblk(o.n, 256)), dim3(256), 0, s, J);
      }
      break;
    }
    case OP_ADJ_SRC: {
      lsf::AdjSrcJob J{o.h_side, o.width, o.m_pts, o.omega, o.dt, (const c32*)o.w, o.n, o.row_ptr, o.target,
                       o.t2, o.scale, (float*)(o.h_side ? e->H : e->E), (const long long*)e->counter, o.n_steps,
                       o.dft_stride};
      if (o.n > 0) hipLaunchKernelGGL(lsf::k_adj_src, dim3(nblk(o.n, 256)), dim3(256), 0, s, J);
      break;
    }
    case OP_TENSOR: {
      lsf::TensorJob<T> J;
      if constexpr (sizeof(T) == sizeof(float)) J = o.tjf;
      else J = o.tjc;
      J.E = (T*)e->E;
      if (!o.d_fused) hipLaunchKernelGGL(lsf::k_tensor_d<T>, dim3(nblk(J.nN, 256)), dim3(256), 0, s, J);
      hipLaunchKernelGGL(lsf::k_tensor_slot_e<T>, dim3(nblk(2 * J.nN, 256)), dim3(256), 0, s, J);
      break;
    }
    case OP_MIXED:
      if (o.n > 0)
        hipLaunchKernelGGL(lsf::k_mixed<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, o.n, o.S, o.maxp, o.target, o.f,
                           o.wmix, o.h0, o.inv_a, o.sm, (T*)o.state, o.has_h, o.has_a, (T*)e->E);
      break;
  }
}

// A table view's only readers are monitors, which read it when they sample: it is evaluated then (run_views), with
// the step counters of its phase, which are the same then.
void run_phase(lsf_engine* e, int phase, hipStream_t s) {
  for (const Op& o : e->ops) {
    if (o.phase != phase || o.kind == OP_VIEW) continue;
    if (e->cplx) run_op<c32>(e, o, s);
    else run_op<float>(e, o, s);
  }
}

void one_step(lsf_engine* e, hipStream_t s) {
  if (e->use_fused) {
    launch_fused_step(e, s);
    for (int ph = LSF_PH_OFFDIAG; ph <= LSF_PH_ADVANCE_E; ++ph) run_phase(e, ph, s);
    return;
  }
  run_phase(e, LSF_PH_BEFORE_H, s);
  launch_h(e, s);
  run_phase(e, LSF_PH_CORRECT_H, s);
  run_phase(e, LSF_PH_ADVANCE_H, s);
  launch_e(e, s);
  for (int ph = LSF_PH_OFFDIAG; ph <= LSF_PH_ADVANCE_E; ++ph) run_phase(e, ph, s);
}

template <class S>
lsf::DftJob<S> dft_job(const lsf_engine* e, const Monitor& m) {
  lsf::DftJob<S> J{};
  J.nf = m.nf;
  J.np = m.np;
  J.omega = m.omega;
  J.dt = e->dt;
  J.counter = e->counter;
  J.esrc = (const S*)(m.esrc ? m.esrc : e->E);
  J.hsrc = (const S*)(m.hsrc ? m.hsrc : e->H);
  J.eidx = m.e;
  J.h1 = m.h1;
  J.h2 = m.h2;
  J.acc_e = (c32*)m.acc_e;
  J.acc_h = (c32*)m.acc_h;
  J.partial = m.group >= 0 ? e->partial + m.poff : nullptr;
  return J;
}

bool mon_complex(const lsf_engine* e, const Monitor& m) { return m.esrc ? m.src_complex != 0 : e->cplx; }

bool has_energy(const lsf_engine* e) {
  if (!e->energy) return false;
  for (const Monitor& m : e->mons)
    if (m.group >= 0) return true;
  return false;
}

lsf::EnergyJob energy_job(const lsf_engine* e) {
  const int nm = (int)e->mons.size();
  return lsf::EnergyJob{nm, e->mon_tab, e->mon_tab + nm + 1, e->partial, e->energy};
}

void sample_time(lsf_engine* e, hipStream_t s) {
  for (const TimeMon& t : e->tmons) {
    if (e->cplx)
      hipLaunchKernelGGL(lsf::k_timemon<c32>, dim3(nblk(t.np, 256)), dim3(256), 0, s, t.np, t.e, (const c32*)e->E,
                         (c32*)t.buf, t.n_max, (const long long*)t.pos);
    else
      hipLaunchKernelGGL(lsf::k_timemon<float>, dim3(nblk(t.np, 256)), dim3(256), 0, s, t.np, t.e,
                         (const float*)e->E, (float*)t.buf, t.n_max, (const long long*)t.pos);
    hipLaunchKernelGGL(lsf::k_bump, dim3(1), dim3(64), 0, s, t.pos);
  }
}

void run_views(lsf_engine* e, hipStream_t s) {
  for (const Op& o : e->ops) {
    if (o.kind != OP_VIEW) continue;
    if (e->cplx) run_op<c32>(e, o, s);
    else run_op<float>(e, o, s);
  }
}

void sample(lsf_engine* e, hipStream_t s) {
  run_views(e, s);
  for (const Monitor& m : e->mons) {
    const size_t shm = sizeof(c32) * 2 * m.nf;
    if (mon_complex(e, m))
      hipLaunchKernelGGL(lsf::k_dft<c32>, dim3(m.nblocks), dim3(256), shm, s, dft_job<c32>(e, m));
    else
      hipLaunchKernelGGL(lsf::k_dft<float>, dim3(m.nblocks), dim3(256), shm, s, dft_job<float>(e, m));
  }
  if (has_energy(e)) hipLaunchKernelGGL(lsf::k_energy, dim3(1), dim3(64), 0, s, energy_job(e));
  sample_time(e, s);
}

int valid(lsf_engine* e) { return e ? LSF_OK : fail(LSF_EINVAL, "null engine handle"); }

// Configuration calls: refused while a batch holds the engine; they invalidate the uploaded arguments.
int configurable(lsf_engine* e) {
  if (!e) return fail(LSF_EINVAL, "null engine handle");
  if (e->held) return fail(LSF_ESTATE, "the engine is held by a batch; destroy the batch before changing it");
  e->dirty = true;
  return LSF_OK;
}

int add_op(lsf_engine* e, const Op& o) {
  if (int rc0 = configurable(e)) return rc0;
  if (o.phase < LSF_PH_CORRECT_H || o.phase > LSF_PH_BEFORE_H) return fail(LSF_EINVAL, "bad phase");
  if (o.field != 0 && o.field != 1) return fail(LSF_EINVAL, "field must be 0 (E) or 1 (H)");
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before adding work");
  if (o.phase == LSF_PH_CORRECT_E || o.phase == LSF_PH_GROUPS)
    for (Op& q : e->ops) {  // a fused table + D launch holds only for the ops it was checked against
      q.fuse_tensor = -1;
      q.d_fused = 0;
    }
  e->ops.push_back(o);
  return LSF_OK;
}

}  // namespace

extern "C" {

const char* lsf_last_error(void) { return g_err.c_str(); }
const char* lsf_version(void) { return "lsf 0.1 gfx950"; }

int lsf_create(const lsf_grid* gd, lsf_engine** out) {
  if (!gd || !out) return fail(LSF_EINVAL, "null argument");
  for (int a = 0; a < 3; ++a)
    if (gd->shape[a] < 3) return fail(LSF_EINVAL, "every axis needs at least 3 nodes");
  if (!gd->dev_E || !gd->dev_H || !gd->dev_counter) return fail(LSF_EINVAL, "fields and counter are required");
  if (!(gd->dt > 0)) return fail(LSF_EINVAL, "dt must be positive");
  lsf_engine* e = new lsf_engine();
  lsf::Grid& g = e->g;
  const bool cubic = gd->spacing[0] == gd->spacing[1] && gd->spacing[1] == gd->spacing[2];
  for (int a = 0; a < 3; ++a) {
    g.n[a] = gd->shape[a];
    g.per[a] = gd->periodic[a] ? 1 : 0;
    g.s[a] = cubic ? 1.f : (float)(1.0 / gd->spacing[a]);
    g.ph[a] = lsf::mk((float)gd->bloch_phase[a][0], (float)gd->bloch_phase[a][1]);
    if (!g.per[a]) g.ph[a] = lsf::mk(1.f, 0.f);
  }
  g.N = (long long)g.n[0] * g.n[1] * g.n[2];
  if (3 * g.N >= (1ll << 31)) {
    delete e;
    return fail(LSF_EINVAL, "grid too large: 3 * cells must stay below 2^31 (32-bit offsets)");
  }
  const double cdx = cubic ? gd->dt / gd->spacing[0] : gd->dt;
  g.cH = (float)cdx;
  g.cdx = (float)cdx;
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a) {
      const int n = g.n[a];
      if (g.per[a]) {
        g.rE[c][a][0] = 1; g.rE[c][a][1] = n - 1;
        g.rH[c][a][0] = 1; g.rH[c][a][1] = n - 1;
      } else {
        g.rE[c][a][0] = (a != c) ? 1 : 0; g.rE[c][a][1] = n - 1;
        g.rH[c][a][0] = 0; g.rH[c][a][1] = (a == c) ? n : n - 1;
      }
    }
  e->cplx = gd->complex_fields != 0;
  e->dt = gd->dt;
  e->E = gd->dev_E;
  e->H = gd->dev_H;
  e->Ebuf[0] = e->E;
  e->Hbuf[0] = e->H;
  e->counter = (long long*)gd->dev_counter;
  e->tick = gd->tick ? 1 : 0;
  *out = e;
  return LSF_OK;
}

void lsf_destroy(lsf_engine* e) {
  if (!e) return;
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->partial) (void)hipFree(e->partial);
  if (e->mon_tab) (void)hipFree(e->mon_tab);
  if (e->dev_args) (void)hipFree(e->dev_args);
  if (e->dev_fused) (void)hipFree(e->dev_fused);
  if (e->dev_tiles) (void)hipFree(e->dev_tiles);
  if (e->exec2) (void)hipGraphExecDestroy(e->exec2);
  if (e->graph2) (void)hipGraphDestroy(e->graph2);
  delete e;
}

int lsf_set_material_dense(lsf_engine* e, const float* cb) {
  if (int rc0 = configurable(e)) return rc0;
  if (!cb) return fail(LSF_EINVAL, "null material");
  e->cbd = cb;
  e->mid = nullptr;
  e->mtab = nullptr;
  return LSF_OK;
}

int lsf_set_material_indexed(lsf_engine* e, const uint32_t* ids, const float* table) {
  if (int rc0 = configurable(e)) return rc0;
  if (!ids || !table) return fail(LSF_EINVAL, "null material");
  e->cbd = nullptr;
  e->mid = ids;
  e->mtab = table;
  return LSF_OK;
}

int lsf_set_pml(lsf_engine* e, int field, int axis, const void* prof, int n_slab, void* p0, void* p1, void* p2,
                int has_k) {
  if (int rc0 = configurable(e)) return rc0;
  if (field < 0 || field > 1 || axis < 0 || axis > 2) return fail(LSF_EINVAL, "bad field or axis");
  if (e->g.per[axis]) return fail(LSF_EINVAL, "a periodic axis cannot have a PML");
  lsf::PmlAx& ax = e->pml[field][axis];
  ax.prof = (const float4*)prof;
  ax.ns = n_slab;
  ax.psi[0] = p0;
  ax.psi[1] = p1;
  ax.psi[2] = p2;
  ax.has_k = has_k ? 1 : 0;
  for (int c = 0; c < 3; ++c)
    if (c != axis && prof && !ax.psi[c]) return fail(LSF_EINVAL, "missing psi array");
  if (field == 1)
    for (int c = 0; c < 3; ++c) e->psiH[e->parity][axis][c] = ax.psi[c];
  return LSF_OK;
}

int lsf_set_ade(lsf_engine* e, const int32_t* chunk_base, void* P, int64_t n, int max_poles, const float* media,
                int n_media) {
  if (int rc0 = configurable(e)) return rc0;
  if (n > 0 && (!chunk_base || !P || !media)) return fail(LSF_EINVAL, "missing ADE arrays");
  if (max_poles < 0 || max_poles > lsf::MAX_POLES) return fail(LSF_EINVAL, "pure D-form media may have at most 4 poles");
  if (n > 0 && max_poles < 1) return fail(LSF_EINVAL, "D-form nodes need at least one pole");
  if (n >= (1ll << 30)) return fail(LSF_EINVAL, "too many D-form nodes for 32-bit indexing");
  if (n > 0 && (n_media * (2 + 3 * max_poles) > lsf::MAX_MEDIA_FLOATS ||
                n_media * (max_poles + 1) > lsf::MAX_MEDIA_VEC4))
    return fail(LSF_EINVAL, "too many D-form media for the E kernel's table (at most 16 media of 4 poles)");
  if (n > 0) {  // the E kernel advances max_poles poles for every node: a medium's unused poles must be all zero
    const int stride = 2 + 3 * max_poles;
    std::vector<float> tab((size_t)n_media * stride);
    const int rc = hip_check(hipMemcpy(tab.data(), media, tab.size() * sizeof(float), hipMemcpyDeviceToHost),
                             "D-form media table read");
    if (rc) return rc;
    for (int m = 0; m < n_media; ++m) {
      const float* t = tab.data() + (size_t)m * stride;
      const int np = (int)t[1];
      if ((float)np != t[1] || np < 0 || np > max_poles)
        return fail(LSF_EINVAL, "D-form media table: a medium's pole count must be an integer in [0, max_poles]");
      for (int q = 2 + 3 * np; q < stride; ++q)
        if (t[q] != 0.f) return fail(LSF_EINVAL, "D-form media table: the coefficients past a medium's poles must be 0");
    }
  }
  e->chunk_base = (const int*)chunk_base;
  e->P = P;
  e->n_ade = n;
  e->max_poles = max_poles;
  e->media = media;
  e->n_media = n_media;
  return LSF_OK;
}

int lsf_add_gather(lsf_engine* e, int phase, int field, int64_t n, int width, const int64_t* target, const void* src,
                   const int64_t* sidx, const void* w, const float* scale) {
  if (valid(e)) return LSF_EINVAL;
  if (n > 0 && (!target || !sidx || !w || width < 1)) return fail(LSF_EINVAL, "missing gather arrays");
  Op o;
  o.kind = OP_GATHER;
  o.phase = phase;
  o.field = field;
  o.n = n;
  o.width = width;
  o.target = (const long long*)target;
  o.src = src;
  o.sidx = (const long long*)sidx;
  o.w = w;
  o.scale = scale;
  return add_op(e, o);
}

int lsf_add_modal(lsf_engine* e, int phase, int field, int64_t n, int nm, const int64_t* target, const void* m,
                  const float* table, int64_t table_len, const float* scale) {
  if (valid(e)) return LSF_EINVAL;
  if (n > 0 && (!target || !m || !table || nm < 1)) return fail(LSF_EINVAL, "missing modal arrays");
  Op o;
  o.kind = OP_MODAL;
  o.phase = phase;
  o.field = field;
  o.n = n;
  o.width = nm;
  o.target = (const long long*)target;
  o.w = m;
  o.table = table;
  o.table_len = table_len;
  o.scale = scale;
  return add_op(e, o);
}

int lsf_add_table(lsf_engine* e, int phase, int field, int64_t n, const int64_t* target, const int32_t* row_ptr,
                  const float* T, int m, int64_t cmax, const int64_t* base, const int64_t* i0, const float* u0,
                  const float* coef, const float* scale) {
  if (valid(e)) return LSF_EINVAL;
  if (m < 1) return fail(LSF_EINVAL, "table stride must be >= 1");
  if (n > 0 && (!target || !row_ptr || !T || !base || !i0 || !u0 || !coef))
    return fail(LSF_EINVAL, "missing table arrays");
  Op o;
  o.kind = OP_TABLE;
  o.phase = phase;
  o.field = field;
  o.n = n;
  o.target = (const long long*)target;
  o.row_ptr = (const int*)row_ptr;
  o.table = T;
  o.m = m;
  o.cmax = cmax;
  o.base = (const long long*)base;
  o.i0 = (const long long*)i0;
  o.u0 = u0;
  o.coef = coef;
  o.scale = scale;
  return add_op(e, o);
}

int lsf_add_table_view(lsf_engine* e, int phase, int n, float* out, const float* T, int m, int64_t cmax,
                       const int64_t* base, const int64_t* i0, const float* u0) {
  if (valid(e)) return LSF_EINVAL;
  if (n < 1 || n > 64 || !out || !T || !base || !i0 || !u0 || m < 1) return fail(LSF_EINVAL, "bad table view");
  if (phase < LSF_PH_OFFDIAG || phase > LSF_PH_ADVANCE_E)
    return fail(LSF_EINVAL, "a table view is evaluated at sampling with the E half's counters: give an E-side phase");
  Op o;
  o.kind = OP_VIEW;
  o.phase = phase;
  o.n = n;
  o.out = out;
  o.table = T;
  o.m = m;
  o.cmax = cmax;
  o.base = (const long long*)base;
  o.i0 = (const long long*)i0;
  o.u0 = u0;
  return add_op(e, o);
}

int lsf_add_line1d(lsf_engine* e, int M, float* E1, float* H1, const float* ca, const float* cb, const float* da,
                   const float* db, int m_src, const float* wave, int64_t wave_len) {
  if (valid(e)) return LSF_EINVAL;
  if (M < 3 || !E1 || !H1 || !ca || !cb || !da || !db || !wave || m_src < 1 || m_src >= M - 1)
    return fail(LSF_EINVAL, "bad 1D line");
  Op o;
  o.n = M;
  o.e1 = E1;
  o.h1 = H1;
  o.ca = ca;
  o.cb = cb;
  o.da = da;
  o.db = db;
  o.m_src = m_src;
  o.table = wave;
  o.table_len = wave_len;
  o.kind = OP_LINE_H;
  o.phase = LSF_PH_ADVANCE_H;
  int rc = add_op(e, o);
  if (rc) return rc;
  o.kind = OP_LINE_E;
  o.phase = LSF_PH_ADVANCE_E;
  return add_op(e, o);
}

int lsf_add_subengine(lsf_engine* e, lsf_engine* sub) {
  if (valid(e) || valid(sub)) return LSF_EINVAL;
  if (sub == e || sub->tick) return fail(LSF_EINVAL, "a sub-engine must be another handle created with tick = 0");
  Op o;
  o.sub = sub;
  o.kind = OP_SUB_H;
  o.phase = LSF_PH_ADVANCE_H;
  int rc = add_op(e, o);
  if (rc) return rc;
  o.kind = OP_SUB_E;
  o.phase = LSF_PH_ADVANCE_E;
  return add_op(e, o);
}

int lsf_add_mixed_group(lsf_engine* e, int64_t n, int S, int max_poles, const int64_t* idx, const float* f,
                        const float* w, const float* h0, const float* inv_a, const float* slot_media, void* state,
                        int has_h, int has_a) {
  if (valid(e)) return LSF_EINVAL;
  if (n > 0 && (!idx || !f || !w || !h0 || !inv_a || !slot_media || !state || S < 1))
    return fail(LSF_EINVAL, "missing mixed-group arrays");
  if (!has_h && !has_a) return fail(LSF_EINVAL, "a mixed group needs a harmonic or an arithmetic path");
  Op o;
  o.kind = OP_MIXED;
  o.phase = LSF_PH_GROUPS;
  o.n = n;
  o.S = S;
  o.maxp = max_poles;
  o.target = (const long long*)idx;
  o.f = f;
  o.wmix = w;
  o.h0 = h0;
  o.inv_a = inv_a;
  o.sm = slot_media;
  o.state = state;
  o.has_h = has_h;
  o.has_a = has_a;
  return add_op(e, o);
}

int lsf_set_adjoint(lsf_engine* e, int on) {
  if (int rc0 = configurable(e)) return rc0;
  if (on) {
    if (e->cplx || any_periodic(e)) return fail(LSF_EINVAL, "the adjoint update takes real fields without periodic axes");
    if (e->n_ade > 0) return fail(LSF_EINVAL, "the adjoint update has no D-form (ADE) nodes");
    if (e->xw != 64) return fail(LSF_EINVAL, "the adjoint kernels take 64 x cells per wavefront row");
    for (const Op& o : e->ops)
      if (o.kind == OP_MIXED || o.kind == OP_TENSOR || o.kind == OP_SUB_H || o.kind == OP_SUB_E || o.kind == OP_LINE_H)
        return fail(LSF_EINVAL, "the adjoint engine takes only adjoint sources and gradient operations");
  }
  e->adjoint = on ? 1 : 0;
  return LSF_OK;
}

int lsf_add_tape(lsf_engine* e, int64_t n, const int64_t* nodes, float* tape, int64_t n_slots, const int64_t* base) {
  if (valid(e)) return LSF_EINVAL;
  if (n < 1 || !nodes || !tape || n_slots < 1 || !base) return fail(LSF_EINVAL, "bad tape");
  if (e->cplx || e->adjoint) return fail(LSF_EINVAL, "the tape records a real forward engine");
  Op o;
  o.kind = OP_TAPE;
  o.phase = LSF_PH_ADVANCE_H;
  o.field = 1;
  o.n = n;
  o.target = (const long long*)nodes;
  o.out = tape;
  o.n_slots = n_slots;
  o.base = (const long long*)base;
  return add_op(e, o);
}

int lsf_add_adjoint_gradient(lsf_engine* e, int64_t n, const int64_t* nodes, const float* cb, const float* tape,
                             int64_t n_slots, const int64_t* base, int64_t n_steps, double* g) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->adjoint) return fail(LSF_ESTATE, "gradient operations belong to an adjoint engine (lsf_set_adjoint)");
  if (n < 1 || !nodes || !cb || !tape || n_slots < 1 || !base || n_steps < 1 || !g)
    return fail(LSF_EINVAL, "bad adjoint gradient");
  Op o;
  o.kind = OP_ADJ_GRAD;
  o.phase = LSF_PH_ADVANCE_H;
  o.field = 0;
  o.n = n;
  o.target = (const long long*)nodes;
  o.coef = cb;
  o.table = tape;
  o.n_slots = n_slots;
  o.base = (const long long*)base;
  o.n_steps = n_steps;
  o.gacc = g;
  return add_op(e, o);
}

int lsf_add_adjoint_ade(lsf_engine* e, int64_t n, const int64_t* nodes, const int32_t* medium, int n_media,
                        int max_poles, const float* media, float* state, float* save) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->adjoint) return fail(LSF_ESTATE, "adjoint D-form updates belong to an adjoint engine (lsf_set_adjoint)");
  if (n < 1 || !nodes || !medium || n_media < 1 || max_poles < 1 || max_poles > 4 || !media || !state || !save)
    return fail(LSF_EINVAL, "bad adjoint D-form group");
  for (int ph = 0; ph < 2; ++ph) {
    Op o;
    o.kind = ph == 0 ? OP_ADJ_ADE_PRE : OP_ADJ_ADE_POST;
    o.phase = ph == 0 ? LSF_PH_BEFORE_H : LSF_PH_ADVANCE_H;
    o.field = 0;
    o.n = n;
    o.target = (const long long*)nodes;
    o.ade_medium = medium;
    o.coef = media;
    o.maxp = max_poles;
    o.S = n_media;
    o.state = state;
    o.out = save;
    if (int rc = add_op(e, o)) return rc;
  }
  return LSF_OK;
}

int lsf_add_adjoint_source(lsf_engine* e, int h_side, int nf, int64_t n_pts, const double* omega, double dt,
                           const void* W, int64_t n_rows, const int32_t* row_ptr, const int64_t* target,
                           const int64_t* point, const float* scale, int64_t n_steps, int stride) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->adjoint) return fail(LSF_ESTATE, "adjoint sources belong to an adjoint engine (lsf_set_adjoint)");
  if (nf < 1 || n_pts < 1 || n_rows < 1 || !omega || !W || !row_ptr || !target || !point || !scale || n_steps < 1 ||
      stride < 1)
    return fail(LSF_EINVAL, "bad adjoint source");
  Op o;
  o.kind = OP_ADJ_SRC;
  o.phase = h_side ? LSF_PH_CORRECT_H : LSF_PH_CORRECT_E;
  o.field = h_side ? 1 : 0;
  o.h_side = h_side ? 1 : 0;
  o.width = nf;
  o.n = n_rows;
  o.m_pts = n_pts;
  o.omega = omega;
  o.dt = dt;
  o.w = W;
  o.row_ptr = (const int*)row_ptr;
  o.target = (const long long*)target;
  o.t2 = (const long long*)point;
  o.scale = scale;
  o.n_steps = n_steps;
  o.dft_stride = stride;
  return add_op(e, o);
}

int lsf_add_tensor_group(lsf_engine* e, int64_t n_nodes, const int64_t* nodes, const int32_t* comp, int64_t n_vert,
                         const int32_t* vert_nodes, const void* vert_phase, const float* normal,
                         const int32_t* slot_vert, int S, int max_poles, const float* frac, const float* h0,
                         const float* inv_a, const float* slot_media, void* state) {
  if (valid(e)) return LSF_EINVAL;
  if (n_nodes < 1 || n_vert < 1 || !nodes || !comp || !vert_nodes || !normal || !slot_vert || !frac || !h0 ||
      !inv_a || !slot_media || !state || S < 1 || max_poles < 0)
    return fail(LSF_EINVAL, "missing tensor-group arrays");
  if (e->cplx != (vert_phase != nullptr))
    return fail(LSF_EINVAL, "a tensor group carries vertex phases exactly when the fields are complex");
  Op o;
  o.kind = OP_TENSOR;
  o.phase = LSF_PH_GROUPS;
  o.n = n_nodes;
  auto fill = [&](auto& J, auto* st) {
    using TT = std::remove_pointer_t<decltype(st)>;
    const long long nN = n_nodes, ns = 2 * nN;
    J.nN = nN;
    J.node = (const long long*)nodes;
    J.comp = comp;
    J.vn = vert_nodes;
    J.vph = (const TT*)vert_phase;
    J.normal = normal;
    J.slot_v = slot_vert;
    J.S = S;
    J.maxp = max_poles;
    J.ms = 2 + 3 * max_poles;
    J.f = frac;
    J.h0 = h0;
    J.inv_a = inv_a;
    J.sm = slot_media;
    J.D = st;
    J.Ep = st + nN;
    J.out = st + 2 * nN;
    J.Ea = st + 2 * nN + ns;
    J.Eh = J.Ea + ns;
    J.Ph = J.Eh + (long long)S * ns;
    J.Pa = J.Ph + (long long)S * max_poles * 2 * ns;
    J.E = nullptr;
  };
  if (e->cplx) fill(o.tjc, (c32*)state);
  else fill(o.tjf, (float*)state);
  return add_op(e, o);
}

int lsf_fuse_table_tensor(lsf_engine* e) {
  if (int rc0 = configurable(e)) return rc0;
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before fusing launches");
  int last_ce = -1, first_gr = -1;
  for (int i = 0; i < (int)e->ops.size(); ++i) {
    const Op& o = e->ops[i];
    if (o.kind == OP_VIEW) continue;
    if (o.phase == LSF_PH_CORRECT_E) last_ce = i;
    if (o.phase == LSF_PH_GROUPS && first_gr < 0) first_gr = i;
  }
  if (last_ce < 0 || e->ops[last_ce].kind != OP_TABLE || e->ops[last_ce].field != 0)
    return fail(LSF_EINVAL, "no E-side table correction at the end of the correction phase");
  if (first_gr < 0 || e->ops[first_gr].kind != OP_TENSOR)
    return fail(LSF_EINVAL, "no vertex-tensor group at the start of the group phase");
  e->ops[last_ce].fuse_tensor = first_gr;
  e->ops[first_gr].d_fused = 1;
  return LSF_OK;
}

int lsf_set_energy(lsf_engine* e, float* energy) {
  if (int rc0 = configurable(e)) return rc0;
  e->energy = energy;
  return LSF_OK;
}

int lsf_add_dft_monitor(lsf_engine* e, int nf, int64_t np, const double* omega, const int64_t* e_idx,
                        const int64_t* h1, const int64_t* h2, void* acc_e, void* acc_h, const void* esrc,
                        const void* hsrc, int src_complex, int energy_group) {
  if (int rc0 = configurable(e)) return rc0;
  if (nf < 1 || nf > 2048 || np < 1 || !omega || !e_idx || !acc_e) return fail(LSF_EINVAL, "bad DFT monitor");
  if (acc_h && (!h1 || !h2)) return fail(LSF_EINVAL, "H accumulators need h1 and h2");
  if (energy_group < 0 || energy_group > 2) return fail(LSF_EINVAL, "energy group must be 0, 1 or 2");
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before adding monitors");
  Monitor m;
  m.nf = nf;
  m.np = np;
  m.omega = omega;
  m.e = (const long long*)e_idx;
  m.h1 = (const long long*)h1;
  m.h2 = (const long long*)h2;
  m.acc_e = acc_e;
  m.acc_h = acc_h;
  m.esrc = esrc;
  m.hsrc = hsrc ? hsrc : esrc;
  m.src_complex = src_complex;
  m.group = energy_group - 1;
  m.nblocks = (int)((np + 255) / 256);
  m.poff = e->n_partial;
  e->n_partial += m.nblocks;
  e->mons.push_back(m);
  // rebuild the library scratch: partial sums and the per-monitor table
  if (e->partial) (void)hipFree(e->partial);
  if (e->mon_tab) (void)hipFree(e->mon_tab);
  e->partial = nullptr;
  e->mon_tab = nullptr;
  int rc = hip_check(hipMalloc(&e->partial, sizeof(float) * e->n_partial), "hipMalloc partials");
  if (rc) return rc;
  const int nm = (int)e->mons.size();
  std::vector<int> tab(2 * nm + 1);
  for (int q = 0; q < nm; ++q) {
    tab[q] = e->mons[q].poff;
    tab[nm + 1 + q] = e->mons[q].group;
  }
  tab[nm] = e->n_partial;
  rc = hip_check(hipMalloc(&e->mon_tab, sizeof(int) * tab.size()), "hipMalloc monitor table");
  if (rc) return rc;
  return hip_check(hipMemcpy(e->mon_tab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice),
                   "monitor table upload");
}

int lsf_add_time_monitor(lsf_engine* e, int64_t np, const int64_t* e_idx, void* buf, int64_t n_max, int64_t* pos) {
  if (int rc0 = configurable(e)) return rc0;
  if (np < 1 || !e_idx || !buf || !pos || n_max < 1) return fail(LSF_EINVAL, "bad time monitor");
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before adding monitors");
  TimeMon t;
  t.np = np;
  t.e = (const long long*)e_idx;
  t.buf = buf;
  t.n_max = n_max;
  t.pos = (long long*)pos;
  e->tmons.push_back(t);
  return LSF_OK;
}

int lsf_set_tiling(lsf_engine* e, int ry, int waves, int xw, int zc) {
  if (int rc0 = configurable(e)) return rc0;
  if (!lsf::tile_shape_ok(ry, waves, xw))
    return fail(LSF_EINVAL, "unsupported tile shape: rows per lane 1 or 2, waves 4 or 8, x cells per row 64 or 32");
  if (zc < 1) return fail(LSF_EINVAL, "z chunk must be >= 1");
  if (e->adjoint && xw != 64) return fail(LSF_EINVAL, "the adjoint kernels take 64 x cells per wavefront row");
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before retiling");
  e->ry = ry;
  e->waves = waves;
  e->xw = xw;
  e->zc = zc;
  return LSF_OK;
}

int lsf_set_tile_order(lsf_engine* e, int reverse_e) {
  if (int rc0 = configurable(e)) return rc0;
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before changing the tile order");
  e->rev_e = reverse_e ? 1 : 0;
  return LSF_OK;
}

int lsf_step(lsf_engine* e, int n, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  if (int rc1 = sync_args(e)) return rc1;
  hipStream_t s = (hipStream_t)stream;
  for (int q = 0; q < n; ++q) one_step(e, s);
  return hip_check(hipGetLastError(), "step launch");
}

}  // extern "C"

namespace {

// lsf_step_noalloc's device scratch: the two kernel-argument blocks of the engine being stepped, staged here on
// the caller's stream instead of in a library allocation. One array per GPU for the whole process; every call on
// a GPU shares it, so calls that overlap on one GPU (different streams, or host threads) race.
constexpr size_t kNoAllocScratchBytes = 8192;
static_assert(2 * sizeof(lsf::StepArgs) <= kNoAllocScratchBytes, "scratch too small for the kernel arguments");
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device (host code never uses the name)

// The device address of `scratch` on the current device, looked up once per device.
int scratch_address(void** out) {
  int dev = 0;
  if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice")) return rc;
  if (dev < 0 || dev >= kMaxDevices) return fail(LSF_EINVAL, "device index beyond the scratch table");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return rc;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LSF_OK;
}

}  // namespace

extern "C" {

int lsf_step_noalloc(lsf_engine* e, int n, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  if (n < 0) return fail(LSF_EINVAL, "negative step count");
  for (const Op& o : e->ops)
    if (o.kind == OP_SUB_H)
      return fail(LSF_EINVAL, "lsf_step_noalloc: an engine with sub-engines needs more kernel-argument scratch "
                              "than the 8 KiB it holds; use lsf_step");
  if (fused_eligible(e))
    return fail(LSF_EINVAL, "lsf_step_noalloc: single-pass stepping is not supported; use lsf_step");
  const size_t need = 2 * sizeof(StepArgs);
  if (need > kNoAllocScratchBytes) return fail(LSF_EINVAL, "lsf_step_noalloc: kernel arguments exceed the scratch");
  void* p = nullptr;
  if (int rc = scratch_address(&p)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Tiling t = own_tiling(e);
  e->host_args[0] = make_args(e, 0, t);
  e->host_args[1] = make_args(e, 1, t);
  for (int f = 0; f < 2; ++f) {
    e->nb[f] = e->host_args[f].nb;
    e->nbp[f] = e->host_args[f].nbp;
  }
  // stream-ordered upload: kernels of an earlier call on this stream have read the previous contents
  if (int rc = hip_check(hipMemcpyAsync(p, e->host_args, need, hipMemcpyHostToDevice, s), "argument staging"))
    return rc;
  e->use_fused = false;
  e->args_override = (const StepArgs*)p;
  for (int q = 0; q < n; ++q) one_step(e, s);
  e->args_override = nullptr;
  return hip_check(hipGetLastError(), "step launch");
}

int lsf_step_timed(lsf_engine* e, int n, void* stream, double* ms) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  if (!ms || n < 1) return fail(LSF_EINVAL, "bad timed step arguments");
  if (int rc1 = sync_args(e)) return rc1;
  hipStream_t s = (hipStream_t)stream;
  std::vector<hipEvent_t> ev(4 * (size_t)n, nullptr);
  int rc = LSF_OK;
  for (auto& x : ev)
    if ((rc = hip_check(hipEventCreate(&x), "event create"))) break;
  if (!rc) {
    for (int q = 0; q < n; ++q) {
      run_phase(e, LSF_PH_BEFORE_H, s);
      (void)hipEventRecord(ev[4 * q + 0], s);
      if (!e->use_fused) launch_h(e, s);
      (void)hipEventRecord(ev[4 * q + 1], s);
      run_phase(e, LSF_PH_CORRECT_H, s);
      run_phase(e, LSF_PH_ADVANCE_H, s);
      (void)hipEventRecord(ev[4 * q + 2], s);
      if (e->use_fused) launch_fused_step(e, s);  // the single pass is reported as the E time
      else launch_e(e, s);
      (void)hipEventRecord(ev[4 * q + 3], s);
      for (int ph = LSF_PH_OFFDIAG; ph <= LSF_PH_ADVANCE_E; ++ph) run_phase(e, ph, s);
    }
    rc = hip_check(hipGetLastError(), "timed step launch");
  }
  if (!rc) rc = hip_check(hipEventSynchronize(ev.back()), "event synchronize");
  if (!rc) {
    double th = 0.0, te = 0.0;
    for (int q = 0; q < n; ++q) {
      float a = 0.f, b = 0.f;
      (void)hipEventElapsedTime(&a, ev[4 * q + 0], ev[4 * q + 1]);
      (void)hipEventElapsedTime(&b, ev[4 * q + 2], ev[4 * q + 3]);
      th += e->use_fused ? 0.0 : a;  // no H kernel in a single pass
      te += b;
    }
    ms[0] += th;
    ms[1] += te;
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  return rc;
}

int lsf_sample(lsf_engine* e, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  sample(e, (hipStream_t)stream);
  return hip_check(hipGetLastError(), "sample launch");
}

int lsf_chunk(lsf_engine* e, int n_chunks, int stride, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  if (stride < 1) return fail(LSF_EINVAL, "stride must be >= 1");
  if (int rc1 = sync_args(e)) return rc1;
  hipStream_t s = (hipStream_t)stream;
  for (int c = 0; c < n_chunks; ++c) {
    for (int q = 0; q < stride; ++q) one_step(e, s);
    sample(e, s);
  }
  return hip_check(hipGetLastError(), "chunk launch");
}

}  // extern "C"

namespace {
int capture_chunk(lsf_engine* e, int stride, hipStream_t s, hipGraph_t* graph, hipGraphExec_t* exec) {
  int rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  for (int q = 0; q < stride; ++q) one_step(e, s);
  sample(e, s);
  hipGraph_t g = nullptr;
  rc = hip_check(hipStreamEndCapture(s, &g), "end capture");
  if (rc) return rc;
  *graph = g;
  rc = hip_check(hipGraphInstantiate(exec, g, nullptr, nullptr, 0), "graph instantiate");
  if (rc) {
    (void)hipGraphDestroy(g);
    *graph = nullptr;
    *exec = nullptr;
  }
  return rc;
}
}  // namespace

extern "C" {

int lsf_graph_capture(lsf_engine* e, int stride, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!stream) return fail(LSF_EINVAL, "graph capture needs a non-default stream");
  if (e->exec) return fail(LSF_ESTATE, "a graph is already captured");
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  if (stride < 1) return fail(LSF_EINVAL, "stride must be >= 1");
  if (int rc1 = sync_args(e)) return rc1;
  hipStream_t s = (hipStream_t)stream;
  const int p0 = e->parity;
  int rc = capture_chunk(e, stride, s, &e->graph, &e->exec);  // the capture advances the host parity only
  if (!rc && e->use_fused && (stride & 1)) rc = capture_chunk(e, stride, s, &e->graph2, &e->exec2);
  set_parity(e, p0);
  e->graph_parity = p0;
  e->graph_stride = stride;
  return rc;
}

int lsf_graph_replay(lsf_engine* e, int n, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->exec) return fail(LSF_ESTATE, "no captured graph");
  for (int q = 0; q < n; ++q) {
    hipGraphExec_t x = (e->exec2 && e->parity != e->graph_parity) ? e->exec2 : e->exec;
    const int rc = hip_check(hipGraphLaunch(x, (hipStream_t)stream), "graph launch");
    if (rc) return rc;
    if (e->use_fused && (e->graph_stride & 1)) set_parity(e, e->parity ^ 1);
  }
  return LSF_OK;
}

int lsf_parity(lsf_engine* e) { return e ? e->parity : 0; }

int lsf_set_fused(lsf_engine* e, void* E2, void* H2, void* const* psiH2) {
  if (int rc0 = configurable(e)) return rc0;
  if (!E2 || !H2) return fail(LSF_EINVAL, "missing second field buffers");
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it first");
  if (e->parity != 0) return fail(LSF_ESTATE, "second buffers are set before stepping");
  e->Ebuf[1] = E2;
  e->Hbuf[1] = H2;
  for (int ax = 0; ax < 3; ++ax) {
    for (int c = 0; c < 3; ++c) e->psiH[1][ax][c] = nullptr;
    if (!e->pml[1][ax].prof) continue;
    if (!psiH2 || !psiH2[ax]) return fail(LSF_EINVAL, "missing second H-side psi set");
    const long long gap = psi_gap(e, 1, ax, 0);
    if (gap < 0) return fail(LSF_EINVAL, "the H-side psi arrays of an axis must be contiguous");
    const int c1 = ax == 0 ? 1 : 0, c2 = ax == 2 ? 1 : 2;
    e->psiH[1][ax][c1] = psiH2[ax];
    e->psiH[1][ax][c2] = (float*)psiH2[ax] + gap;
  }
  e->fused_set = true;
  return LSF_OK;
}

int lsf_graph_reset(lsf_engine* e) {
  if (valid(e)) return LSF_EINVAL;
  int rc = LSF_OK;
  if (e->exec || e->graph) rc = hip_check(hipDeviceSynchronize(), "synchronize before graph reset");
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->exec2) (void)hipGraphExecDestroy(e->exec2);
  if (e->graph2) (void)hipGraphDestroy(e->graph2);
  e->exec = nullptr;
  e->graph = nullptr;
  e->exec2 = nullptr;
  e->graph2 = nullptr;
  return rc;
}

}  // extern "C"

// =============================================================================================
// Batches: many independent engines stepped by shared launches. Each half-step is one launch of the bulk
// kernel per kernel configuration over every member's tiles; the members' per-step operations are fused
// by position (the q-th operation of a phase over all members is one slot: gathers and modal sources of a
// slot become one launch each, sub-engines of a slot become a batch level of their own), and the DFT
// monitors of all members are sampled by one launch per source type.
// =============================================================================================

struct lsf_batch;

namespace {

struct KGroup {
  int cplx = 0, per = 0, mat = 0, np = 0, adj = 0;
  std::vector<lsf_engine*> mem;
  lsf::KLaunch L{};
};

struct Fused {
  int kind = 0;  // OP_GATHER or OP_MODAL
  int cplx = 0;
  void* jobs = nullptr;
  lsf::Sched sc{};
};

struct Level;

struct Slot {
  std::vector<Fused> fused;
  std::vector<std::pair<lsf_engine*, Op>> single;  // run one by one
  std::unique_ptr<Level> sub;                     // the sub-engines of SUB_H / SUB_E operations
};

struct DftLaunch {
  int cplx = 0;
  void* jobs = nullptr;
  lsf::Sched sc{};
  size_t shm = 0;
};

struct Level {
  std::vector<lsf_engine*> mem;
  std::vector<KGroup> hg, eg;
  std::vector<Slot> slots[LSF_PH_BEFORE_H + 1];
  std::vector<DftLaunch> dft;
  lsf::EnergyJob* energy = nullptr;
  int n_energy = 0;
};

}  // namespace

struct lsf_batch {
  std::vector<lsf_engine*> members;
  Level root;
  Tiling tiling{2, 4, 64, 16};
  int grid_blocks = 0;  // 0: one block per tile; else a persistent grid of this many blocks
  std::vector<void*> owned;         // device allocations of the plan
  std::vector<void*> owned_groups;  // device allocations of the root kernel groups (rebuilt on retiling)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int dev_upload(const void* src, size_t bytes, void** dst, std::vector<void*>& owned) {
  *dst = nullptr;
  if (bytes == 0) return LSF_OK;
  int rc = hip_check(hipMalloc(dst, bytes), "hipMalloc batch plan");
  if (rc) return rc;
  owned.push_back(*dst);
  return hip_check(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "batch plan upload");
}

// A work list over members with the given block counts: block -> member, member -> first block.
int make_sched(const std::vector<int>& counts, lsf::Sched* sc, std::vector<void*>& owned) {
  std::vector<int> inst, first;
  for (size_t m = 0; m < counts.size(); ++m) {
    first.push_back((int)inst.size());
    inst.insert(inst.end(), counts[m], (int)m);
  }
  sc->n_tiles = (int)inst.size();
  sc->n_pad = (sc->n_tiles + 7) / 8 * 8;
  void *di = nullptr, *df = nullptr;
  int rc = dev_upload(inst.data(), sizeof(int) * inst.size(), &di, owned);
  if (!rc) rc = dev_upload(first.data(), sizeof(int) * first.size(), &df, owned);
  sc->inst_of = (const int*)di;
  sc->first = (const int*)df;
  return rc;
}

// Bulk-kernel groups of a level under tiling t: members sharing a kernel instantiation share a launch.
int build_groups(Level& L, const Tiling& t, int grid_blocks, std::vector<void*>& owned) {
  L.hg.clear();
  L.eg.clear();
  for (int field = 0; field < 2; ++field) {
    std::vector<KGroup>& groups = field ? L.hg : L.eg;
    for (lsf_engine* e : L.mem) {
      KGroup key;
      key.cplx = e->cplx ? 1 : 0;
      key.per = any_periodic(e) ? 1 : 0;
      key.mat = field ? 0 : (e->mid ? 1 : 0);
      key.np = field ? 0 : pole_regs(e);
      key.adj = e->adjoint ? 1 : 0;
      KGroup* g = nullptr;
      for (KGroup& x : groups)
        if (x.cplx == key.cplx && x.per == key.per && x.mat == key.mat && x.np == key.np && x.adj == key.adj) g = &x;
      if (!g) {
        groups.push_back(key);
        g = &groups.back();
      }
      g->mem.push_back(e);
    }
    for (KGroup& g : groups) {
      // the adjoint kernels take 64-cell rows and one block per tile
      Tiling tg = t;
      if (g.adj) tg.xw = 64;
      std::vector<StepArgs> args;
      std::vector<int> counts;
      for (lsf_engine* e : g.mem) {
        args.push_back(make_args(e, field, tg));
        counts.push_back(args.back().nb);
      }
      void* da = nullptr;
      int rc = dev_upload(args.data(), sizeof(StepArgs) * args.size(), &da, owned);
      if (rc) return rc;
      g.L.args = (const StepArgs*)da;
      if ((rc = make_sched(counts, &g.L.sc, owned))) return rc;
      g.L.grid = g.L.sc.n_pad;
      if (grid_blocks > 0 && !g.adj) g.L.grid = std::min(g.L.sc.n_pad, (grid_blocks + 7) / 8 * 8);
      g.L.ry = tg.ry;
      g.L.waves = tg.waves;
      g.L.xw = tg.xw;
      g.L.cplx = g.cplx;
      g.L.per = g.per;
      g.L.mat = g.mat;
      g.L.np = g.np;
    }
  }
  return LSF_OK;
}

template <class T>
int fuse_ops(const std::vector<std::pair<lsf_engine*, Op>>& jobs, int kind, Slot& sl, std::vector<void*>& owned) {
  std::vector<int> counts;
  std::vector<lsf::GatherJob<T>> gj;
  std::vector<lsf::ModalJob<T>> mj;
  for (const auto& j : jobs) {
    counts.push_back((int)nblk(j.second.n, 256));
    if (kind == OP_GATHER) gj.push_back(gather_job<T>(j.first, j.second));
    else mj.push_back(modal_job<T>(j.first, j.second));
  }
  Fused f;
  f.kind = kind;
  f.cplx = sizeof(T) == sizeof(c32);
  int rc = kind == OP_GATHER ? dev_upload(gj.data(), sizeof(gj[0]) * gj.size(), &f.jobs, owned)
                             : dev_upload(mj.data(), sizeof(mj[0]) * mj.size(), &f.jobs, owned);
  if (!rc) rc = make_sched(counts, &f.sc, owned);
  if (!rc) sl.fused.push_back(f);
  return rc;
}

int build_level(Level& L, const Tiling& t, int grid_blocks, bool sample, std::vector<void*>& owned,
                std::vector<void*>& owned_groups);

int build_slot(const std::vector<std::pair<lsf_engine*, Op>>& jobs, Slot& sl, std::vector<void*>& owned) {
  // fusable kinds by field type; sub-engines as a level; the rest one by one
  std::vector<std::pair<lsf_engine*, Op>> part[2][2];  // [gather, modal][float, c32]
  std::vector<lsf_engine*> subs;
  for (const auto& j : jobs) {
    const Op& o = j.second;
    if ((o.kind == OP_GATHER || o.kind == OP_MODAL) && o.n > 0)
      part[o.kind == OP_MODAL][j.first->cplx ? 1 : 0].push_back(j);
    else if (o.kind == OP_SUB_H || o.kind == OP_SUB_E)
      subs.push_back(o.sub);
    else if (!((o.kind == OP_GATHER || o.kind == OP_MODAL) && o.n <= 0))
      sl.single.push_back(j);
  }
  int rc = LSF_OK;
  for (int k = 0; k < 2 && !rc; ++k) {
    const int kind = k ? OP_MODAL : OP_GATHER;
    if (!part[k][0].empty()) rc = fuse_ops<float>(part[k][0], kind, sl, owned);
    if (!rc && !part[k][1].empty()) rc = fuse_ops<c32>(part[k][1], kind, sl, owned);
  }
  if (!rc && !subs.empty()) {
    sl.sub = std::make_unique<Level>();
    sl.sub->mem = subs;
    rc = build_level(*sl.sub, Tiling{1, 4, 64, 16}, 0, false, owned, owned);
  }
  return rc;
}

int build_sampling(Level& L, std::vector<void*>& owned) {
  for (int cx = 0; cx < 2; ++cx) {
    std::vector<lsf::DftJob<float>> jf;
    std::vector<lsf::DftJob<c32>> jc;
    std::vector<int> counts;
    int max_nf = 0;
    for (lsf_engine* e : L.mem)
      for (const Monitor& m : e->mons) {
        if ((int)mon_complex(e, m) != cx) continue;
        if (cx) jc.push_back(dft_job<c32>(e, m));
        else jf.push_back(dft_job<float>(e, m));
        counts.push_back(m.nblocks);
        max_nf = std::max(max_nf, m.nf);
      }
    if (counts.empty()) continue;
    DftLaunch d;
    d.cplx = cx;
    d.shm = sizeof(c32) * 2 * max_nf;
    int rc = cx ? dev_upload(jc.data(), sizeof(jc[0]) * jc.size(), &d.jobs, owned)
                : dev_upload(jf.data(), sizeof(jf[0]) * jf.size(), &d.jobs, owned);
    if (!rc) rc = make_sched(counts, &d.sc, owned);
    if (rc) return rc;
    L.dft.push_back(d);
  }
  std::vector<lsf::EnergyJob> ej;
  for (lsf_engine* e : L.mem)
    if (has_energy(e)) ej.push_back(energy_job(e));
  L.n_energy = (int)ej.size();
  void* d = nullptr;
  const int rc = dev_upload(ej.data(), sizeof(lsf::EnergyJob) * ej.size(), &d, owned);
  L.energy = (lsf::EnergyJob*)d;
  return rc;
}

// Kernel groups of the level go to owned_groups (the root level's are rebuilt on retiling), the rest to owned.
int build_level(Level& L, const Tiling& t, int grid_blocks, bool sample, std::vector<void*>& owned,
                std::vector<void*>& owned_groups) {
  int rc = build_groups(L, t, grid_blocks, owned_groups);
  for (int ph = LSF_PH_CORRECT_H; ph <= LSF_PH_BEFORE_H && !rc; ++ph) {
    std::vector<std::vector<std::pair<lsf_engine*, Op>>> by_pos;
    for (lsf_engine* e : L.mem) {
      size_t q = 0;
      for (const Op& o : e->ops) {
        if (o.phase != ph || o.kind == OP_VIEW) continue;  // views run when the batch samples
        if (by_pos.size() <= q) by_pos.resize(q + 1);
        by_pos[q++].push_back({e, o});
      }
    }
    for (const auto& jobs : by_pos) {
      L.slots[ph].emplace_back();
      if ((rc = build_slot(jobs, L.slots[ph].back(), owned))) break;
    }
  }
  if (!rc && sample) rc = build_sampling(L, owned);
  return rc;
}

void level_h(Level& L, hipStream_t s);
void level_e(Level& L, hipStream_t s);

void run_slot(Slot& sl, bool h_half, hipStream_t s) {
  for (const Fused& f : sl.fused) {
    const dim3 grid(f.sc.n_tiles), block(256);
    if (f.kind == OP_GATHER) {
      if (f.cplx) hipLaunchKernelGGL(lsf::k_gather_multi<c32>, grid, block, 0, s, (const lsf::GatherJob<c32>*)f.jobs, f.sc);
      else hipLaunchKernelGGL(lsf::k_gather_multi<float>, grid, block, 0, s, (const lsf::GatherJob<float>*)f.jobs, f.sc);
    } else {
      if (f.cplx) hipLaunchKernelGGL(lsf::k_modal_multi<c32>, grid, block, 0, s, (const lsf::ModalJob<c32>*)f.jobs, f.sc);
      else hipLaunchKernelGGL(lsf::k_modal_multi<float>, grid, block, 0, s, (const lsf::ModalJob<float>*)f.jobs, f.sc);
    }
  }
  for (auto& j : sl.single) {
    if (j.first->cplx) run_op<c32>(j.first, j.second, s);
    else run_op<float>(j.first, j.second, s);
  }
  if (sl.sub) {
    if (h_half) level_h(*sl.sub, s);
    else level_e(*sl.sub, s);
  }
}

void level_h_kernels(Level& L, hipStream_t s) {
  for (KGroup& g : L.hg) {
    if (g.adj) lsf::launch_adjoint_h(g.L, s);
    else lsf::launch_update_h(g.L, s);
  }
}

void level_e_kernels(Level& L, hipStream_t s) {
  for (KGroup& g : L.eg) {
    lsf::KLaunch K = g.L;
    K.sc.rev = 1;  // each XCD from the end of its range, where the H launch ended (as a single engine's E)
    if (g.adj) lsf::launch_adjoint_e(K, s);
    else lsf::launch_update_e(K, s);
  }
}

void level_h_ops(Level& L, hipStream_t s) {
  for (int ph = LSF_PH_CORRECT_H; ph <= LSF_PH_ADVANCE_H; ++ph)
    for (Slot& sl : L.slots[ph]) run_slot(sl, true, s);
}

void level_e_ops(Level& L, hipStream_t s) {
  for (int ph = LSF_PH_OFFDIAG; ph <= LSF_PH_ADVANCE_E; ++ph)
    for (Slot& sl : L.slots[ph]) run_slot(sl, false, s);
}

void level_h(Level& L, hipStream_t s) {
  for (Slot& sl : L.slots[LSF_PH_BEFORE_H]) run_slot(sl, true, s);
  level_h_kernels(L, s);
  level_h_ops(L, s);
}

void level_e(Level& L, hipStream_t s) {
  level_e_kernels(L, s);
  level_e_ops(L, s);
}

void batch_sample(lsf_batch* b, hipStream_t s) {
  Level& L = b->root;
  for (lsf_engine* e : L.mem) run_views(e, s);
  for (const DftLaunch& d : L.dft) {
    const dim3 grid(d.sc.n_tiles), block(256);
    if (d.cplx)
      hipLaunchKernelGGL(lsf::k_dft_multi<c32>, grid, block, d.shm, s, (const lsf::DftJob<c32>*)d.jobs, d.sc);
    else
      hipLaunchKernelGGL(lsf::k_dft_multi<float>, grid, block, d.shm, s, (const lsf::DftJob<float>*)d.jobs, d.sc);
  }
  if (L.n_energy) hipLaunchKernelGGL(lsf::k_energy_multi, dim3(L.n_energy), dim3(64), 0, s, L.energy);
  for (lsf_engine* e : L.mem) sample_time(e, s);
}

void batch_step(lsf_batch* b, hipStream_t s) {
  level_h(b->root, s);
  level_e(b->root, s);
}

void hold(lsf_engine* e, int d) {
  e->held += d;
  for (const Op& o : e->ops)
    if (o.kind == OP_SUB_H) hold(o.sub, d);
}

void free_all(std::vector<void*>& v) {
  for (void* p : v) (void)hipFree(p);
  v.clear();
}

int valid_batch(lsf_batch* b) { return b ? LSF_OK : fail(LSF_EINVAL, "null batch handle"); }

}  // namespace

extern "C" {

int lsf_batch_create(lsf_engine* const* members, int n, lsf_batch** out) {
  if (!members || !out || n < 1) return fail(LSF_EINVAL, "a batch needs at least one engine");
  for (int q = 0; q < n; ++q) {
    lsf_engine* e = members[q];
    if (!e) return fail(LSF_EINVAL, "null engine in batch");
    if (!e->tick) return fail(LSF_EINVAL, "batch members are top-level engines (tick = 1)");
    if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "a batch member has no material");
    for (int r = 0; r < q; ++r)
      if (members[r] == e || members[r]->counter == e->counter)
        return fail(LSF_EINVAL, "batch members must be distinct engines with their own step counters");
  }
  lsf_batch* b = new lsf_batch();
  b->members.assign(members, members + n);
  b->root.mem = b->members;
  const int rc = build_level(b->root, b->tiling, b->grid_blocks, true, b->owned, b->owned_groups);
  if (rc) {
    free_all(b->owned);
    free_all(b->owned_groups);
    delete b;
    return rc;
  }
  for (lsf_engine* e : b->members) hold(e, 1);
  *out = b;
  return LSF_OK;
}

void lsf_batch_destroy(lsf_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipDeviceSynchronize();
  free_all(b->owned);
  free_all(b->owned_groups);
  for (lsf_engine* e : b->members) hold(e, -1);
  delete b;
}

int lsf_batch_set_schedule(lsf_batch* b, int ry, int waves, int xw, int zc, int grid_blocks) {
  if (valid_batch(b)) return LSF_EINVAL;
  if (!lsf::tile_shape_ok(ry, waves, xw))
    return fail(LSF_EINVAL, "unsupported tile shape: rows per lane 1 or 2, waves 4 or 8, x cells per row 64 or 32");
  if (zc < 1 || grid_blocks < 0) return fail(LSF_EINVAL, "z chunk must be >= 1 and grid blocks >= 0");
  if (b->exec) return fail(LSF_ESTATE, "the batch has a captured graph; reset it before rescheduling");
  int rc = hip_check(hipDeviceSynchronize(), "synchronize before rescheduling");
  if (rc) return rc;
  b->tiling = Tiling{ry, waves, xw, zc};
  b->grid_blocks = grid_blocks;
  free_all(b->owned_groups);
  return build_groups(b->root, b->tiling, b->grid_blocks, b->owned_groups);
}

int lsf_batch_tiles(lsf_batch* b, int64_t* n_tiles) {
  if (valid_batch(b)) return LSF_EINVAL;
  if (!n_tiles) return fail(LSF_EINVAL, "null output");
  int64_t t = 0;
  for (const KGroup& g : b->root.eg) t += g.L.sc.n_tiles;
  *n_tiles = t;
  return LSF_OK;
}

int lsf_batch_step(lsf_batch* b, int n, void* stream) {
  if (valid_batch(b)) return LSF_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  for (int q = 0; q < n; ++q) batch_step(b, s);
  return hip_check(hipGetLastError(), "batch step launch");
}

int lsf_batch_step_timed(lsf_batch* b, int n, void* stream, double* ms) {
  if (valid_batch(b)) return LSF_EINVAL;
  if (!ms || n < 1) return fail(LSF_EINVAL, "bad timed step arguments");
  hipStream_t s = (hipStream_t)stream;
  std::vector<hipEvent_t> ev(4 * (size_t)n, nullptr);
  int rc = LSF_OK;
  for (auto& x : ev)
    if ((rc = hip_check(hipEventCreate(&x), "event create"))) break;
  if (!rc) {
    for (int q = 0; q < n; ++q) {
      for (Slot& sl : b->root.slots[LSF_PH_BEFORE_H]) run_slot(sl, true, s);
      (void)hipEventRecord(ev[4 * q + 0], s);
      level_h_kernels(b->root, s);
      (void)hipEventRecord(ev[4 * q + 1], s);
      level_h_ops(b->root, s);
      (void)hipEventRecord(ev[4 * q + 2], s);
      level_e_kernels(b->root, s);
      (void)hipEventRecord(ev[4 * q + 3], s);
      level_e_ops(b->root, s);
    }
    rc = hip_check(hipGetLastError(), "timed batch step launch");
  }
  if (!rc) rc = hip_check(hipEventSynchronize(ev.back()), "event synchronize");
  if (!rc) {
    double th = 0.0, te = 0.0;
    for (int q = 0; q < n; ++q) {
      float x = 0.f, y = 0.f;
      (void)hipEventElapsedTime(&x, ev[4 * q + 0], ev[4 * q + 1]);
      (void)hipEventElapsedTime(&y, ev[4 * q + 2], ev[4 * q + 3]);
      th += x;
      te += y;
    }
    ms[0] += th;
    ms[1] += te;
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  return rc;
}

int lsf_batch_sample(lsf_batch* b, void* stream) {
  if (valid_batch(b)) return LSF_EINVAL;
  batch_sample(b, (hipStream_t)stream);
  return hip_check(hipGetLastError(), "batch sample launch");
}

int lsf_batch_chunk(lsf_batch* b, int n_chunks, int stride, void* stream) {
  if (valid_batch(b)) return LSF_EINVAL;
  if (stride < 1) return fail(LSF_EINVAL, "stride must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  for (int c = 0; c < n_chunks; ++c) {
    for (int q = 0; q < stride; ++q) batch_step(b, s);
    batch_sample(b, s);
  }
  return hip_check(hipGetLastError(), "batch chunk launch");
}

int lsf_batch_graph_capture(lsf_batch* b, int stride, void* stream) {
  if (valid_batch(b)) return LSF_EINVAL;
  if (!stream) return fail(LSF_EINVAL, "graph capture needs a non-default stream");
  if (b->exec) return fail(LSF_ESTATE, "a graph is already captured");
  if (stride < 1) return fail(LSF_EINVAL, "stride must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  int rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  for (int q = 0; q < stride; ++q) batch_step(b, s);
  batch_sample(b, s);
  hipGraph_t g = nullptr;
  rc = hip_check(hipStreamEndCapture(s, &g), "end capture");
  if (rc) return rc;
  b->graph = g;
  rc = hip_check(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
  if (rc) {
    (void)hipGraphDestroy(g);
    b->graph = nullptr;
    b->exec = nullptr;
  }
  return rc;
}

int lsf_batch_graph_replay(lsf_batch* b, int n, void* stream) {
  if (valid_batch(b)) return LSF_EINVAL;
  if (!b->exec) return fail(LSF_ESTATE, "no captured graph");
  for (int q = 0; q < n; ++q) {
    const int rc = hip_check(hipGraphLaunch(b->exec, (hipStream_t)stream), "graph launch");
    if (rc) return rc;
  }
  return LSF_OK;
}

int lsf_batch_graph_reset(lsf_batch* b) {
  if (valid_batch(b)) return LSF_EINVAL;
  int rc = LSF_OK;
  if (b->exec || b->graph) rc = hip_check(hipDeviceSynchronize(), "synchronize before graph reset");
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  return rc;
}

}  // extern "C"
