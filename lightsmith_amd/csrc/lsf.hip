// C ABI of the FDTD time-stepping library: engine handles, per-step launch sequence, HIP graphs.
// Declarations and conventions: include/lsf.h.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lsf.h"
#include "lsf_sparse.h"
#include "lsf_step.h"

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

enum OpKind { OP_GATHER, OP_MODAL, OP_TABLE, OP_VIEW, OP_LINE_H, OP_LINE_E, OP_SUB_H, OP_SUB_E, OP_MIXED };

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
  int ry = 2, waves = 4, zc = 16;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

using lsf::StepArgs;

int nchunk_of(const lsf_engine* e) { return (e->g.n[0] + 63) / 64; }

StepArgs make_args(const lsf_engine* e, int field) {
  StepArgs a{};
  a.g = e->g;
  a.E = e->E;
  a.H = e->H;
  a.cbd = e->cbd;
  a.mid = e->mid;
  a.mtab = e->mtab;
  for (int ax = 0; ax < 3; ++ax) a.pml[ax] = e->pml[field][ax];
  a.chunk_base = e->chunk_base;
  a.P = e->P;
  a.n_ade = (int)e->n_ade;
  a.nchunk = nchunk_of(e);
  a.max_poles = e->max_poles;
  a.media_stride = 2 + 3 * e->max_poles;
  a.media = e->media;
  a.n_media_floats = e->n_ade > 0 ? e->n_media * (2 + 3 * e->max_poles) : 0;
  a.counter = e->counter;
  a.tick = e->tick;
  const int rows = e->ry * e->waves;
  a.tiles_x = (e->g.n[0] + 63) / 64;
  a.tiles_y = (e->g.n[1] + rows - 1) / rows;
  a.zc = e->zc;
  const int tiles_z = (e->g.n[2] + e->zc - 1) / e->zc;
  a.nb = a.tiles_x * a.tiles_y * tiles_z;
  a.nbp = (a.nb + 7) / 8 * 8;
  return a;
}

template <class T, int RY, int W, bool PER>
void launch_h_t(const StepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((lsf::k_update_h<T, RY, W, PER>), dim3(a.nbp), dim3(64 * W), 0, s, a);
}

bool any_periodic(const lsf_engine* e) { return e->g.per[0] || e->g.per[1] || e->g.per[2]; }

template <class T, bool PER>
void launch_h_p(const lsf_engine* e, hipStream_t s) {
  const StepArgs a = make_args(e, 1);
  const bool w8 = e->waves == 8;
  if (e->ry == 2) {
    if (w8) launch_h_t<T, 2, 8, PER>(a, s);
    else launch_h_t<T, 2, 4, PER>(a, s);
  } else {
    if (w8) launch_h_t<T, 1, 8, PER>(a, s);
    else launch_h_t<T, 1, 4, PER>(a, s);
  }
}

template <class T>
void launch_h(const lsf_engine* e, hipStream_t s) {
  if (any_periodic(e)) launch_h_p<T, true>(e, s);
  else launch_h_p<T, false>(e, s);
}

template <class T, int MAT, bool PER, int NP>
void launch_e_np(const lsf_engine* e, hipStream_t s) {
  const StepArgs a = make_args(e, 0);
  const dim3 grid(a.nbp);
  const bool w8 = e->waves == 8;
  if (e->ry == 2) {
    if (w8) hipLaunchKernelGGL((lsf::k_update_e<T, MAT, 2, 8, PER, NP>), grid, dim3(512), 0, s, a);
    else hipLaunchKernelGGL((lsf::k_update_e<T, MAT, 2, 4, PER, NP>), grid, dim3(256), 0, s, a);
  } else {
    if (w8) hipLaunchKernelGGL((lsf::k_update_e<T, MAT, 1, 8, PER, NP>), grid, dim3(512), 0, s, a);
    else hipLaunchKernelGGL((lsf::k_update_e<T, MAT, 1, 4, PER, NP>), grid, dim3(256), 0, s, a);
  }
}

template <class T, int MAT, bool PER>
void launch_e_m(const lsf_engine* e, hipStream_t s) {
  const int np = e->n_ade > 0 ? e->max_poles : 0;
  if (np == 0) launch_e_np<T, MAT, PER, 0>(e, s);
  else if (np <= 2) launch_e_np<T, MAT, PER, 2>(e, s);
  else if (np == 3) launch_e_np<T, MAT, PER, 3>(e, s);
  else launch_e_np<T, MAT, PER, 4>(e, s);
}

template <class T>
void launch_e(const lsf_engine* e, hipStream_t s) {
  const bool per = any_periodic(e);
  if (e->mid) {
    if (per) launch_e_m<T, 1, true>(e, s);
    else launch_e_m<T, 1, false>(e, s);
  } else {
    if (per) launch_e_m<T, 0, true>(e, s);
    else launch_e_m<T, 0, false>(e, s);
  }
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void run_phase(lsf_engine* e, int phase, hipStream_t s);

template <class T>
void run_op(lsf_engine* e, const Op& o, hipStream_t s) {
  const int slot = o.phase <= LSF_PH_ADVANCE_H ? 0 : 1;
  T* F = (T*)(o.field == 0 ? e->E : e->H);
  switch (o.kind) {
    case OP_GATHER:
      if (o.n > 0)
        hipLaunchKernelGGL(lsf::k_gather<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, o.n, o.width, o.target, F,
                           (const T*)o.src, o.sidx, (const T*)o.w, o.scale);
      break;
    case OP_MODAL:
      if (o.n > 0)
        hipLaunchKernelGGL(lsf::k_modal<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, o.n, o.width, o.target, F,
                           (const T*)o.w, o.table, o.table_len, (const long long*)e->counter, slot, o.scale);
      break;
    case OP_TABLE:
      if (o.n > 0)
        hipLaunchKernelGGL(lsf::k_table<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, o.n, o.target, o.row_ptr, F,
                           o.table, o.m, o.cmax, o.base, o.i0, o.u0, o.coef, (const long long*)e->counter, slot,
                           o.scale);
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
      if (o.sub->cplx) launch_h<c32>(o.sub, s);
      else launch_h<float>(o.sub, s);
      run_phase(o.sub, LSF_PH_CORRECT_H, s);
      run_phase(o.sub, LSF_PH_ADVANCE_H, s);
      break;
    case OP_SUB_E:
      if (o.sub->cplx) launch_e<c32>(o.sub, s);
      else launch_e<float>(o.sub, s);
      for (int ph = LSF_PH_OFFDIAG; ph <= LSF_PH_ADVANCE_E; ++ph) run_phase(o.sub, ph, s);
      break;
    case OP_MIXED:
      if (o.n > 0)
        hipLaunchKernelGGL(lsf::k_mixed<T>, dim3(nblk(o.n, 256)), dim3(256), 0, s, o.n, o.S, o.maxp, o.target, o.f,
                           o.wmix, o.h0, o.inv_a, o.sm, (T*)o.state, o.has_h, o.has_a, (T*)e->E);
      break;
  }
}

void run_phase(lsf_engine* e, int phase, hipStream_t s) {
  for (const Op& o : e->ops) {
    if (o.phase != phase) continue;
    if (e->cplx) run_op<c32>(e, o, s);
    else run_op<float>(e, o, s);
  }
}

void one_step(lsf_engine* e, hipStream_t s) {
  if (e->cplx) launch_h<c32>(e, s);
  else launch_h<float>(e, s);
  run_phase(e, LSF_PH_CORRECT_H, s);
  run_phase(e, LSF_PH_ADVANCE_H, s);
  if (e->cplx) launch_e<c32>(e, s);
  else launch_e<float>(e, s);
  for (int ph = LSF_PH_OFFDIAG; ph <= LSF_PH_ADVANCE_E; ++ph) run_phase(e, ph, s);
}

void sample(lsf_engine* e, hipStream_t s) {
  for (const Monitor& m : e->mons) {
    const size_t shm = sizeof(c32) * 2 * m.nf;
    float* part = m.group >= 0 ? e->partial + m.poff : nullptr;
    const void* es = m.esrc ? m.esrc : e->E;
    const void* hs = m.hsrc ? m.hsrc : e->H;
    const bool cx = (m.esrc ? m.src_complex : e->cplx);
    if (cx)
      hipLaunchKernelGGL(lsf::k_dft<c32>, dim3(m.nblocks), dim3(256), shm, s, m.nf, m.np, m.omega, e->dt,
                         (const long long*)e->counter, (const c32*)es, (const c32*)hs, m.e, m.h1, m.h2,
                         (c32*)m.acc_e, (c32*)m.acc_h, part);
    else
      hipLaunchKernelGGL(lsf::k_dft<float>, dim3(m.nblocks), dim3(256), shm, s, m.nf, m.np, m.omega,
                         e->dt, (const long long*)e->counter, (const float*)es, (const float*)hs, m.e,
                         m.h1, m.h2, (c32*)m.acc_e, (c32*)m.acc_h, part);
  }
  bool any_energy = false;
  for (const Monitor& m : e->mons) any_energy |= m.group >= 0;
  if (any_energy && e->energy)
    hipLaunchKernelGGL(lsf::k_energy, dim3(1), dim3(64), 0, s, (int)e->mons.size(), e->mon_tab,
                       e->mon_tab + e->mons.size() + 1, e->partial, e->energy);
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

int valid(lsf_engine* e) { return e ? LSF_OK : fail(LSF_EINVAL, "null engine handle"); }

int add_op(lsf_engine* e, const Op& o) {
  if (o.phase < LSF_PH_CORRECT_H || o.phase > LSF_PH_ADVANCE_E) return fail(LSF_EINVAL, "bad phase");
  if (o.field != 0 && o.field != 1) return fail(LSF_EINVAL, "field must be 0 (E) or 1 (H)");
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before adding work");
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
  delete e;
}

int lsf_set_material_dense(lsf_engine* e, const float* cb) {
  if (valid(e)) return LSF_EINVAL;
  if (!cb) return fail(LSF_EINVAL, "null material");
  e->cbd = cb;
  e->mid = nullptr;
  e->mtab = nullptr;
  return LSF_OK;
}

int lsf_set_material_indexed(lsf_engine* e, const uint32_t* ids, const float* table) {
  if (valid(e)) return LSF_EINVAL;
  if (!ids || !table) return fail(LSF_EINVAL, "null material");
  e->cbd = nullptr;
  e->mid = ids;
  e->mtab = table;
  return LSF_OK;
}

int lsf_set_pml(lsf_engine* e, int field, int axis, const void* prof, int n_slab, void* p0, void* p1, void* p2,
                int has_k) {
  if (valid(e)) return LSF_EINVAL;
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
  return LSF_OK;
}

int lsf_set_ade(lsf_engine* e, const int32_t* chunk_base, void* P, int64_t n, int max_poles, const float* media,
                int n_media) {
  if (valid(e)) return LSF_EINVAL;
  if (n > 0 && (!chunk_base || !P || !media)) return fail(LSF_EINVAL, "missing ADE arrays");
  if (max_poles < 0 || max_poles > lsf::MAX_POLES) return fail(LSF_EINVAL, "pure D-form media may have at most 4 poles");
  if (n > 0 && max_poles < 1) return fail(LSF_EINVAL, "D-form nodes need at least one pole");
  if (n >= (1ll << 30)) return fail(LSF_EINVAL, "too many D-form nodes for 32-bit indexing");
  if (n > 0 && n_media * (2 + 3 * max_poles) > lsf::MAX_MEDIA_FLOATS)
    return fail(LSF_EINVAL, "too many D-form media for the E kernel's table (at most 16 media of 4 poles)");
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
  if (n > 0 && (!target || !src || !sidx || !w || width < 1)) return fail(LSF_EINVAL, "missing gather arrays");
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

int lsf_set_energy(lsf_engine* e, float* energy) {
  if (valid(e)) return LSF_EINVAL;
  e->energy = energy;
  return LSF_OK;
}

int lsf_add_dft_monitor(lsf_engine* e, int nf, int64_t np, const double* omega, const int64_t* e_idx,
                        const int64_t* h1, const int64_t* h2, void* acc_e, void* acc_h, const void* esrc,
                        const void* hsrc, int src_complex, int energy_group) {
  if (valid(e)) return LSF_EINVAL;
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
  if (valid(e)) return LSF_EINVAL;
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

int lsf_set_tiling(lsf_engine* e, int ry, int waves, int zc) {
  if (valid(e)) return LSF_EINVAL;
  if (!(ry == 1 || ry == 2)) return fail(LSF_EINVAL, "rows per lane must be 1 or 2");
  if (!(waves == 4 || waves == 8)) return fail(LSF_EINVAL, "waves per block must be 4 or 8");
  if (zc < 1) return fail(LSF_EINVAL, "z chunk must be >= 1");
  if (e->exec) return fail(LSF_ESTATE, "the engine has a captured graph; reset it before retiling");
  e->ry = ry;
  e->waves = waves;
  e->zc = zc;
  return LSF_OK;
}

int lsf_step(lsf_engine* e, int n, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  hipStream_t s = (hipStream_t)stream;
  for (int q = 0; q < n; ++q) one_step(e, s);
  return hip_check(hipGetLastError(), "step launch");
}

int lsf_step_timed(lsf_engine* e, int n, void* stream, double* ms) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  if (!ms || n < 1) return fail(LSF_EINVAL, "bad timed step arguments");
  hipStream_t s = (hipStream_t)stream;
  std::vector<hipEvent_t> ev(4 * (size_t)n, nullptr);
  int rc = LSF_OK;
  for (auto& x : ev)
    if ((rc = hip_check(hipEventCreate(&x), "event create"))) break;
  if (!rc) {
    for (int q = 0; q < n; ++q) {
      (void)hipEventRecord(ev[4 * q + 0], s);
      if (e->cplx) launch_h<c32>(e, s);
      else launch_h<float>(e, s);
      (void)hipEventRecord(ev[4 * q + 1], s);
      run_phase(e, LSF_PH_CORRECT_H, s);
      run_phase(e, LSF_PH_ADVANCE_H, s);
      (void)hipEventRecord(ev[4 * q + 2], s);
      if (e->cplx) launch_e<c32>(e, s);
      else launch_e<float>(e, s);
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
      th += a;
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
  hipStream_t s = (hipStream_t)stream;
  for (int c = 0; c < n_chunks; ++c) {
    for (int q = 0; q < stride; ++q) one_step(e, s);
    sample(e, s);
  }
  return hip_check(hipGetLastError(), "chunk launch");
}

int lsf_graph_capture(lsf_engine* e, int stride, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!stream) return fail(LSF_EINVAL, "graph capture needs a non-default stream");
  if (e->exec) return fail(LSF_ESTATE, "a graph is already captured");
  if (!e->cbd && !e->mid) return fail(LSF_ESTATE, "no material set");
  hipStream_t s = (hipStream_t)stream;
  int rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  for (int q = 0; q < stride; ++q) one_step(e, s);
  sample(e, s);
  hipGraph_t g = nullptr;
  rc = hip_check(hipStreamEndCapture(s, &g), "end capture");
  if (rc) return rc;
  e->graph = g;
  rc = hip_check(hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0), "graph instantiate");
  if (rc) {
    (void)hipGraphDestroy(g);
    e->graph = nullptr;
    e->exec = nullptr;
  }
  return rc;
}

int lsf_graph_replay(lsf_engine* e, int n, void* stream) {
  if (valid(e)) return LSF_EINVAL;
  if (!e->exec) return fail(LSF_ESTATE, "no captured graph");
  for (int q = 0; q < n; ++q) {
    int rc = hip_check(hipGraphLaunch(e->exec, (hipStream_t)stream), "graph launch");
    if (rc) return rc;
  }
  return LSF_OK;
}

int lsf_graph_reset(lsf_engine* e) {
  if (valid(e)) return LSF_EINVAL;
  int rc = LSF_OK;
  if (e->exec || e->graph) rc = hip_check(hipDeviceSynchronize(), "synchronize before graph reset");
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  return rc;
}

}  // extern "C"
