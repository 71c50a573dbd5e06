// The bulk Yee update: one H kernel and one E kernel per time step, each fusing the curl, the CPML
// psi recursions and (E kernel) the pure-ADE D-form update for its cells.
//
// Mapping: a wavefront's 64 lanes take XW consecutive x cells (XW = 64: one row; XW = 32: two rows) and
// each lane holds RY consecutive y rows in registers and marches up a chunk of z planes, carrying the
// neighbouring plane in registers. x neighbours come from the adjacent lane (cross-lane shuffle; the edge
// lane loads its halo cell), y neighbours from the lane's own rows (the edge row loads one halo row), z
// neighbours from the marching registers. Every z iteration issues all of its loads before any store, so
// one memory round trip covers the iteration (stores to H or E would otherwise order each following load
// behind them).
//
// A launch covers the tiles of one engine or of many (a batch): the engines' arguments sit in a device
// array, a work list maps each tile to its engine, and each block walks the list with a grid stride (one
// tile per block when the grid covers the list; a persistent loop when it is smaller).
#pragma once

#include "lsf_common.h"

namespace lsf {

struct PmlAx {
  const float4* prof;  // per position along the axis: {b, c, 1/kappa, slab index or -1}; null = no PML
  void* psi[3];        // psi of component c (null for c == axis)
  int ns;              // slab extent along the axis
  int has_k;           // kappa_max > 1 somewhere on this axis
};

struct StepArgs {
  Grid g;
  void* E;
  void* H;
  const float* cbd;      // dense material [3][N], or null
  const uint32_t* mid;   // indexed material [N], or null
  const float* mtab;     // indexed material table [256]
  PmlAx pml[3];          // this kernel's field
  PmlAx xpml[3];         // the other field's CPML (read by the adjoint kernels)
  const int* chunk_base; // first D-form slot of each 32-cell chunk [3][N2][N1][nchunk]
  void* P;
  int n_ade;
  int nchunk;            // 32-cell chunks per x row
  int max_poles;
  int media_stride;
  const float* media;
  int n_media_floats;
  long long* counter;
  int tick;
  int tiles_x, tiles_y, zc, nb, nbp;
};

// The tiles of one launch: every member's tiles laid end to end.
struct Sched {
  const int* inst_of;  // [n_tiles] member of each tile; null: one member (index 0)
  const int* first;    // [members] first tile of each member
  int n_tiles;
  int n_pad;           // n_tiles rounded up to a multiple of 8 (the period of the XCD-aware order)
  int rev = 0;         // each XCD walks its tile range backwards (xcd_tile)
};

// D-form media table in device memory: n_media * (2 + 3 * max_poles) floats
// ({1/eps_inf, poles, c1 c2 c3 of each pole}, zeros past a medium's poles)
constexpr int MAX_MEDIA_FLOATS = 16 * (2 + 3 * 4);
// The bulk E kernel's LDS copy: NP + 1 float4 per medium, {1/eps_inf, poles, 0, 0} then {c1, c2, c3, 0} per
// pole. Every table that MAX_MEDIA_FLOATS admits fits (44 media of 1 pole: 88 float4).
constexpr int MAX_MEDIA_VEC4 = 88;

template <class T>
struct Pair {
  T p, q;  // P^n, P^{n-1}
};

// pure-ADE media: at most this many poles each (the E kernel keeps every pole pair of a row in registers)
constexpr int MAX_POLES = 4;

__device__ __forceinline__ float4 no_pml() { return make_float4(1.f, 0.f, 1.f, -1.f); }

// A load whose address is the same for every lane, through the constant address space: it becomes a scalar
// load into SGPRs (the data is never written by the kernels that read it this way).
typedef const __attribute__((address_space(4))) float* cfloat_ptr;
typedef const __attribute__((address_space(4))) int* cint_ptr;
__device__ __forceinline__ float4 ld_uniform(const float4* p, unsigned idx) {
  cfloat_ptr q = (cfloat_ptr)(p + idx);
  return make_float4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ int ld_uniform(const int* p, unsigned idx) { return ((cint_ptr)p)[idx]; }

__device__ __forceinline__ unsigned cell32(unsigned nx, unsigned ny, int i, int j, int k) {
  return ((unsigned)k * ny + (unsigned)j) * nx + (unsigned)i;
}

// psi index of a cell in the slab array of `axis` (slab extent ns along the axis, full extent elsewhere)
__device__ __forceinline__ unsigned psi_index(int axis, int p, int ns, unsigned nx, unsigned ny, int i, int j, int k) {
  if (axis == 0) return ((unsigned)k * ny + (unsigned)j) * (unsigned)ns + (unsigned)p;
  if (axis == 1) return ((unsigned)k * (unsigned)ns + (unsigned)p) * nx + (unsigned)i;
  return ((unsigned)p * ny + (unsigned)j) * nx + (unsigned)i;
}

// F(i, j, k) where an index may be a periodic axis' high ghost (n - 1): its image at 1 times exp(+i k L)
// (engine.py:595-611, low=False). Without periodic axes (PER false) a plain load.
template <bool PER, class T>
__device__ __forceinline__ T load_hi(const T* __restrict__ F, const Grid& g, int i, int j, int k) {
  if (!PER) return F[cell32(g.n[0], g.n[1], i, j, k)];
  return ld_hi(F, g, i, j, k);
}

// F(i, j, k) where an index may be a periodic axis' low ghost (0): its image at n - 2 times exp(-i k L).
template <bool PER, class T>
__device__ __forceinline__ T load_lo(const T* __restrict__ F, const Grid& g, int i, int j, int k) {
  if (!PER) return F[cell32(g.n[0], g.n[1], i, j, k)];
  return ld_lo(F, g, i, j, k);
}

// psi = b psi + cp der; returns the field increment kp der (when kappa > 1) + psi_new
// (engine.py:636-643, 656-663) with cp = coef_s c and kp = coef_s (1/kappa - 1).
template <class T>
__device__ __forceinline__ T pml_apply(const float4 pr, int has_k, float coef_s, T der, T& psi) {
  psi = fmad(pr.x, psi, scl(coef_s * pr.y, der));
  return has_k ? fmad(coef_s * (pr.z - 1.f), der, psi) : psi;
}

// The stepped boxes follow from the grid and its periodic axes (engine.py:47-66, no mirror planes; the host builds
// Grid::rE/rH the same way): a periodic axis steps [1, n - 1); otherwise E_c spans [0, n - 1) along c and
// [1, n - 1) across, H_c [0, n) along c and [0, n - 1) across. Computing them here keeps the 36 bounds out of the
// scalar registers. PER false: no axis is periodic. Indices are >= 0.
template <bool PER>
__device__ __forceinline__ bool e_box(const Grid& g, int c, int i, int j, int k) {
  const int v[3] = {i, j, k};
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = g.n[a];
    const bool per = PER && g.per[a];
    ok = ok && (per || a != c ? (v[a] >= 1 && v[a] < n - 1) : v[a] < n - 1);
  }
  return ok;
}
template <bool PER>
__device__ __forceinline__ bool h_box(const Grid& g, int c, int i, int j, int k) {
  const int v[3] = {i, j, k};
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = g.n[a];
    const bool per = PER && g.per[a];
    ok = ok && (per ? (v[a] >= 1 && v[a] < n - 1) : (a == c ? v[a] < n : v[a] < n - 1));
  }
  return ok;
}

// The y profile of a lane's row r: rows jw + r (lanes of the first x row) or jw + RY + r (the second x row
// when XW = 32). Both are wavefront-uniform loads.
template <int XW, int RY>
__device__ __forceinline__ float4 row_prof(const float4* prof, int jw, int yl, int r, int ny) {
  if (!prof) return no_pml();
  const float4 pa = jw + r < ny ? ld_uniform(prof, jw + r) : no_pml();
  if (XW == 64) return pa;
  const float4 pb = jw + RY + r < ny ? ld_uniform(prof, jw + RY + r) : no_pml();
  return yl ? pb : pa;
}

// Member and member-local tile of work-list entry t.
__device__ __forceinline__ void member_of(const Sched& sc, int t, int& m, int& lt) {
  m = 0;
  lt = t;
  if (sc.inst_of) {
    m = ld_uniform(sc.inst_of, t);
    lt = t - ld_uniform(sc.first, m);
  }
}

// ---------------------------------------------------------------------------------------------
// H update (engine.py:630-643): H_c -= cH [s_d1 (E_d2(+e_d1) - E_d2) - s_d2 (E_d1(+e_d2) - E_d1)] + CPML.
// ---------------------------------------------------------------------------------------------
template <class T, int RY, int WAVES, int XW, bool PER>
__device__ __forceinline__ void h_tile(const StepArgs& a, int t) {
  const Grid& g = a.g;
  const int lane = threadIdx.x & 63;
  const int xl = lane & (XW - 1);
  const int yl = XW == 64 ? 0 : (lane >> 5);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int jw = (ty * WAVES + wv) * (64 / XW) * RY;  // the wavefront's first row
  if (jw >= g.n[1]) return;
  const int i = tx * XW + xl;
  const int j0 = jw + yl * RY;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, g.n[2]);
  const unsigned N = (unsigned)g.N, nx = g.n[0], ny = g.n[1];
  const gptr<const T> Eg = as_global((const T*)a.E);
  const gptr<T> Hg = as_global((T*)a.H);
  const bool inx = i < g.n[0];
  const int ih = tx * XW + XW;
  const bool hx_ok = (xl == XW - 1) && ih < g.n[0];
  gptr<T> psi[3][3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
#pragma unroll
    for (int c = 0; c < 3; ++c) psi[ax][c] = as_global((T*)a.pml[ax].psi[c]);

  // CPML profiles: x per lane and y per row are constant over the march
  const float4 px = (a.pml[0].prof && inx) ? ((const float4*)as_global(a.pml[0].prof))[i] : no_pml();
  float4 py[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) py[r] = row_prof<XW, RY>(a.pml[1].prof, jw, yl, r, g.n[1]);

  T cEx[RY + 1], cEy[RY], cEz[RY + 1], hEy[RY], hEz[RY];
  T nEx[RY + 1], nEy[RY], nEz[RY + 1], nhEy[RY], nhEz[RY];

  auto load_plane = [&](int k, T* xE, T* yE, T* zE, T* hy, T* hz) {
    const T* __restrict__ Ex = (const T*)Eg;
    const T* __restrict__ Ey = Ex + N;
    const T* __restrict__ Ez = Ex + 2 * N;
#pragma unroll
    for (int r = 0; r <= RY; ++r) {
      const int j = j0 + r;
      const bool ok = inx && j < g.n[1] && k < g.n[2];
      xE[r] = ok ? load_hi<PER>(Ex, g, i, j, k) : zero<T>();
      zE[r] = ok ? load_hi<PER>(Ez, g, i, j, k) : zero<T>();
      if (r < RY) {
        yE[r] = ok ? load_hi<PER>(Ey, g, i, j, k) : zero<T>();
        const bool hok = hx_ok && j < g.n[1] && k < g.n[2];
        hy[r] = hok ? load_hi<PER>(Ey, g, ih, j, k) : zero<T>();
        hz[r] = hok ? load_hi<PER>(Ez, g, ih, j, k) : zero<T>();
      }
    }
  };

  load_plane(z0, cEx, cEy, cEz, hEy, hEz);
  for (int k = z0; k < z1; ++k) {
    const float4 pz = a.pml[2].prof ? ld_uniform(a.pml[2].prof, k) : no_pml();
    // ---- loads: the next E plane, old H, psi
    load_plane(k + 1, nEx, nEy, nEz, nhEy, nhEz);
    T* __restrict__ H = (T*)Hg;
    bool st[RY][3];
    T hv[RY][3];
    T ps[RY][3][2];  // psi of component c along axes d1 = c+1, d2 = c+2 (mod 3)
    float4 pr[RY][3][2];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const unsigned cell = cell32(nx, ny, i, j, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        st[r][c] = inx && j < g.n[1] && h_box<PER>(g, c, i, j, k);
        hv[r][c] = st[r][c] ? H[(unsigned)c * N + cell] : zero<T>();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          pr[r][c][s] = ax == 0 ? px : (ax == 1 ? py[r] : pz);
          const int p = (int)pr[r][c][s].w;
          ps[r][c][s] = zero<T>();
          if (st[r][c] && p >= 0)
            ps[r][c][s] = ((const T*)psi[ax][c])[psi_index(ax, p, a.pml[ax].ns, nx, ny, i, j, k)];
        }
      }
    }
    // ---- compute
    T ezp[RY], eyp[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      ezp[r] = shfl_dn(cEz[r]);
      eyp[r] = shfl_dn(cEy[r]);
      if (xl == XW - 1) {
        ezp[r] = hEz[r];
        eyp[r] = hEy[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // da: difference along d1 of E_d2; db: difference along d2 of E_d1
        T da, db;
        if (c == 0) {
          da = sub(cEz[r + 1], cEz[r]);
          db = sub(nEy[r], cEy[r]);
        } else if (c == 1) {
          da = sub(nEx[r], cEx[r]);
          db = sub(ezp[r], cEz[r]);
        } else {
          da = sub(eyp[r], cEy[r]);
          db = sub(cEx[r + 1], cEx[r]);
        }
        const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        T acc = scl(-g.cH, fmad(g.s[d1], da, scl(-g.s[d2], db)));
        // CPML terms in ascending axis order (the reference's term order)
        const int first = d1 < d2 ? 0 : 1;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int s = q == 0 ? first : 1 - first;
          const int ax = (c + 1 + s) % 3;
          if (a.pml[ax].prof && (int)pr[r][c][s].w >= 0)
            acc = add(acc, pml_apply(pr[r][c][s], a.pml[ax].has_k, s == 0 ? -g.cH * g.s[d1] : g.cH * g.s[d2],
                                     s == 0 ? da : db, ps[r][c][s]));
        }
        hv[r][c] = add(hv[r][c], acc);
      }
    }
    // ---- stores
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const unsigned cell = cell32(nx, ny, i, j, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!st[r][c]) continue;
        H[(unsigned)c * N + cell] = hv[r][c];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          const int p = (int)pr[r][c][s].w;
          if (p >= 0) ((T*)psi[ax][c])[psi_index(ax, p, a.pml[ax].ns, nx, ny, i, j, k)] = ps[r][c][s];
        }
      }
    }
#pragma unroll
    for (int r = 0; r <= RY; ++r) {
      cEx[r] = nEx[r];
      cEz[r] = nEz[r];
      if (r < RY) {
        cEy[r] = nEy[r];
        hEy[r] = nhEy[r];
        hEz[r] = nhEz[r];
      }
    }
  }
}

// LOOP false: one tile per block (the grid covers the work list). LOOP true: a persistent grid whose blocks
// stride over the work list (a separate instantiation: the loop costs registers the one-tile form does not).
template <class T, int RY, int WAVES, int XW, bool PER, bool LOOP>
__global__ void __launch_bounds__(64 * WAVES) k_update_h(const StepArgs* __restrict__ args, Sched sc) {
  for (int b = blockIdx.x; b < sc.n_pad; b += gridDim.x) {
    const int t = xcd_tile(b, sc.n_pad, sc.rev);
    if (t < sc.n_tiles) {
      int m, lt;
      member_of(sc, t, m, lt);
      h_tile<T, RY, WAVES, XW, PER>(args[m], lt);
    }
    if (!LOOP) break;
  }
}

// ---------------------------------------------------------------------------------------------
// E update (engine.py:645-663) + pure-ADE groups (engine.py:769-793):
// E_c += cb [s_d1 (H_d2 - H_d2(-e_d1)) - s_d2 (H_d1 - H_d1(-e_d2))] + CPML. A D-form node (NP poles at most)
// advances its poles from E^n, P^{n+1} = c1 P^n + c2 P^{n-1} + c3 E^n, and takes
// E^{n+1} = E^n + (dE - sum_p (P_p^{n+1} - P_p^n)) / eps_inf, which is (D^{n+1} - sum P^{n+1}) / eps_inf with
// D^{n+1} = D^n + dE and D^n = eps_inf E^n + sum P^n (the reference's D-form, without storing D).
// Iteration k loads the H plane of plane k + 1; without D-form poles also the cell state (old E, psi) of plane
// k + 1 and the material of plane k + 2, so that the latency of every load overlaps a plane's arithmetic.
// ---------------------------------------------------------------------------------------------
// A D-form node's NP pole pairs {P^n, P^{n-1}} are adjacent in memory: move them with 16-byte (and one 8-byte)
// accesses instead of one access per pair. A float node's slot takes 8 NP bytes, so for odd NP it is 8-byte but
// not always 16-byte aligned; the vector types declare 4-byte alignment, which gfx950 global accesses accept.
// Fewer, wider memory instructions for the same bytes cut the E kernel's time by ~8 % at 10 % D-form cells.
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
template <int NP>
__device__ __forceinline__ void load_slot(const Pair<float>* s, Pair<float> (&pp)[NP]) {
  const float* f = (const float*)s;
#pragma unroll
  for (int q = 0; q + 1 < NP; q += 2) {
    const f4u v = *(const f4u*)(f + 2 * q);
    pp[q] = Pair<float>{v.x, v.y};
    pp[q + 1] = Pair<float>{v.z, v.w};
  }
  if (NP % 2) {
    const f2u v = *(const f2u*)(f + 2 * (NP - 1));
    pp[NP - 1] = Pair<float>{v.x, v.y};
  }
}
template <int NP>
__device__ __forceinline__ void store_slot(Pair<float>* s, const Pair<float> (&pp)[NP]) {
  float* f = (float*)s;
#pragma unroll
  for (int q = 0; q + 1 < NP; q += 2) *(f4u*)(f + 2 * q) = f4u{pp[q].p, pp[q].q, pp[q + 1].p, pp[q + 1].q};
  if (NP % 2) *(f2u*)(f + 2 * (NP - 1)) = f2u{pp[NP - 1].p, pp[NP - 1].q};
}
template <int NP>
__device__ __forceinline__ void load_slot(const Pair<c32>* s, Pair<c32> (&pp)[NP]) {
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const f4u v = *(const f4u*)((const float*)s + 4 * q);
    pp[q] = Pair<c32>{mk(v.x, v.y), mk(v.z, v.w)};
  }
}
template <int NP>
__device__ __forceinline__ void store_slot(Pair<c32>* s, const Pair<c32> (&pp)[NP]) {
#pragma unroll
  for (int q = 0; q < NP; ++q) *(f4u*)((float*)s + 4 * q) = f4u{pp[q].p.x, pp[q].p.y, pp[q].q.x, pp[q].q.y};
}

template <class T, int MAT, int RY, int WAVES, int XW, bool PER, int NP>
__device__ __forceinline__ void e_tile(const StepArgs& a, int t, const float* tab, const float4* med_tab) {
  const Grid& g = a.g;
  const int lane = threadIdx.x & 63;
  const int xl = lane & (XW - 1);
  const int yl = XW == 64 ? 0 : (lane >> 5);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int jw = (ty * WAVES + wv) * (64 / XW) * RY;
  if (jw >= g.n[1]) return;
  const int i = tx * XW + xl;
  const int j0 = jw + yl * RY;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, g.n[2]);
  const unsigned N = (unsigned)g.N, nx = g.n[0], ny = g.n[1];
  const gptr<const T> Hg = as_global((const T*)a.H);
  const gptr<T> Eg = as_global((T*)a.E);
  const gptr<const float> cbdg = as_global(a.cbd);
  const gptr<const uint32_t> midg = as_global(a.mid);
  gptr<T> psi[3][3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
#pragma unroll
    for (int c = 0; c < 3; ++c) psi[ax][c] = as_global((T*)a.pml[ax].psi[c]);
  // CPML slab widths in vector registers: read as kernel arguments inside the march they are reloaded from memory
  // (the scalar registers are full), and each reload stalls the plane's load issue until it returns
  const int nsv[3] = {vregi(a.pml[0].ns), vregi(a.pml[1].ns), vregi(a.pml[2].ns)};
  const bool inx = i < g.n[0];
  const int il = tx * XW - 1;
  const bool xl_ok = (xl == 0) && il >= 0;
  // lanes of this lane's x row below it
  // lanes of this lane's 32-cell chunk below it (the rank of a D-form node within its chunk)
  const int xc = xl & 31;
  const unsigned long long below = xc == 0 ? 0ull : (((~0ull) >> (64 - xc)) << ((yl + (xl >> 5)) * 32));
  const float4 px = (a.pml[0].prof && inx) ? ((const float4*)as_global(a.pml[0].prof))[i] : no_pml();
  float4 py[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) py[r] = row_prof<XW, RY>(a.pml[1].prof, jw, yl, r, g.n[1]);

  // H rows r = -1 .. RY-1 of H_x and H_z are stored at r + 1
  T cHx[RY + 1], cHy[RY], cHz[RY + 1], xHy[RY], xHz[RY];
  T nHx[RY + 1], nHy[RY], nHz[RY + 1], nxHy[RY], nxHz[RY];
  T pHx[RY], pHy[RY];
  // material of planes k, k + 1, k + 2: packed 8-bit ids (MAT 1) or the three coefficients (MAT 0)
  uint32_t id_c[RY], id_n[RY], id_nn[RY];
  float cbd_c[RY][MAT == 0 ? 3 : 1], cbd_n[RY][MAT == 0 ? 3 : 1], cbd_nn[RY][MAT == 0 ? 3 : 1];
  int ab_c[RY][3], ab_n[RY][3], ab_nn[RY][3];  // first D-form slot of each row's x chunk per component
  // first D-form slot of this lane's 32-cell x chunk per component: chunk_base[((k ny + j) nchunk + x / 32) 3 + c]
  // (slots ordered by plane, row, chunk, component, x: a wavefront row's pole pairs of all three components are one
  // contiguous block), one vector load per row and component (the index sits in a VGPR), so that these loads join
  // the plane's other loads instead of waiting one by one in the scalar cache
  const gptr<const int> cbase = as_global(a.chunk_base);
  const unsigned cb_k = ny * (unsigned)a.nchunk;  // per plane
  // (a lane past the row's last chunk, in a tile that ends within 32 cells of x0, reads the last chunk's entry:
  // it holds no node, and the index stays inside the table at the grid's last row and plane)
  const unsigned cb_j = (unsigned)vregi((int)((unsigned)j0 * (unsigned)a.nchunk +
                                               (unsigned)min(tx * (XW / 32) + (xl >> 5), a.nchunk - 1)));
  auto load_bases = [&](int k, int (*ab)[3]) {
#pragma unroll
    for (int r = 0; r < RY; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ab[r][c] = 0;
        if (NP > 0 && k < g.n[2] && j0 + r < g.n[1])
          ab[r][c] = ((const int*)cbase)[(cb_j + (unsigned)r * (unsigned)a.nchunk + (unsigned)k * cb_k) * 3u + (unsigned)c];
      }
  };
  auto load_mat = [&](int k, uint32_t* ids, float (*cb)[MAT == 0 ? 3 : 1]) {
    const uint32_t* __restrict__ mid = (const uint32_t*)midg;
    const float* __restrict__ cbd = (const float*)cbdg;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const bool ok = inx && j < g.n[1] && k < g.n[2];
      const unsigned cell = cell32(nx, ny, i, ok ? j : 0, ok ? k : 0);
      if (MAT == 1) {
        ids[r] = ok ? mid[cell] : 0u;
      } else {
#pragma unroll
        for (int c = 0; c < (MAT == 0 ? 3 : 1); ++c) cb[r][c] = ok ? cbd[(unsigned)c * N + cell] : 0.f;
      }
    }
  };
  auto load_h = [&](int k, T* xH, T* yH, T* zH, T* hy, T* hz) {
    const T* __restrict__ Hx = (const T*)Hg;
    const T* __restrict__ Hy = Hx + N;
    const T* __restrict__ Hz = Hx + 2 * N;
#pragma unroll
    for (int r = -1; r < RY; ++r) {
      const int j = j0 + r;
      const bool ok = inx && j >= 0 && j < g.n[1] && k < g.n[2];
      xH[r + 1] = ok ? load_lo<PER>(Hx, g, i, j, k) : zero<T>();
      zH[r + 1] = ok ? load_lo<PER>(Hz, g, i, j, k) : zero<T>();
      if (r >= 0) {
        yH[r] = ok ? load_lo<PER>(Hy, g, i, j, k) : zero<T>();
        const bool lok = xl_ok && j < g.n[1] && k < g.n[2];
        hy[r] = lok ? load_lo<PER>(Hy, g, il, j, k) : zero<T>();
        hz[r] = lok ? load_lo<PER>(Hz, g, il, j, k) : zero<T>();
      }
    }
  };
  auto coef_of = [&](const uint32_t* ids, const float (*cb)[MAT == 0 ? 3 : 1], int r, int c) -> float {
    if (MAT == 1) return tab[(ids[r] >> (8 * c)) & 255u];
    return cb[r][MAT == 0 ? c : 0];
  };
  auto coef = [&](int r, int c) -> float { return coef_of(id_c, cbd_c, r, c); };

  const gptr<Pair<T>> Pg = as_global((Pair<T>*)a.P);
  // The cell state of a plane: stepped flags, old E, E-side psi, D-form slots and pole pairs. It is loaded one
  // plane ahead (from the material loaded two planes ahead), so its latency overlaps a plane's arithmetic.
  bool st[RY][3], st_n[RY][3];
  T ev[RY][3], ev_n[RY][3];
  T ps[RY][3][2], ps_n[RY][3][2];
  int aq[RY][3], aq_n[RY][3];
  Pair<T> pp[RY][3][NP > 0 ? NP : 1], pp_n[RY][3][NP > 0 ? NP : 1];
  auto load_cells = [&](int k, const uint32_t* ids, const float (*cb)[MAT == 0 ? 3 : 1], const int (*ab)[3],
                        bool (*stq)[3], T (*evq)[3], T (*psq)[3][2], int (*aqq)[3],
                        Pair<T> (*ppq)[3][NP > 0 ? NP : 1]) {
    const T* __restrict__ E = (const T*)Eg;
    const Pair<T>* __restrict__ P = (const Pair<T>*)Pg;
    const float4 pz = (a.pml[2].prof && k < g.n[2]) ? ld_uniform(a.pml[2].prof, k) : no_pml();
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const unsigned cell = cell32(nx, ny, i, j, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        stq[r][c] = inx && j < g.n[1] && k < g.n[2] && e_box<PER>(g, c, i, j, k);
        evq[r][c] = stq[r][c] ? E[(unsigned)c * N + cell] : zero<T>();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          const float4 pr = ax == 0 ? px : (ax == 1 ? py[r] : pz);
          const int p = (int)pr.w;
          psq[r][c][s] = zero<T>();
          if (stq[r][c] && p >= 0)
            psq[r][c][s] = ((const T*)psi[ax][c])[psi_index(ax, p, nsv[ax], nx, ny, i, j, k)];
        }
        aqq[r][c] = -1;
        if (NP > 0) {
          // D-form node: its slot in the compact pole arrays (chunk base + rank among its row's lanes)
          const bool dsp = stq[r][c] && coef_of(ids, cb, r, c) < 0.f;
          const unsigned long long m = __ballot(dsp);
          if (dsp) aqq[r][c] = ab[r][c] + __popcll(m & below);
          // a node's NP pole pairs are consecutive (NP is the engine's max_poles)
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p) ppq[r][c][p] = Pair<T>{zero<T>(), zero<T>()};
          if (aqq[r][c] >= 0) load_slot(P + (unsigned)aqq[r][c] * (unsigned)NP, ppq[r][c]);
        }
      }
    }
  };

#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r, k = z0 - 1;
    const bool ok = inx && j < g.n[1] && k >= 0;
    const T* __restrict__ Hx = (const T*)Hg;
    pHx[r] = ok ? load_lo<PER>(Hx, g, i, j, k) : zero<T>();
    pHy[r] = ok ? load_lo<PER>(Hx + N, g, i, j, k) : zero<T>();
  }
  // Without D-form nodes the cell state is loaded one plane ahead; with them the pole pairs of two planes would
  // cost more registers (occupancy) than the earlier loads gain, so the cell state of plane k is loaded in
  // iteration k, behind the prefetched H plane and material.
  constexpr bool AHEAD = NP == 0;
  load_h(z0, cHx, cHy, cHz, xHy, xHz);
  load_mat(z0, id_c, cbd_c);
  load_bases(z0, ab_c);
  if (AHEAD) {
    load_mat(z0 + 1, id_n, cbd_n);
    load_bases(z0 + 1, ab_n);
    load_cells(z0, id_c, cbd_c, ab_c, st, ev, ps, aq, pp);
  }

  for (int k = z0; k < z1; ++k) {
    T* __restrict__ E = (T*)Eg;
    Pair<T>* __restrict__ P = (Pair<T>*)Pg;
    const float4 pz = a.pml[2].prof ? ld_uniform(a.pml[2].prof, k) : no_pml();
    // ---- loads: H of plane k + 1; the cells of k + 1 and the material of k + 2 (AHEAD), or the material of k + 1
    // and the cells of k
    if (k + 1 < z1) {
      load_h(k + 1, nHx, nHy, nHz, nxHy, nxHz);
      if (AHEAD) {
        load_cells(k + 1, id_n, cbd_n, ab_n, st_n, ev_n, ps_n, aq_n, pp_n);
        if (k + 2 < z1) {
          load_mat(k + 2, id_nn, cbd_nn);
          load_bases(k + 2, ab_nn);
        }
      } else {
        load_mat(k + 1, id_n, cbd_n);
        load_bases(k + 1, ab_n);
      }
    }
    if (!AHEAD) load_cells(k, id_c, cbd_c, ab_c, st, ev, ps, aq, pp);
    // ---- compute
    T hzm[RY], hym[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      hzm[r] = shfl_up(cHz[r + 1]);
      hym[r] = shfl_up(cHy[r]);
      if (xl == 0) {
        hzm[r] = xHz[r];
        hym[r] = xHy[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T da, db;
        if (c == 0) {
          da = sub(cHz[r + 1], cHz[r]);
          db = sub(cHy[r], pHy[r]);
        } else if (c == 1) {
          da = sub(cHx[r + 1], pHx[r]);
          db = sub(cHz[r + 1], hzm[r]);
        } else {
          da = sub(cHy[r], hym[r]);
          db = sub(cHx[r + 1], cHx[r]);
        }
        const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const float cbv = coef(r, c);
        const float cb = cbv < 0.f ? g.cdx : cbv;
        T dE = scl(cb, fmad(g.s[d1], da, scl(-g.s[d2], db)));
        const int first = d1 < d2 ? 0 : 1;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int s = q == 0 ? first : 1 - first;
          const int ax = (c + 1 + s) % 3;
          const float4 pr = ax == 0 ? px : (ax == 1 ? py[r] : pz);
          if (a.pml[ax].prof && (int)pr.w >= 0)
            dE = add(dE, pml_apply(pr, a.pml[ax].has_k, s == 0 ? cb * g.s[d1] : -cb * g.s[d2], s == 0 ? da : db,
                                   ps[r][c][s]));
        }
        const T eold = ev[r][c];
        T enew = add(eold, dE);
        if (NP > 0) {
          // branch-free over the wavefront: every lane advances NP poles with its medium's coefficients (medium 0
          // for lanes without a D-form node, whose registers are then discarded); a pole past the medium's count
          // has zero coefficients and stays zero, adding exactly 0 to dp
          const bool d = aq[r][c] >= 0;
          const float4* mp = med_tab + (d ? (int)(-cbv) - 1 : 0) * (NP + 1);
          T dp = zero<T>();  // sum of the poles' increments
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p) {
            const float4 q = mp[1 + p];
            const T pn = fmad(q.x, pp[r][c][p].p, fmad(q.y, pp[r][c][p].q, scl(q.z, eold)));
            dp = add(dp, sub(pn, pp[r][c][p].p));
            pp[r][c][p].q = pp[r][c][p].p;
            pp[r][c][p].p = pn;
          }
          if (d) enew = fmad(mp[0].x, sub(dE, dp), eold);
        }
        ev[r][c] = enew;
      }
    }
    // ---- stores
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const unsigned cell = cell32(nx, ny, i, j, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!st[r][c]) continue;
        E[(unsigned)c * N + cell] = ev[r][c];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          const float4 pr = ax == 0 ? px : (ax == 1 ? py[r] : pz);
          const int p = (int)pr.w;
          if (p >= 0) ((T*)psi[ax][c])[psi_index(ax, p, nsv[ax], nx, ny, i, j, k)] = ps[r][c][s];
        }
        if (NP > 0 && aq[r][c] >= 0) store_slot(P + (unsigned)aq[r][c] * (unsigned)NP, pp[r][c]);
      }
    }
    // ---- the next plane becomes the current one
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      pHx[r] = cHx[r + 1];
      pHy[r] = cHy[r];
      id_c[r] = id_n[r];
      if (AHEAD) id_n[r] = id_nn[r];
#pragma unroll
      for (int c = 0; c < (MAT == 0 ? 3 : 1); ++c) {
        cbd_c[r][c] = cbd_n[r][c];
        if (AHEAD) cbd_n[r][c] = cbd_nn[r][c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ab_c[r][c] = ab_n[r][c];
        if (!AHEAD) continue;
        ab_n[r][c] = ab_nn[r][c];
        st[r][c] = st_n[r][c];
        ev[r][c] = ev_n[r][c];
        ps[r][c][0] = ps_n[r][c][0];
        ps[r][c][1] = ps_n[r][c][1];
        aq[r][c] = aq_n[r][c];
#pragma unroll
        for (int p = 0; p < (NP > 0 ? NP : 1); ++p) pp[r][c][p] = pp_n[r][c][p];
      }
    }
#pragma unroll
    for (int r = 0; r <= RY; ++r) {
      cHx[r] = nHx[r];
      cHz[r] = nHz[r];
      if (r < RY) {
        cHy[r] = nHy[r];
        xHy[r] = nxHy[r];
        xHz[r] = nxHz[r];
      }
    }
  }
}

// Each member's tile 0 advances its step counter: counter[1] = n, counter[0] = n + 1 (engine.py:678). The LDS
// tables (indexed material, D-form media) are loaded again whenever the block moves to another member.
template <class T, int MAT, int RY, int WAVES, int XW, bool PER, int NP, bool LOOP>
__global__ void __launch_bounds__(64 * WAVES) k_update_e(const StepArgs* __restrict__ args, Sched sc) {
  __shared__ float tab[MAT == 1 ? 256 : 1];
  __shared__ float4 med_tab[NP > 0 ? MAX_MEDIA_VEC4 : 1];
  int cur = -1;
  for (int b = blockIdx.x; b < sc.n_pad; b += gridDim.x) {
    const int t = xcd_tile(b, sc.n_pad, sc.rev);
    if (t >= sc.n_tiles) {
      if (!LOOP) break;
      continue;
    }
    int m, lt;
    member_of(sc, t, m, lt);
    const StepArgs& a = args[m];
    if (a.tick && lt == 0 && threadIdx.x == 0) {
      long long* cnt = (long long*)as_global(a.counter);
      const long long n = cnt[0];
      cnt[1] = n;
      cnt[0] = n + 1;
    }
    if ((MAT == 1 || NP > 0) && m != cur) {
      if (cur >= 0) __syncthreads();  // the previous member's readers are done
      if (MAT == 1)
        for (int q = threadIdx.x; q < 256; q += 64 * WAVES) tab[q] = ((const float*)as_global(a.mtab))[q];
      if (NP > 0) {
        const float* src = (const float*)as_global(a.media);
        const int per = NP + 1, n = a.n_media_floats / a.media_stride;
        for (int q = threadIdx.x; q < n * per; q += 64 * WAVES) {
          const float* s = src + (q / per) * a.media_stride;
          const int p = q % per - 1;
          med_tab[q] = p < 0 ? make_float4(s[0], s[1], 0.f, 0.f)
                             : make_float4(s[2 + 3 * p], s[3 + 3 * p], s[4 + 3 * p], 0.f);
        }
      }
      __syncthreads();
      cur = m;
    }
    e_tile<T, MAT, RY, WAVES, XW, PER, NP>(a, lt, tab, med_tab);
    if (!LOOP) break;
  }
}

}  // namespace lsf
