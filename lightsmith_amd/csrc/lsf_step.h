// The bulk Yee update: one H kernel and one E kernel per time step, each fusing the curl, the CPML
// psi recursions and (E kernel) the pure-ADE D-form update for its cells.
//
// Mapping: a wavefront's 64 lanes take 64 consecutive x cells; each lane holds RY consecutive y rows in
// registers and marches up a chunk of z planes, carrying the previous/next plane in registers. x
// neighbours come from the adjacent lane (cross-lane shuffle; the edge lane loads its halo cell), y
// neighbours from the lane's own rows (the edge row loads one halo row), z neighbours from the marching
// registers. Each field value is read from HBM about once per kernel.
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
  const int* chunk_base;
  void* D;
  void* P;
  long long n_ade;
  int nchunk;
  int max_poles;
  int media_stride;
  const float* media;
  long long* counter;
  int tick;
  int tiles_x, tiles_y, zc, nb, nbp;
};

template <class T>
struct Pair {
  T p, q;  // P^n, P^{n-1}
};

template <int MAT>
__device__ __forceinline__ float load_cb(const StepArgs& a, const float* tab, int c, long long cell, uint32_t ids) {
  if (MAT == 0) return a.cbd[c * a.g.N + cell];
  return tab[(ids >> (8 * c)) & 255u];
}

// One CPML term of a component (engine.py:636-643 for H, 656-663 for E): psi = b psi + cp der,
// field += kp der (when kappa > 1) + psi, with cp = sgn coef s c and kp = sgn coef s (1/kappa - 1).
template <class T>
__device__ __forceinline__ T pml_term(const PmlAx& ax, int c, int axis, int pos, int i, int j, int k, const Grid& g,
                                      T der, float coef_s, T acc) {
  const float4 pr = ax.prof[pos];
  const int p = (int)pr.w;
  if (p < 0) return acc;
  long long q;
  if (axis == 0) q = ((long long)k * g.n[1] + j) * ax.ns + p;
  else if (axis == 1) q = ((long long)k * ax.ns + p) * g.n[0] + i;
  else q = ((long long)p * g.n[1] + j) * g.n[0] + i;
  T* psi = (T*)ax.psi[c];
  T v = psi[q];
  v = add(scl(pr.x, v), scl(coef_s * pr.y, der));
  psi[q] = v;
  if (ax.has_k) acc = add(acc, scl(coef_s * (pr.z - 1.f), der));
  return add(acc, v);
}

// ---------------------------------------------------------------------------------------------
// H update (engine.py:630-643): H_c -= cH [s_d1 (E_d2(+e_d1) - E_d2) - s_d2 (E_d1(+e_d2) - E_d1)] + CPML.
// ---------------------------------------------------------------------------------------------
template <class T, int RY, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_update_h(StepArgs a) {
  const Grid& g = a.g;
  const int t = xcd_tile(blockIdx.x, a.nbp);
  if (t >= a.nb) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int i = tx * 64 + lane;
  const int j0 = (ty * WAVES + wv) * RY;
  if (j0 >= g.n[1]) return;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, g.n[2]);
  const long long N = g.N;
  const T* __restrict__ Ex = (const T*)a.E;
  const T* __restrict__ Ey = Ex + N;
  const T* __restrict__ Ez = Ex + 2 * N;
  T* __restrict__ H = (T*)a.H;
  const bool inx = i < g.n[0];
  const int ih = tx * 64 + 64;
  const bool hx_ok = (lane == 63) && ih < g.n[0];

  T cEx[RY + 1], cEy[RY], cEz[RY + 1], hEy[RY], hEz[RY];
  T nEx[RY + 1], nEy[RY], nEz[RY + 1], nhEy[RY], nhEz[RY];

  auto load = [&](int k, T* xE, T* yE, T* zE, T* hy, T* hz) {
#pragma unroll
    for (int r = 0; r <= RY; ++r) {
      const int j = j0 + r;
      const bool ok = inx && j < g.n[1] && k < g.n[2];
      xE[r] = ok ? ld_hi(Ex, g, i, j, k) : zero<T>();
      zE[r] = ok ? ld_hi(Ez, g, i, j, k) : zero<T>();
      if (r < RY) {
        yE[r] = ok ? ld_hi(Ey, g, i, j, k) : zero<T>();
        const bool hok = hx_ok && j < g.n[1] && k < g.n[2];
        hy[r] = hok ? ld_hi(Ey, g, ih, j, k) : zero<T>();
        hz[r] = hok ? ld_hi(Ez, g, ih, j, k) : zero<T>();
      }
    }
  };

  load(z0, cEx, cEy, cEz, hEy, hEz);
  for (int k = z0; k < z1; ++k) {
    load(k + 1, nEx, nEy, nEz, nhEy, nhEz);
    T ezp[RY], eyp[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      ezp[r] = shfl_dn(cEz[r]);
      eyp[r] = shfl_dn(cEy[r]);
      if (lane == 63) {
        ezp[r] = hEz[r];
        eyp[r] = hEy[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      if (!inx || j >= g.n[1]) continue;
      const long long cell = cidx(g, i, j, k);
      // H_x: d1 = y, d2 = z
      if (i >= g.rH[0][0][0] && i < g.rH[0][0][1] && j >= g.rH[0][1][0] && j < g.rH[0][1][1] &&
          k >= g.rH[0][2][0] && k < g.rH[0][2][1]) {
        const T da = sub(cEz[r + 1], cEz[r]);
        const T db = sub(nEy[r], cEy[r]);
        T acc = scl(-g.cH, sub(scl(g.s[1], da), scl(g.s[2], db)));
        if (a.pml[1].prof) acc = pml_term(a.pml[1], 0, 1, j, i, j, k, g, da, -g.cH * g.s[1], acc);
        if (a.pml[2].prof) acc = pml_term(a.pml[2], 0, 2, k, i, j, k, g, db, g.cH * g.s[2], acc);
        H[cell] = add(H[cell], acc);
      }
      // H_y: d1 = z, d2 = x
      if (i >= g.rH[1][0][0] && i < g.rH[1][0][1] && j >= g.rH[1][1][0] && j < g.rH[1][1][1] &&
          k >= g.rH[1][2][0] && k < g.rH[1][2][1]) {
        const T da = sub(nEx[r], cEx[r]);
        const T db = sub(ezp[r], cEz[r]);
        T acc = scl(-g.cH, sub(scl(g.s[2], da), scl(g.s[0], db)));
        if (a.pml[0].prof) acc = pml_term(a.pml[0], 1, 0, i, i, j, k, g, db, g.cH * g.s[0], acc);
        if (a.pml[2].prof) acc = pml_term(a.pml[2], 1, 2, k, i, j, k, g, da, -g.cH * g.s[2], acc);
        H[N + cell] = add(H[N + cell], acc);
      }
      // H_z: d1 = x, d2 = y
      if (i >= g.rH[2][0][0] && i < g.rH[2][0][1] && j >= g.rH[2][1][0] && j < g.rH[2][1][1] &&
          k >= g.rH[2][2][0] && k < g.rH[2][2][1]) {
        const T da = sub(eyp[r], cEy[r]);
        const T db = sub(cEx[r + 1], cEx[r]);
        T acc = scl(-g.cH, sub(scl(g.s[0], da), scl(g.s[1], db)));
        if (a.pml[0].prof) acc = pml_term(a.pml[0], 2, 0, i, i, j, k, g, da, -g.cH * g.s[0], acc);
        if (a.pml[1].prof) acc = pml_term(a.pml[1], 2, 1, j, i, j, k, g, db, g.cH * g.s[1], acc);
        H[2 * N + cell] = add(H[2 * N + cell], acc);
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

// ---------------------------------------------------------------------------------------------
// E update (engine.py:645-663) + pure-ADE groups (engine.py:769-793):
// E_c += cb [s_d1 (H_d2 - H_d2(-e_d1)) - s_d2 (H_d1 - H_d1(-e_d2))] + CPML; D-form nodes then take
// D += dE, P^{n+1} = c1 P^n + c2 P^{n-1} + c3 E^n per pole, E = (D - sum P) / eps_inf.
// ---------------------------------------------------------------------------------------------
template <class T, int MAT, int RY, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_update_e(StepArgs a) {
  __shared__ float tab[256];
  if (a.tick && blockIdx.x == 0 && threadIdx.x == 0) {
    const long long n = a.counter[0];
    a.counter[1] = n;
    a.counter[0] = n + 1;
  }
  if (MAT == 1) {
    for (int q = threadIdx.x; q < 256; q += 64 * WAVES) tab[q] = a.mtab[q];
    __syncthreads();
  }
  const Grid& g = a.g;
  const int t = xcd_tile(blockIdx.x, a.nbp);
  if (t >= a.nb) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int i = tx * 64 + lane;
  const int j0 = (ty * WAVES + wv) * RY;
  if (j0 >= g.n[1]) return;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, g.n[2]);
  const long long N = g.N;
  const T* __restrict__ Hx = (const T*)a.H;
  const T* __restrict__ Hy = Hx + N;
  const T* __restrict__ Hz = Hx + 2 * N;
  T* __restrict__ E = (T*)a.E;
  const bool inx = i < g.n[0];
  const int il = tx * 64 - 1;
  const bool xl_ok = (lane == 0) && il >= 0;
  const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));

  // rows r = -1 .. RY-1 of H_x and H_z are stored at r + 1
  T cHx[RY + 1], cHy[RY], cHz[RY + 1], pHx[RY], pHy[RY], xHy[RY], xHz[RY];

  auto load = [&](int k) {
#pragma unroll
    for (int r = -1; r < RY; ++r) {
      const int j = j0 + r;
      const bool ok = inx && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2];
      cHx[r + 1] = ok ? ld_lo(Hx, g, i, j, k) : zero<T>();
      cHz[r + 1] = ok ? ld_lo(Hz, g, i, j, k) : zero<T>();
      if (r >= 0) {
        cHy[r] = ok ? ld_lo(Hy, g, i, j, k) : zero<T>();
        const bool lok = xl_ok && j < g.n[1] && k >= 0 && k < g.n[2];
        xHy[r] = lok ? ld_lo(Hy, g, il, j, k) : zero<T>();
        xHz[r] = lok ? ld_lo(Hz, g, il, j, k) : zero<T>();
      }
    }
  };

  // previous plane (H_x, H_y at k - 1)
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r, k = z0 - 1;
    const bool ok = inx && j < g.n[1] && k >= 0;
    pHx[r] = ok ? ld_lo(Hx, g, i, j, k) : zero<T>();
    pHy[r] = ok ? ld_lo(Hy, g, i, j, k) : zero<T>();
  }

  for (int k = z0; k < z1; ++k) {
    load(k);
    T hzm[RY], hym[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      hzm[r] = shfl_up(cHz[r + 1]);
      hym[r] = shfl_up(cHy[r]);
      if (lane == 0) {
        hzm[r] = xHz[r];
        hym[r] = xHy[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const bool row = inx && j < g.n[1];
      const long long cell = cidx(g, i, row ? j : 0, k);
      uint32_t ids = 0;
      if (MAT == 1 && row) ids = a.mid[cell];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const bool st = row && i >= g.rE[c][0][0] && i < g.rE[c][0][1] && j >= g.rE[c][1][0] &&
                        j < g.rE[c][1][1] && k >= g.rE[c][2][0] && k < g.rE[c][2][1];
        float cbv = 0.f;
        T dE = zero<T>(), eold = zero<T>();
        if (st) {
          cbv = load_cb<MAT>(a, tab, c, cell, ids);
          const float cb = cbv < 0.f ? g.cdx : cbv;
          T da, db;
          int d1, d2;
          if (c == 0) {
            da = sub(cHz[r + 1], cHz[r]);
            db = sub(cHy[r], pHy[r]);
            d1 = 1;
            d2 = 2;
          } else if (c == 1) {
            da = sub(cHx[r + 1], pHx[r]);
            db = sub(cHz[r + 1], hzm[r]);
            d1 = 2;
            d2 = 0;
          } else {
            da = sub(cHy[r], hym[r]);
            db = sub(cHx[r + 1], cHx[r]);
            d1 = 0;
            d2 = 1;
          }
          dE = scl(cb, sub(scl(g.s[d1], da), scl(g.s[d2], db)));
          // CPML terms in ascending axis order (the reference's term list order)
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            if (ax == c || !a.pml[ax].prof) continue;
            const int pos = ax == 0 ? i : (ax == 1 ? j : k);
            if (ax == d1) dE = pml_term(a.pml[ax], c, ax, pos, i, j, k, g, da, cb * g.s[d1], dE);
            else dE = pml_term(a.pml[ax], c, ax, pos, i, j, k, g, db, -cb * g.s[d2], dE);
          }
          eold = E[c * N + cell];
        }
        const bool dsp = st && cbv < 0.f;
        const unsigned long long m = __ballot(dsp);
        T enew = add(eold, dE);
        if (m) {
          if (dsp) {
            const long long base = a.chunk_base[(((long long)c * g.n[2] + k) * g.n[1] + j) * a.nchunk + tx];
            const long long q = base + __popcll(m & below);
            const int med = (int)(-cbv) - 1;
            const float* mp = a.media + med * a.media_stride;
            const float ie = mp[0];
            const int np = (int)mp[1];
            T* D = (T*)a.D;
            Pair<T>* P = (Pair<T>*)a.P;
            const T Dn = add(D[q], dE);
            T tot = zero<T>();
            for (int p = 0; p < a.max_poles; ++p) {
              if (p < np) {
                const float c1 = mp[2 + 3 * p], c2 = mp[3 + 3 * p], c3 = mp[4 + 3 * p];
                Pair<T> pp = P[(long long)p * a.n_ade + q];
                const T pn = add(add(scl(c1, pp.p), scl(c2, pp.q)), scl(c3, eold));
                pp.q = pp.p;
                pp.p = pn;
                P[(long long)p * a.n_ade + q] = pp;
                tot = (p == 0) ? pn : add(tot, pn);
              }
            }
            D[q] = Dn;
            enew = scl(ie, sub(Dn, tot));
          }
        }
        if (st) E[c * N + cell] = enew;
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      pHx[r] = cHx[r + 1];
      pHy[r] = cHy[r];
    }
  }
}

}  // namespace lsf
