// Adjoint (time-reversed, transposed) Yee + CPML update.
//
// The forward step is H' = H - cH curl+ E (+ H-side CPML), E' = E + cb (curl- H' + E-side CPML). For the
// objective L, the adjoint fields Ebar = dL/dE, Hbar = dL/dH obey the transposed recursion, run backwards.
// In the variables e~ = cb Ebar and h~ = -cH Hbar it is the forward Yee update itself:
//   h~_h -= cH [s_{h+1} D+_{h+1} Q_{h+2} - s_{h+2} D+_{h+2} Q_{h+1}]      (adjoint of the E half-step)
//   e~_e += cb [s_{e+1} D-_{e+1} R_{e+2} - s_{e+2} D-_{e+2} R_{e+1}]       (adjoint of the H half-step)
// where outside a CPML slab Q = e~ and R = h~. The transposed CPML terms carry the coefficients of the node
// that owned the psi term, so inside a slab along axis a the differenced values are
//   Q_{e,a} = (k + c_a) e~_e + c_a chi~E_{e,a}   (E-side profile at the E node; k = 1/kappa when kappa is used,
//   R_{h,a} = (k + c_a) h~_h + c_a chi~H_{h,a}    else 1; H-side profile at the H node)
// and the adjoint psi variables advance where their node is owned: chi~E <- b (e~_old + chi~E) in the E
// kernel, chi~H <- b (h~_old + chi~H) in the H kernel (completing the previous step's H half). The chi~
// arrays have the layout of the forward psi arrays. Real fields, no periodic axes, no D-form nodes.
#pragma once

#include "lsf_step.h"

namespace lsf {

// E-side (or H-side) profile term f = k + c and c of a slab position; identity outside the slab
__device__ __forceinline__ float2 adj_fc(const float4 pr, int has_k) {
  if ((int)pr.w < 0) return make_float2(1.f, 0.f);
  return make_float2((has_k ? pr.z : 1.f) + pr.y, pr.y);
}

// ---------------------------------------------------------------------------------------------
// adjoint H half-step (transpose of the forward E update, engine.py:645-663)
// ---------------------------------------------------------------------------------------------
template <int RY, int WAVES>
__device__ __forceinline__ void ah_tile(const StepArgs& a, int t) {
  const Grid& g = a.g;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int j0 = (ty * WAVES + wv) * RY;
  if (j0 >= g.n[1]) return;
  const int i = tx * 64 + lane;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, g.n[2]);
  const unsigned N = (unsigned)g.N, nx = g.n[0], ny = g.n[1];
  const gptr<const float> Eg = as_global((const float*)a.E);
  const gptr<float> Hg = as_global((float*)a.H);
  const PmlAx* pe = a.xpml;  // E-side CPML: profiles and chi~E (read here)
  const PmlAx* ph = a.pml;   // H-side CPML: profiles and chi~H (advanced here)
  const bool inx = i < g.n[0];
  const int ih = tx * 64 + 64;
  const bool hx_ok = (lane == 63) && ih < g.n[0];

  const float4 pex = (pe[0].prof && inx) ? ((const float4*)as_global(pe[0].prof))[i] : no_pml();
  const float4 pexh = (pe[0].prof && hx_ok) ? ((const float4*)as_global(pe[0].prof))[ih] : no_pml();
  const float4 phx = (ph[0].prof && inx) ? ((const float4*)as_global(ph[0].prof))[i] : no_pml();
  float4 pey[RY + 1], phy[RY];
#pragma unroll
  for (int r = 0; r <= RY; ++r) pey[r] = (pe[1].prof && j0 + r < g.n[1]) ? ld_uniform(pe[1].prof, j0 + r) : no_pml();
#pragma unroll
  for (int r = 0; r < RY; ++r) phy[r] = (ph[1].prof && j0 + r < g.n[1]) ? ld_uniform(ph[1].prof, j0 + r) : no_pml();

  // chi~E of component e along axis ax at (x, j, k) when that position is in the slab (p >= 0)
  auto chi_e = [&](int e, int ax, int p, int x, int j, int k) -> float {
    if (p < 0) return 0.f;
    const float* c = (const float*)as_global((const float*)pe[ax].psi[e]);
    return c[psi_index(ax, p, pe[ax].ns, nx, ny, x, j, k)];
  };
  // Q of component e along axis ax from e~ and the profile of the E node
  auto q_of = [&](float ev, float4 pr, int has_k, float chi) -> float {
    const float2 fc = adj_fc(pr, has_k);
    return fc.x * ev + fc.y * chi;
  };

  // per plane: Qxy[r] r=0..RY, Qxz[r], Qyx[r], Qyz[r], Qzx[r] r<RY, Qzy[r] r=0..RY; halos hQyx, hQzx
  float cQxy[RY + 1], cQxz[RY], cQyx[RY], cQyz[RY], cQzx[RY], cQzy[RY + 1], chQyx[RY], chQzx[RY];
  float nQxy[RY + 1], nQxz[RY], nQyx[RY], nQyz[RY], nQzx[RY], nQzy[RY + 1], nhQyx[RY], nhQzx[RY];

  auto load_plane = [&](int k, float* Qxy, float* Qxz, float* Qyx, float* Qyz, float* Qzx, float* Qzy, float* hQyx,
                        float* hQzx) {
    const float* __restrict__ Ex = (const float*)Eg;
    const float* __restrict__ Ey = Ex + N;
    const float* __restrict__ Ez = Ex + 2 * N;
    const bool kin = k < g.n[2];
    const float4 pez = (pe[2].prof && kin) ? ld_uniform(pe[2].prof, k) : no_pml();
    const int pzs = (int)pez.w, pxs = (int)pex.w, pxh = (int)pexh.w;
#pragma unroll
    for (int r = 0; r <= RY; ++r) {
      const int j = j0 + r;
      const bool ok = inx && j < g.n[1] && kin;
      const unsigned cell = cell32(nx, ny, i, ok ? j : 0, ok ? k : 0);
      const float ex = ok ? Ex[cell] : 0.f;
      const float ez = ok ? Ez[cell] : 0.f;
      const int pys = (int)pey[r].w;
      Qxy[r] = q_of(ex, pey[r], pe[1].has_k, ok ? chi_e(0, 1, pys, i, j, k) : 0.f);
      Qzy[r] = q_of(ez, pey[r], pe[1].has_k, ok ? chi_e(2, 1, pys, i, j, k) : 0.f);
      if (r < RY) {
        const float ey = ok ? Ey[cell] : 0.f;
        Qxz[r] = q_of(ex, pez, pe[2].has_k, ok ? chi_e(0, 2, pzs, i, j, k) : 0.f);
        Qyz[r] = q_of(ey, pez, pe[2].has_k, ok ? chi_e(1, 2, pzs, i, j, k) : 0.f);
        Qyx[r] = q_of(ey, pex, pe[0].has_k, ok ? chi_e(1, 0, pxs, i, j, k) : 0.f);
        Qzx[r] = q_of(ez, pex, pe[0].has_k, ok ? chi_e(2, 0, pxs, i, j, k) : 0.f);
        const bool hok = hx_ok && j < g.n[1] && kin;
        const unsigned hc = cell32(nx, ny, hok ? ih : 0, hok ? j : 0, hok ? k : 0);
        hQyx[r] = hok ? q_of(Ey[hc], pexh, pe[0].has_k, chi_e(1, 0, pxh, ih, j, k)) : 0.f;
        hQzx[r] = hok ? q_of(Ez[hc], pexh, pe[0].has_k, chi_e(2, 0, pxh, ih, j, k)) : 0.f;
      }
    }
  };

  load_plane(z0, cQxy, cQxz, cQyx, cQyz, cQzx, cQzy, chQyx, chQzx);
  for (int k = z0; k < z1; ++k) {
    const float4 phz = ph[2].prof ? ld_uniform(ph[2].prof, k) : no_pml();
    load_plane(k + 1, nQxy, nQxz, nQyx, nQyz, nQzx, nQzy, nhQyx, nhQzx);
    float* __restrict__ H = (float*)Hg;
    bool st[RY][3];
    float hv[RY][3];
    float ch[RY][3][2];  // chi~H of component c along axes c+1, c+2
    float4 pr[RY][3][2];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const unsigned cell = cell32(nx, ny, i, j, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        st[r][c] = inx && j < g.n[1] && h_box<false>(g, c, i, j, k);
        hv[r][c] = st[r][c] ? H[(unsigned)c * N + cell] : 0.f;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          pr[r][c][s] = ax == 0 ? phx : (ax == 1 ? phy[r] : phz);
          const int p = (int)pr[r][c][s].w;
          ch[r][c][s] = 0.f;
          if (st[r][c] && p >= 0)
            ch[r][c][s] = ((const float*)as_global((float*)ph[ax].psi[c]))[psi_index(ax, p, ph[ax].ns, nx, ny, i, j, k)];
        }
      }
    }
    float qzp[RY], qyp[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      qzp[r] = shfl_dn(cQzx[r]);
      qyp[r] = shfl_dn(cQyx[r]);
      if (lane == 63) {
        qzp[r] = chQzx[r];
        qyp[r] = chQyx[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float da, db;
        if (c == 0) {
          da = cQzy[r + 1] - cQzy[r];  // along y of Q_{z,y}
          db = nQyz[r] - cQyz[r];      // along z of Q_{y,z}
        } else if (c == 1) {
          da = nQxz[r] - cQxz[r];      // along z of Q_{x,z}
          db = qzp[r] - cQzx[r];       // along x of Q_{z,x}
        } else {
          da = qyp[r] - cQyx[r];       // along x of Q_{y,x}
          db = cQxy[r + 1] - cQxy[r];  // along y of Q_{x,y}
        }
        const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const float hold = hv[r][c];
        // chi~H <- b (h~_old + chi~H): the transposed H-side psi recursion of the previous step
#pragma unroll
        for (int s = 0; s < 2; ++s)
          if ((int)pr[r][c][s].w >= 0) ch[r][c][s] = pr[r][c][s].x * (hold + ch[r][c][s]);
        hv[r][c] = hold - g.cH * (g.s[d1] * da - g.s[d2] * db);
      }
    }
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
          if (p >= 0) ((float*)as_global((float*)ph[ax].psi[c]))[psi_index(ax, p, ph[ax].ns, nx, ny, i, j, k)] = ch[r][c][s];
        }
      }
    }
#pragma unroll
    for (int r = 0; r <= RY; ++r) {
      cQxy[r] = nQxy[r];
      cQzy[r] = nQzy[r];
      if (r < RY) {
        cQxz[r] = nQxz[r];
        cQyz[r] = nQyz[r];
        cQyx[r] = nQyx[r];
        cQzx[r] = nQzx[r];
        chQyx[r] = nhQyx[r];
        chQzx[r] = nhQzx[r];
      }
    }
  }
}

template <int RY, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_adjoint_h(const StepArgs* __restrict__ args, Sched sc) {
  const int t = xcd_tile(blockIdx.x, sc.n_pad);
  if (t >= sc.n_tiles) return;
  int m, lt;
  member_of(sc, t, m, lt);
  ah_tile<RY, WAVES>(args[m], lt);
}

// ---------------------------------------------------------------------------------------------
// adjoint E half-step (transpose of the forward H update, engine.py:630-643)
// ---------------------------------------------------------------------------------------------
template <int MAT, int RY, int WAVES>
__device__ __forceinline__ void ae_tile(const StepArgs& a, int t, const float* tab) {
  const Grid& g = a.g;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int j0 = (ty * WAVES + wv) * RY;
  if (j0 >= g.n[1]) return;
  const int i = tx * 64 + lane;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, g.n[2]);
  const unsigned N = (unsigned)g.N, nx = g.n[0], ny = g.n[1];
  const gptr<const float> Hg = as_global((const float*)a.H);
  const gptr<float> Eg = as_global((float*)a.E);
  const gptr<const float> cbdg = as_global(a.cbd);
  const gptr<const uint32_t> midg = as_global(a.mid);
  const PmlAx* pe = a.pml;   // E-side CPML: profiles and chi~E (advanced here)
  const PmlAx* ph = a.xpml;  // H-side CPML: profiles and chi~H (read here)
  const bool inx = i < g.n[0];
  const int il = tx * 64 - 1;
  const bool xl_ok = (lane == 0) && il >= 0;

  const float4 pex = (pe[0].prof && inx) ? ((const float4*)as_global(pe[0].prof))[i] : no_pml();
  const float4 phx = (ph[0].prof && inx) ? ((const float4*)as_global(ph[0].prof))[i] : no_pml();
  const float4 phxl = (ph[0].prof && xl_ok) ? ((const float4*)as_global(ph[0].prof))[il] : no_pml();
  float4 pey[RY], phy[RY + 1];  // phy[r + 1] is row j0 + r, r = -1 .. RY-1
#pragma unroll
  for (int r = 0; r < RY; ++r) pey[r] = (pe[1].prof && j0 + r < g.n[1]) ? ld_uniform(pe[1].prof, j0 + r) : no_pml();
#pragma unroll
  for (int r = -1; r < RY; ++r)
    phy[r + 1] = (ph[1].prof && j0 + r >= 0 && j0 + r < g.n[1]) ? ld_uniform(ph[1].prof, j0 + r) : no_pml();

  auto chi_h = [&](int h, int ax, int p, int x, int j, int k) -> float {
    if (p < 0) return 0.f;
    const float* c = (const float*)as_global((const float*)ph[ax].psi[h]);
    return c[psi_index(ax, p, ph[ax].ns, nx, ny, x, j, k)];
  };
  auto r_of = [&](float hv, float4 pr, int has_k, float chi) -> float {
    const float2 fc = adj_fc(pr, has_k);
    return fc.x * hv + fc.y * chi;
  };

  // per plane: Rxz[r], Ryz[r], Ryx[r] (r < RY), Rxy[r + 1], Rzy[r + 1] (r = -1 .. RY-1), Rzx[r] (r < RY);
  // halos (x - 1): hRyx, hRzx
  float cRxz[RY], cRyz[RY], cRyx[RY], cRzx[RY], cRxy[RY + 1], cRzy[RY + 1], chRyx[RY], chRzx[RY];
  float nRxz[RY], nRyz[RY], nRyx[RY], nRzx[RY], nRxy[RY + 1], nRzy[RY + 1], nhRyx[RY], nhRzx[RY];
  float pRxz[RY], pRyz[RY];  // previous plane's R along z
  uint32_t id_c[RY], id_n[RY];
  float cbd_c[RY][MAT == 0 ? 3 : 1], cbd_n[RY][MAT == 0 ? 3 : 1];

  auto load_plane = [&](int k, float* Rxz, float* Ryz, float* Ryx, float* Rzx, float* Rxy, float* Rzy, float* hRyx,
                        float* hRzx, uint32_t* ids, float (*cb)[MAT == 0 ? 3 : 1]) {
    const float* __restrict__ Hx = (const float*)Hg;
    const float* __restrict__ Hy = Hx + N;
    const float* __restrict__ Hz = Hx + 2 * N;
    const uint32_t* __restrict__ mid = (const uint32_t*)midg;
    const float* __restrict__ cbd = (const float*)cbdg;
    const bool kin = k >= 0 && k < g.n[2];
    const float4 phz = (ph[2].prof && kin) ? ld_uniform(ph[2].prof, k) : no_pml();
    const int pzs = (int)phz.w, pxs = (int)phx.w, pxl = (int)phxl.w;
#pragma unroll
    for (int r = -1; r < RY; ++r) {
      const int j = j0 + r;
      const bool ok = inx && j >= 0 && j < g.n[1] && kin;
      const unsigned cell = cell32(nx, ny, i, ok ? j : 0, ok ? k : 0);
      const float hx = ok ? Hx[cell] : 0.f;
      const float hz = ok ? Hz[cell] : 0.f;
      const int pys = (int)phy[r + 1].w;
      Rxy[r + 1] = r_of(hx, phy[r + 1], ph[1].has_k, ok ? chi_h(0, 1, pys, i, j, k) : 0.f);
      Rzy[r + 1] = r_of(hz, phy[r + 1], ph[1].has_k, ok ? chi_h(2, 1, pys, i, j, k) : 0.f);
      if (r >= 0) {
        const float hy = ok ? Hy[cell] : 0.f;
        Rxz[r] = r_of(hx, phz, ph[2].has_k, ok ? chi_h(0, 2, pzs, i, j, k) : 0.f);
        Ryz[r] = r_of(hy, phz, ph[2].has_k, ok ? chi_h(1, 2, pzs, i, j, k) : 0.f);
        Ryx[r] = r_of(hy, phx, ph[0].has_k, ok ? chi_h(1, 0, pxs, i, j, k) : 0.f);
        Rzx[r] = r_of(hz, phx, ph[0].has_k, ok ? chi_h(2, 0, pxs, i, j, k) : 0.f);
        const bool lok = xl_ok && j < g.n[1] && kin;
        const unsigned lc = cell32(nx, ny, lok ? il : 0, lok ? j : 0, lok ? k : 0);
        hRyx[r] = lok ? r_of(Hy[lc], phxl, ph[0].has_k, chi_h(1, 0, pxl, il, j, k)) : 0.f;
        hRzx[r] = lok ? r_of(Hz[lc], phxl, ph[0].has_k, chi_h(2, 0, pxl, il, j, k)) : 0.f;
        if (ids != nullptr) {
          if (MAT == 1) {
            ids[r] = ok ? mid[cell] : 0u;
          } else {
#pragma unroll
            for (int c = 0; c < (MAT == 0 ? 3 : 1); ++c) cb[r][c] = ok ? cbd[(unsigned)c * N + cell] : 0.f;
          }
        }
      }
    }
  };
  auto coef = [&](int r, int c) -> float {
    if (MAT == 1) return tab[(id_c[r] >> (8 * c)) & 255u];
    return cbd_c[r][MAT == 0 ? c : 0];
  };

  {
    float t0[RY], t1[RY], t2[RY], t3[RY], t4[RY + 1], t5[RY + 1], t6[RY], t7[RY];
    load_plane(z0 - 1, pRxz, pRyz, t0, t1, t4, t5, t2, t3, nullptr, nullptr);
    (void)t6;
    (void)t7;
  }
  load_plane(z0, cRxz, cRyz, cRyx, cRzx, cRxy, cRzy, chRyx, chRzx, id_c, cbd_c);

  for (int k = z0; k < z1; ++k) {
    const float4 pez = pe[2].prof ? ld_uniform(pe[2].prof, k) : no_pml();
    if (k + 1 < z1) load_plane(k + 1, nRxz, nRyz, nRyx, nRzx, nRxy, nRzy, nhRyx, nhRzx, id_n, cbd_n);
    float* __restrict__ E = (float*)Eg;
    bool st[RY][3];
    float ev[RY][3];
    float ce[RY][3][2];
    float4 pr[RY][3][2];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int j = j0 + r;
      const unsigned cell = cell32(nx, ny, i, j, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        st[r][c] = inx && j < g.n[1] && e_box<false>(g, c, i, j, k);
        ev[r][c] = st[r][c] ? E[(unsigned)c * N + cell] : 0.f;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          pr[r][c][s] = ax == 0 ? pex : (ax == 1 ? pey[r] : pez);
          const int p = (int)pr[r][c][s].w;
          ce[r][c][s] = 0.f;
          if (st[r][c] && p >= 0)
            ce[r][c][s] = ((const float*)as_global((float*)pe[ax].psi[c]))[psi_index(ax, p, pe[ax].ns, nx, ny, i, j, k)];
        }
      }
    }
    float rzm[RY], rym[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      rzm[r] = shfl_up(cRzx[r]);
      rym[r] = shfl_up(cRyx[r]);
      if (lane == 0) {
        rzm[r] = chRzx[r];
        rym[r] = chRyx[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float da, db;
        if (c == 0) {
          da = cRzy[r + 1] - cRzy[r];  // along y of R_{z,y}
          db = cRyz[r] - pRyz[r];      // along z of R_{y,z}
        } else if (c == 1) {
          da = cRxz[r] - pRxz[r];      // along z of R_{x,z}
          db = cRzx[r] - rzm[r];       // along x of R_{z,x}
        } else {
          da = cRyx[r] - rym[r];       // along x of R_{y,x}
          db = cRxy[r + 1] - cRxy[r];  // along y of R_{x,y}
        }
        const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const float eold = ev[r][c];
        // chi~E <- b (e~_old + chi~E): the transposed E-side psi recursion
#pragma unroll
        for (int s = 0; s < 2; ++s)
          if ((int)pr[r][c][s].w >= 0) ce[r][c][s] = pr[r][c][s].x * (eold + ce[r][c][s]);
        ev[r][c] = eold + coef(r, c) * (g.s[d1] * da - g.s[d2] * db);
      }
    }
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
          const int p = (int)pr[r][c][s].w;
          if (p >= 0) ((float*)as_global((float*)pe[ax].psi[c]))[psi_index(ax, p, pe[ax].ns, nx, ny, i, j, k)] = ce[r][c][s];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      pRxz[r] = cRxz[r];
      pRyz[r] = cRyz[r];
      id_c[r] = id_n[r];
#pragma unroll
      for (int c = 0; c < (MAT == 0 ? 3 : 1); ++c) cbd_c[r][c] = cbd_n[r][c];
    }
#pragma unroll
    for (int r = 0; r <= RY; ++r) {
      cRxy[r] = nRxy[r];
      cRzy[r] = nRzy[r];
      if (r < RY) {
        cRxz[r] = nRxz[r];
        cRyz[r] = nRyz[r];
        cRyx[r] = nRyx[r];
        cRzx[r] = nRzx[r];
        chRyx[r] = nhRyx[r];
        chRzx[r] = nhRzx[r];
      }
    }
  }
}

template <int MAT, int RY, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_adjoint_e(const StepArgs* __restrict__ args, Sched sc) {
  __shared__ float tab[MAT == 1 ? 256 : 1];
  const int t = xcd_tile(blockIdx.x, sc.n_pad, sc.rev);
  if (t >= sc.n_tiles) return;
  int m, lt;
  member_of(sc, t, m, lt);
  const StepArgs& a = args[m];
  if (a.tick && lt == 0 && threadIdx.x == 0) {
    long long* cnt = (long long*)as_global(a.counter);
    const long long n = cnt[0];
    cnt[1] = n;
    cnt[0] = n + 1;
  }
  if (MAT == 1) {
    for (int q = threadIdx.x; q < 256; q += 64 * WAVES) tab[q] = ((const float*)as_global(a.mtab))[q];
    __syncthreads();
  }
  ae_tile<MAT, RY, WAVES>(a, lt, tab);
}

}  // namespace lsf
