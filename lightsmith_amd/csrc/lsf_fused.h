// One time step in a single pass: H^{n+1/2} and E^{n+1} from E^n and H^{n-1/2}, for real fields without periodic
// axes (engine.py:630-663 with the pure-ADE update of engine.py:769-793, in the order of the two-kernel step).
//
// The fields are double-buffered: the kernel reads E_in, H_in and writes E_out, H_out (the H-side psi likewise),
// so a block never reads another block's new values and no store is ordered before a load. A block is 64 x cells
// by WAVES rows (one row per wavefront) marching up a chunk of z planes. Per plane every wavefront computes
// H^{n+1/2} of its row and publishes H_x, H_z to LDS (double-buffered by plane, one barrier per plane), then
// updates E^{n+1} of its row from its own H (the node and, by shuffle, x - 1), the row below (LDS) and the
// previous plane (registers). Values owned by neighbouring blocks are recomputed with their owners' arithmetic:
// the row below the block (first wavefront), the column x - 1 of the tile (lane 0) and the plane below the chunk
// (a first pass through the march), so the result is bit-identical to the two-kernel step. E-side psi and the pole
// state are updated in place (one owner each). Per cell the pass moves E and H once in and once out (48 B) plus
// the material, against 72 B for the two kernels.
#pragma once

#include "lsf_step.h"

namespace lsf {

// Arguments of the single-pass kernel. The CPML psi arrays of one (side, axis) are contiguous: the component
// c != axis has slot c - (c > axis) and starts at psi + slot * psz (elements); the H side has an input and an
// output set.
struct FusedArgs {
  Grid g;
  const float* E_in;
  const float* H_in;
  float* E_out;
  float* H_out;
  const float* cbd;
  const uint32_t* mid;
  const float* mtab;
  const float4* prof[2][3];  // [E side, H side][axis], null without CPML on the axis
  float* psiE[3];
  const float* psiHi[3];
  float* psiHo[3];
  long long psz[2][3];
  int ns[2][3];
  int has_k[2][3];
  const int* chunk_base;
  void* P;
  int nchunk, max_poles, media_stride, n_media_floats;
  const float* media;
  long long* counter;
  int tick;
  int tiles_x, tiles_y, zc;
};

__host__ __device__ __forceinline__ constexpr int psi_slot(int c, int ax) { return c - (c > ax ? 1 : 0); }

template <int MAT, int WAVES, int NP>
__global__ void __launch_bounds__(64 * WAVES) k_step_fused(const FusedArgs* __restrict__ args, Sched sc) {
  __shared__ float tab[MAT == 1 ? 256 : 1];
  __shared__ float med_tab[NP > 0 ? MAX_MEDIA_FLOATS : 1];
  __shared__ float lhx[2][WAVES + 1][64];  // H_x, H_z of each row per plane parity; [0] is the row below the block
  __shared__ float lhz[2][WAVES + 1][64];
  const int t = xcd_tile(blockIdx.x, sc.n_pad);
  if (t >= sc.n_tiles) return;  // whole block
  const FusedArgs& a = args[0];
  if (a.tick && t == 0 && threadIdx.x == 0) {
    long long* cnt = (long long*)as_global(a.counter);
    const long long n = cnt[0];
    cnt[1] = n;
    cnt[0] = n + 1;
  }
  if (MAT == 1)
    for (int q = threadIdx.x; q < 256; q += 64 * WAVES) tab[q] = ((const float*)as_global(a.mtab))[q];
  if (NP > 0)
    for (int q = threadIdx.x; q < a.n_media_floats; q += 64 * WAVES) med_tab[q] = ((const float*)as_global(a.media))[q];

  const Grid& g = a.g;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int j = ty * WAVES + wv;  // this wavefront's row
  const bool jin = j < g.n[1];     // rows past the grid still take part in the barriers
  const int i = tx * 64 + lane;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, g.n[2]);
  const unsigned N = (unsigned)g.N, nx = g.n[0], ny = g.n[1];
  const bool inx = i < g.n[0];
  const bool on = inx && jin;
  const int ih = tx * 64 + 64;
  const bool hx_ok = (lane == 63) && ih < g.n[0] && jin;
  const int il = tx * 64 - 1;
  const bool xl_ok = (lane == 0) && il >= 0 && jin;
  const bool yhalo = wv == 0 && j >= 1;  // the first wavefront recomputes the row below the block
  const int jm = j - 1;
  const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const float* __restrict__ Ei = (const float*)as_global(a.E_in);
  const float* __restrict__ Hi = (const float*)as_global(a.H_in);
  float* __restrict__ Eo = (float*)as_global(a.E_out);
  float* __restrict__ Ho = (float*)as_global(a.H_out);

  const float4 pex = (a.prof[0][0] && inx) ? ((const float4*)as_global(a.prof[0][0]))[i] : no_pml();
  const float4 phx = (a.prof[1][0] && inx) ? ((const float4*)as_global(a.prof[1][0]))[i] : no_pml();
  const float4 phxl = (a.prof[1][0] && xl_ok) ? ((const float4*)as_global(a.prof[1][0]))[il] : no_pml();
  const float4 pey = (a.prof[0][1] && jin) ? ld_uniform(a.prof[0][1], j) : no_pml();
  const float4 phy = (a.prof[1][1] && jin) ? ld_uniform(a.prof[1][1], j) : no_pml();
  const float4 phym = (a.prof[1][1] && yhalo) ? ld_uniform(a.prof[1][1], jm) : no_pml();

  auto cell = [&](int x, int y, int k) -> unsigned { return cell32(nx, ny, x, y, k); };
  auto psi_hi = [&](int c, int ax, int p, int x, int y, int k) -> float {
    if (p < 0) return 0.f;
    return ((const float*)as_global(a.psiHi[ax]))[psi_slot(c, ax) * a.psz[1][ax] +
                                                  psi_index(ax, p, a.ns[1][ax], nx, ny, x, y, k)];
  };
  // h_new = h_old - cH (s_d1 da - s_d2 db) + CPML terms in ascending axis order (the H kernel's arithmetic)
  auto h_new = [&](int c, float hold, float da, float db, float4 pr1, float4 pr2, float& ps1, float& ps2) -> float {
    const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
    float acc = -g.cH * (g.s[d1] * da - g.s[d2] * db);
    if (d1 < d2) {
      if ((int)pr1.w >= 0) acc += pml_apply(pr1, a.has_k[1][d1], -g.cH * g.s[d1], da, ps1);
      if ((int)pr2.w >= 0) acc += pml_apply(pr2, a.has_k[1][d2], g.cH * g.s[d2], db, ps2);
    } else {
      if ((int)pr2.w >= 0) acc += pml_apply(pr2, a.has_k[1][d2], g.cH * g.s[d2], db, ps2);
      if ((int)pr1.w >= 0) acc += pml_apply(pr1, a.has_k[1][d1], -g.cH * g.s[d1], da, ps1);
    }
    return hold + acc;
  };
  auto ld = [&](const float* F, bool ok, int x, int y, int k) -> float { return ok ? F[cell(x, y, k)] : 0.f; };

  const float* Exi = Ei;
  const float* Eyi = Ei + N;
  const float* Ezi = Ei + 2 * N;
  const int k0 = z0 > 0 ? z0 - 1 : z0;
  // E^n of the row at plane k (carried) and x + 1 halos (lane 63)
  float ex = ld(Exi, on && k0 < g.n[2], i, j, k0), ey = ld(Eyi, on && k0 < g.n[2], i, j, k0),
        ez = ld(Ezi, on && k0 < g.n[2], i, j, k0);
  float eyh = ld(Eyi, hx_ok && k0 < g.n[2], ih, j, k0), ezh = ld(Ezi, hx_ok && k0 < g.n[2], ih, j, k0);
  float hxp = 0.f, hyp = 0.f;  // H_x, H_y of the row at the previous plane
  // material / chunk bases, prefetched a plane ahead
  uint32_t idc = 0u, idn = 0u;
  float cbc[MAT == 0 ? 3 : 1] = {}, cbn[MAT == 0 ? 3 : 1] = {};
  int abc[3] = {0, 0, 0}, abn[3] = {0, 0, 0};
  auto load_mat = [&](int k, uint32_t& id, float* cb, int* ab) {
    const bool ok = on && k < g.n[2];
    const unsigned c = cell(ok ? i : 0, ok ? j : 0, ok ? k : 0);
    if (MAT == 1) id = ok ? ((const uint32_t*)as_global(a.mid))[c] : 0u;
    else
      for (int q = 0; q < (MAT == 0 ? 3 : 1); ++q) cb[q] = ok ? ((const float*)as_global(a.cbd))[(unsigned)q * N + c] : 0.f;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      ab[q] = (NP > 0 && jin && k < g.n[2])
                  ? ld_uniform(a.chunk_base, (((unsigned)q * g.n[2] + (unsigned)k) * ny + (unsigned)j) * a.nchunk +
                                                 (unsigned)tx * 2u)
                  : 0;
  };
  load_mat(z0, idc, cbc, abc);
  __syncthreads();  // LDS tables ready
  Pair<float>* P = (Pair<float>*)as_global((Pair<float>*)a.P);

  for (int k = k0; k < z1; ++k) {
    const bool own = k >= z0;
    const int b = k & 1;
    const bool kin1 = k + 1 < g.n[2];
    const float4 phz = a.prof[1][2] ? ld_uniform(a.prof[1][2], k) : no_pml();
    const float4 pez = a.prof[0][2] ? ld_uniform(a.prof[0][2], k) : no_pml();
    // ---- loads for H(k): E(k + 1) of the row, E(k) of the row above, old H, H-side psi
    const float exn = ld(Exi, on && kin1, i, j, k + 1), eyn = ld(Eyi, on && kin1, i, j, k + 1),
                ezn = ld(Ezi, on && kin1, i, j, k + 1);
    const float eyhn = ld(Eyi, hx_ok && kin1, ih, j, k + 1), ezhn = ld(Ezi, hx_ok && kin1, ih, j, k + 1);
    const bool up = on && j + 1 < g.n[1];
    const float exu = ld(Exi, up, i, j + 1, k), ezu = ld(Ezi, up, i, j + 1, k);
    const unsigned cc = cell(on ? i : 0, on ? j : 0, k);
    const bool stx = on && in_box(g.rH[0], i, j, k), sty = on && in_box(g.rH[1], i, j, k),
               stz = on && in_box(g.rH[2], i, j, k);
    const float hox = stx ? Hi[cc] : 0.f, hoy = sty ? Hi[N + cc] : 0.f, hoz = stz ? Hi[2 * N + cc] : 0.f;
    float qx0 = stx ? psi_hi(0, 1, (int)phy.w, i, j, k) : 0.f, qx1 = stx ? psi_hi(0, 2, (int)phz.w, i, j, k) : 0.f;
    float qy0 = sty ? psi_hi(1, 2, (int)phz.w, i, j, k) : 0.f, qy1 = sty ? psi_hi(1, 0, (int)phx.w, i, j, k) : 0.f;
    float qz0 = stz ? psi_hi(2, 0, (int)phx.w, i, j, k) : 0.f, qz1 = stz ? psi_hi(2, 1, (int)phy.w, i, j, k) : 0.f;
    // lane 0: the column x - 1 (H_y, H_z)
    const bool sly = xl_ok && in_box(g.rH[1], il, j, k), slz = xl_ok && in_box(g.rH[2], il, j, k);
    const float exl = ld(Exi, xl_ok, il, j, k), exln = ld(Exi, xl_ok && kin1, il, j, k + 1);
    const float exlu = ld(Exi, xl_ok && j + 1 < g.n[1], il, j + 1, k);
    const float eyl = ld(Eyi, xl_ok, il, j, k), ezl = ld(Ezi, xl_ok, il, j, k);
    const unsigned lc = cell(xl_ok ? il : 0, xl_ok ? j : 0, k);
    const float holy = sly ? Hi[N + lc] : 0.f, holz = slz ? Hi[2 * N + lc] : 0.f;
    float qly0 = sly ? psi_hi(1, 2, (int)phz.w, il, j, k) : 0.f, qly1 = sly ? psi_hi(1, 0, (int)phxl.w, il, j, k) : 0.f;
    float qlz0 = slz ? psi_hi(2, 0, (int)phxl.w, il, j, k) : 0.f, qlz1 = slz ? psi_hi(2, 1, (int)phy.w, il, j, k) : 0.f;
    // ---- loads for E(k): old E, E-side psi, pole state, next material
    bool ste[3];
    float ev[3], pse[3][2];
    int aq[3];
    Pair<float> pp[3][NP > 0 ? NP : 1];
    if (own) {
      if (k + 1 < z1) load_mat(k + 1, idn, cbn, abn);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ste[c] = on && in_box(g.rE[c], i, j, k);
        ev[c] = ste[c] ? Ei[(unsigned)c * N + cc] : 0.f;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          const float4 pr = ax == 0 ? pex : (ax == 1 ? pey : pez);
          const int p = (int)pr.w;
          pse[c][s] = (ste[c] && p >= 0) ? ((const float*)as_global(a.psiE[ax]))[psi_slot(c, ax) * a.psz[0][ax] +
                                                                                   psi_index(ax, p, a.ns[0][ax], nx, ny, i, j, k)]
                                         : 0.f;
        }
        aq[c] = -1;
        if (NP > 0) {
          const float cv = MAT == 1 ? tab[(idc >> (8 * c)) & 255u] : cbc[MAT == 0 ? c : 0];
          const bool dsp = ste[c] && cv < 0.f;
          const unsigned long long m = __ballot(dsp);
          if (m && dsp) aq[c] = abc[c] + __popcll(m & below);
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p) {
            pp[c][p] = Pair<float>{0.f, 0.f};
            if (aq[c] >= 0 && p < a.max_poles) pp[c][p] = P[(unsigned)aq[c] * (unsigned)a.max_poles + (unsigned)p];
          }
        }
      }
    }

    // ---- H^{n+1/2}(k) of the row
    float eyp = shfl_dn(ey), ezp = shfl_dn(ez);
    if (lane == 63) {
      eyp = eyh;
      ezp = ezh;
    }
    const float hx = stx ? h_new(0, hox, ezu - ez, eyn - ey, phy, phz, qx0, qx1) : hox;
    const float hy = sty ? h_new(1, hoy, exn - ex, ezp - ez, phz, phx, qy0, qy1) : hoy;
    const float hz = stz ? h_new(2, hoz, eyp - ey, exu - ex, phx, phy, qz0, qz1) : hoz;
    const float hly = sly ? h_new(1, holy, exln - exl, ez - ezl, phz, phxl, qly0, qly1) : holy;
    const float hlz = slz ? h_new(2, holz, ey - eyl, exlu - exl, phxl, phy, qlz0, qlz1) : holz;
    lhx[b][wv + 1][lane] = hx;
    lhz[b][wv + 1][lane] = hz;
    if (wv == 0) {
      // the first wavefront: H_x, H_z of the row below the block
      float hmx = 0.f, hmz = 0.f;
      if (yhalo) {
        const bool ym = inx;
        const bool smx = ym && in_box(g.rH[0], i, jm, k), smz = ym && in_box(g.rH[2], i, jm, k);
        const float exm = ld(Exi, ym, i, jm, k), eym = ld(Eyi, ym, i, jm, k), ezm = ld(Ezi, ym, i, jm, k);
        const float eymn = ld(Eyi, ym && kin1, i, jm, k + 1);
        const float eymh = ld(Eyi, lane == 63 && ih < g.n[0], ih, jm, k);
        const unsigned mc = cell(ym ? i : 0, jm, k);
        const float homx = smx ? Hi[mc] : 0.f, homz = smz ? Hi[2 * N + mc] : 0.f;
        float qmx0 = smx ? psi_hi(0, 1, (int)phym.w, i, jm, k) : 0.f;
        float qmx1 = smx ? psi_hi(0, 2, (int)phz.w, i, jm, k) : 0.f;
        float qmz0 = smz ? psi_hi(2, 0, (int)phx.w, i, jm, k) : 0.f;
        float qmz1 = smz ? psi_hi(2, 1, (int)phym.w, i, jm, k) : 0.f;
        float eymp = shfl_dn(eym);
        if (lane == 63) eymp = eymh;
        hmx = smx ? h_new(0, homx, ez - ezm, eymn - eym, phym, phz, qmx0, qmx1) : homx;
        hmz = smz ? h_new(2, homz, eymp - eym, ex - exm, phx, phym, qmz0, qmz1) : homz;
      }
      lhx[b][0][lane] = hmx;
      lhz[b][0][lane] = hmz;
    }
    __syncthreads();

    if (own) {
      // ---- E^{n+1}(k) of the row
      const float hxm = lhx[b][wv][lane], hzmy = lhz[b][wv][lane];  // the row below
      float hym = shfl_up(hy), hzm = shfl_up(hz);                    // x - 1
      if (lane == 0) {
        hym = hly;
        hzm = hlz;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float da, db;
        if (c == 0) {
          da = hz - hzmy;
          db = hy - hyp;
        } else if (c == 1) {
          da = hx - hxp;
          db = hz - hzm;
        } else {
          da = hy - hym;
          db = hx - hxm;
        }
        const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const float cbv = MAT == 1 ? tab[(idc >> (8 * c)) & 255u] : cbc[MAT == 0 ? c : 0];
        const float cb = cbv < 0.f ? g.cdx : cbv;
        float dE = cb * (g.s[d1] * da - g.s[d2] * db);
        const int first = d1 < d2 ? 0 : 1;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int s = q == 0 ? first : 1 - first;
          const int ax = (c + 1 + s) % 3;
          const float4 pr = ax == 0 ? pex : (ax == 1 ? pey : pez);
          if ((int)pr.w >= 0)
            dE += pml_apply(pr, a.has_k[0][ax], s == 0 ? cb * g.s[d1] : -cb * g.s[d2], s == 0 ? da : db, pse[c][s]);
        }
        const float eold = ev[c];
        float enew = eold + dE;
        if (NP > 0 && aq[c] >= 0) {
          const float* mp = med_tab + ((int)(-cbv) - 1) * a.media_stride;
          const int np = (int)mp[1];
          float dp = 0.f;
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p) {
            if (p < np) {
              const float pn = mp[2 + 3 * p] * pp[c][p].p + mp[3 + 3 * p] * pp[c][p].q + mp[4 + 3 * p] * eold;
              dp = dp + (pn - pp[c][p].p);
              pp[c][p].q = pp[c][p].p;
              pp[c][p].p = pn;
            }
          }
          enew = eold + mp[0] * (dE - dp);
        }
        ev[c] = enew;
      }
      // ---- stores
      auto put_psi = [&](int c, int ax, float4 pr, float v) {
        const int p = (int)pr.w;
        if (p >= 0)
          ((float*)as_global(a.psiHo[ax]))[psi_slot(c, ax) * a.psz[1][ax] + psi_index(ax, p, a.ns[1][ax], nx, ny, i, j, k)] = v;
      };
      if (stx) {
        Ho[cc] = hx;
        put_psi(0, 1, phy, qx0);
        put_psi(0, 2, phz, qx1);
      }
      if (sty) {
        Ho[N + cc] = hy;
        put_psi(1, 2, phz, qy0);
        put_psi(1, 0, phx, qy1);
      }
      if (stz) {
        Ho[2 * N + cc] = hz;
        put_psi(2, 0, phx, qz0);
        put_psi(2, 1, phy, qz1);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!ste[c]) continue;
        Eo[(unsigned)c * N + cc] = ev[c];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          const float4 pr = ax == 0 ? pex : (ax == 1 ? pey : pez);
          const int p = (int)pr.w;
          if (p >= 0)
            ((float*)as_global(a.psiE[ax]))[psi_slot(c, ax) * a.psz[0][ax] + psi_index(ax, p, a.ns[0][ax], nx, ny, i, j, k)] =
                pse[c][s];
        }
        if (NP > 0 && aq[c] >= 0) {
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p)
            if (p < a.max_poles) P[(unsigned)aq[c] * (unsigned)a.max_poles + (unsigned)p] = pp[c][p];
        }
      }
      idc = idn;
#pragma unroll
      for (int q = 0; q < (MAT == 0 ? 3 : 1); ++q) cbc[q] = cbn[q];
#pragma unroll
      for (int q = 0; q < 3; ++q) abc[q] = abn[q];
    }
    hxp = hx;
    hyp = hy;
    ex = exn;
    ey = eyn;
    ez = ezn;
    eyh = eyhn;
    ezh = ezhn;
  }
}

}  // namespace lsf
