// One time step in a single pass: H^{n+1/2} and E^{n+1} from E^n and H^{n-1/2}, for real fields without periodic
// axes (engine.py:630-663 with the pure-ADE update of engine.py:769-793, in the order of the two-kernel step).
//
// The fields are double-buffered: the kernel reads E_in, H_in and writes E_out, H_out (the H-side psi likewise),
// so no wavefront reads another's new values and no store is ordered before a load. Each wavefront owns one row of
// 64 x cells and marches up a chunk of z planes on its own, with no barrier and no exchange through LDS: per plane
// it computes H^{n+1/2} of its row (stored), recomputes H_x, H_z of the row below and H_y, H_z of the column x - 1
// with their owners' arithmetic and inputs (not stored), and then updates E^{n+1} of its row from those, its own H
// (the node and, by a lane shift, x - 1) and the previous plane's H (registers). The plane below the chunk is a
// first pass of the march (H of the row only). E-side psi and the pole state are updated in place (one owner each).
// Per cell the pass moves E and H once in and once out (48 B) plus the material, against 72 B for the two kernels;
// the rows above and below that it reads as well are the rows the neighbouring wavefronts read in the same planes,
// so those loads are served by the caches.
//
// Instruction economy: field loads use clamped, in-grid addresses given as 32-bit byte offsets from a wave-uniform
// base (no branches around them, no 64-bit address arithmetic); values outside a component's stepped box are masked
// with selects; tests on the row and the plane are wave-uniform. The x - 1 and x + 64 halo cells of the row are
// wave-uniform addresses read by vector loads: as scalar loads (constant address space) they share the lgkm
// counter with the LDS table lookups and the lane shuffles, whose waits then wait for them too (slower).
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
  unsigned pso[2][3][3];  // byte offset of component c's slot in the (side, axis) psi block (c != axis)
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

namespace fused {

// base + a 32-bit byte offset: the loads compile to global loads with a scalar base and a 32-bit vector offset
__device__ __forceinline__ float ld(gptr<const float> base, unsigned off) {
  return *(gptr<const float>)((gptr<const char>)base + off);
}
__device__ __forceinline__ void st(gptr<float> base, unsigned off, float v) { *(gptr<float>)((gptr<char>)base + off) = v; }
__device__ __forceinline__ int ldi(gptr<const int> base, unsigned off) {
  return *(gptr<const int>)((gptr<const char>)base + off);
}
}  // namespace fused

// PML: bit a set when the tiles given may hold CPML cells of axis a (their halos included); the psi paths of the
// other axes drop out. NP: the engine's pole count per D-form node (the stride of the pole slots).
template <int MAT, int WAVES, int NP, int PML>
__global__ void __launch_bounds__(64 * WAVES) k_step_fused(const FusedArgs* __restrict__ args, Sched sc,
                                                          const int* __restrict__ tiles) {
  using namespace fused;
  __shared__ float tab[MAT == 1 ? 256 : 1];
  __shared__ float med_tab[NP > 0 ? MAX_MEDIA_FLOATS : 1];
  const int tb = xcd_tile(blockIdx.x, sc.n_pad);
  if (tb >= sc.n_tiles) return;  // whole block
  const int t = ld_uniform(tiles, tb);  // logical tile (x fastest, then y, then z chunk)
  const FusedArgs& a = args[0];
  if (a.tick && t == 0 && threadIdx.x == 0) {  // tile 0 is in exactly one list
    long long* cnt = (long long*)as_global(a.counter);
    const long long n = cnt[0];
    cnt[1] = n;
    cnt[0] = n + 1;
  }
  if (MAT == 1)
    for (int q = threadIdx.x; q < 256; q += 64 * WAVES) tab[q] = ((const float*)as_global(a.mtab))[q];
  if (NP > 0)
    for (int q = threadIdx.x; q < a.n_media_floats; q += 64 * WAVES) med_tab[q] = ((const float*)as_global(a.media))[q];
  __syncthreads();  // LDS tables ready: the kernel's only barrier

  const Grid& g = a.g;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int j = ty * WAVES + wv;  // the row this wavefront owns
  if (j >= ny) return;
  const float cH = vreg(g.cH), cdx = vreg(g.cdx);
  const float sv[3] = {vreg(g.s[0]), vreg(g.s[1]), vreg(g.s[2])};
  const int mstr = vregi(a.media_stride);  // floats per medium
  const int x0 = tx * 64;
  const int i = x0 + lane;
  const int ic = i < nx ? i : nx - 1;
  const bool hb = j >= 1;  // the row below exists (its H_x, H_z are recomputed)
  const int jb = hb ? j - 1 : j;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, nz);
  const unsigned N4 = (unsigned)g.N * 4u, P4 = (unsigned)(nx * ny) * 4u, R4 = (unsigned)nx * 4u;
  const unsigned ol = (unsigned)j * R4 + (unsigned)ic * 4u;   // this lane's byte offset within a plane
  const unsigned olb = (unsigned)jb * R4 + (unsigned)ic * 4u;  // the same in the row below (clamped)
  const unsigned dup = j + 1 < ny ? R4 : 0u;                   // to the row above (clamped)
  const bool has_l = x0 >= 1;                                  // the column x - 1 exists
  const bool has_r = x0 + 64 < nx;                             // the column x + 64 exists
  const unsigned xl0 = (unsigned)(has_l ? x0 - 1 : 0) * 4u, xr0 = (unsigned)(has_r ? x0 + 64 : 0) * 4u;
  const unsigned oxl = (unsigned)j * R4 + xl0, oxr = (unsigned)j * R4 + xr0, oxrb = (unsigned)jb * R4 + xr0;
  // lanes of this lane's 32-cell chunk below it (the rank of a D-form node within its chunk)
  const unsigned long long below = (lane & 31) == 0 ? 0ull : (((~0ull) >> (64 - (lane & 31))) << ((lane >> 5) * 32));

  gptr<const float> __restrict__ Ex = as_global(a.E_in);
  gptr<const float> __restrict__ Ey = as_global(a.E_in + g.N);
  gptr<const float> __restrict__ Ez = as_global(a.E_in + 2 * g.N);
  gptr<const float> __restrict__ Hx = as_global(a.H_in);
  gptr<const float> __restrict__ Hy = as_global(a.H_in + g.N);
  gptr<const float> __restrict__ Hz = as_global(a.H_in + 2 * g.N);
  gptr<float> __restrict__ Eo = as_global(a.E_out);
  gptr<float> __restrict__ Ho = as_global(a.H_out);

  // stepped boxes without periodic axes or mirror planes (engine.py:47-66): E_c spans [0, n - 1) along c and
  // [1, n - 1) across; H_c spans [0, n) along c and [0, n - 1) across. Per lane along x, per wavefront along y,
  // per plane along z. The row below (j - 1 < ny - 1) and the column x - 1 (x0 - 1 < nx - 1) lie in the H boxes
  // across.
  bool bxE[3], bxH[3], byE[3], byH[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    bxE[c] = c == 0 ? i < nx - 1 : (i >= 1 && i < nx - 1);
    bxH[c] = c == 0 ? i < nx : i < nx - 1;
    byE[c] = c == 1 ? j < ny - 1 : (j >= 1 && j < ny - 1);
    byH[c] = c == 1 || j < ny - 1;
  }
  const bool lxH1 = has_l && byH[1];
  const bool lxH2 = has_l && byH[2];

  // CPML profiles {b, c, 1/kappa, slab index or -1}
  const float4 pex = ((PML & 1) && a.prof[0][0]) ? ((const float4*)as_global(a.prof[0][0]))[ic] : no_pml();
  const float4 phx = ((PML & 1) && a.prof[1][0]) ? ((const float4*)as_global(a.prof[1][0]))[ic] : no_pml();
  const float4 phxl = ((PML & 1) && a.prof[1][0] && has_l) ? ld_uniform(a.prof[1][0], x0 - 1) : no_pml();
  const float4 pey = ((PML & 2) && a.prof[0][1]) ? ld_uniform(a.prof[0][1], j) : no_pml();
  const float4 phy = ((PML & 2) && a.prof[1][1]) ? ld_uniform(a.prof[1][1], j) : no_pml();
  const float4 phyb = ((PML & 2) && a.prof[1][1] && hb) ? ld_uniform(a.prof[1][1], jb) : no_pml();
  const int pxE = (int)pex.w, pxH = (int)phx.w, pxl = (int)phxl.w, pyE = (int)pey.w, pyH = (int)phy.w;
  const int pyB = (int)phyb.w;
  // whether any lane of the tile lies in the x slabs (wave-uniform)
  const bool xslab = (PML & 1) && __ballot(pxE >= 0 || pxH >= 0) != 0ull;

  // byte offset of a psi value: slab axis 0 by lane, axes 1 and 2 by row / plane
  auto psi_off = [&](int f, int c, int ax, int p, int k, int row, unsigned lane4) -> unsigned {
    const unsigned base = a.pso[f][ax][c];
    if (ax == 0) return base + (((unsigned)k * ny + (unsigned)row) * (unsigned)a.ns[f][0] + (unsigned)p) * 4u;
    if (ax == 1) return base + ((unsigned)k * (unsigned)a.ns[f][1] + (unsigned)p) * R4 + lane4;
    return base + ((unsigned)p * ny + (unsigned)row) * R4 + lane4;
  };
  const unsigned ic4 = (unsigned)ic * 4u;

  // h_new = h_old - cH (s_d1 da - s_d2 db) + CPML terms in ascending axis order (the H kernel's arithmetic);
  // use1 / use2: the term of axis d1 / d2 applies (in its slab)
  auto h_upd = [&](int c, float hold, float da, float db, float4 pr1, float4 pr2, bool use1, bool use2, float& ps1,
                   float& ps2) -> float {
    const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
    float acc = -cH * (sv[d1] * da - sv[d2] * db);
    if (d1 < d2) {
      if (use1) acc += pml_apply(pr1, 1, -cH * sv[d1], da, ps1);
      if (use2) acc += pml_apply(pr2, 1, cH * sv[d2], db, ps2);
    } else {
      if (use2) acc += pml_apply(pr2, 1, cH * sv[d2], db, ps2);
      if (use1) acc += pml_apply(pr1, 1, -cH * sv[d1], da, ps1);
    }
    return hold + acc;
  };

  // material and chunk bases of the plane being updated, prefetched one plane ahead
  uint32_t idc = 0u, idn = 0u;
  float cbc[MAT == 0 ? 3 : 1] = {}, cbn[MAT == 0 ? 3 : 1] = {};
  int abc[3] = {0, 0, 0}, abn[3] = {0, 0, 0};
  // the lane's 32-cell chunk (clamped: a lane past the row's last chunk holds no node)
  const unsigned chk = (unsigned)min(tx * 2 + (lane >> 5), a.nchunk - 1);
  auto load_mat = [&](int k, uint32_t& id, float* cb, int* ab) {
    const unsigned o = (unsigned)k * P4 + ol;
    if (MAT == 1) id = *(gptr<const uint32_t>)((gptr<const char>)as_global(a.mid) + o);
    else
#pragma unroll
      for (int q = 0; q < 3; ++q) cb[q] = ld(as_global(a.cbd), (unsigned)q * N4 + o);
    if (NP > 0)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        ab[q] = ldi(as_global(a.chunk_base), ((((unsigned)k * ny + (unsigned)j) * a.nchunk + chk) * 3u + (unsigned)q) * 4u);
  };

  const int k0 = z0 > 0 ? z0 - 1 : z0;
  // carried from plane to plane: E of the row at k, E_y of the row below at k, H_x and H_y of the row at k - 1
  float ex, ey, ez, eyb;
  {
    const unsigned o = (unsigned)k0 * P4 + ol;
    ex = ld(Ex, o);
    ey = ld(Ey, o);
    ez = ld(Ez, o);
    eyb = ld(Ey, (unsigned)z0 * P4 + olb);
  }
  float hxp = 0.f, hyp = 0.f;
  load_mat(z0, idc, cbc, abc);
  const gptr<Pair<float>> Pg = as_global((Pair<float>*)a.P);
  Pair<float>* __restrict__ P = (Pair<float>*)Pg;  // pole pairs {P^n, P^{n-1}}, NP per D-form node

  for (int k = k0; k < z1; ++k) {
    const bool own = k >= z0;  // wave-uniform: the first pass below the chunk computes H of the row only
    const unsigned ok = (unsigned)k * P4;
    const unsigned o = ok + ol, ob = ok + olb;
    const unsigned o1 = k + 1 < nz ? ok + P4 : ok;  // plane k + 1, clamped
    const float4 phz = ((PML & 4) && a.prof[1][2]) ? ld_uniform(a.prof[1][2], k) : no_pml();
    const float4 pez = ((PML & 4) && a.prof[0][2]) ? ld_uniform(a.prof[0][2], k) : no_pml();
    const int pzH = (int)phz.w, pzE = (int)pez.w;
    bool bzE[3], bzH[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bzE[c] = c == 2 ? k < nz - 1 : (k >= 1 && k < nz - 1);
      bzH[c] = c == 2 || k < nz - 1;
    }
    const bool stx = bxH[0] && byH[0] && bzH[0], sty = bxH[1] && byH[1] && bzH[1], stz = bxH[2] && byH[2] && bzH[2];

    // ---- the plane's field values: E of the row at k + 1, of the row above at k, old H of the row; the row below
    // and the halo columns
    const float exn = ld(Ex, o1 + ol), eyn = ld(Ey, o1 + ol), ezn = ld(Ez, o1 + ol);
    const float exu = ld(Ex, o + dup), ezu = ld(Ez, o + dup);
    const float hox = ld(Hx, o), hoy = ld(Hy, o), hoz = ld(Hz, o);
    const float eyr = ld(Ey, ok + oxr), ezr = ld(Ez, ok + oxr);  // the column x + 64 (wave-uniform; used if has_r)
    float exb = 0.f, ezb = 0.f, eybn = 0.f, hoxb = 0.f, hozb = 0.f, eyrb = 0.f;
    if (own) {
      exb = ld(Ex, ob);
      ezb = ld(Ez, ob);
      eybn = ld(Ey, o1 + olb);
      hoxb = ld(Hx, ob);
      hozb = ld(Hz, ob);
      eyrb = ld(Ey, ok + oxrb);
    }

    // H-side psi of the row (slabs only): [component][term]
    float qx0 = 0.f, qx1 = 0.f, qy0 = 0.f, qy1 = 0.f, qz0 = 0.f, qz1 = 0.f;
    if (pyH >= 0) {  // y slab: H_x (d1 = y) and H_z (d2 = y)
      qx0 = ld(as_global(a.psiHi[1]), psi_off(1, 0, 1, pyH, k, j, ic4));
      qz1 = ld(as_global(a.psiHi[1]), psi_off(1, 2, 1, pyH, k, j, ic4));
    }
    if (pzH >= 0) {  // z slab: H_x (d2 = z) and H_y (d1 = z)
      qx1 = ld(as_global(a.psiHi[2]), psi_off(1, 0, 2, pzH, k, j, ic4));
      qy0 = ld(as_global(a.psiHi[2]), psi_off(1, 1, 2, pzH, k, j, ic4));
    }
    if (xslab && pxH >= 0) {  // x slab: H_y (d2 = x) and H_z (d1 = x), in the slab's lanes only
      qy1 = ld(as_global(a.psiHi[0]), psi_off(1, 1, 0, pxH, k, j, 0u));
      qz0 = ld(as_global(a.psiHi[0]), psi_off(1, 2, 0, pxH, k, j, 0u));
    }
    // the same for H_x, H_z of the row below (read only: their owner advances them)
    float rx0 = 0.f, rx1 = 0.f, rz0 = 0.f, rz1 = 0.f;
    if (own) {
      if (pyB >= 0) {
        rx0 = ld(as_global(a.psiHi[1]), psi_off(1, 0, 1, pyB, k, jb, ic4));
        rz1 = ld(as_global(a.psiHi[1]), psi_off(1, 2, 1, pyB, k, jb, ic4));
      }
      if (pzH >= 0) rx1 = ld(as_global(a.psiHi[2]), psi_off(1, 0, 2, pzH, k, jb, ic4));
      if (xslab && pxH >= 0) rz0 = ld(as_global(a.psiHi[0]), psi_off(1, 2, 0, pxH, k, jb, 0u));
    }

    // ---- the cell state of E(k): E-side psi, pole state, next material
    bool ste[3] = {false, false, false};
    float pse[3][2] = {};
    int aq[3] = {-1, -1, -1};
    bool anyd[3] = {false, false, false};  // some lane of the row holds a D-form node of the component
    Pair<float> pp[3][NP > 0 ? NP : 1];
    if (own) {
      if (k + 1 < z1) load_mat(k + 1, idn, cbn, abn);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ste[c] = bxE[c] && byE[c] && bzE[c];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          if (ax == 0 && xslab && pxE >= 0)
            pse[c][s] = ld(as_global((const float*)a.psiE[0]), psi_off(0, c, 0, pxE, k, j, 0u));
          if (ax == 1 && pyE >= 0) pse[c][s] = ld(as_global((const float*)a.psiE[1]), psi_off(0, c, 1, pyE, k, j, ic4));
          if (ax == 2 && pzE >= 0) pse[c][s] = ld(as_global((const float*)a.psiE[2]), psi_off(0, c, 2, pzE, k, j, ic4));
        }
        if (NP > 0) {
          const float cv = MAT == 1 ? tab[(idc >> (8 * c)) & 255u] : cbc[MAT == 0 ? c : 0];
          const bool dsp = ste[c] && cv < 0.f;
          const unsigned long long m = __ballot(dsp);
          anyd[c] = m != 0ull;
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p) pp[c][p] = Pair<float>{0.f, 0.f};
          if (dsp) {
            aq[c] = abc[c] + __popcll(m & below);
            load_slot(P + (unsigned)aq[c] * (unsigned)NP, pp[c]);
          }
        }
      }
    }

    // ---- H^{n+1/2}(k) of the row
    float eyp = shfl_dn(ey), ezp = shfl_dn(ez);  // x + 1
    if (lane == 63) {
      eyp = eyr;
      ezp = ezr;
    }
    const float hx = stx ? h_upd(0, hox, ezu - ez, eyn - ey, phy, phz, pyH >= 0, pzH >= 0, qx0, qx1) : hox;
    const float hy = sty ? h_upd(1, hoy, exn - ex, ezp - ez, phz, phx, pzH >= 0, xslab && pxH >= 0, qy0, qy1) : hoy;
    const float hz = stz ? h_upd(2, hoz, eyp - ey, exu - ex, phx, phy, xslab && pxH >= 0, pyH >= 0, qz0, qz1) : hoz;

    if (own) {
      // ---- H_x, H_z of the row below (as their owner computes them)
      float eybp = shfl_dn(eyb);
      if (lane == 63) eybp = eyrb;
      const bool sbx = hb && bxH[0] && bzH[0], sbz = hb && bxH[2] && bzH[2];
      const float hxm = sbx ? h_upd(0, hoxb, ez - ezb, eybn - eyb, phyb, phz, pyB >= 0, pzH >= 0, rx0, rx1) : hoxb;
      const float hzmy =
          sbz ? h_upd(2, hozb, eybp - eyb, ex - exb, phx, phyb, xslab && pxH >= 0, pyB >= 0, rz0, rz1) : hozb;

      // ---- H_y, H_z of the column x - 1 (wave-uniform)
      float hly = 0.f, hlz = 0.f;
      if (has_l) {
        const unsigned ol0 = ok + oxl, oln = o1 + oxl;
        const float exl = ld(Ex, ol0), exln = ld(Ex, oln), exlu = ld(Ex, ol0 + dup);
        const float eyl = ld(Ey, ol0), ezl = ld(Ez, ol0);
        const float holy = ld(Hy, ol0), holz = ld(Hz, ol0);
        const float ezx0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ez), 0));
        const float eyx0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ey), 0));
        const bool sly = lxH1 && bzH[1], slz = lxH2 && bzH[2];
        float ly0 = 0.f, ly1 = 0.f, lz0 = 0.f, lz1 = 0.f;
        if (pzH >= 0) ly0 = ld(as_global(a.psiHi[2]), psi_off(1, 1, 2, pzH, k, j, xl0));
        if (pxl >= 0) {
          ly1 = ld(as_global(a.psiHi[0]), psi_off(1, 1, 0, pxl, k, j, 0u));
          lz0 = ld(as_global(a.psiHi[0]), psi_off(1, 2, 0, pxl, k, j, 0u));
        }
        if (pyH >= 0) lz1 = ld(as_global(a.psiHi[1]), psi_off(1, 2, 1, pyH, k, j, xl0));
        hly = sly ? h_upd(1, holy, exln - exl, ezx0 - ezl, phz, phxl, pzH >= 0, pxl >= 0, ly0, ly1) : holy;
        hlz = slz ? h_upd(2, holz, eyx0 - eyl, exlu - exl, phxl, phy, pxl >= 0, pyH >= 0, lz0, lz1) : holz;
      }

      // ---- E^{n+1}(k) of the row
      float hym = shfl_up(hy), hzm = shfl_up(hz);  // x - 1
      if (lane == 0) {
        hym = hly;
        hzm = hlz;
      }
      const float eold[3] = {ex, ey, ez};
      float ev[3];
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
        const float cb = cbv < 0.f ? cdx : cbv;
        float dE = cb * (sv[d1] * da - sv[d2] * db);
        const int first = d1 < d2 ? 0 : 1;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int s = q == 0 ? first : 1 - first;
          const int ax = (c + 1 + s) % 3;
          const float4 pr = ax == 0 ? pex : (ax == 1 ? pey : pez);
          const bool use = ax == 0 ? (xslab && pxE >= 0) : (ax == 1 ? pyE >= 0 : pzE >= 0);
          if (use) dE += pml_apply(pr, 1, s == 0 ? cb * sv[d1] : -cb * sv[d2], s == 0 ? da : db, pse[c][s]);
        }
        const float eo = eold[c];
        float enew = eo + dE;
        if (NP > 0 && anyd[c]) {  // branch-free over the row: lanes without a D-form node keep enew
          const bool d = aq[c] >= 0;
          const float* mp = med_tab + (d ? (int)(-cbv) - 1 : 0) * mstr;
          const int npc = d ? (int)mp[1] : 0;  // poles of the lane's medium
          float dp = 0.f;
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p) {
            const bool use = p < npc;
            const float pn = mp[2 + 3 * p] * pp[c][p].p + mp[3 + 3 * p] * pp[c][p].q + mp[4 + 3 * p] * eo;
            dp = use ? dp + (pn - pp[c][p].p) : dp;
            pp[c][p].q = use ? pp[c][p].p : pp[c][p].q;
            pp[c][p].p = use ? pn : pp[c][p].p;
          }
          enew = d ? eo + mp[0] * (dE - dp) : enew;
        }
        ev[c] = enew;
      }
      // ---- stores
      if (stx) {
        st(Ho, o, hx);
        if (pyH >= 0) st(as_global(a.psiHo[1]), psi_off(1, 0, 1, pyH, k, j, ic4), qx0);
        if (pzH >= 0) st(as_global(a.psiHo[2]), psi_off(1, 0, 2, pzH, k, j, ic4), qx1);
      }
      if (sty) {
        st(Ho, N4 + o, hy);
        if (pzH >= 0) st(as_global(a.psiHo[2]), psi_off(1, 1, 2, pzH, k, j, ic4), qy0);
        if (xslab && pxH >= 0) st(as_global(a.psiHo[0]), psi_off(1, 1, 0, pxH, k, j, 0u), qy1);
      }
      if (stz) {
        st(Ho, 2 * N4 + o, hz);
        if (xslab && pxH >= 0) st(as_global(a.psiHo[0]), psi_off(1, 2, 0, pxH, k, j, 0u), qz0);
        if (pyH >= 0) st(as_global(a.psiHo[1]), psi_off(1, 2, 1, pyH, k, j, ic4), qz1);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!ste[c]) continue;
        st(Eo, (unsigned)c * N4 + o, ev[c]);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          if (ax == 0 && xslab && pxE >= 0) st(as_global(a.psiE[0]), psi_off(0, c, 0, pxE, k, j, 0u), pse[c][s]);
          if (ax == 1 && pyE >= 0) st(as_global(a.psiE[1]), psi_off(0, c, 1, pyE, k, j, ic4), pse[c][s]);
          if (ax == 2 && pzE >= 0) st(as_global(a.psiE[2]), psi_off(0, c, 2, pzE, k, j, ic4), pse[c][s]);
        }
        if (NP > 0 && aq[c] >= 0) store_slot(P + (unsigned)aq[c] * (unsigned)NP, pp[c]);
      }
      eyb = eybn;
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
  }
}

}  // namespace lsf
