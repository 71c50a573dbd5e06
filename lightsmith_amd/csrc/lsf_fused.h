// One time step in a single pass: H^{n+1/2} and E^{n+1} from E^n and H^{n-1/2}, for real fields without periodic
// axes (engine.py:630-663 with the pure-ADE update of engine.py:769-793, in the order of the two-kernel step).
//
// The fields are double-buffered: the kernel reads E_in, H_in and writes E_out, H_out (the H-side psi likewise),
// so a block never reads another block's new values and no store is ordered before a load. A block is 64 x cells
// by WAVES rows (one row per wavefront) marching up a chunk of z planes; it owns the upper WAVES - 1 rows and
// recomputes the row below them (blocks overlap by one row). Per plane every wavefront computes H^{n+1/2} of its
// row and publishes H_x, H_z to LDS (double-buffered by plane, one barrier per plane); the owning wavefronts then
// update E^{n+1} of their rows from their own H (the node and, by a lane shift, x - 1), the row below (LDS) and the
// previous plane (registers). The column x - 1 of the tile and the plane below the chunk (a first pass through the
// march) are recomputed too, with their owners' arithmetic. E-side psi and the pole state are updated in place
// (one owner each). Per cell the pass moves E and H once in and once out (48 B) plus the material, against 72 B for
// the two kernels.
//
// Instruction economy: every lane loads unconditionally from clamped, in-grid addresses given as 32-bit byte
// offsets from a wave-uniform base (no branches around loads, no 64-bit address arithmetic); values outside a
// component's stepped box are masked with selects; the x - 1 and x + 1 halo cells of the tile are wave-uniform and
// come through scalar loads (E_in and H_in are read-only here); tests on the row and the plane are wave-uniform.
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
// other axes drop out.
template <int MAT, int WAVES, int NP, int PML>
__global__ void __launch_bounds__(64 * WAVES) k_step_fused(const FusedArgs* __restrict__ args, Sched sc,
                                                          const int* __restrict__ tiles) {
  using namespace fused;
  __shared__ float tab[MAT == 1 ? 256 : 1];
  __shared__ float med_tab[NP > 0 ? MAX_MEDIA_FLOATS : 1];
  __shared__ float lhx[2][WAVES][64];  // H_x, H_z of each row per plane parity
  __shared__ float lhz[2][WAVES][64];
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

  const Grid& g = a.g;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const float cH = vreg(g.cH), cdx = vreg(g.cdx);
  const float sv[3] = {vreg(g.s[0]), vreg(g.s[1]), vreg(g.s[2])};
  const int mpol = vregi(a.max_poles), mstr = vregi(a.media_stride);  // pole pairs per node; floats per medium
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tx = t % a.tiles_x, rest = t / a.tiles_x;
  const int ty = rest % a.tiles_y, tz = rest / a.tiles_y;
  const int x0 = tx * 64;
  // Wavefront w computes H on row j0 - 1 + w (j0 = ty (WAVES - 1)); wavefronts 1.. also update E there and own the
  // row. Wavefront 0 recomputes the row below the block (owned by the block below), H only, so that every
  // wavefront does the same H arithmetic between barriers. Rows outside the grid still take part in the barriers.
  const int j = ty * (WAVES - 1) + wv - 1;
  const bool jin = j >= 0 && j < ny;
  const bool owner = wv >= 1;
  const int jc = j < 0 ? 0 : (j < ny ? j : ny - 1);  // the row loads read
  const int i = x0 + lane;
  const int ic = i < nx ? i : nx - 1;
  const int z0 = tz * a.zc, z1 = min(z0 + a.zc, nz);
  const unsigned N4 = (unsigned)g.N * 4u, P4 = (unsigned)(nx * ny) * 4u, R4 = (unsigned)nx * 4u;
  const unsigned ol = (unsigned)jc * R4 + (unsigned)ic * 4u;  // this lane's byte offset within a plane
  const unsigned dup = jc + 1 < ny ? R4 : 0u;                  // to the row above (clamped)
  const bool has_l = x0 >= 1 && jin;                            // the column x - 1 exists
  const bool has_r = x0 + 64 < nx && jin;                       // the column x + 64 exists
  const unsigned oxl = ((unsigned)jc * nx + (unsigned)(has_l ? x0 - 1 : 0)) * 4u;
  const unsigned oxr = ((unsigned)jc * nx + (unsigned)(has_r ? x0 + 64 : 0)) * 4u;
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
  // per plane along z.
  bool bxE[3], bxH[3], byE[3], byH[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    bxE[c] = c == 0 ? i < nx - 1 : (i >= 1 && i < nx - 1);
    bxH[c] = c == 0 ? i < nx : i < nx - 1;
    byE[c] = jin && (c == 1 ? j < ny - 1 : (j >= 1 && j < ny - 1));
    byH[c] = jin && (c == 1 || j < ny - 1);
  }
  // the row below the block (j - 1 < ny - 1) and the column x - 1 (x0 - 1 < nx - 1) lie in the H boxes across
  const bool lxH1 = has_l && byH[1];
  const bool lxH2 = has_l && byH[2];

  // CPML profiles {b, c, 1/kappa, slab index or -1}
  const float4 pex = ((PML & 1) && a.prof[0][0]) ? ((const float4*)as_global(a.prof[0][0]))[ic] : no_pml();
  const float4 phx = ((PML & 1) && a.prof[1][0]) ? ((const float4*)as_global(a.prof[1][0]))[ic] : no_pml();
  const float4 phxl = ((PML & 1) && a.prof[1][0] && has_l) ? ld_uniform(a.prof[1][0], x0 - 1) : no_pml();
  const float4 pey = ((PML & 2) && a.prof[0][1]) ? ld_uniform(a.prof[0][1], jc) : no_pml();
  const float4 phy = ((PML & 2) && a.prof[1][1]) ? ld_uniform(a.prof[1][1], jc) : no_pml();
  const int pxE = (int)pex.w, pxH = (int)phx.w, pxl = (int)phxl.w, pyE = (int)pey.w, pyH = (int)phy.w;
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

  const int k0 = z0 > 0 ? z0 - 1 : z0;
  // The march is software-pipelined: the field values a plane needs (E of the row at k and k + 1, E of the row
  // above, old H, the x + 64 halo column) are loaded one iteration ahead, so their latency overlaps the previous
  // plane's arithmetic, barrier and stores.
  float ex, ey, ez, eyr, ezr;       // E(k) of the row; E_y, E_z of the column x + 64 (wave-uniform)
  float exn, eyn, ezn, eyrn, ezrn;  // the same at plane k + 1 (clamped)
  float exu, ezu, hox, hoy, hoz;    // E_x, E_z of the row above at k; old H at k
  {
    const unsigned o = (unsigned)k0 * P4, o1 = (unsigned)min(k0 + 1, nz - 1) * P4;
    ex = ld(Ex, o + ol);
    ey = ld(Ey, o + ol);
    ez = ld(Ez, o + ol);
    exn = ld(Ex, o1 + ol);
    eyn = ld(Ey, o1 + ol);
    ezn = ld(Ez, o1 + ol);
    exu = ld(Ex, o + ol + dup);
    ezu = ld(Ez, o + ol + dup);
    hox = ld(Hx, o + ol);
    hoy = ld(Hy, o + ol);
    hoz = ld(Hz, o + ol);
    eyr = ld(Ey, o + oxr);  // all lanes read the same (in-grid) address; used when has_r
    ezr = ld(Ez, o + oxr);
    eyrn = ld(Ey, o1 + oxr);
    ezrn = ld(Ez, o1 + oxr);
  }
  float hxp = 0.f, hyp = 0.f;  // H_x, H_y of the row at the previous plane
  // material and chunk bases of the plane being updated, prefetched one plane ahead
  uint32_t idc = 0u, idn = 0u;
  float cbc[MAT == 0 ? 3 : 1] = {}, cbn[MAT == 0 ? 3 : 1] = {};
  int abc[3] = {0, 0, 0}, abn[3] = {0, 0, 0};
  auto load_mat = [&](int k, uint32_t& id, float* cb, int* ab) {
    const unsigned o = (unsigned)k * P4 + ol;
    if (MAT == 1) id = *(gptr<const uint32_t>)((gptr<const char>)as_global(a.mid) + o);
    else
#pragma unroll
      for (int q = 0; q < 3; ++q) cb[q] = ld(as_global(a.cbd), (unsigned)q * N4 + o);
    if (NP > 0)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        ab[q] = ldi(as_global(a.chunk_base),
                    ((((unsigned)k * ny + (unsigned)jc) * a.nchunk + (unsigned)tx * 2u + (unsigned)(lane >> 5)) * 3u +
                     (unsigned)q) * 4u);
  };
  load_mat(z0, idc, cbc, abc);
  __syncthreads();  // LDS tables ready
  gptr<float> P = as_global((float*)a.P);  // pole pairs {P^n, P^{n-1}}

  for (int k = k0; k < z1; ++k) {
    const bool own = k >= z0;
    const int b = k & 1;
    const unsigned ok = (unsigned)k * P4;
    const unsigned o = ok + ol;
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

    // H-side psi (slabs only): [component][term]
    float qx0 = 0.f, qx1 = 0.f, qy0 = 0.f, qy1 = 0.f, qz0 = 0.f, qz1 = 0.f;
    if (pyH >= 0) {  // y slab: H_x (d1 = y) and H_z (d2 = y)
      qx0 = ld(as_global(a.psiHi[1]), psi_off(1, 0, 1, pyH, k, jc, ic4));
      qz1 = ld(as_global(a.psiHi[1]), psi_off(1, 2, 1, pyH, k, jc, ic4));
    }
    if (pzH >= 0) {  // z slab: H_x (d2 = z) and H_y (d1 = z)
      qx1 = ld(as_global(a.psiHi[2]), psi_off(1, 0, 2, pzH, k, jc, ic4));
      qy0 = ld(as_global(a.psiHi[2]), psi_off(1, 1, 2, pzH, k, jc, ic4));
    }
    if (xslab) {  // x slab: H_y (d2 = x) and H_z (d1 = x), per lane
      const int p = pxH >= 0 ? pxH : 0;
      qy1 = ld(as_global(a.psiHi[0]), psi_off(1, 1, 0, p, k, jc, 0u));
      qz0 = ld(as_global(a.psiHi[0]), psi_off(1, 2, 0, p, k, jc, 0u));
    }
    // ---- the column x - 1 (wave-uniform): H_y, H_z there
    float hly = 0.f, hlz = 0.f;
    if (has_l) {
      const unsigned ol0 = ok + oxl;
      const unsigned oln = (k + 1 < nz ? ok + P4 : ok) + oxl;
      const float exl = ld(Ex, ol0), exln = ld(Ex, oln), exlu = ld(Ex, ol0 + dup);
      const float eyl = ld(Ey, ol0), ezl = ld(Ez, ol0);
      const float holy = ld(Hy, ol0), holz = ld(Hz, ol0);
      const float ezx0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ez), 0));
      const float eyx0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ey), 0));
      const bool sly = lxH1 && bzH[1], slz = lxH2 && bzH[2];
      const unsigned xl4 = (unsigned)(x0 - 1) * 4u;
      float ly0 = 0.f, ly1 = 0.f, lz0 = 0.f, lz1 = 0.f;
      if (pzH >= 0) ly0 = ld(as_global(a.psiHi[2]), psi_off(1, 1, 2, pzH, k, jc, xl4));
      if (pxl >= 0) {
        ly1 = ld(as_global(a.psiHi[0]), psi_off(1, 1, 0, pxl, k, jc, 0u));
        lz0 = ld(as_global(a.psiHi[0]), psi_off(1, 2, 0, pxl, k, jc, 0u));
      }
      if (pyH >= 0) lz1 = ld(as_global(a.psiHi[1]), psi_off(1, 2, 1, pyH, k, jc, xl4));
      hly = sly ? h_upd(1, holy, exln - exl, ezx0 - ezl, phz, phxl, pzH >= 0, pxl >= 0, ly0, ly1) : holy;
      hlz = slz ? h_upd(2, holz, eyx0 - eyl, exlu - exl, phxl, phy, pxl >= 0, pyH >= 0, lz0, lz1) : holz;
    }

    // ---- loads for E(k): E-side psi, pole state, next material
    bool ste[3];
    float pse[3][2] = {};
    int aq[3] = {-1, -1, -1};
    bool anyd[3] = {false, false, false};  // some lane of the row holds a D-form node of the component
    Pair<float> pp[3][NP > 0 ? NP : 1];
    if (own && owner) {
      if (k + 1 < z1) load_mat(k + 1, idn, cbn, abn);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ste[c] = bxE[c] && byE[c] && bzE[c];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          if (ax == 0 && xslab) pse[c][s] = ld(as_global((const float*)a.psiE[0]),
                                               psi_off(0, c, 0, pxE >= 0 ? pxE : 0, k, jc, 0u));
          if (ax == 1 && pyE >= 0) pse[c][s] = ld(as_global((const float*)a.psiE[1]), psi_off(0, c, 1, pyE, k, jc, ic4));
          if (ax == 2 && pzE >= 0) pse[c][s] = ld(as_global((const float*)a.psiE[2]), psi_off(0, c, 2, pzE, k, jc, ic4));
        }
        if (NP > 0) {
          const float cv = MAT == 1 ? tab[(idc >> (8 * c)) & 255u] : cbc[MAT == 0 ? c : 0];
          const bool dsp = ste[c] && cv < 0.f;
          const unsigned long long m = __ballot(dsp);
          anyd[c] = m != 0ull;
          if (anyd[c]) {  // every lane loads: its node's pole pairs, or node 0's (in range) where it has none
            const int slot = abc[c] + __popcll(m & below);
            aq[c] = dsp ? slot : -1;
            const unsigned po = (unsigned)(dsp ? slot : 0) * (unsigned)mpol * 8u;
#pragma unroll
            for (int p = 0; p < (NP > 0 ? NP : 1); ++p) {
              const unsigned pq = (unsigned)min(p, mpol - 1) * 8u;
              pp[c][p].p = ld((gptr<const float>)P, po + pq);
              pp[c][p].q = ld((gptr<const float>)P, po + pq + 4u);
            }
          }
        }
      }
    }

    // ---- the next plane's field values (used in the next iteration), issued after this
    // plane's loads so that waiting for those leaves these in flight
    const unsigned p1 = (unsigned)min(k + 1, nz - 1) * P4, p2 = (unsigned)min(k + 2, nz - 1) * P4;
    const float exnn = ld(Ex, p2 + ol), eynn = ld(Ey, p2 + ol), eznn = ld(Ez, p2 + ol);
    const float exun = ld(Ex, p1 + ol + dup), ezun = ld(Ez, p1 + ol + dup);
    const float hoxn = ld(Hx, p1 + ol), hoyn = ld(Hy, p1 + ol), hozn = ld(Hz, p1 + ol);
    const float eyrnn = ld(Ey, p2 + oxr), ezrnn = ld(Ez, p2 + oxr);
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
      lhx[b][wv][lane] = hx;
      lhz[b][wv][lane] = hz;
      __syncthreads();
    }
    if (own && owner) {
      // ---- E^{n+1}(k) of the row
      const float hxm = lhx[b][wv - 1][lane], hzmy = lhz[b][wv - 1][lane];  // the row below
      float hym = shfl_up(hy), hzm = shfl_up(hz);                    // x - 1
      if (lane == 0) {
        hym = hly;
        hzm = hlz;
      }
      const float eold[3] = {ex, ey, ez};
      float ev[3];
      int npc[3] = {0, 0, 0};  // poles of each lane's medium (0 without a D-form node)
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
          if (use)
            dE += pml_apply(pr, 1, s == 0 ? cb * sv[d1] : -cb * sv[d2], s == 0 ? da : db, pse[c][s]);
        }
        const float eo = eold[c];
        float enew = eo + dE;
        if (NP > 0 && anyd[c]) {  // branch-free over the row: lanes without a D-form node keep enew
          const bool d = aq[c] >= 0;
          const float* mp = med_tab + (d ? (int)(-cbv) - 1 : 0) * mstr;
          npc[c] = d ? (int)mp[1] : 0;
          float dp = 0.f;
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p) {
            const bool use = p < npc[c];
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
        if (pyH >= 0) st(as_global(a.psiHo[1]), psi_off(1, 0, 1, pyH, k, jc, ic4), qx0);
        if (pzH >= 0) st(as_global(a.psiHo[2]), psi_off(1, 0, 2, pzH, k, jc, ic4), qx1);
      }
      if (sty) {
        st(Ho, N4 + o, hy);
        if (pzH >= 0) st(as_global(a.psiHo[2]), psi_off(1, 1, 2, pzH, k, jc, ic4), qy0);
        if (xslab && pxH >= 0) st(as_global(a.psiHo[0]), psi_off(1, 1, 0, pxH, k, jc, 0u), qy1);
      }
      if (stz) {
        st(Ho, 2 * N4 + o, hz);
        if (xslab && pxH >= 0) st(as_global(a.psiHo[0]), psi_off(1, 2, 0, pxH, k, jc, 0u), qz0);
        if (pyH >= 0) st(as_global(a.psiHo[1]), psi_off(1, 2, 1, pyH, k, jc, ic4), qz1);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!ste[c]) continue;
        st(Eo, (unsigned)c * N4 + o, ev[c]);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ax = (c + 1 + s) % 3;
          if (ax == 0 && xslab && pxE >= 0) st(as_global(a.psiE[0]), psi_off(0, c, 0, pxE, k, jc, 0u), pse[c][s]);
          if (ax == 1 && pyE >= 0) st(as_global(a.psiE[1]), psi_off(0, c, 1, pyE, k, jc, ic4), pse[c][s]);
          if (ax == 2 && pzE >= 0) st(as_global(a.psiE[2]), psi_off(0, c, 2, pzE, k, jc, ic4), pse[c][s]);
        }
        if (NP > 0 && aq[c] >= 0) {  // the medium's poles (the others are unchanged)
#pragma unroll
          for (int p = 0; p < (NP > 0 ? NP : 1); ++p)
            if (p < npc[c]) {
              const unsigned po = ((unsigned)aq[c] * (unsigned)mpol + (unsigned)p) * 8u;
              st((gptr<float>)P, po, pp[c][p].p);
              st((gptr<float>)P, po + 4u, pp[c][p].q);
            }
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
    exn = exnn;
    eyn = eynn;
    ezn = eznn;
    exu = exun;
    ezu = ezun;
    hox = hoxn;
    hoy = hoyn;
    hoz = hozn;
    eyr = eyrn;
    ezr = ezrn;
    eyrn = eyrnn;
    ezrn = ezrnn;
  }
}

}  // namespace lsf
