// Sparse per-step work: source corrections, auxiliary grids, interface ADE groups and monitors.
// Each kernel is one thread per output row; rows that write the same node never coexist in one
// launch (the host merges them), so no atomics are needed and results are deterministic.
#pragma once

#include "lsf_step.h"

namespace lsf {

// A correction v for E (or H) node t. The reference adds it to E before the D-form group update, where
// D += E_now - E_prev picks it up and the new E is (D - sum P) / eps_inf: for a pure D-form node the
// correction therefore moves E by v / eps_inf. scale[r] holds that factor (1 for other nodes).
template <class T>
__device__ __forceinline__ void apply_corr(T* F, long long t, T v, const float* scale, long long r) {
  F[t] = add(F[t], scale ? scl(scale[r], v) : v);
}

template <class T>
struct GatherJob {
  long long n;
  int width;
  const long long* target;
  T* F;
  const T* src;
  const long long* sidx;
  const T* w;
  const float* scale;
};

template <class T>
__device__ __forceinline__ void gather_row(const GatherJob<T>& J, long long r) {
  if (r >= J.n) return;
  T acc = zero<T>();
  for (int q = 0; q < J.width; ++q) {
    const long long e = r * J.width + q;
    acc = add(acc, mul(J.w[e], J.src[J.sidx[e]]));
  }
  apply_corr(J.F, J.target[r], acc, J.scale, r);
}

template <class T>
__global__ void k_gather(GatherJob<T> J) {
  gather_row(J, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <class T>
struct ModalJob {
  long long n;
  int nm;
  const long long* target;
  T* F;
  const T* m;
  const float* table;
  long long table_len;
  const long long* counter;
  int slot;
  const float* scale;
};

template <class T>
__device__ __forceinline__ void modal_row(const ModalJob<T>& J, long long r) {
  if (r >= J.n) return;
  const long long t = J.counter[J.slot];
  if (t < 0 || t >= J.table_len) return;
  T acc = zero<T>();
  for (int q = 0; q < J.nm; ++q) acc = add(acc, scl(J.table[t * J.nm + q], J.m[r * J.nm + q]));
  apply_corr(J.F, J.target[r], acc, J.scale, r);
}

template <class T>
__global__ void k_modal(ModalJob<T> J) {
  modal_row(J, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// One launch over the rows of many jobs (blocks of 256 rows; sc maps each block to its job).
template <class T>
__global__ void k_gather_multi(const GatherJob<T>* __restrict__ jobs, Sched sc) {
  int m, lt;
  member_of(sc, blockIdx.x, m, lt);
  gather_row(jobs[m], (long long)lt * blockDim.x + threadIdx.x);
}

template <class T>
__global__ void k_modal_multi(const ModalJob<T>* __restrict__ jobs, Sched sc) {
  int m, lt;
  member_of(sc, blockIdx.x, m, lt);
  modal_row(jobs[m], (long long)lt * blockDim.x + threadIdx.x);
}

// Quintic interpolation of tables sampled every m steps (layered.py:164-187), float32 as on the GPU
// path of the reference: (t + i0) / m = c + u with c clamped at cmax.
__device__ __forceinline__ float quintic(const float* __restrict__ T, long long t, long long i0, float u0, long long base,
                                         int m, long long cmax) {
  const long long q = t + i0;
  long long r = q % m;
  if (r < 0) r += m;
  long long c = (q - r) / m;
  if (c > cmax) c = cmax;
  const float u = ((float)r + u0) / (float)m;
  const float d0 = u + 2.f, d1 = u + 1.f, d2 = u, d3 = u - 1.f, d4 = u - 2.f, d5 = u - 3.f;
  const long long ix = base + c;
  // weights prod_{k != j} d_k / prod_{k != j} (x_j - x_k), stencil x = -2 .. 3
  const float w0 = (d1 * d2 * d3 * d4 * d5) / -120.f;
  const float w1 = (d0 * d2 * d3 * d4 * d5) / 24.f;
  const float w2 = (d0 * d1 * d3 * d4 * d5) / -12.f;
  const float w3 = (d0 * d1 * d2 * d4 * d5) / 12.f;
  const float w4 = (d0 * d1 * d2 * d3 * d5) / -24.f;
  const float w5 = (d0 * d1 * d2 * d3 * d4) / 120.f;
  return w0 * T[ix] + w1 * T[ix + 1] + w2 * T[ix + 2] + w3 * T[ix + 3] + w4 * T[ix + 4] + w5 * T[ix + 5];
}

template <class T>
__device__ __forceinline__ void table_row(long long r, const long long* __restrict__ target,
                                          const int* __restrict__ row_ptr, T* F, const float* __restrict__ tab, int m,
                                          long long cmax, const long long* __restrict__ base,
                                          const long long* __restrict__ i0, const float* __restrict__ u0,
                                          const float* __restrict__ coef, const long long* counter, int slot,
                                          const float* scale) {
  const long long t = counter[slot];
  float acc = 0.f;
  for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p) acc += coef[p] * quintic(tab, t, i0[p], u0[p], base[p], m, cmax);
  T v = zero<T>();
  if constexpr (sizeof(T) == sizeof(float)) v = acc;
  else v = mk(acc, 0.f);
  apply_corr(F, target[r], v, scale, r);
}

template <class T>
__global__ void k_table(long long n_rows, const long long* __restrict__ target, const int* __restrict__ row_ptr, T* F,
                        const float* __restrict__ tab, int m, long long cmax, const long long* __restrict__ base,
                        const long long* __restrict__ i0, const float* __restrict__ u0, const float* __restrict__ coef,
                        const long long* counter, int slot, const float* scale) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  table_row(r, target, row_ptr, F, tab, m, cmax, base, i0, u0, coef, counter, slot, scale);
}

__global__ void k_table_view(int n, float* out, const float* __restrict__ tab, int m, long long cmax,
                             const long long* __restrict__ base, const long long* __restrict__ i0,
                             const float* __restrict__ u0, const long long* counter, int slot) {
  const int r = threadIdx.x;
  if (r >= n) return;
  out[r] = quintic(tab, counter[slot], i0[r], u0[r], base[r], m, cmax);
}

// Auxiliary 1D grid of a TFSF plane wave (tfsf.py:200-211). One block strides over the line.
__global__ void k_line_h(int M, const float* __restrict__ E1, float* H1, const float* __restrict__ da,
                         const float* __restrict__ db) {
  for (int m = threadIdx.x; m < M - 1; m += blockDim.x) H1[m] = da[m] * H1[m] - db[m] * (E1[m + 1] - E1[m]);
}

__global__ void k_line_e(int M, float* E1, const float* __restrict__ H1, const float* __restrict__ ca,
                         const float* __restrict__ cb, int m_src, const float* __restrict__ wave, long long wave_len,
                         const long long* counter, int slot) {
  const long long t = counter[slot];
  const float w = (t >= 0 && t < wave_len) ? wave[t] : 0.f;
  for (int m = 1 + threadIdx.x; m < M - 1; m += blockDim.x) {
    float e = ca[m] * E1[m] - cb[m] * (H1[m] - H1[m - 1]);
    if (m == m_src) e += w;
    E1[m] = e;
  }
}

// Mixed (interface) ADE group, engine.py:743-791. State per node: D, E_prev, E_a, E_h[S], then per slot s and
// pole p the pairs (P, P_prev) of the harmonic path and of the arithmetic path.
template <class T>
__global__ void k_mixed(long long n, int S, int maxp, const long long* __restrict__ idx, const float* __restrict__ f,
                        const float* __restrict__ w, const float* __restrict__ h0, const float* __restrict__ inv_a,
                        const float* __restrict__ sm, T* st, int has_h, int has_a, T* E) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int ms = 2 + 3 * maxp;
  T* D = st;
  T* Ep = st + n;
  T* Ea = st + 2 * n;
  T* Eh = st + 3 * n;                              // [S][n]
  T* Ph = st + (3 + S) * n;                        // [S][maxp][2][n]
  T* Pa = st + (3 + S + 2 * S * maxp) * n;         // [S][maxp][2][n]
  const T e_now = E[idx[r]];
  const T d = add(D[r], sub(e_now, Ep[r]));
  D[r] = d;
  T out = zero<T>();
  bool first = true;
  if (has_h) {
    T harm = scl(h0[r], d);
    for (int s = 0; s < S; ++s) {
      const float* mp = sm + s * ms;
      const int np = (int)mp[1];
      if (np == 0) continue;
      T tot = zero<T>();
      const T eh = Eh[s * n + r];
      for (int p = 0; p < np; ++p) {
        T* P0 = Ph + ((long long)(s * maxp + p) * 2) * n;
        T* P1 = P0 + n;
        const T pn = add(add(scl(mp[2 + 3 * p], P0[r]), scl(mp[3 + 3 * p], P1[r])), scl(mp[4 + 3 * p], eh));
        P1[r] = P0[r];
        P0[r] = pn;
        tot = p == 0 ? pn : add(tot, pn);
      }
      const T ehn = scl(mp[0], sub(d, tot));
      Eh[s * n + r] = ehn;
      harm = add(harm, scl(f[s * n + r], ehn));
    }
    out = scl(w[r], harm);
    first = false;
  }
  if (has_a) {
    T acc = d;
    const T ea = Ea[r];
    for (int s = 0; s < S; ++s) {
      const float* mp = sm + s * ms;
      const int np = (int)mp[1];
      if (np == 0) continue;
      T tot = zero<T>();
      for (int p = 0; p < np; ++p) {
        T* P0 = Pa + ((long long)(s * maxp + p) * 2) * n;
        T* P1 = P0 + n;
        const T pn = add(add(scl(mp[2 + 3 * p], P0[r]), scl(mp[3 + 3 * p], P1[r])), scl(mp[4 + 3 * p], ea));
        P1[r] = P0[r];
        P0[r] = pn;
        tot = p == 0 ? pn : add(tot, pn);
      }
      acc = sub(acc, scl(f[s * n + r], tot));
    }
    const T ean = scl(inv_a[r], acc);
    Ea[r] = ean;
    const T part = scl(1.f - w[r], ean);
    out = first ? part : add(out, part);
  }
  E[idx[r]] = out;
  Ep[r] = out;
}

// Conformal vertex tensor (engine.py:839-923). Filters of a slot r (of n slots) for a medium table sm[S][ms]:
// harmonic y_h = h0 u_h + sum_s f_s E_s, E_s = (u_h - sum_p P_s,p) / eps_inf_s with each medium's poles driven by
// its own previous E_s (engine.py:796-816); arithmetic (when `arith`) y_a = (u_a - sum_s f_s sum_p P_s,p) inv_a
// with all poles driven by the previous y_a (engine.py:819-836). A medium's pole pairs of both filters (four
// poles at a time) are loaded before any of them is stored, and their addresses do not wait for f_s, so the
// slot's state costs one memory round trip per dispersive medium; each filter's sums keep the reference's order.
template <class T>
__device__ __forceinline__ void tensor_filters(T u_h, T u_a, bool arith, long long r, long long n, int S, int maxp,
                                               const float* __restrict__ sm, int ms, const float* __restrict__ f,
                                               float h0, float inv_a, T* __restrict__ Eh, T* __restrict__ Ph,
                                               T* __restrict__ Ea, T* __restrict__ Pa, T& y_h, T& y_a) {
  T yh = scl(h0, u_h);
  T acc = u_a;
  const T ea = arith ? Ea[r] : zero<T>();
  for (int s = 0; s < S; ++s) {
    const float* mp = sm + s * ms;
    const int np = (int)mp[1];
    if (np == 0) continue;
    const float fs = f[s * n + r];
    const T eh = Eh[s * n + r];
    T th = zero<T>(), ta = zero<T>();
    for (int p0 = 0; p0 < np; p0 += 4) {
      T hp[4], hq[4], ap[4], aq[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long o = ((long long)(s * maxp + p0 + j) * 2) * n + r;
        if (p0 + j < np) {
          hp[j] = Ph[o];
          hq[j] = Ph[o + n];
          if (arith) {
            ap[j] = Pa[o];
            aq[j] = Pa[o + n];
          }
        }
      }
      if (fs == 0.f) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = p0 + j;
        if (p >= np) break;
        const float c1 = mp[2 + 3 * p], c2 = mp[3 + 3 * p], c3 = mp[4 + 3 * p];
        const long long o = ((long long)(s * maxp + p) * 2) * n + r;
        const T pn = add(add(scl(c1, hp[j]), scl(c2, hq[j])), scl(c3, eh));
        Ph[o + n] = hp[j];
        Ph[o] = pn;
        th = p == 0 ? pn : add(th, pn);
        if (arith) {
          const T qn = add(add(scl(c1, ap[j]), scl(c2, aq[j])), scl(c3, ea));
          Pa[o + n] = ap[j];
          Pa[o] = qn;
          ta = p == 0 ? qn : add(ta, qn);
        }
      }
    }
    if (fs == 0.f) continue;
    const T ehn = scl(mp[0], sub(u_h, th));
    Eh[s * n + r] = ehn;
    yh = add(yh, scl(fs, ehn));
    if (arith) acc = sub(acc, scl(fs, ta));
  }
  y_h = yh;
  if (arith) {
    y_a = scl(inv_a, acc);
    Ea[r] = y_a;
  }
}

template <class T>
struct TensorJob {
  long long nN;             // nodes; slots 2 nN (slot k nN + node: k = 0 the vertex at the node index, 1 at + e_c)
  const long long* node;    // native E index of each node
  const int* comp;          // component of each node
  const int* vn;            // [nV][3][2] node numbers of each vertex's (minus, plus) d-nodes
  const T* vph;             // [nV][3][2] Bloch phase of those node positions, or null
  const float* normal;      // [nV][3]
  const int* slot_v;        // [2 nN] tensor vertex of each slot, -1 for a plain slot
  int S, maxp, ms;
  const float* f;           // [S][2 nN] medium fractions (one-hot for plain slots)
  const float* h0;          // [2 nN]
  const float* inv_a;       // [2 nN]
  const float* sm;          // [S][ms] {1/eps_inf, n_poles, (c1, c2, c3) per pole}
  T* D;                     // [nN]
  T* Ep;                    // [nN]
  T* out;                   // [2 nN]
  T* Ea;                    // [2 nN]
  T* Eh;                    // [S][2 nN]
  T* Ph;                    // [S][maxp][2][2 nN]
  T* Pa;                    // [S][maxp][2][2 nN]
  T* E;
};

// D += E_now - E_prev per node
template <class T>
__global__ void k_tensor_d(TensorJob<T> J) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= J.nN) return;
  J.D[q] = add(J.D[q], sub(J.E[J.node[q]], J.Ep[q]));
}

// The E-side table correction and the tensor nodes' D update in one launch: blocks [0, table_blocks) take the
// table rows, the rest the tensor nodes. Only valid when no table target is a tensor node (the host checks), so
// that the D update reads the same E as after a separate correction launch.
template <class T>
__global__ void k_table_tensor_d(long long n_rows, int table_blocks, const long long* __restrict__ target,
                                 const int* __restrict__ row_ptr, T* F, const float* __restrict__ tab, int m,
                                 long long cmax, const long long* __restrict__ base, const long long* __restrict__ i0,
                                 const float* __restrict__ u0, const float* __restrict__ coef,
                                 const long long* counter, int slot, const float* scale, TensorJob<T> J) {
  if ((int)blockIdx.x < table_blocks) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_rows) table_row(r, target, row_ptr, F, tab, m, cmax, base, i0, u0, coef, counter, slot, scale);
    return;
  }
  const long long q = (long long)(blockIdx.x - table_blocks) * blockDim.x + threadIdx.x;
  if (q >= J.nN) return;
  J.D[q] = add(J.D[q], sub(J.E[J.node[q]], J.Ep[q]));
}

// One slot per thread, a node's two slots on a lane pair (lane 2m: slot k = 0 at the node's index, lane 2m + 1:
// k = 1 at + e_c): the tensor vertex's share, or half the plain medium's; the even lane then sets E_p = the sum of
// the node's two slots (slot 0 + slot 1, the order of the reference's sum).
template <class T>
__device__ __forceinline__ T tensor_slot(const TensorJob<T>& J, long long q, int k, long long node) {
  const long long n = 2 * J.nN;
  const int v = J.slot_v[q];
  const float h0 = J.h0[q];
  const float inv_a = J.inv_a[q];
  T u_h, u_a = zero<T>();
  float nc = 0.f;
  int side = 0, c = 0;
  if (v >= 0) {
    T dv[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      dv[t] = J.D[J.vn[(long long)v * 6 + t]];
      if (J.vph) dv[t] = mul(dv[t], J.vph[(long long)v * 6 + t]);
    }
    T dbar[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) dbar[d] = scl(0.5f, add(dv[2 * d], dv[2 * d + 1]));
    const float* nv = J.normal + (long long)v * 3;
    const T sv = add(add(scl(nv[0], dbar[0]), scl(nv[1], dbar[1])), scl(nv[2], dbar[2]));
    c = J.comp[node];
    side = 1 - k;
    T dp = dv[0], db = dbar[0];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d == c) {
        dp = side ? dv[2 * d + 1] : dv[2 * d];
        db = dbar[d];
      }
    nc = nv[c];
    u_h = add(sv, scl(nc, sub(dp, db)));  // sigma
    u_a = sub(dp, scl(nc, u_h));
  } else {
    u_h = J.D[node];
  }
  T a, b = zero<T>();
  tensor_filters(u_h, u_a, v >= 0, q, n, J.S, J.maxp, J.sm, J.ms, J.f, h0, inv_a, J.Eh, J.Ph, J.Ea, J.Pa, a, b);
  if (v < 0) return scl(0.5f, a);
  T contrib = scl(0.5f, add(b, scl(nc, a)));
  if (J.vph) contrib = mul(contrib, conjg_t(J.vph[(long long)v * 6 + c * 2 + side]));
  return contrib;
}

template <class T>
__global__ void k_tensor_slot_e(TensorJob<T> J) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long node = t >> 1;
  const int k = (int)(t & 1);
  const bool in = node < J.nN;  // lanes past the end still take part in the pair exchange
  T contrib = zero<T>();
  if (in) {
    const long long q = k * J.nN + node;
    contrib = tensor_slot(J, q, k, node);
    J.out[q] = contrib;
  }
  const T other = shfl_xor1(contrib);
  if (in && k == 0) {
    const T e = add(contrib, other);
    J.E[J.node[node]] = e;
    J.Ep[node] = e;
  }
}

// Running DFT monitor (engine.py:564-592). Block size 256; phases for all frequencies in LDS, computed in
// double from the step counter and cast to complex64 as the reference does.
template <class S>
struct DftJob {
  int nf;
  long long np;
  const double* omega;
  double dt;
  const long long* counter;
  const S* esrc;
  const S* hsrc;
  const long long* eidx;
  const long long* h1;
  const long long* h2;
  c32* acc_e;
  c32* acc_h;
  float* partial;  // per-block energy partials of this monitor, or null
};

template <class S>
__device__ __forceinline__ void dft_block(const DftJob<S>& J, int blk, c32* ph, float* red) {
  const double te = (double)J.counter[0] * J.dt;
  const double th = te - 0.5 * J.dt;
  const int nf = J.nf;
  const long long np = J.np;
  for (int q = threadIdx.x; q < nf; q += blockDim.x) {
    const double we = J.omega[q] * te, wh = J.omega[q] * th;
    ph[q] = mk((float)cos(we), (float)sin(we));
    ph[nf + q] = mk((float)cos(wh), (float)sin(wh));
  }
  __syncthreads();
  const long long p = (long long)blk * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (p < np) {
    const S se = J.esrc[J.eidx[p]];
    e2 += norm2(se);
    for (int q = 0; q < nf; ++q) {
      const c32 v = mul(ph[q], to_c(se));
      c32 a = J.acc_e[q * np + p];
      J.acc_e[q * np + p] = add(a, v);
    }
    if (J.acc_h) {
      const S sh = scl(0.5f, add(J.hsrc[J.h1[p]], J.hsrc[J.h2[p]]));
      e2 += norm2(sh);
      for (int q = 0; q < nf; ++q) {
        const c32 v = mul(ph[nf + q], to_c(sh));
        c32 a = J.acc_h[q * np + p];
        J.acc_h[q * np + p] = add(a, v);
      }
    }
  }
  if (J.partial) {  // the block's sum of |E|^2 (+ |H|^2): a butterfly within each wavefront, then the wavefronts'
                    // sums in order (a fixed order: the same partials on every run)
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) e2 += __shfl_xor(e2, s, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = e2;
    __syncthreads();
    if (threadIdx.x == 0) {
      float s = 0.f;
      for (int q = 0; q < nw; ++q) s += red[q];
      J.partial[blk] = s;
    }
  }
}

template <class S>
__global__ void k_dft(DftJob<S> J) {
  extern __shared__ c32 ph[];  // [2 * nf]
  __shared__ float red[256];
  dft_block(J, blockIdx.x, ph, red);
}

template <class S>
__global__ void k_dft_multi(const DftJob<S>* __restrict__ jobs, Sched sc) {
  extern __shared__ c32 ph[];  // [2 * max nf]
  __shared__ float red[256];
  int m, lt;
  member_of(sc, blockIdx.x, m, lt);
  dft_block(jobs[m], lt, ph, red);
}

// Sum each monitor's block partials in a fixed order and set the energy slots (engine.py:580-589).
// mon_off[m] .. mon_off[m+1] are monitor m's partials, mon_group[m] its slot (0, 1), -1 for none.
struct EnergyJob {
  int n_mon;
  const int* mon_off;
  const int* mon_group;
  const float* partial;
  float* energy;
};

__device__ __forceinline__ void energy_sum(const EnergyJob& J) {
  if (threadIdx.x != 0) return;
  double tot[2] = {0.0, 0.0};
  int have[2] = {0, 0};
  for (int m = 0; m < J.n_mon; ++m) {
    const int g = J.mon_group[m];
    if (g < 0) continue;
    double s = 0.0;
    for (int q = J.mon_off[m]; q < J.mon_off[m + 1]; ++q) s += J.partial[q];
    tot[g] += s;
    have[g] = 1;
  }
  for (int g = 0; g < 2; ++g)
    if (have[g]) J.energy[g] = (float)tot[g];
}

__global__ void k_energy(EnergyJob J) { energy_sum(J); }

// one block per engine
__global__ void k_energy_multi(const EnergyJob* __restrict__ jobs) { energy_sum(jobs[blockIdx.x]); }

template <class T>
__global__ void k_timemon(long long np, const long long* __restrict__ eidx, const T* __restrict__ E, T* buf,
                          long long n_max, const long long* pos) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long s = pos[0];
  if (p >= np || s < 0 || s >= n_max) return;
  buf[s * np + p] = E[eidx[p]];
}

__global__ void k_bump(long long* pos) {
  if (threadIdx.x == 0 && blockIdx.x == 0) pos[0] += 1;
}

// ---------------------------------------------------------------------------------------------
// adjoint operations (the bulk adjoint kernels: lsf_adjoint.h)
// ---------------------------------------------------------------------------------------------

// Forward tape: X = curl- H at design E nodes (the factor multiplying cb in the E update), one row per step.
// slot = counter[0] - base[0] (the forward step n of the update that follows); rows outside [0, n_slots) are
// not recorded.
struct TapeJob {
  long long n;           // design nodes
  const long long* node; // native flat E indices
  const float* H;
  int n0, n1;
  long long N;
  float s[3];
  float* tape;           // [n_slots][n]
  long long n_slots;
  const long long* base;
  const long long* counter;
};

__global__ void k_tape(TapeJob J) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= J.n) return;
  const long long slot = J.counter[0] - J.base[0];
  if (slot < 0 || slot >= J.n_slots) return;
  const long long nd = J.node[q];
  const int c = (int)(nd / J.N);
  const long long cell = nd - (long long)c * J.N;
  const long long st[3] = {1, J.n0, (long long)J.n0 * J.n1};
  const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
  const float* Hd2 = J.H + (long long)d2 * J.N;
  const float* Hd1 = J.H + (long long)d1 * J.N;
  const float da = Hd2[cell] - Hd2[cell - st[d1]];
  const float db = Hd1[cell] - Hd1[cell - st[d2]];
  J.tape[slot * J.n + q] = J.s[d1] * da - J.s[d2] * db;
}

// Gradient accumulation of an adjoint step: g[q] += (e~ / cb)[node q] * tape[n][q], n = n_steps - 1 -
// counter[0] the forward step being reversed (tape row n - base).
struct GradJob {
  long long n;
  const long long* node;
  const float* E;       // e~
  const float* cb;      // [n] cb of each design node
  const float* tape;
  long long n_slots;
  const long long* base;
  const long long* counter;
  long long n_steps;
  double* g;
};

__global__ void k_adj_grad(GradJob J) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= J.n) return;
  const long long n = J.n_steps - 1 - J.counter[0];
  const long long slot = n - J.base[0];
  if (n < 0 || slot < 0 || slot >= J.n_slots) return;
  J.g[q] += (double)(J.E[J.node[q]] / J.cb[q]) * (double)J.tape[slot * J.n + q];
}

// Adjoint of the D-form (ADE) update of pure-group nodes (engine.py:759-793 transposed). Forward, per node with
// eps_inf and poles (c1, c2, c3): P'_p = c1 P_p + c2 P_p^prev + c3 E, E' = E + (X - sum_p (P'_p - P_p)) / eps_inf,
// X = cb (curl- H + CPML terms). With the adjoint pole pair (a_p, b_p) = (dL/dP'_p, dL/dP_p^prev') and
// t_p = a_p - E'bar / eps_inf, the transposed step is
//   Ebar = E'bar + sum_p c3 t_p,  a_p <- E'bar / eps_inf + b_p + c1 t_p,  b_p <- c2 t_p,  Xbar = E'bar / eps_inf.
// In the adjoint engine's variable e~ = cb Ebar (cb constant, dt/dx, at D-form nodes) every term scales alike.
// pre (before the adjoint H kernel): keep Ebar's new value aside and leave cb Xbar = e~ / eps_inf in the field,
// which is what the H kernel's transposed curl (and CPML) must read there; post (after it): restore Ebar.
struct AdjAdeJob {
  long long n;
  const long long* node;   // native flat E index, sorted
  const int* medium;       // per node
  const float* media;      // per medium: {1/eps_inf, poles, (c1, c2, c3) x max_poles}
  int stride;              // 2 + 3 max_poles
  int max_poles;
  float* state;            // [n][max_poles][2]: (a, b) per pole
  float* save;             // [n]
  float* E;
};

__global__ void k_adj_ade_pre(AdjAdeJob J) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= J.n) return;
  const float* m = J.media + (long long)J.medium[q] * J.stride;
  const float inv = m[0];
  const int np = (int)m[1];
  const long long nd = J.node[q];
  const float e1 = J.E[nd];
  const float x = e1 * inv;
  float acc = 0.f;
  float* st = J.state + q * 2 * J.max_poles;
  for (int p = 0; p < np; ++p) {
    const float c1 = m[2 + 3 * p], c2 = m[3 + 3 * p], c3 = m[4 + 3 * p];
    const float t = st[2 * p] - x;
    acc = fmaf(c3, t, acc);
    st[2 * p] = fmaf(c1, t, x + st[2 * p + 1]);
    st[2 * p + 1] = c2 * t;
  }
  J.save[q] = e1 + acc;
  J.E[nd] = x;
}

__global__ void k_adj_ade_post(AdjAdeJob J) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= J.n) return;
  J.E[J.node[q]] = J.save[q];
}

// Adjoint DFT source: F[node] += sum over the points p that touch the node of scale * sum_f Re(W[f, p] e^{i omega_f t})
// at sampling steps, where W = conj(dL/dacc) in torch's gradient convention for the monitor accumulator acc[f, p].
// E side (h_side = 0): after an adjoint E half-step the state is E^n, n = n_steps - 1 - counter[1]; sampled when
// n > 0 and n % stride == 0, t = n dt. H side (h_side = 1): after an adjoint H half-step the state is H^{n-1/2} with
// n = n_steps - counter[0], sampled at n with t = n dt - dt/2; each H point is the average of two nodes.
// Points that share a node (edges and corners of a flux box) meet in one CSR row built on the host; a thread sums
// its row in a fixed order, so the result does not depend on scheduling (no atomics).
struct AdjSrcJob {
  int h_side;
  int nf;
  long long np;          // monitor points (columns of W)
  const double* omega;
  double dt;
  const c32* W;          // [nf][np]
  long long n_rows;      // distinct target nodes
  const int* row_ptr;    // [n_rows + 1] into point / scale
  const long long* target;  // [n_rows] native flat node
  const long long* point;   // contributing monitor point per entry
  const float* scale;    // per entry: cb at the E node, or -cH / 2 for an H node
  float* F;
  const long long* counter;
  long long n_steps;
  int stride;
};

__global__ void k_adj_src(AdjSrcJob J) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= J.n_rows) return;
  const long long n = J.h_side ? J.n_steps - J.counter[0] : J.n_steps - 1 - J.counter[1];
  if (n <= 0 || n % J.stride != 0) return;
  const double t = (double)n * J.dt - (J.h_side ? 0.5 * J.dt : 0.0);
  double acc = 0.0;
  for (int q = J.row_ptr[r]; q < J.row_ptr[r + 1]; ++q) {
    const long long p = J.point[q];
    double v = 0.0;
    for (int f = 0; f < J.nf; ++f) {
      const double ph = J.omega[f] * t;
      const c32 w = J.W[(long long)f * J.np + p];
      v += (double)w.x * cos(ph) - (double)w.y * sin(ph);
    }
    acc += v * (double)J.scale[q];
  }
  J.F[J.target[r]] += (float)acc;
}

}  // namespace lsf
