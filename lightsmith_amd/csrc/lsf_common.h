// Shared device helpers for the FDTD kernels: complex arithmetic on {re, im} pairs, wrapped
// (periodic / Bloch) neighbour loads, XCD-aware block remapping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsf {

struct c32 {
  float x, y;
};

__host__ __device__ __forceinline__ c32 mk(float a, float b) { return c32{a, b}; }

// Arithmetic on the field type T (float or c32). Coefficients are real unless named otherwise.
__device__ __forceinline__ float add(float a, float b) { return a + b; }
__device__ __forceinline__ c32 add(c32 a, c32 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float sub(float a, float b) { return a - b; }
__device__ __forceinline__ c32 sub(c32 a, c32 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float scl(float s, float a) { return s * a; }
__device__ __forceinline__ c32 scl(float s, c32 a) { return mk(s * a.x, s * a.y); }
// s a + b with one rounding. The library compiles with floating-point contraction off and fuses only here, so
// that a kernel rounds the same way in every instantiation (tile shape, CPML class, one tile per block or
// persistent): a batch member then matches its run alone bit for bit.
__device__ __forceinline__ float fmad(float s, float a, float b) { return __builtin_fmaf(s, a, b); }
__device__ __forceinline__ c32 fmad(float s, c32 a, c32 b) {
  return mk(__builtin_fmaf(s, a.x, b.x), __builtin_fmaf(s, a.y, b.y));
}
// a * b with a, b of the field type (complex product for c32)
__device__ __forceinline__ float mul(float a, float b) { return a * b; }
__device__ __forceinline__ c32 mul(c32 a, c32 b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// rotate by a phase (a no-op for real fields, whose periodic axes carry no Bloch phase)
__device__ __forceinline__ float rot(float a, c32) { return a; }
__device__ __forceinline__ c32 rot(c32 a, c32 p) { return mul(a, p); }
__device__ __forceinline__ c32 conjg(c32 a) { return mk(a.x, -a.y); }
// complex conjugate on the field type (identity for real values)
__device__ __forceinline__ float conjg_t(float a) { return a; }
__device__ __forceinline__ c32 conjg_t(c32 a) { return conjg(a); }
__device__ __forceinline__ c32 to_c(float a) { return mk(a, 0.f); }
__device__ __forceinline__ c32 to_c(c32 a) { return a; }
__device__ __forceinline__ float norm2(float a) { return a * a; }
__device__ __forceinline__ float norm2(c32 a) { return a.x * a.x + a.y * a.y; }
template <class T> __device__ __forceinline__ T zero();
template <> __device__ __forceinline__ float zero<float>() { return 0.f; }
template <> __device__ __forceinline__ c32 zero<c32>() { return mk(0.f, 0.f); }

__device__ __forceinline__ float shfl_dn(float v) { return __shfl_down(v, 1, 64); }
__device__ __forceinline__ c32 shfl_dn(c32 v) { return mk(__shfl_down(v.x, 1, 64), __shfl_down(v.y, 1, 64)); }
__device__ __forceinline__ float shfl_up(float v) { return __shfl_up(v, 1, 64); }
__device__ __forceinline__ float shfl_xor1(float v) { return __shfl_xor(v, 1, 64); }
__device__ __forceinline__ c32 shfl_xor1(c32 v) { return mk(__shfl_xor(v.x, 1, 64), __shfl_xor(v.y, 1, 64)); }
__device__ __forceinline__ c32 shfl_up(c32 v) { return mk(__shfl_up(v.x, 1, 64), __shfl_up(v.y, 1, 64)); }

// Grid geometry shared by every kernel of an engine.
struct Grid {
  int n[3];          // N0 (x), N1 (y), N2 (z)
  int per[3];        // periodic axes
  long long N;       // cells per component
  float s[3];        // per-axis difference scale: 1 on a cubic grid, 1/d otherwise (engine.py:274-276, 622-625)
  float cH;          // H coefficient: dt/dx (cubic) or dt (engine.py:300)
  float cdx;         // E coefficient of a D-form node (inv_eps = 1)
  c32 ph[3];         // exp(+i k_a L_a) per axis: the high ghost's phase; the low ghost takes its conjugate
  int rE[3][3][2];   // stepped box of E_c along axis a: [lo, hi) (engine.py:47-60)
  int rH[3][3][2];   // stepped box of H_c (engine.py:63-66)
};

__device__ __forceinline__ long long cidx(const Grid& g, int i, int j, int k) {
  return ((long long)k * g.n[1] + j) * g.n[0] + i;
}

// Load F(i, j, k) where an index may be the high ghost (n - 1) of a periodic axis: that ghost holds the
// periodic image at index 1 times exp(+i k L) (engine.py:595-611, low=False).
template <class T>
__device__ __forceinline__ T ld_hi(const T* __restrict__ F, const Grid& g, int i, int j, int k) {
  int w = 0;
  if (g.per[0] && i == g.n[0] - 1) { i = 1; w |= 1; }
  if (g.per[1] && j == g.n[1] - 1) { j = 1; w |= 2; }
  if (g.per[2] && k == g.n[2] - 1) { k = 1; w |= 4; }
  T v = F[cidx(g, i, j, k)];
  if (w & 1) v = rot(v, g.ph[0]);
  if (w & 2) v = rot(v, g.ph[1]);
  if (w & 4) v = rot(v, g.ph[2]);
  return v;
}

// Load F(i, j, k) where an index may be the low ghost (0) of a periodic axis: its image at n - 2 times
// exp(-i k L) (engine.py:595-611, low=True).
template <class T>
__device__ __forceinline__ T ld_lo(const T* __restrict__ F, const Grid& g, int i, int j, int k) {
  int w = 0;
  if (g.per[0] && i == 0) { i = g.n[0] - 2; w |= 1; }
  if (g.per[1] && j == 0) { j = g.n[1] - 2; w |= 2; }
  if (g.per[2] && k == 0) { k = g.n[2] - 2; w |= 4; }
  T v = F[cidx(g, i, j, k)];
  if (w & 1) v = rot(v, conjg(g.ph[0]));
  if (w & 2) v = rot(v, conjg(g.ph[1]));
  if (w & 4) v = rot(v, conjg(g.ph[2]));
  return v;
}

// Pointers into global memory. A pointer loaded from memory is generic (flat) to the compiler; held in a
// gptr and converted back to the generic type where it is used, its loads and stores compile to global
// instructions (a generic pointer captured by a lambda, or cast to global and straight back, loses that).
template <class T>
using gptr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> as_global(T* p) {
  return (gptr<T>)p;
}

// A copy of a wave-uniform constant in a vector register. The bulk kernels keep their arithmetic constants there:
// the scalar register file then holds the bases and loop state without spilling, and no constant is reloaded from
// the argument block inside the z march (each reload would wait on the counter it shares with the LDS).
__device__ __forceinline__ float vreg(float x) {
  float r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}
__device__ __forceinline__ int vregi(int x) {
  int r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Blocks are dealt round-robin to the 8 XCDs (each with its own L2). Give each XCD a contiguous range of
// logical tiles so that tiles sharing halo planes share an L2. The grid is padded to a multiple of 8. With `rev`
// each XCD walks its range from the end, so a launch that follows a forward one starts on the tiles whose data the
// previous launch touched last.
__device__ __forceinline__ int xcd_tile(int b, int nblocks_padded, int rev = 0) {
  const int per = nblocks_padded >> 3;
  const int k = b >> 3;
  return (b & 7) * per + (rev ? per - 1 - k : k);
}

}  // namespace lsf
