// Line-pass building blocks shared by the alternating-axis streaming kernels (adjoint_stream.hip, jvp.hip):
// the column exchange buffer, the geometry of a thread's line in an n x n wave, and the transmission window
// along that line.
#pragma once
#include "common.hpp"

namespace pty {

template <int NC>
struct ColBufBlk {
  float2* base;
  __device__ __forceinline__ float2 ld(int e) const { return base[pad16(e) * NC]; }
  __device__ __forceinline__ void st(int e, float2 v) const { base[pad16(e) * NC] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return base[pe * NC]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { base[pe * NC] = v; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};

// Line geometry of a pass: rows (COL = false; line li is row y, point q at x = t + T q) or columns (line li is
// column x, point q at y = t + T q).  Point q of the line sits lo + q * QS bytes into an n x n array.
template <int N, bool COL>
struct LineGeo {
  static constexpr int T = N / 16;
  static constexpr unsigned QS = COL ? (unsigned)(T * N * sizeof(float2)) : (unsigned)(T * sizeof(float2));
  int t, li;
  unsigned lo;
  __device__ __forceinline__ LineGeo(int t_, int li_)
      : t(t_), li(li_), lo(COL ? (unsigned)((t_ * N + li_) * sizeof(float2)) : (unsigned)((li_ * N + t_) * sizeof(float2))) {}
  __device__ __forceinline__ int ky(int q) const { return COL ? t + T * q : li; }
  __device__ __forceinline__ int kx(int q) const { return COL ? li : t + T * q; }
  __device__ __forceinline__ void load(float2 (&v)[16], const float2* arr) const {
    const BufView b(arr, (unsigned long long)N * N * sizeof(float2));
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = b.ld2(lo, QS * q);
  }
  __device__ __forceinline__ void store(const float2 (&v)[16], float2* arr) const {
    const BufView b(arr, (unsigned long long)N * N * sizeof(float2));
#pragma unroll
    for (int q = 0; q < 16; ++q) b.st2(v[q], lo, QS * q);
  }
  // slab offset (within a slice) of point q of the window at (oy, ox), wrapped periodically
  __device__ __forceinline__ long long slab_offset(const WaveCtx& g, int q) const {
    int gy = g.oy + ky(q), gx = g.ox + kx(q);
    if (gy >= g.slab_h) gy -= (int)g.slab_h;
    if (gx >= g.slab_w) gx -= (int)g.slab_w;
    return (long long)gy * g.slab_w + gx;
  }
};

// Points Q0 .. Q0 + NQ - 1 of the window of slice s of a slab laid out like the transmission (the transmission
// itself, or its tangent) along the thread's line; tv[i] holds point Q0 + i.
template <int N, bool COL, int Q0, int NQ>
__device__ __forceinline__ void load_window_points(float2 (&tv)[NQ], const float2* slab, const WaveCtx& g, int s,
                                                   const LineGeo<N, COL>& L) {
  constexpr int T = N / 16;
  const float2* slice = slab + (long long)s * g.stride;
  if constexpr (!COL) {
    int gy = g.oy + L.li;
    if (gy >= g.slab_h) gy -= (int)g.slab_h;
    const float2* row = slice + (long long)gy * g.slab_w;
    if (g.ox + N <= g.slab_w) {
      const float2* r = row + g.ox + L.t;
#pragma unroll
      for (int i = 0; i < NQ; ++i) tv[i] = r[T * (Q0 + i)];
    } else {
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        int x = g.ox + L.t + T * (Q0 + i);
        if (x >= g.slab_w) x -= (int)g.slab_w;
        tv[i] = row[x];
      }
    }
  } else {
    int gx = g.ox + L.li;
    if (gx >= g.slab_w) gx -= (int)g.slab_w;
    const long long w = g.slab_w;
    const unsigned long long slice_bytes = (unsigned long long)g.stride * sizeof(float2);
    if (g.oy + N <= g.slab_h && slice_bytes <= 0xFFFFFFFFull) {
      const BufView tb(slice, slice_bytes);
      const unsigned lo = (unsigned)(((long long)(g.oy + L.t) * w + gx) * sizeof(float2));
      const unsigned qs = (unsigned)(T * w * sizeof(float2));
#pragma unroll
      for (int i = 0; i < NQ; ++i) tv[i] = tb.ld2(lo, qs * (Q0 + i));
    } else {
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        int gy = g.oy + L.t + T * (Q0 + i);
        if (gy >= g.slab_h) gy -= (int)g.slab_h;
        tv[i] = slice[(long long)gy * w + gx];
      }
    }
  }
}

// All sixteen points of the thread's window line.
template <int N, bool COL>
__device__ __forceinline__ void load_window(float2 (&tv)[16], const float2* slab, const WaveCtx& g, int s,
                                            const LineGeo<N, COL>& L) {
  load_window_points<N, COL, 0, 16>(tv, slab, g, s, L);
}

}  // namespace pty
