// Shared device helpers for the multislice kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft.hpp"
#include "forward.hpp"

namespace pty {

// Buffer view of `bytes` (below 4 GiB) of global memory at a wave-uniform base.  An access takes the lane's byte
// offset in a VGPR and a wave-uniform byte offset in an SGPR, so the 16 strided accesses a thread makes to one
// column (rows t + (n / 16) q) cost no vector address arithmetic; a plain pointer needs a 64-bit add per access.
// Word 3 of the resource is the gfx9 raw-buffer format (32-bit data, no swizzle).
struct BufView {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ BufView(const void* base, unsigned long long bytes)
      : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                             (int)(unsigned)(bytes < 0xFFFFFFFFull ? bytes : 0xFFFFFFFFull),
                                             0x00020000)) {}
  __device__ __forceinline__ float2 ld2(unsigned voff, unsigned soff) const {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
  }
  __device__ __forceinline__ void st2(float2 v, unsigned voff, unsigned soff) const {
    typedef __attribute__((__vector_size__(2 * sizeof(unsigned)))) unsigned u2;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, voff, soff, 0);
  }
  __device__ __forceinline__ void st1(float v, unsigned voff, unsigned soff) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, voff, soff, 0);
  }
};

// Padded index inside a line: one spare slot every 16 points keeps the
// Stockham exchange free of LDS bank conflicts.
__device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }

// exp(-2 pi i k s / n) / n for the FFT bin `kidx` (fftfreq order, Nyquist
// negative) and the shift s = m + d pixels; the 1/n folds the inverse-FFT
// normalisation of the shifted probe into the ramp.
__device__ __forceinline__ float2 shift_ramp(int kidx, int n, int m, float d) {
  const int k = kidx < n / 2 ? kidx : kidx - n;
  long long km = ((long long)k * m) % n;
  if (km < 0) km += n;
  const float t = (float)km / (float)n + (float)k * d / (float)n;
  float s, c;
  sincospif(-2.0f * t, &s, &c);
  const float inv = 1.0f / (float)n;
  return make_float2(c * inv, s * inv);
}

// Row buffer: a padded line in LDS, owned by a group inside one wavefront.
// (unit, column block) of a column-pass workgroup.  Workgroups are dispatched
// round-robin over the 8 XCDs.  With at most 2 columns per workgroup (n >= 2048)
// the grid is remapped so that each XCD takes a contiguous run of column blocks:
// workgroups that share the 128-byte lines of a row segment then meet in one
// L2.  Otherwise consecutive blocks walk the waves first.  Both are bijections of
// [0, nb * blocks_per_unit) when nb * blocks_per_unit is a multiple of 8.
template <int NC>
__device__ __forceinline__ void column_block(int b, int nb, int blocks_per_unit, int& unit, int& cblk) {
  if constexpr (NC <= 2) {
    const int total = nb * blocks_per_unit;
    const int lin = (b % 8) * (total / 8) + b / 8;
    unit = lin / blocks_per_unit;
    cblk = lin % blocks_per_unit;
  } else {
    unit = b % nb;
    cblk = b / nb;
  }
}

// v[q] *= row[(ox + t + (N / 16) q) mod w] for the 16 points a thread holds of one window row: a window
// row that does not wrap (the common case, uniform per workgroup) reads with constant offsets from one
// base, the wrapped case computes each index.
template <int N>
__device__ __forceinline__ void multiply_window_row(float2 (&v)[16], const float2* row, int ox, int t, int w) {
  constexpr int T = N / 16;
  if (ox + N <= w) {
    const float2* r = row + ox + t;
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], r[T * q]);
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      int x = ox + t + T * q;
      if (x >= w) x -= w;
      v[q] = cmul(v[q], row[x]);
    }
  }
}

struct RowBufWave {
  float2* row;
  __device__ __forceinline__ float2 ld(int e) const { return row[pad16(e)]; }
  __device__ __forceinline__ void st(int e, float2 v) const { row[pad16(e)] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return row[pe]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { row[pe] = v; }
  __device__ __forceinline__ void sync() const { wave_sync(); }
};

// The same row buffer for lines whose N / 16 threads span several wavefronts
// (N >= 2048): every exchange needs a workgroup barrier.
struct RowBufBlock {
  float2* row;
  __device__ __forceinline__ float2 ld(int e) const { return row[pad16(e)]; }
  __device__ __forceinline__ void st(int e, float2 v) const { row[pad16(e)] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return row[pe]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { row[pe] = v; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};

template <bool ONE_WAVE>
struct RowBufFor {
  using type = RowBufWave;
};
template <>
struct RowBufFor<false> {
  using type = RowBufBlock;
};

// Deterministic FP64 sum over a workgroup of NT threads (NT a multiple of 64): a fixed xor
// butterfly inside each wavefront, then thread 0 adds the wavefront sums in order.  `scratch`
// holds NT / 64 doubles of LDS; every thread of the group must call it.  The total is returned
// on thread 0.
template <int NT>
__device__ __forceinline__ double block_sum_fp64(double v, double* scratch) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int tid = threadIdx.x;
  __syncthreads();  // scratch may still be read from a previous call
  if ((tid & 63) == 0) scratch[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0)
    for (int w = 0; w < NT / 64; ++w) s += scratch[w];
  return s;
}

}  // namespace pty
