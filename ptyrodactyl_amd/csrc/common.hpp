// Shared device helpers for the multislice kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft.hpp"

namespace pty {

// Padded index inside a line: one spare slot every 16 points keeps the
// Stockham exchange free of LDS bank conflicts.
__device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }

// Per-wave geometry resolved from the caller's origin/shift arrays.
struct WaveGeom {
  int oy, ox;   // window origin inside the slab, reduced to [0, H) x [0, W)
  int my, mx;   // integer part of the probe shift, reduced modulo n
  float dy, dx; // fractional part of the probe shift, in [-0.5, 0.5]
};

__device__ __forceinline__ int posmod(long long a, long long n) {
  long long r = a % n;
  return (int)(r < 0 ? r + n : r);
}

__device__ __forceinline__ void split_shift(double s, int n, int& m, float& d) {
  if (!(fabs(s) < 1e15)) s = 0.0;  // non-finite or absurd shifts are rejected upstream
  double r = floor(s + 0.5);
  d = (float)(s - r);
  m = posmod((long long)fmod(r, (double)n), n);
}

__device__ __forceinline__ WaveGeom load_geom(const int* origins, const double* shifts, int p, int n,
                                              long long slab_h, long long slab_w) {
  WaveGeom g;
  g.oy = posmod(origins[2 * p], slab_h);
  g.ox = posmod(origins[2 * p + 1], slab_w);
  split_shift(shifts[2 * p], n, g.my, g.dy);
  split_shift(shifts[2 * p + 1], n, g.mx, g.dx);
  return g;
}

// exp(-2 pi i k s / n) / n for the FFT bin `kidx` (fftfreq order, Nyquist
// negative) and the shift s = m + d pixels; the 1/n folds the inverse-FFT
// normalisation of the shifted probe into the ramp.
__device__ __forceinline__ float2 shift_ramp(int kidx, int n, int m, float d) {
  const int k = kidx < n / 2 ? kidx : kidx - n;
  const int km = posmod((long long)k * m, n);
  const float t = (float)km / (float)n + (float)k * d / (float)n;
  float s, c;
  sincospif(-2.0f * t, &s, &c);
  const float inv = 1.0f / (float)n;
  return make_float2(c * inv, s * inv);
}

// Row buffer: a padded line in LDS, owned by a group inside one wavefront.
struct RowBufWave {
  float2* row;
  __device__ __forceinline__ float2 ld(int e) const { return row[pad16(e)]; }
  __device__ __forceinline__ void st(int e, float2 v) const { row[pad16(e)] = v; }
  __device__ __forceinline__ void sync() const { wave_sync(); }
};

}  // namespace pty
