// Shared device helpers for the multislice kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft.hpp"
#include "forward.hpp"

namespace pty {

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
struct RowBufWave {
  float2* row;
  __device__ __forceinline__ float2 ld(int e) const { return row[pad16(e)]; }
  __device__ __forceinline__ void st(int e, float2 v) const { row[pad16(e)] = v; }
  __device__ __forceinline__ void sync() const { wave_sync(); }
};

}  // namespace pty
