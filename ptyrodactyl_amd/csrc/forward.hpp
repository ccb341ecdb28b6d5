// Host-side declarations of the multislice forward launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace pty {

struct FwdParams {
  int n, S, M;
  long long P;
  const float2* slab;        // [S][slab_h][slab_w] complex64 transmission
  long long slab_h, slab_w;  // slab extent (each >= n)
  long long slice_stride;    // elements between slices (slab_h * slab_w)
  const float2* probe_k;     // [M][n][n] fft2 of each real-space probe mode
  const float* weights;      // [M] normalised incoherent mode weights
  const int* origins;        // [P][2] window origin (row, col) in the slab
  const double* shifts;      // [P][2] probe shift (row, col) in pixels
  const float2* prop_y;      // [n] exp(-i pi lambda dz qy^2) / n
  const float2* prop_x;      // [n] exp(-i pi lambda dz qx^2) / n
  float* patterns;           // [P][n][n] fftshifted intensities, or null
  float2* amplitude;         // [P][M][n][n] fftshifted amplitudes, or null
  float2* ckpt;              // [P*M][C-1][n][n] wave entering slices k, 2k, ... (C = ceil(S/k)), or null
  int ckpt_every;            // checkpoint interval k (slices)
};

// Number of stored checkpoints per wave: slices k, 2k, ... below S.
__host__ __device__ inline int num_checkpoints(int S, int k) { return k > 0 ? (S - 1) / k : 0; }

size_t forward_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves);
hipError_t launch_forward(const FwdParams& a, void* ws, size_t ws_bytes, hipStream_t st);

}  // namespace pty
