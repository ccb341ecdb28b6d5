// Multislice forward and adjoint on grids of any size h x w (1 to 4096 per axis,
// square or not, any factorisation).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace pty {

constexpr int GRID_MAX_EDGE = 4096;
constexpr int GRID_MAX_STAGES = 12;  // prime factors of an edge <= 4096, twos paired into fours

struct GridParams {
  int h, w, S, M;
  long long P;
  const float2* slab;        // [S][slab_h][slab_w] complex64 transmission
  long long slab_h, slab_w;  // slab extent (>= h, >= w)
  const float2* probe_k;     // [M][h][w] fft2 of each real-space probe mode
  const float* weights;      // [M] incoherent mode weights
  const int* origins;        // [P][2] window origin (row, col) in the slab
  const double* shifts;      // [P][2] probe shift (row, col) in pixels
  const float2* prop_y;      // [h] exp(-i pi lambda dz qy^2) / h
  const float2* prop_x;      // [w] exp(-i pi lambda dz qx^2) / w
  float* patterns;           // [P][h][w] fftshifted sum_m w_m |A_m|^2, or null
  float2* amplitude;         // [P][M][h][w] fftshifted A_m, or null
  float2* ckpt;              // [P*M][(S-1)/k][h][w] wave entering slices k, 2k, ..., or null
  int ckpt_every;            // k
  const float* grad_patterns;    // adjoint seed dL/dI [P][h][w] (fftshifted), or null
  const float2* grad_amplitude;  // adjoint seed dL/dA [P][M][h][w] (fftshifted), or null
  float2* grad_slab;             // [S][slab_h][slab_w], accumulated
  float2* grad_probe_k;          // [M][h][w], accumulated, or null
  double* grad_shifts;           // [P][2], accumulated, or null
  // set by the launcher
  int ry[GRID_MAX_STAGES], rx[GRID_MAX_STAGES];
  int ny, nx;                // stage counts along y and x
  int gy, gx;                // lines per LDS group in column / row passes
  int lsy, lsx;              // padded line strides in LDS
  int gcap;                  // elements of each LDS ping-pong buffer
};

// Bytes of device scratch per workgroup (one position at a time).
size_t grid_block_bytes(int h, int w, int ckpt_every, bool adjoint);
// Workspace for a launch with the default number of workgroups.
size_t grid_workspace_bytes(int h, int w, int ckpt_every, long long num_positions, bool adjoint);

// Forward (patterns and/or amplitudes, optional checkpoints).  The caller has validated the operands.
hipError_t launch_grid_forward(GridParams a, void* ws, size_t ws_bytes, hipStream_t st);
// Adjoint seeded by grad_patterns or grad_amplitude.  The per-position kernel accumulates dL/dT and dL/dprobe_k
// with float32 atomics; the streaming kernels gather dL/dT in a fixed order and add dL/dprobe_k with atomics.
hipError_t launch_grid_adjoint(GridParams a, void* ws, size_t ws_bytes, hipStream_t st);

}  // namespace pty
