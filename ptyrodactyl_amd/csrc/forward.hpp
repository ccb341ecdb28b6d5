// Host-side declarations of the multislice forward launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace pty {

// Per-position descriptor of a batched launch that mixes positions of many
// requests (same window n and slice count S).  Offsets are in elements of
// the arena each pointer of FwdParams addresses.  Layout shared with
// pty_position_desc in include/ptyrodactyl_amd.h.
struct PosDesc {
  long long slab_offset;     // complex64 elements: request's slab in the slab arena
  long long slab_h, slab_w;  // request's slab extent (>= n)
  long long pattern_offset;  // float32 elements: this position's pattern (and dL/dI) in the pattern arena
  long long probe_offset;    // complex64 elements: request's probe_k (one mode) in the probe arena
  long long prop_offset;     // complex64 elements: request's prop_y[n] then prop_x[n]
  long long grad_offset;     // complex64 elements: request's slab gradient in the gradient arena
  long long gprobe_offset;   // complex64 elements: request's probe_k gradient
  int origin_y, origin_x;
  double shift_y, shift_x;
};

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
  const PosDesc* desc;       // [P] per-position descriptors (batched launches, M == 1), or null
  const unsigned char* det_masks;  // [D][n][n] 0/1 virtual-detector masks (fftshifted layout), or null
  double* det_out;                 // [P][D] sum over each mask of the pattern intensity, or null
  int det_count;                   // D, at most MAX_DETECTORS
  int max_lanes;                   // streaming launch lanes allowed (0 = the launcher's default, 1 = one stream)
};

constexpr int MAX_DETECTORS = 4;

// Number of stored checkpoints per wave: slices k, 2k, ... below S.
__host__ __device__ inline int num_checkpoints(int S, int k) { return k > 0 ? (S - 1) / k : 0; }

// Per-wave view of the operands: the shared arrays, or the request's slice
// of each arena when the launch is batched through descriptors.
struct WaveCtx {
  const float2* slab;
  long long slab_h, slab_w, stride;
  const float2* probe_k;  // mode 0 of this position's probe
  const float2* prop_y;
  const float2* prop_x;
  long long pat_off;      // float32 offset of this position's pattern
  long long grad_off;     // complex64 offset of the request's slab gradient
  long long gprobe_off;   // complex64 offset of the request's probe_k gradient
  int oy, ox;             // window origin, reduced into the slab
  int my, mx;             // integer probe shift modulo n
  float dy, dx;           // fractional probe shift
};

__device__ __forceinline__ int wrap_index(long long a, long long n) {
  if (a >= 0 && a < n) return (int)a;
  long long r = a % n;
  return (int)(r < 0 ? r + n : r);
}

__device__ __forceinline__ void split_shift_px(double s, int n, int& m, float& d) {
  if (!(fabs(s) < 1e15)) s = 0.0;  // non-finite or absurd shifts are rejected upstream
  const double r = floor(s + 0.5);
  d = (float)(s - r);
  m = wrap_index((long long)fmod(r, (double)n), n);
}

// SHIFTS = false leaves my, mx, dy, dx unset (passes that do not build the incident wave skip the FP64 split).
template <bool SHIFTS = true>
__device__ __forceinline__ WaveCtx resolve_wave(const FwdParams& a, long long p) {
  WaveCtx c;
  const int n = a.n;
  int oy, ox;
  double sy, sx;
  if (a.desc) {
    const PosDesc& d = a.desc[p];
    c.slab = a.slab + d.slab_offset;
    c.slab_h = d.slab_h;
    c.slab_w = d.slab_w;
    c.probe_k = a.probe_k + d.probe_offset;
    c.prop_y = a.prop_y + d.prop_offset;
    c.prop_x = a.prop_y + d.prop_offset + n;
    c.pat_off = d.pattern_offset;
    c.grad_off = d.grad_offset;
    c.gprobe_off = d.gprobe_offset;
    oy = d.origin_y;
    ox = d.origin_x;
    sy = d.shift_y;
    sx = d.shift_x;
  } else {
    c.slab = a.slab;
    c.slab_h = a.slab_h;
    c.slab_w = a.slab_w;
    c.probe_k = a.probe_k;
    c.prop_y = a.prop_y;
    c.prop_x = a.prop_x;
    c.pat_off = p * (long long)n * n;
    c.grad_off = 0;
    c.gprobe_off = 0;
    oy = a.origins[2 * p];
    ox = a.origins[2 * p + 1];
    sy = a.shifts[2 * p];
    sx = a.shifts[2 * p + 1];
  }
  c.stride = c.slab_h * c.slab_w;
  c.oy = wrap_index(oy, c.slab_h);
  c.ox = wrap_index(ox, c.slab_w);
  if constexpr (SHIFTS) {
    split_shift_px(sy, n, c.my, c.dy);
    split_shift_px(sx, n, c.mx, c.dx);
  } else {
    (void)sy;
    (void)sx;
  }
  return c;
}

size_t forward_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves);
hipError_t launch_forward(const FwdParams& a, void* ws, size_t ws_bytes, hipStream_t st);

}  // namespace pty
