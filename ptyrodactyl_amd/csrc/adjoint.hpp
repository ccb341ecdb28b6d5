// Host-side declarations of the multislice adjoint (reverse-mode) launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "forward.hpp"

namespace pty {

struct AdjParams {
  FwdParams f;              // forward operands (geometry, slab, probe_k, weights, propagators)
  const float* grad_i;      // [P][n][n] dL/dI in the fftshifted pattern layout, or null with grad_amp
  const float2* grad_amp;   // [P][M][n][n] dL/dA of the fftshifted amplitudes (replaces 2 w g A), or null
  const float2* ckpt;       // [P*M][nc][n][n] checkpoints written by the forward (nc = (S-1)/k)
  int k;                    // checkpoint interval used by the forward
  float2* grad_slab;        // [S][Hs][Ws] += dL/dRe + i dL/dIm of the transmission
  float2* grad_probe_k;     // [M][n][n]   += gradient w.r.t. probe_k, or null
  double* grad_shifts;      // [P][2] += dL/d(probe shift in pixels), or null (gathered mode only)
  // Set by the launchers.  When gt_contrib is non-null the per-wave contributions to the slab
  // gradient are stored (window coordinates, [wave - w_begin][contrib_slices][n][n]) and gathered
  // in a fixed order by k_grad_reduce; otherwise they are added with float32 atomics.  gp_contrib
  // plays the same role for the probe gradient ([wave - w_begin][n][n]).
  float2* gt_contrib;
  float2* gp_contrib;
  int contrib_slices;
  long long w_begin, w_end;  // waves handled by one launch of the resident kernel
};

// Device scratch for the deterministic (gathered) adjoint of `waves` waves with S slices; a
// workspace down to adjoint_min_workspace_bytes runs with atomics instead.
size_t adjoint_workspace_bytes(int n, int S, int k, long long waves);
size_t adjoint_min_workspace_bytes(int n, int k);
hipError_t launch_adjoint(const AdjParams& a, void* ws, size_t ws_bytes, hipStream_t st);

// The streaming adjoint (n = 256 ... 4096, adjoint_stream.hip) and the deterministic gathers it shares with the
// resident adjoint: slab contributions of slices s_first .. s_first + s_count - 1 and (probe) the probe and shift
// gradients of waves w0 .. w0 + cnt - 1.
hipError_t launch_adj_streaming(const AdjParams& a, void* ws, size_t ws_bytes, hipStream_t st);
hipError_t launch_adj_reduces(const AdjParams& a, long long w0, int cnt, int s_first, int s_count, bool probe,
                              hipStream_t st);
// Waves of one streaming batch (n >= 256).
long long adjoint_streaming_batch(int n);

}  // namespace pty
