// Host-side declarations of the multislice adjoint (reverse-mode) launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "forward.hpp"

namespace pty {

struct AdjParams {
  FwdParams f;              // forward operands (geometry, slab, probe_k, weights, propagators)
  const float* grad_i;      // [P][n][n] dL/dI in the fftshifted pattern layout
  const float2* ckpt;       // [P*M][nc][n][n] checkpoints written by the forward (nc = (S-1)/k)
  int k;                    // checkpoint interval used by the forward
  float2* grad_slab;        // [S][Hs][Ws] += dL/dRe + i dL/dIm of the transmission (atomic)
  float2* grad_probe_k;     // [M][n][n]   += gradient w.r.t. probe_k (atomic), or null
};

// Device scratch the adjoint needs for `slots` waves in flight.
size_t adjoint_workspace_bytes(int n, int k, long long waves);
hipError_t launch_adjoint(const AdjParams& a, void* ws, size_t ws_bytes, hipStream_t st);

}  // namespace pty
