// Host-side declarations of the forward-mode (Jacobian-vector product) multislice launcher.
#pragma once
#include "forward.hpp"

namespace pty {

struct JvpParams {
  FwdParams f;               // primal operands; f.patterns and f.amplitude are optional outputs
  const float2* d_slab;      // [S][slab_h][slab_w] tangent of the transmission slab, or null (zero)
  const float2* d_probe_k;   // [M][n][n] tangent of probe_k, or null (zero)
  const double* d_shifts;    // [P][2] tangent of the probe shifts (pixels), or null (zero)
  float* d_patterns;         // [P][n][n] fftshifted tangent of the intensities, or null
  float2* d_amplitude;       // [P][M][n][n] fftshifted tangent of the amplitudes, or null
  long long w0;              // first wave of the current batch (set by the launcher)
};

size_t jvp_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves);
hipError_t launch_jvp(const JvpParams& a, void* ws, size_t ws_bytes, hipStream_t st);

}  // namespace pty
