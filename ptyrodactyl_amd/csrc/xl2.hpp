// Host-side declarations of the persistent XCD-local streaming forward.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "forward.hpp"

namespace pty {

size_t xl2_state_bytes();
size_t xl2_workspace_bytes(int n, int group_waves);
hipError_t launch_forward_xl2(const FwdParams& a, void* ws, size_t ws_bytes, int* state, int group_waves, int grid,
                              hipStream_t st);

}  // namespace pty
