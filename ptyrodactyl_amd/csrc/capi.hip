// C-ABI boundary of the library: plain pointers, sizes and the caller's
// stream; no framework types cross this line.  Every entry point returns a
// status (PTY_OK, PTY_ERR_INVALID, PTY_ERR_DEVICE) and records a message
// readable with pty_last_error() on the calling thread.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "../../include/ptyrodactyl_amd.h"
#include "adjoint.hpp"
#include "aux_kernels.hpp"
#include "forward.hpp"
#include "grid.hpp"
#include "jvp.hpp"
#include "optim.hpp"
#include "potential.hpp"
#include "xl2.hpp"

// Streaming-wave scratch of pty_forward_noalloc.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` is only a stand-in for the symbol: kernels get
// the per-device address that hipGetSymbolAddress returns (noalloc_scratch below), never the name.
static __device__ __align__(256) unsigned char scratch[PTY_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local char g_err[512] = "";

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int device_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return PTY_OK;
  return fail(PTY_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

bool valid_n(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096;
}

// Windows the batched (descriptor) paths support.
bool batched_n(int n) { return valid_n(n) && n <= 512; }

int check_common(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                 long long slab_h, long long slab_w, const void* probe_k, const float* weights,
                 const int32_t* origins, const double* shifts, const void* prop_y, const void* prop_x) {
  if (!valid_n(n))
    return fail(PTY_ERR_INVALID, "window size n=%d is not one of 32, 64, ..., 4096 (powers of two)", n);
  if (num_slices < 1) return fail(PTY_ERR_INVALID, "num_slices must be >= 1 (got %d)", num_slices);
  if (num_positions < 1 || num_positions > (1LL << 30))
    return fail(PTY_ERR_INVALID, "num_positions must be in [1, 2^30] (got %lld)", num_positions);
  if (num_modes < 1 || num_modes > 4096) return fail(PTY_ERR_INVALID, "num_modes must be in [1, 4096]");
  if (slab_h < n || slab_w < n || slab_h > (1LL << 30) || slab_w > (1LL << 30))
    return fail(PTY_ERR_INVALID, "slab extent %lldx%lld must be at least the window %dx%d", slab_h, slab_w, n, n);
  if (!slab || !probe_k || !weights || !origins || !shifts || !prop_y || !prop_x)
    return fail(PTY_ERR_INVALID, "null input pointer");
  return PTY_OK;
}

pty::FwdParams make_params(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                           long long slab_h, long long slab_w, const void* probe_k, const float* weights,
                           const int32_t* origins, const double* shifts, const void* prop_y,
                           const void* prop_x) {
  pty::FwdParams a;
  a.n = n;
  a.S = num_slices;
  a.M = num_modes;
  a.P = num_positions;
  a.slab = (const float2*)slab;
  a.slab_h = slab_h;
  a.slab_w = slab_w;
  a.slice_stride = slab_h * slab_w;
  a.probe_k = (const float2*)probe_k;
  a.weights = weights;
  a.origins = origins;
  a.shifts = shifts;
  a.prop_y = (const float2*)prop_y;
  a.prop_x = (const float2*)prop_x;
  a.patterns = nullptr;
  a.amplitude = nullptr;
  a.ckpt = nullptr;
  a.ckpt_every = 0;
  a.desc = nullptr;
  a.det_masks = nullptr;
  a.det_out = nullptr;
  a.det_count = 0;
  a.max_lanes = 0;
  return a;
}

int check_grid(int h, int w, int num_slices, long long num_positions, int num_modes, const void* slab,
               long long slab_h, long long slab_w, const void* probe_k, const float* weights, const int32_t* origins,
               const double* shifts, const void* prop_y, const void* prop_x) {
  if (h < 1 || w < 1 || h > pty::GRID_MAX_EDGE || w > pty::GRID_MAX_EDGE)
    return fail(PTY_ERR_INVALID, "grid %dx%d: each edge must be in [1, %d]", h, w, pty::GRID_MAX_EDGE);
  if (num_slices < 1) return fail(PTY_ERR_INVALID, "num_slices must be >= 1 (got %d)", num_slices);
  if (num_positions < 1 || num_positions > (1LL << 30))
    return fail(PTY_ERR_INVALID, "num_positions must be in [1, 2^30] (got %lld)", num_positions);
  if (num_modes < 1 || num_modes > 4096) return fail(PTY_ERR_INVALID, "num_modes must be in [1, 4096]");
  if (slab_h < h || slab_w < w || slab_h > (1LL << 30) || slab_w > (1LL << 30))
    return fail(PTY_ERR_INVALID, "slab extent %lldx%lld must be at least the window %dx%d", slab_h, slab_w, h, w);
  if (!slab || !probe_k || !weights || !origins || !shifts || !prop_y || !prop_x)
    return fail(PTY_ERR_INVALID, "null input pointer");
  return PTY_OK;
}

pty::GridParams make_grid_params(int h, int w, int num_slices, long long num_positions, int num_modes,
                                 const void* slab, long long slab_h, long long slab_w, const void* probe_k,
                                 const float* weights, const int32_t* origins, const double* shifts,
                                 const void* prop_y, const void* prop_x) {
  pty::GridParams a;
  memset(&a, 0, sizeof(a));
  a.h = h;
  a.w = w;
  a.S = num_slices;
  a.M = num_modes;
  a.P = num_positions;
  a.slab = (const float2*)slab;
  a.slab_h = slab_h;
  a.slab_w = slab_w;
  a.probe_k = (const float2*)probe_k;
  a.weights = weights;
  a.origins = origins;
  a.shifts = shifts;
  a.prop_y = (const float2*)prop_y;
  a.prop_x = (const float2*)prop_x;
  a.ckpt_every = 1;
  return a;
}

int grid_ckpt(int ckpt_every, int num_slices) { return ckpt_every < num_slices ? ckpt_every : num_slices; }

// Device address of the current device's `scratch`, looked up once per device (with that device current).
int noalloc_scratch(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return device_status(e, "pty_forward_noalloc: hipGetDevice");
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(PTY_ERR_DEVICE, "pty_forward_noalloc: device %d is beyond the %d devices supported", dev,
                NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return device_status(e, "pty_forward_noalloc: hipGetSymbolAddress(scratch)");
    if (!p) return fail(PTY_ERR_DEVICE, "pty_forward_noalloc: hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PTY_OK;
}

}  // namespace

extern "C" {

int pty_abi_version(void) { return PTY_ABI_VERSION; }

const char* pty_last_error(void) { return g_err; }

size_t pty_forward_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves) {
  if (!valid_n(n) || num_positions < 1 || num_modes < 1) return 0;
  return pty::forward_workspace_bytes(n, num_positions, num_modes, max_batch_waves);
}

int pty_forward(int n, int num_slices, long long num_positions, int num_modes, const void* slab, long long slab_h,
                long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
                const double* shifts_px, const void* prop_y, const void* prop_x, float* patterns, void* amplitude,
                void* workspace, size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                        origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (!patterns && !amplitude) return fail(PTY_ERR_INVALID, "nothing to compute: patterns and amplitude are null");
  const size_t need = pty::forward_workspace_bytes(n, 1, num_modes, 0);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one position needs", workspace_bytes,
                need);
  pty::FwdParams a = make_params(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k,
                                 mode_weights, origins, shifts_px, prop_y, prop_x);
  a.patterns = patterns;
  a.amplitude = (float2*)amplitude;
  return device_status(pty::launch_forward(a, workspace, workspace_bytes, (hipStream_t)stream), "pty_forward");
}

int pty_forward_noalloc(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                        long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                        const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                        float* patterns, void* amplitude, void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                        origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (!patterns && !amplitude) return fail(PTY_ERR_INVALID, "nothing to compute: patterns and amplitude are null");
  const size_t need = pty::forward_workspace_bytes(n, 1, num_modes, 0);
  if (need > PTY_NOALLOC_SCRATCH_BYTES)
    return fail(PTY_ERR_INVALID,
                "pty_forward_noalloc: one position of %d mode(s) at n=%d needs %zu bytes of scratch, above the "
                "%zu-byte static scratch (use pty_forward with a workspace)",
                num_modes, n, need, (size_t)PTY_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  size_t ws_bytes = 0;
  if (need > 0) {
    rc = noalloc_scratch(&ws);
    if (rc) return rc;
    ws_bytes = PTY_NOALLOC_SCRATCH_BYTES;
  }
  pty::FwdParams a = make_params(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k,
                                 mode_weights, origins, shifts_px, prop_y, prop_x);
  a.patterns = patterns;
  a.amplitude = (float2*)amplitude;
  a.max_lanes = 1;  // the caller's stream only: no library streams or events are created
  return device_status(pty::launch_forward(a, ws, ws_bytes, (hipStream_t)stream), "pty_forward_noalloc");
}

size_t pty_jvp_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves) {
  if (!batched_n(n) || num_positions < 1 || num_modes < 1) return 0;
  return pty::jvp_workspace_bytes(n, num_positions, num_modes, max_batch_waves);
}

int pty_jvp(int n, int num_slices, long long num_positions, int num_modes, const void* slab, long long slab_h,
            long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
            const double* shifts_px, const void* prop_y, const void* prop_x, const void* d_slab,
            const void* d_probe_k, const double* d_shifts_px, float* patterns, void* amplitude, float* d_patterns,
            void* d_amplitude, void* workspace, size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                        origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (!batched_n(n)) return fail(PTY_ERR_INVALID, "pty_jvp takes windows n = 32 ... 512 (got n=%d)", n);
  if (!d_patterns && !d_amplitude)
    return fail(PTY_ERR_INVALID, "nothing to compute: d_patterns and d_amplitude are null");
  const size_t need = pty::jvp_workspace_bytes(n, 1, num_modes, 0);
  if (!workspace || workspace_bytes < need)
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one position needs", workspace_bytes,
                need);
  pty::JvpParams a;
  a.f = make_params(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights, origins,
                    shifts_px, prop_y, prop_x);
  a.f.patterns = patterns;
  a.f.amplitude = (float2*)amplitude;
  a.d_slab = (const float2*)d_slab;
  a.d_probe_k = (const float2*)d_probe_k;
  a.d_shifts = d_shifts_px;
  a.d_patterns = d_patterns;
  a.d_amplitude = (float2*)d_amplitude;
  a.w0 = 0;
  return device_status(pty::launch_jvp(a, workspace, workspace_bytes, (hipStream_t)stream), "pty_jvp");
}

int pty_forward_detectors(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                          long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                          const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                          float* patterns, int num_detectors, const unsigned char* masks, double* detector_out,
                          void* workspace, size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                        origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (n > 1024) return fail(PTY_ERR_INVALID, "virtual detectors take windows up to 1024 (got n=%d)", n);
  if (num_detectors < 1 || num_detectors > PTY_MAX_DETECTORS)
    return fail(PTY_ERR_INVALID, "num_detectors must be in [1, %d] (got %d)", PTY_MAX_DETECTORS, num_detectors);
  if (!masks || !detector_out) return fail(PTY_ERR_INVALID, "masks and detector_out must be non-null");
  const size_t need = pty::forward_workspace_bytes(n, 1, num_modes, 0);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one position needs", workspace_bytes,
                need);
  pty::FwdParams a = make_params(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k,
                                 mode_weights, origins, shifts_px, prop_y, prop_x);
  a.patterns = patterns;
  a.det_masks = masks;
  a.det_out = detector_out;
  a.det_count = num_detectors;
  return device_status(pty::launch_forward(a, workspace, workspace_bytes, (hipStream_t)stream),
                       "pty_forward_detectors");
}

int pty_forward_checkpointed(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                             long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                             const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                             float* patterns, void* amplitude, void* checkpoints, int ckpt_every, void* workspace,
                             size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                        origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (ckpt_every < 1) return fail(PTY_ERR_INVALID, "ckpt_every must be >= 1 (got %d)", ckpt_every);
  if (pty::num_checkpoints(num_slices, ckpt_every) > 0 && !checkpoints)
    return fail(PTY_ERR_INVALID, "checkpoint buffer is null but %d checkpoints are needed",
                pty::num_checkpoints(num_slices, ckpt_every));
  if (!patterns && !amplitude) return fail(PTY_ERR_INVALID, "patterns and amplitude are both null");
  const size_t need = pty::forward_workspace_bytes(n, 1, num_modes, 0);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one position needs", workspace_bytes,
                need);
  pty::FwdParams a = make_params(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k,
                                 mode_weights, origins, shifts_px, prop_y, prop_x);
  a.patterns = patterns;
  a.amplitude = (float2*)amplitude;
  a.ckpt = (float2*)checkpoints;
  a.ckpt_every = ckpt_every;
  return device_status(pty::launch_forward(a, workspace, workspace_bytes, (hipStream_t)stream),
                       "pty_forward_checkpointed");
}

size_t pty_adjoint_workspace_bytes(int n, int num_slices, int ckpt_every, long long num_positions, int num_modes) {
  if (!valid_n(n) || num_slices < 1 || ckpt_every < 1 || num_positions < 1 || num_modes < 1) return 0;
  return pty::adjoint_workspace_bytes(n, num_slices, ckpt_every, num_positions * num_modes);
}

int pty_adjoint(int n, int num_slices, long long num_positions, int num_modes, const void* slab, long long slab_h,
                long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
                const double* shifts_px, const void* prop_y, const void* prop_x, const float* grad_patterns,
                const void* checkpoints, int ckpt_every, void* grad_slab, void* grad_probe_k, double* grad_shifts,
                void* workspace, size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                        origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (ckpt_every < 1) return fail(PTY_ERR_INVALID, "ckpt_every must be >= 1 (got %d)", ckpt_every);
  if (pty::num_checkpoints(num_slices, ckpt_every) > 0 && !checkpoints)
    return fail(PTY_ERR_INVALID, "checkpoint buffer is null but %d checkpoints are needed",
                pty::num_checkpoints(num_slices, ckpt_every));
  if (!grad_patterns || !grad_slab) return fail(PTY_ERR_INVALID, "grad_patterns and grad_slab must be non-null");
  const size_t need = pty::adjoint_min_workspace_bytes(n, ckpt_every);
  if (!workspace || workspace_bytes < need)
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one wave needs", workspace_bytes,
                need);
  if (grad_shifts) {
    const size_t full = pty::adjoint_workspace_bytes(n, num_slices, ckpt_every, num_positions * num_modes);
    if (workspace_bytes < full)
      return fail(PTY_ERR_INVALID, "position gradients need the gathered workspace of %zu bytes (got %zu)", full,
                  workspace_bytes);
  }
  pty::AdjParams a{};
  a.f = make_params(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights, origins,
                    shifts_px, prop_y, prop_x);
  a.grad_i = grad_patterns;
  a.ckpt = (const float2*)checkpoints;
  a.k = ckpt_every;
  a.grad_slab = (float2*)grad_slab;
  a.grad_probe_k = (float2*)grad_probe_k;
  a.grad_shifts = grad_shifts;
  return device_status(pty::launch_adjoint(a, workspace, workspace_bytes, (hipStream_t)stream), "pty_adjoint");
}

int pty_adjoint_amplitude(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                          long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                          const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                          const void* grad_amplitude, const void* checkpoints, int ckpt_every, void* grad_slab,
                          void* grad_probe_k, double* grad_shifts, void* workspace, size_t workspace_bytes,
                          void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                        origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (ckpt_every < 1) return fail(PTY_ERR_INVALID, "ckpt_every must be >= 1 (got %d)", ckpt_every);
  if (pty::num_checkpoints(num_slices, ckpt_every) > 0 && !checkpoints)
    return fail(PTY_ERR_INVALID, "checkpoint buffer is null but %d checkpoints are needed",
                pty::num_checkpoints(num_slices, ckpt_every));
  if (!grad_amplitude || !grad_slab) return fail(PTY_ERR_INVALID, "grad_amplitude and grad_slab must be non-null");
  const size_t need = pty::adjoint_min_workspace_bytes(n, ckpt_every);
  if (!workspace || workspace_bytes < need)
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one wave needs", workspace_bytes,
                need);
  if (grad_shifts) {
    const size_t full = pty::adjoint_workspace_bytes(n, num_slices, ckpt_every, num_positions * num_modes);
    if (workspace_bytes < full)
      return fail(PTY_ERR_INVALID, "position gradients need the gathered workspace of %zu bytes (got %zu)", full,
                  workspace_bytes);
  }
  pty::AdjParams a{};
  a.f = make_params(n, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights, origins,
                    shifts_px, prop_y, prop_x);
  a.grad_i = nullptr;
  a.grad_amp = (const float2*)grad_amplitude;
  a.ckpt = (const float2*)checkpoints;
  a.k = ckpt_every;
  a.grad_slab = (float2*)grad_slab;
  a.grad_probe_k = (float2*)grad_probe_k;
  a.grad_shifts = grad_shifts;
  return device_status(pty::launch_adjoint(a, workspace, workspace_bytes, (hipStream_t)stream),
                       "pty_adjoint_amplitude");
}

int pty_transmission(const double* potential, double sigma, void* transmission, long long count, void* stream) {
  g_err[0] = '\0';
  if (!potential || !transmission) return fail(PTY_ERR_INVALID, "null pointer");
  if (count < 0) return fail(PTY_ERR_INVALID, "count must be >= 0");
  if (count == 0) return PTY_OK;
  return device_status(pty::launch_transmission(potential, sigma, (float2*)transmission, count, (hipStream_t)stream),
                       "pty_transmission");
}

size_t pty_sqrt_intensity_loss_workspace_bytes(void) { return (size_t)pty::SQRT_LOSS_BLOCKS * sizeof(double); }

int pty_sqrt_intensity_loss(long long count, const float* patterns, const float* measured, float eps, float* grad,
                            double* loss, void* workspace, size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  if (!patterns || !measured) return fail(PTY_ERR_INVALID, "null pointer");
  if (count < 0) return fail(PTY_ERR_INVALID, "count must be >= 0");
  if (!(eps >= 0.f)) return fail(PTY_ERR_INVALID, "eps must be >= 0");
  if (!workspace || workspace_bytes < pty_sqrt_intensity_loss_workspace_bytes())
    return fail(PTY_ERR_INVALID, "workspace smaller than pty_sqrt_intensity_loss_workspace_bytes()");
  return device_status(pty::launch_sqrt_loss(patterns, measured, eps, count, grad, loss, (double*)workspace,
                                             (hipStream_t)stream),
                       "pty_sqrt_intensity_loss");
}

size_t pty_normal_equations_workspace_bytes(long long num_positions, int n, int num_params) {
  if (num_positions < 1 || n < 1 || num_params < 1 || num_params > PTY_MAX_NORMAL_PARAMS) return 0;
  const int nprod = num_params * (num_params + 1) / 2 + num_params + 1;
  return (size_t)pty::normal_partial_blocks(num_positions, n * n) * nprod * sizeof(double);
}

int pty_normal_equations(const void* amplitude, const float* data, long long num_positions, int n, int num_params,
                         void* workspace, size_t workspace_bytes, double* jtj, double* jtr, double* rss,
                         void* stream) {
  g_err[0] = '\0';
  if (!amplitude || !data || !jtj || !jtr || !rss) return fail(PTY_ERR_INVALID, "null pointer");
  if (num_params < 1 || num_params > PTY_MAX_NORMAL_PARAMS)
    return fail(PTY_ERR_INVALID, "num_params must be in [1, %d] (got %d)", PTY_MAX_NORMAL_PARAMS, num_params);
  if (num_positions < 1 || n < 1) return fail(PTY_ERR_INVALID, "empty problem");
  const size_t need = pty_normal_equations_workspace_bytes(num_positions, n, num_params);
  if (!workspace || workspace_bytes < need)
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes needed", workspace_bytes, need);
  return device_status(pty::launch_normal_equations((const float2*)amplitude, data, num_positions, n * n, num_params,
                                                    (double*)workspace, jtj, jtr, rss, (hipStream_t)stream),
                       "pty_normal_equations");
}

int pty_damped_solve(const double* matrix, const double* rhs, double damping, int num_params, double* solution,
                     double* model_reduction, int* status, void* stream) {
  g_err[0] = '\0';
  if (!matrix || !rhs || !solution || !model_reduction || !status) return fail(PTY_ERR_INVALID, "null pointer");
  if (num_params < 1 || num_params > 64) return fail(PTY_ERR_INVALID, "num_params must be in [1, 64]");
  if (!(damping >= 0.0)) return fail(PTY_ERR_INVALID, "damping must be >= 0");
  return device_status(pty::launch_damped_solve(matrix, rhs, damping, num_params, solution, model_reduction, status,
                                                (hipStream_t)stream),
                       "pty_damped_solve");
}

int pty_annular_integrate(const float* patterns, long long num_patterns, long long pixels_per_pattern,
                          const unsigned char* mask, double* out, void* stream) {
  g_err[0] = '\0';
  if (num_patterns < 0 || pixels_per_pattern < 1)
    return fail(PTY_ERR_INVALID, "num_patterns must be >= 0 and pixels_per_pattern >= 1");
  if (num_patterns > 0x7fffffffLL) return fail(PTY_ERR_INVALID, "at most 2^31 - 1 patterns per call");
  if (num_patterns > 0 && (!patterns || !mask || !out)) return fail(PTY_ERR_INVALID, "null pointer");
  return device_status(pty::launch_annular(patterns, num_patterns, pixels_per_pattern, mask, out,
                                           (hipStream_t)stream),
                       "pty_annular_integrate");
}

int pty_optimizer_update(int rule, int fp64, int is_complex, long long count, void* params, const void* grads,
                         void* mu, void* nu, long long step, double learning_rate, void* stream) {
  g_err[0] = '\0';
  if (rule < PTY_OPTAX_ADAM || rule > PTY_OPTAX_RMSPROP) return fail(PTY_ERR_INVALID, "unknown update rule %d", rule);
  if (count < 0) return fail(PTY_ERR_INVALID, "count must be >= 0");
  if (count > 0 && (!params || !grads || !nu)) return fail(PTY_ERR_INVALID, "null pointer");
  if (count > 0 && rule == PTY_OPTAX_ADAM && !mu) return fail(PTY_ERR_INVALID, "adam needs the first-moment buffer");
  if (rule == PTY_OPTAX_ADAM && step < 1) return fail(PTY_ERR_INVALID, "adam step count must be >= 1");
  pty::OptaxStep s{};
  s.neg_lr = -learning_rate;
  s.eps = 1e-8;
  if (rule == PTY_OPTAX_ADAM) {
    s.b1 = 0.9;
    s.b2 = 0.999;
    s.bc1 = 1.0 - pow(0.9, (double)step);
    s.bc2 = 1.0 - pow(0.999, (double)step);
  } else if (rule == PTY_OPTAX_RMSPROP) {
    s.b1 = 0.9;  // decay
  }
  s.one_minus_b1 = 1.0 - s.b1;
  s.one_minus_b2 = 1.0 - s.b2;
  return device_status(pty::launch_optax(rule, fp64 != 0, is_complex != 0, count, params, grads, mu, nu, s,
                                         (hipStream_t)stream),
                       "pty_optimizer_update");
}

int pty_random_normal(unsigned key0, unsigned key1, long long count, double* out, void* stream) {
  g_err[0] = '\0';
  if (count < 0 || count > (1LL << 40)) return fail(PTY_ERR_INVALID, "count must be in [0, 2^40]");
  if (count > 0 && !out) return fail(PTY_ERR_INVALID, "null pointer");
  return device_status(pty::launch_random_normal(key0, key1, count, out, (hipStream_t)stream), "pty_random_normal");
}

int pty_projected_potential(int parameterization, const double* row, long long count, const double* r, double* out,
                            void* stream) {
  g_err[0] = '\0';
  if (parameterization != PTY_PARAM_LOBATO && parameterization != PTY_PARAM_KIRKLAND)
    return fail(PTY_ERR_INVALID, "parameterization must be PTY_PARAM_LOBATO or PTY_PARAM_KIRKLAND");
  if (count < 0 || count > (1LL << 40)) return fail(PTY_ERR_INVALID, "count must be in [0, 2^40]");
  if (!row || (count > 0 && (!r || !out))) return fail(PTY_ERR_INVALID, "null pointer");
  return device_status(pty::launch_projected_radii(parameterization, row, count, r, out, (hipStream_t)stream),
                       "pty_projected_potential");
}

int pty_atom_potential_grid(int parameterization, int num_types, const double* rows, double pixel_size, int height,
                            int width, int supersampling, const double* centers, double* out, void* stream) {
  g_err[0] = '\0';
  if (parameterization != PTY_PARAM_LOBATO && parameterization != PTY_PARAM_KIRKLAND)
    return fail(PTY_ERR_INVALID, "parameterization must be PTY_PARAM_LOBATO or PTY_PARAM_KIRKLAND");
  if (num_types < 0 || num_types > 65535) return fail(PTY_ERR_INVALID, "num_types must be in [0, 65535]");
  if (height < 1 || width < 1 || height > 32768 || width > 32768)
    return fail(PTY_ERR_INVALID, "height and width must be in [1, 32768]");
  if (supersampling < 1 || supersampling > 64) return fail(PTY_ERR_INVALID, "supersampling must be in [1, 64]");
  if (!(pixel_size > 0.0) || !isfinite(pixel_size)) return fail(PTY_ERR_INVALID, "pixel_size must be positive");
  if (num_types > 0 && (!rows || !out)) return fail(PTY_ERR_INVALID, "null pointer");
  return device_status(pty::launch_atom_grid(parameterization, num_types, rows, pixel_size, height, width,
                                             supersampling, centers, out, (hipStream_t)stream),
                       "pty_atom_potential_grid");
}

size_t pty_atom_center_grad_workspace_bytes(int num_types) { return pty::atom_center_grad_workspace_bytes(num_types); }

int pty_atom_potential_grid_center_grad(int parameterization, int num_types, const double* rows, double pixel_size,
                                        int height, int width, int supersampling, const double* centers,
                                        const double* grad_out, double* grad_centers, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  if (parameterization != PTY_PARAM_LOBATO && parameterization != PTY_PARAM_KIRKLAND)
    return fail(PTY_ERR_INVALID, "parameterization must be PTY_PARAM_LOBATO or PTY_PARAM_KIRKLAND");
  if (num_types < 0 || num_types > 65535) return fail(PTY_ERR_INVALID, "num_types must be in [0, 65535]");
  if (height < 1 || width < 1 || height > 32768 || width > 32768)
    return fail(PTY_ERR_INVALID, "height and width must be in [1, 32768]");
  if (supersampling < 1 || supersampling > 64) return fail(PTY_ERR_INVALID, "supersampling must be in [1, 64]");
  if (!(pixel_size > 0.0) || !isfinite(pixel_size)) return fail(PTY_ERR_INVALID, "pixel_size must be positive");
  if (num_types > 0 && (!rows || !grad_out || !grad_centers || !workspace))
    return fail(PTY_ERR_INVALID, "null pointer");
  if (workspace_bytes < pty::atom_center_grad_workspace_bytes(num_types))
    return fail(PTY_ERR_INVALID, "workspace smaller than pty_atom_center_grad_workspace_bytes(num_types)");
  return device_status(pty::launch_atom_center_grad(parameterization, num_types, rows, pixel_size, height, width,
                                                    supersampling, centers, grad_out, grad_centers, workspace,
                                                    (hipStream_t)stream),
                       "pty_atom_potential_grid_center_grad");
}

int pty_potential_slices_fourier(int num_slices, int num_types, int height, int width, long long num_atoms,
                                 const double* shifts, const long long* segments, const void* kernel_k, void* out_k,
                                 void* stream) {
  g_err[0] = '\0';
  if (num_slices < 0 || num_slices > 65535) return fail(PTY_ERR_INVALID, "num_slices must be in [0, 65535]");
  if (num_types < 0 || num_types > 4096) return fail(PTY_ERR_INVALID, "num_types must be in [0, 4096]");
  if (height < 1 || width < 1 || height > 32768 || width > 32768)
    return fail(PTY_ERR_INVALID, "height and width must be in [1, 32768]");
  if (num_atoms < 0) return fail(PTY_ERR_INVALID, "num_atoms must be non-negative");
  if (num_slices > 0 && (!segments || !out_k || (num_types > 0 && !kernel_k) || (num_atoms > 0 && !shifts)))
    return fail(PTY_ERR_INVALID, "null pointer");
  return device_status(pty::launch_slice_structure(num_slices, num_types, height, width, num_atoms, shifts,
                                                   segments, (const double2*)kernel_k, (double2*)out_k,
                                                   (hipStream_t)stream),
                       "pty_potential_slices_fourier");
}

size_t pty_potential_shift_grad_workspace_bytes(int height, int width, long long num_atoms) {
  if (height < 1 || width < 1 || num_atoms < 0) return 0;
  return pty::slice_shift_grad_workspace_bytes(height, width, num_atoms);
}

int pty_potential_slices_shift_grad(int num_slices, int num_types, int height, int width, long long num_atoms,
                                    const double* shifts, const long long* segments, const void* kernel_k,
                                    const void* adj_k, double* grad_shifts, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  g_err[0] = '\0';
  if (num_slices < 0 || num_slices > 65535) return fail(PTY_ERR_INVALID, "num_slices must be in [0, 65535]");
  if (num_types < 0 || num_types > 4096) return fail(PTY_ERR_INVALID, "num_types must be in [0, 4096]");
  if (height < 1 || width < 1 || height > 32768 || width > 32768)
    return fail(PTY_ERR_INVALID, "height and width must be in [1, 32768]");
  if (num_atoms < 0) return fail(PTY_ERR_INVALID, "num_atoms must be non-negative");
  if (num_slices > 0 && num_atoms > 0 &&
      (!segments || !shifts || !grad_shifts || !adj_k || (num_types > 0 && !kernel_k) || !workspace))
    return fail(PTY_ERR_INVALID, "null pointer");
  if (num_slices > 0 && num_atoms > 0 && workspace_bytes < pty::slice_shift_grad_workspace_bytes(height, width, num_atoms))
    return fail(PTY_ERR_INVALID, "workspace smaller than pty_potential_shift_grad_workspace_bytes");
  return device_status(pty::launch_slice_shift_grad(num_slices, num_types, height, width, num_atoms, shifts,
                                                    segments, (const double2*)kernel_k, (const double2*)adj_k,
                                                    grad_shifts, workspace, (hipStream_t)stream),
                       "pty_potential_slices_shift_grad");
}

size_t pty_persistent_state_bytes(void) { return pty::xl2_state_bytes(); }

size_t pty_persistent_workspace_bytes(int n, int group_waves) {
  if (n != 256 || group_waves < 1) return 0;
  return pty::xl2_workspace_bytes(n, group_waves);
}

int pty_forward_persistent(int n, int num_slices, long long num_positions, const void* slab, long long slab_h,
                           long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
                           const double* shifts_px, const void* prop_y, const void* prop_x, float* patterns,
                           int group_waves, int grid, void* workspace, size_t workspace_bytes, int* state,
                           void* stream) {
  g_err[0] = '\0';
  int rc = check_common(n, num_slices, num_positions, 1, slab, slab_h, slab_w, probe_k, mode_weights, origins,
                        shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (n != 256) return fail(PTY_ERR_INVALID, "the persistent forward supports n = 256 (got %d)", n);
  if (!patterns || !state) return fail(PTY_ERR_INVALID, "patterns and state must be non-null");
  if (group_waves < 1 || group_waves > 16) return fail(PTY_ERR_INVALID, "group_waves must be in [1, 16]");
  if (grid < 8 || grid > 65536) return fail(PTY_ERR_INVALID, "grid must be in [8, 65536]");
  if (!workspace || workspace_bytes < pty::xl2_workspace_bytes(n, group_waves))
    return fail(PTY_ERR_INVALID, "workspace below pty_persistent_workspace_bytes");
  pty::FwdParams a = make_params(n, num_slices, num_positions, 1, slab, slab_h, slab_w, probe_k, mode_weights,
                                 origins, shifts_px, prop_y, prop_x);
  a.patterns = patterns;
  return device_status(pty::launch_forward_xl2(a, workspace, workspace_bytes, state, group_waves, grid,
                                               (hipStream_t)stream),
                       "pty_forward_persistent");
}

static_assert(PTY_MAX_DETECTORS == pty::MAX_DETECTORS, "detector limit");
static_assert(sizeof(pty_position_desc) == sizeof(pty::PosDesc), "descriptor layout mismatch");

static int check_batched(int n, int num_slices, long long num_positions, const void* descs, const void* slab_arena,
                         const void* probe_arena, const void* prop_arena) {
  if (!batched_n(n))
    return fail(PTY_ERR_INVALID, "window size n=%d: the batched paths take 32, 64, 128, 256, 512", n);
  if (num_slices < 1) return fail(PTY_ERR_INVALID, "num_slices must be >= 1 (got %d)", num_slices);
  if (num_positions < 1 || num_positions > (1LL << 30))
    return fail(PTY_ERR_INVALID, "num_positions must be in [1, 2^30] (got %lld)", num_positions);
  if (!descs || !slab_arena || !probe_arena || !prop_arena) return fail(PTY_ERR_INVALID, "null input pointer");
  return PTY_OK;
}

static pty::FwdParams batched_params(int n, int num_slices, long long num_positions, const void* descs,
                                     const void* slab_arena, const void* probe_arena, const void* prop_arena) {
  pty::FwdParams a = make_params(n, num_slices, num_positions, 1, slab_arena, n, n, probe_arena, nullptr, nullptr,
                                 nullptr, prop_arena, prop_arena);
  a.desc = (const pty::PosDesc*)descs;
  return a;
}

int pty_forward_batched(int n, int num_slices, long long num_positions, const pty_position_desc* descs,
                        const void* slab_arena, const void* probe_arena, const void* prop_arena, float* pattern_arena,
                        void* amplitudes, void* checkpoints, int ckpt_every, void* workspace, size_t workspace_bytes,
                        void* stream) {
  g_err[0] = '\0';
  int rc = check_batched(n, num_slices, num_positions, descs, slab_arena, probe_arena, prop_arena);
  if (rc) return rc;
  if (!pattern_arena) return fail(PTY_ERR_INVALID, "pattern arena is null");
  if (checkpoints && ckpt_every < 1) return fail(PTY_ERR_INVALID, "ckpt_every must be >= 1");
  const size_t need = pty::forward_workspace_bytes(n, 1, 1, 0);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one position needs", workspace_bytes,
                need);
  pty::FwdParams a = batched_params(n, num_slices, num_positions, descs, slab_arena, probe_arena, prop_arena);
  a.patterns = pattern_arena;
  a.amplitude = (float2*)amplitudes;
  if (checkpoints && pty::num_checkpoints(num_slices, ckpt_every) > 0) {
    a.ckpt = (float2*)checkpoints;
    a.ckpt_every = ckpt_every;
  }
  return device_status(pty::launch_forward(a, workspace, workspace_bytes, (hipStream_t)stream), "pty_forward_batched");
}

int pty_adjoint_batched(int n, int num_slices, long long num_positions, const pty_position_desc* descs,
                        const void* slab_arena, const void* probe_arena, const void* prop_arena,
                        const float* grad_pattern_arena, const void* checkpoints, int ckpt_every,
                        void* grad_slab_arena, void* grad_probe_arena, void* workspace, size_t workspace_bytes,
                        void* stream) {
  g_err[0] = '\0';
  int rc = check_batched(n, num_slices, num_positions, descs, slab_arena, probe_arena, prop_arena);
  if (rc) return rc;
  if (ckpt_every < 1) return fail(PTY_ERR_INVALID, "ckpt_every must be >= 1 (got %d)", ckpt_every);
  if (pty::num_checkpoints(num_slices, ckpt_every) > 0 && !checkpoints)
    return fail(PTY_ERR_INVALID, "checkpoint buffer is null but checkpoints are needed");
  if (!grad_pattern_arena || !grad_slab_arena) return fail(PTY_ERR_INVALID, "gradient arenas must be non-null");
  const size_t need = pty::adjoint_min_workspace_bytes(n, ckpt_every);
  if (!workspace || workspace_bytes < need)
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one wave needs", workspace_bytes,
                need);
  pty::AdjParams a{};
  a.f = batched_params(n, num_slices, num_positions, descs, slab_arena, probe_arena, prop_arena);
  a.grad_i = grad_pattern_arena;
  a.ckpt = (const float2*)checkpoints;
  a.k = ckpt_every;
  a.grad_slab = (float2*)grad_slab_arena;
  a.grad_probe_k = (float2*)grad_probe_arena;
  return device_status(pty::launch_adjoint(a, workspace, workspace_bytes, (hipStream_t)stream),
                       "pty_adjoint_batched");
}

size_t pty_grid_workspace_bytes(int h, int w, int num_slices, int ckpt_every, long long num_positions, int adjoint) {
  if (h < 1 || w < 1 || h > pty::GRID_MAX_EDGE || w > pty::GRID_MAX_EDGE || num_slices < 1 || ckpt_every < 1 ||
      num_positions < 1)
    return 0;
  return pty::grid_workspace_bytes(h, w, grid_ckpt(ckpt_every, num_slices), num_positions, adjoint != 0);
}

int pty_forward_grid(int h, int w, int num_slices, long long num_positions, int num_modes, const void* slab,
                     long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                     const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                     float* patterns, void* amplitude, void* checkpoints, int ckpt_every, void* workspace,
                     size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  int rc = check_grid(h, w, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                      origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (!patterns && !amplitude) return fail(PTY_ERR_INVALID, "nothing to compute: patterns and amplitude are null");
  if (checkpoints && ckpt_every < 1) return fail(PTY_ERR_INVALID, "ckpt_every must be >= 1 (got %d)", ckpt_every);
  const size_t need = pty::grid_block_bytes(h, w, 1, false);
  if (!workspace || workspace_bytes < need)
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one position needs", workspace_bytes,
                need);
  pty::GridParams a = make_grid_params(h, w, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k,
                                       mode_weights, origins, shifts_px, prop_y, prop_x);
  a.patterns = patterns;
  a.amplitude = (float2*)amplitude;
  a.ckpt = (float2*)checkpoints;
  a.ckpt_every = checkpoints ? ckpt_every : 1;
  return device_status(pty::launch_grid_forward(a, workspace, workspace_bytes, (hipStream_t)stream),
                       "pty_forward_grid");
}

int pty_adjoint_grid(int h, int w, int num_slices, long long num_positions, int num_modes, const void* slab,
                     long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                     const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                     const float* grad_patterns, const void* grad_amplitude, const void* checkpoints, int ckpt_every,
                     void* grad_slab, void* grad_probe_k, double* grad_shifts, void* workspace,
                     size_t workspace_bytes, void* stream) {
  g_err[0] = '\0';
  int rc = check_grid(h, w, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k, mode_weights,
                      origins, shifts_px, prop_y, prop_x);
  if (rc) return rc;
  if (ckpt_every < 1) return fail(PTY_ERR_INVALID, "ckpt_every must be >= 1 (got %d)", ckpt_every);
  if ((grad_patterns == nullptr) == (grad_amplitude == nullptr))
    return fail(PTY_ERR_INVALID, "exactly one of grad_patterns and grad_amplitude must be given");
  if (pty::num_checkpoints(num_slices, ckpt_every) > 0 && !checkpoints)
    return fail(PTY_ERR_INVALID, "checkpoint buffer is null but %d checkpoints are needed",
                pty::num_checkpoints(num_slices, ckpt_every));
  if (!grad_slab) return fail(PTY_ERR_INVALID, "grad_slab must be non-null");
  const size_t need = pty::grid_block_bytes(h, w, grid_ckpt(ckpt_every, num_slices), true);
  if (!workspace || workspace_bytes < need)
    return fail(PTY_ERR_INVALID, "workspace of %zu bytes is below the %zu bytes one position needs", workspace_bytes,
                need);
  pty::GridParams a = make_grid_params(h, w, num_slices, num_positions, num_modes, slab, slab_h, slab_w, probe_k,
                                       mode_weights, origins, shifts_px, prop_y, prop_x);
  a.grad_patterns = grad_patterns;
  a.grad_amplitude = (const float2*)grad_amplitude;
  a.ckpt = (float2*)checkpoints;
  a.ckpt_every = ckpt_every;
  a.grad_slab = (float2*)grad_slab;
  a.grad_probe_k = (float2*)grad_probe_k;
  a.grad_shifts = grad_shifts;
  return device_status(pty::launch_grid_adjoint(a, workspace, workspace_bytes, (hipStream_t)stream),
                       "pty_adjoint_grid");
}

}  // extern "C"
