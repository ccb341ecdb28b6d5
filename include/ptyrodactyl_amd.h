/*
 * ptyrodactyl_amd -- C-ABI of the MI355X (gfx950) multislice 4D-STEM library.
 *
 * All array arguments are DEVICE pointers (complex64 = two float32, row
 * major); `stream` is the caller's hipStream_t (NULL = legacy default
 * stream).  Calls are stream-ordered and never synchronise the device.
 * Every entry point returns PTY_OK or an error code and leaves a message in
 * pty_last_error() (thread local).  The library never frees caller memory.
 *
 * The reference (ptyrodactyl, pure JAX) has no FFI of its own; the entry
 * points below replace the XLA executables of the Python functions cited
 * beside each one (paths relative to the reference repository root).
 */
#ifndef PTYRODACTYL_AMD_H
#define PTYRODACTYL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTY_ABI_VERSION 8

#define PTY_OK 0
#define PTY_ERR_INVALID 1 /* bad argument: maps to ValueError */
#define PTY_ERR_DEVICE 2  /* HIP runtime / launch failure: maps to RuntimeError */

int pty_abi_version(void);
const char* pty_last_error(void);

/*
 * Workspace (device bytes) that pty_forward needs to process `num_positions`
 * windows of n x n (n a power of two from 32 to 4096; the batched entry points take n <= 512, virtual
 * detectors n <= 1024) with `num_modes` probe modes in one batch, capped at
 * `max_batch_waves` waves per batch (0 = no cap).  Windows n <= 128 run
 * LDS-resident and need none (returns 0).
 */
size_t pty_forward_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves);

/*
 * Multislice forward over a batch of scan positions.
 *
 * Replaces the per-position scan of `_cbed_amplitude_from_slice_provider`
 * (src/ptyrodactyl/multislice/simulations.py:501-603) as mapped over
 * positions by `stem_4d` (simulations.py:1330-1443, via `cbed_image`
 * :1130-1233 and `shift_beam_fourier` :1236-1327) and reduced by
 * `apply_distribution` (src/ptyrodactyl/multislice/reduce.py:29-89).
 *
 *   slab        complex64 [num_slices][slab_h][slab_w] transmission exp(i sigma V)
 *   probe_k     complex64 [num_modes][n][n] fft2 of each real-space probe mode
 *   mode_weights float32 [num_modes] incoherent weights (already normalised)
 *   origins     int32 [num_positions][2] window origin (row, col) in the slab;
 *               the window wraps periodically
 *   shifts_px   float64 [num_positions][2] probe shift (row, col) in pixels,
 *               applied as a Fourier phase ramp exactly as shift_beam_fourier
 *   prop_y/x    complex64 [n] separable propagator exp(-i pi lambda dz q^2)/n
 *   patterns    float32 [num_positions][n][n] fftshifted sum_m w_m |A_m|^2 (or NULL)
 *   amplitude   complex64 [num_positions][num_modes][n][n] fftshifted A_m (or NULL)
 *   workspace   device scratch of workspace_bytes (see pty_forward_workspace_bytes)
 */
int pty_forward(int n, int num_slices, long long num_positions, int num_modes, const void* slab, long long slab_h,
                long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
                const double* shifts_px, const void* prop_y, const void* prop_x, float* patterns, void* amplitude,
                void* workspace, size_t workspace_bytes, void* stream);

/*
 * Forward that also stores the wave entering slices k, 2k, ... (< S) for the
 * adjoint: checkpoints complex64 [num_positions * num_modes][(S-1)/k][n][n].
 * Same multislice as pty_forward, with the patterns and/or the amplitudes
 * (each may be NULL, not both); the reference has no checkpointing (XLA
 * keeps every slice's residuals under jax.value_and_grad,
 * src/ptyrodactyl/multislice/multislice_recon.py:1915-1951).  With k = 1 and
 * n >= 256 the checkpoints are the forward's own wave buffers (the column pass
 * writes into them), so storing every slice adds no memory traffic.
 */
int pty_forward_checkpointed(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                             long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                             const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                             float* patterns, void* amplitude, void* checkpoints, int ckpt_every, void* workspace,
                             size_t workspace_bytes, void* stream);

/*
 * pty_forward without a workspace argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference: it is this library's own addition for those callers.  The
 * streaming waves (n >= 256) live in a static device array of PTY_NOALLOC_SCRATCH_BYTES that
 * every GPU the library runs on holds until the process ends; positions are processed in
 * batches that fit it.  One position's num_modes waves must fit (n * n * 8 * num_modes bytes:
 * n <= 1024 for any mode count up to 8, n = 2048 with at most 2 modes, never n = 4096);
 * larger requests fail with PTY_ERR_INVALID.  The call allocates nothing.  All calls on one
 * GPU share that scratch, so calls that overlap on one GPU (different streams or host
 * threads without ordering) race: order them on one stream.  Results equal pty_forward's.
 */
#define PTY_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

int pty_forward_noalloc(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                        long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                        const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                        float* patterns, void* amplitude, void* stream);

/*
 * Workspace (device bytes) of pty_jvp: a primal and a tangent wave, n * n * 16 bytes, for each of the
 * num_positions * num_modes waves in one batch, capped at max_batch_waves waves (0 = no cap).
 */
size_t pty_jvp_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves);

/*
 * Forward-mode multislice: the Jacobian-vector product of pty_forward's patterns (and amplitudes).
 * Replaces the jax.jvp that jvp_operator and jtj_operator build through the forward model for the
 * Gauss-Newton and Levenberg-Marquardt solvers (src/ptyrodactyl/jacobian/operators.py:30-185,
 * solvers.py:387-521).  Every wave carries its tangent through the slices:
 * d phi_s = d psi_s T_s + psi_s d T_s, d psi_s+1 = Prop(d phi_s), dI = sum_m w_m 2 Re(conj(A_m) dA_m).
 *   d_slab      complex64 tangent of the slab (the slab's shape), or NULL (zero)
 *   d_probe_k   complex64 [num_modes][n][n] tangent of probe_k, or NULL (zero)
 *   d_shifts_px float64 [num_positions][2] tangent of the probe shifts (pixels), or NULL (zero)
 *   patterns, amplitude     primal outputs as in pty_forward (each may be NULL); pty_forward's to rounding
 *   d_patterns  float32 [num_positions][n][n] fftshifted dI, or NULL
 *   d_amplitude complex64 [num_positions][num_modes][n][n] fftshifted dA_m, or NULL (not both NULL)
 * Windows n = 32 ... 512 (powers of two); workspace of at least one position (pty_jvp_workspace_bytes).
 */
int pty_jvp(int n, int num_slices, long long num_positions, int num_modes, const void* slab, long long slab_h,
            long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
            const double* shifts_px, const void* prop_y, const void* prop_x, const void* d_slab,
            const void* d_probe_k, const double* d_shifts_px, float* patterns, void* amplitude, float* d_patterns,
            void* d_amplitude, void* workspace, size_t workspace_bytes, void* stream);

#define PTY_MAX_DETECTORS 4

/*
 * pty_forward with virtual detectors fused into the |psi|^2 epilogue: for each of num_detectors
 * 0/1 masks ([num_detectors][n][n], fftshifted pattern layout, built from the collection angles as
 * annular_detector does, src/ptyrodactyl/multislice/simulations.py:1540-1651) detector_out
 * [num_positions][num_detectors] receives the sum of the pattern intensity over the mask (FP64,
 * fixed order).  patterns may be NULL: the pattern stack is then never written.  Workspace as for
 * pty_forward.
 */
int pty_forward_detectors(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                          long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                          const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                          float* patterns, int num_detectors, const unsigned char* masks, double* detector_out,
                          void* workspace, size_t workspace_bytes, void* stream);

/*
 * Device scratch for pty_adjoint: with at least this many bytes the gradients
 * are gathered in a fixed order (bitwise reproducible).  A smaller workspace,
 * down to the scratch of one wave, runs with float32 atomics instead.
 */
size_t pty_adjoint_workspace_bytes(int n, int num_slices, int ckpt_every, long long num_positions, int num_modes);

/*
 * Reverse-mode multislice: given dL/dI (float32 [num_positions][n][n], the
 * fftshifted pattern layout), ACCUMULATES into grad_slab (complex64, the
 * slab's shape) and grad_probe_k (complex64 [num_modes][n][n], may be NULL)
 * the gradients dL/dRe + i dL/dIm w.r.t. the transmission slab and probe_k.
 * Replaces the VJP that jax.value_and_grad builds through stem_4d in the
 * reconstructors (src/ptyrodactyl/multislice/multislice_recon.py:406-434,
 * 849-885, 1400-1436, 1915-1951).  Each k-slice segment is recomputed from
 * the checkpoints of pty_forward_checkpointed.  Each wave's contributions are
 * stored and then summed per slab pixel (and probe element) in wave order,
 * so the gradients are bitwise reproducible when the workspace is at least
 * pty_adjoint_workspace_bytes; a smaller one falls back to float32 atomics.
 * grad_shifts (float64 [num_positions][2] as (y, x), may be NULL) ACCUMULATES
 * dL/d(shifts_px), the gradient w.r.t. the probe positions that the
 * position-correcting reconstructors optimise
 * (src/ptyrodactyl/multislice/multislice_recon.py:849-885, 1915-1951); it is
 * summed in FP64 in a fixed order from the stored probe contributions and
 * therefore needs the full pty_adjoint_workspace_bytes.
 */
int pty_adjoint(int n, int num_slices, long long num_positions, int num_modes, const void* slab, long long slab_h,
                long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
                const double* shifts_px, const void* prop_y, const void* prop_x, const float* grad_patterns,
                const void* checkpoints, int ckpt_every, void* grad_slab, void* grad_probe_k, double* grad_shifts,
                void* workspace, size_t workspace_bytes, void* stream);

/*
 * pty_adjoint seeded by the gradient of the complex amplitudes instead of the intensities:
 * grad_amplitude is complex64 [num_positions][num_modes][n][n] in the fftshifted layout of
 * pty_forward's amplitude output, in the dL/dRe + i dL/dIm convention.  The VJP of
 * cbed_amplitude (src/ptyrodactyl/multislice/simulations.py:606-679) under jax.grad.  Same
 * workspace and accumulation rules as pty_adjoint.
 */
int pty_adjoint_amplitude(int n, int num_slices, long long num_positions, int num_modes, const void* slab,
                          long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                          const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                          const void* grad_amplitude, const void* checkpoints, int ckpt_every, void* grad_slab,
                          void* grad_probe_k, double* grad_shifts, void* workspace, size_t workspace_bytes,
                          void* stream);

/*
 * Transmission T = exp(i sigma V) from a float64 potential (any shape, `count`
 * values), written complex64; replaces transmission_func
 * (src/ptyrodactyl/multislice/simulations.py:105-159) for a whole slab at once.
 */
int pty_transmission(const double* potential, double sigma, void* transmission, long long count, void* stream);

/* FP64 scratch for pty_sqrt_intensity_loss (one partial sum per workgroup of its fixed grid). */
size_t pty_sqrt_intensity_loss_workspace_bytes(void);

/*
 * The square-root-intensity loss of the gradient workloads and its gradient, in one pass over `count` pixels:
 *   loss = sum (sqrt(I + eps) - sqrt(I_meas))^2        (FP64, fixed summation order; loss may be NULL)
 *   grad = (sqrt(I + eps) - sqrt(I_meas)) / sqrt(I + eps)   (float32, may be NULL)
 * The reference's reconstructors reduce MAE / MSE / RMSE losses (_reduce_loss,
 * src/ptyrodactyl/multislice/multislice_recon.py:127-163); this sqrt-intensity loss has no counterpart there.
 * It replaces the elementwise tensor evaluation of that loss between the forward and the adjoint of a gradient
 * step.  patterns, measured and grad are device float32 arrays; loss is one device double; workspace holds
 * pty_sqrt_intensity_loss_workspace_bytes() bytes.
 */
int pty_sqrt_intensity_loss(long long count, const float* patterns, const float* measured, float eps, float* grad,
                            double* loss, void* workspace, size_t workspace_bytes, void* stream);

#define PTY_MAX_NORMAL_PARAMS 21

/* FP64 scratch for pty_normal_equations. */
size_t pty_normal_equations_workspace_bytes(long long num_positions, int n, int num_params);

/*
 * Gauss-Newton / Levenberg-Marquardt normal equations in FP64 for Q =
 * num_params real parameters of the probe.  amplitude is complex64
 * [num_positions][1 + Q][n][n] as written by pty_forward with 1 + Q modes
 * (mode 0 the probe, mode j the probe's derivative w.r.t. parameter j; the
 * multislice is linear in the probe), data float32 [num_positions][n][n].
 * With r = |A_0|^2 - data and J_j = 2 Re(conj(A_0) A_j):
 *   jtj = J^T J (Q x Q), jtr = J^T r (Q), rss = sum r^2   (all FP64, device)
 * The reference forms these implicitly through jvp/vjp
 * (src/ptyrodactyl/jacobian/operators.py:29-185) inside
 * levenberg_marquardt_step (src/ptyrodactyl/jacobian/solvers.py:387-521).
 * Reduction order is fixed: results are bitwise reproducible.
 */
int pty_normal_equations(const void* amplitude, const float* data, long long num_positions, int n, int num_params,
                         void* workspace, size_t workspace_bytes, double* jtj, double* jtr, double* rss,
                         void* stream);

/*
 * Damped SPD solve in FP64 on the device: (matrix + damping I) solution = rhs
 * (Cholesky, num_params <= 64), plus model_reduction = <x, rhs> - 1/2 <x, matrix x>,
 * the predicted reduction of the LM gain ratio (solvers.py:480-489).
 * status = 1 if the damped matrix was positive definite, else 0.
 */
int pty_damped_solve(const double* matrix, const double* rhs, double damping, int num_params, double* solution,
                     double* model_reduction, int* status, void* stream);

/*
 * Per-position descriptor of a batched launch that mixes scan positions of
 * several requests with the same window n and slice count S (one probe mode
 * each).  Offsets are element offsets into the arenas passed alongside.
 */
typedef struct pty_position_desc {
  long long slab_offset;    /* complex64: the request's slab [S][slab_h][slab_w] in slab_arena */
  long long slab_h, slab_w; /* the request's slab extent (each >= n) */
  long long pattern_offset; /* float32: this position's [n][n] pattern (and dL/dI) in the pattern arena */
  long long probe_offset;   /* complex64: the request's probe_k [n][n] in probe_arena */
  long long prop_offset;    /* complex64: the request's prop_y[n] followed by prop_x[n] in prop_arena */
  long long grad_offset;    /* complex64: the request's slab gradient in grad_slab_arena (adjoint only) */
  long long gprobe_offset;  /* complex64: the request's probe_k gradient in grad_probe_arena (adjoint only) */
  int32_t origin_y, origin_x;
  double shift_y, shift_x;  /* probe shift in pixels */
} pty_position_desc;

/*
 * pty_forward over positions of several requests in ONE launch sequence
 * (continuous batching).  descs: DEVICE array [num_positions]; the caller
 * validates offsets against its arenas.  amplitudes (optional, complex64
 * [num_positions][n][n], fftshifted) and checkpoints (optional, k = ckpt_every)
 * are indexed by the call-local position as in pty_forward_checkpointed.
 * Each position's result is bitwise identical to running its request alone.
 * The ensemble-axis path of cbed_image / stem_4d
 * (src/ptyrodactyl/multislice/simulations.py:769-887, 1130-1233) runs one
 * descriptor per (position, ensemble sample) through this entry point.
 */
int pty_forward_batched(int n, int num_slices, long long num_positions, const pty_position_desc* descs,
                        const void* slab_arena, const void* probe_arena, const void* prop_arena, float* pattern_arena,
                        void* amplitudes, void* checkpoints, int ckpt_every, void* workspace, size_t workspace_bytes,
                        void* stream);

/* pty_adjoint for a batched launch: gradients accumulate into each request's slice of the gradient arenas. */
int pty_adjoint_batched(int n, int num_slices, long long num_positions, const pty_position_desc* descs,
                        const void* slab_arena, const void* probe_arena, const void* prop_arena,
                        const float* grad_pattern_arena, const void* checkpoints, int ckpt_every,
                        void* grad_slab_arena, void* grad_probe_arena, void* workspace, size_t workspace_bytes,
                        void* stream);

/*
 * Virtual annular detector over stored patterns: out[p] = sum_i mask[i] * patterns[p][i] in FP64
 * (fixed summation order, bitwise reproducible).  mask is 0/1 per pixel of the fftshifted pattern;
 * the caller builds it from the collection angles exactly as annular_detector does
 * (src/ptyrodactyl/multislice/simulations.py:1540-1651).
 */
int pty_annular_integrate(const float* patterns, long long num_patterns, long long pixels_per_pattern,
                          const unsigned char* mask, double* out, void* stream);

/*
 * One in-place step of the Optax update rules the reconstructors use
 * (optax 0.2.3 as configured in src/ptyrodactyl/multislice/multislice_recon.py:70-83,
 * applied by _apply_optimizer_step :166-210 and the position updates :914-1000):
 *   PTY_OPTAX_ADAM     adam(lr): b1 0.9, b2 0.999, eps 1e-8
 *   PTY_OPTAX_ADAGRAD  adagrad(lr, initial_accumulator_value=0, eps=1e-8)
 *   PTY_OPTAX_RMSPROP  rmsprop(lr, decay=0.9, eps=1e-8, initial_scale=0)
 * params, grads and mu hold `count` elements of float32 (fp64 = 0) or float64 (fp64 = 1),
 * interleaved (re, im) when is_complex; nu holds `count` real second moments (|g|^2 for complex
 * elements, as Optax's abs_sq); mu is used by adam only.  grads is torch's dL/dRe + i dL/dIm,
 * which equals the conj(jax.grad) the reference passes to Optax.  step is the 1-based count
 * after this update (adam's bias correction).  Moments start at zero (optimizer init).
 */
#define PTY_OPTAX_ADAM 0
#define PTY_OPTAX_ADAGRAD 1
#define PTY_OPTAX_RMSPROP 2
int pty_optimizer_update(int rule, int fp64, int is_complex, long long count, void* params, const void* grads,
                         void* mu, void* nu, long long step, double learning_rate, void* stream);

/*
 * count standard normal float64 variates equal to jax.random.normal(key, (count,), float64) for the
 * raw key (key0, key1) (partitionable Threefry-2x32 counters, 64-bit uniforms, sqrt(2) erfinv):
 * the random mode basis of decompose_beam_to_modes (src/ptyrodactyl/multislice/simulations.py:1446-1536).
 */
int pty_random_normal(unsigned key0, unsigned key1, long long count, double* out, void* stream);

/*
 * Independent-atom potentials, FP64.  A coefficient row is 12 doubles:
 * PTY_PARAM_LOBATO    a1, b1, ..., a5, b5 (last two unused) — Lobato & Van Dyck (2014);
 * PTY_PARAM_KIRKLAND  a1, b1, a2, b2, a3, b3, c1, d1, c2, d2, c3, d3 — Kirkland (legacy 0.5292 * 14.4 scale).
 * Bessel K0/K1 are evaluated as the reference evaluates them (Abramowitz & Stegun 9.8.5-9.8.8).
 */
#define PTY_PARAM_LOBATO 0
#define PTY_PARAM_KIRKLAND 1

/*
 * Projected potential (V A) of one element at `count` radii r (A):
 * lobato_projected_potential / kirkland_projected_potential
 * (src/ptyrodactyl/multislice/form_factors.py:323-439).
 */
int pty_projected_potential(int parameterization, const double* row, long long count, const double* r, double* out,
                            void* stream);

/*
 * Potentials of num_types elements (rows [T][12]) on a height x width grid of pixel_size A, each pixel the mean of
 * supersampling^2 samples of the supersampled grid whose origin is sample (height*ss/2, width*ss/2); element t is
 * centred at centers[t] = (x, y) A (NULL: the origin).  out [T][height][width].  single_atom_potential
 * (src/ptyrodactyl/multislice/atom_potentials.py:588-681) for several elements in one launch.
 */
int pty_atom_potential_grid(int parameterization, int num_types, const double* rows, double pixel_size, int height,
                            int width, int supersampling, const double* centers, double* out, void* stream);

/*
 * Centre gradient of pty_atom_potential_grid: grad_centers[t] = (dL/dx_t, dL/dy_t) for L = sum_ij grad_out[t][i][j]
 * out[t][i][j] (overwritten, not accumulated), from dV/dr by the Bessel identities K0' = -K1 and (x K1)' = -x K0 on
 * the same A&S evaluations.  The reference takes this gradient with jax.grad of single_atom_potential in
 * center_coords (src/ptyrodactyl/multislice/atom_potentials.py:588-681, tested in
 * tests/test_ptyrodactyl/test_multislice/test_atom_potentials.py:91-151).  The sum has a fixed order (per-block
 * partial sums in `workspace`, at least pty_atom_center_grad_workspace_bytes(num_types) bytes, added in order).
 */
size_t pty_atom_center_grad_workspace_bytes(int num_types);
int pty_atom_potential_grid_center_grad(int parameterization, int num_types, const double* rows, double pixel_size,
                                        int height, int width, int supersampling, const double* centers,
                                        const double* grad_out, double* grad_centers, void* workspace,
                                        size_t workspace_bytes, void* stream);

/*
 * Fourier-space potential slices: for every slice s,
 *   out_k[s] = sum_t kernel_k[t] * sum_{a in [segments[s*T+t], segments[s*T+t+1])} exp(2 pi i (kx sx_a + ky sy_a))
 * with (sx_a, sy_a) = shifts[a] in pixels and kx, ky numpy's fftfreq(width), fftfreq(height).  Atoms are sorted by
 * slice, then element; segments has num_slices * num_types + 1 offsets into the num_atoms shifts (values are
 * clamped to [0, num_atoms]).  kernel_k [T][H][W] and out_k [S][H][W] are complex128.  Re ifft2(out_k[s]) is the
 * reference's slice: the per-atom Fourier shift of each element's centred potential in _process_all_slices
 * (src/ptyrodactyl/multislice/atom_potentials.py:826-993), summed by linearity.
 */
int pty_potential_slices_fourier(int num_slices, int num_types, int height, int width, long long num_atoms,
                                 const double* shifts, const long long* segments, const void* kernel_k, void* out_k,
                                 void* stream);

/*
 * Shift gradient of pty_potential_slices_fourier.  adj_k [S][H][W] complex128 is ifft2 (numpy's normalisation) of
 * dL/dV_s, the loss gradient w.r.t. the real slice Re ifft2(out_k[s]) (zero where the slice is cropped away).  For
 * every atom a of the given slices (element t, slice s):
 *   grad_shifts[a] = Re sum_k 2 pi i kernel_k[t](k) adj_k[s](k) (kx, ky) exp(2 pi i (kx sx_a + ky sy_a)),
 * i.e. (dL/dsx_a, dL/dsy_a) in pixels; entries of atoms outside the slices are left untouched.  The reference gets
 * this gradient from jax.grad through _process_all_slices (src/ptyrodactyl/multislice/atom_potentials.py:826-993).
 * Per-tile partial sums go to `workspace` (pty_potential_shift_grad_workspace_bytes) and are added in tile order.
 */
size_t pty_potential_shift_grad_workspace_bytes(int height, int width, long long num_atoms);
int pty_potential_slices_shift_grad(int num_slices, int num_types, int height, int width, long long num_atoms,
                                    const double* shifts, const long long* segments, const void* kernel_k,
                                    const void* adj_k, double* grad_shifts, void* workspace, size_t workspace_bytes,
                                    void* stream);

/* Device bytes of the state block and workspace of pty_forward_persistent. */
size_t pty_persistent_state_bytes(void);
size_t pty_persistent_workspace_bytes(int n, int group_waves);

/*
 * Persistent variant of pty_forward for n = 256, one probe mode: one launch
 * runs every slice; groups of `group_waves` waves stay in one XCD's L2 while
 * that XCD's workgroups take row and column tasks from a per-XCD ticket
 * counter.  `grid` workgroups of 256 threads (2 per CU recommended).  `state`
 * (pty_persistent_state_bytes, reset by the call) holds at int index 0 an
 * error flag (bit 0: a wait timed out; bit 1: an XCD left tasks undone),
 * then the per-XCD counters.  group_waves is at most 16.
 * Results are bitwise equal to pty_forward.  Same multislice as pty_forward
 * (src/ptyrodactyl/multislice/simulations.py:501-603).
 */
int pty_forward_persistent(int n, int num_slices, long long num_positions, const void* slab, long long slab_h,
                           long long slab_w, const void* probe_k, const float* mode_weights, const int32_t* origins,
                           const double* shifts_px, const void* prop_y, const void* prop_x, float* patterns,
                           int group_waves, int grid, void* workspace, size_t workspace_bytes, int* state,
                           void* stream);

/*
 * Grids of any size: h x w probe windows with 1 <= h, w <= 4096, square or not and of any
 * factorisation (odd, prime, 15 x 32, 100 x 100, ...).  The reference's multislice takes any
 * grid shape (src/ptyrodactyl/multislice/simulations.py:501-603 with jnp.fft.fft2; its tests
 * use 8 x 8, 15 x 32 and 31 x 32 grids); the power-of-two entry points above cover only the
 * square power-of-two windows.  Grids below 2^14 pixels run one workgroup per position, with a
 * mixed-radix transform in LDS, and their gradients accumulate with float32 atomics (not bitwise
 * reproducible across runs).  Larger grids run one launch per pass (streaming); there the slab
 * gradient is stored per wave and gathered in a fixed order (bitwise reproducible), and the probe
 * gradient uses float32 atomics.
 *
 * Device scratch for pty_forward_grid (adjoint = 0) or pty_adjoint_grid (adjoint = 1) with the
 * default number of workgroups; any workspace of at least the bytes of one workgroup works
 * (h * w * 8 for the forward, h * w * 8 * (min(ckpt_every, num_slices) + 2) for the adjoint),
 * with fewer positions in flight (the streaming adjoint needs one more wave of h * w * 8 per
 * position and mode, for the slab-gradient contributions).  Pass num_positions * num_modes: a forward without a pattern
 * output runs every (position, mode) wave as its own task, and the streaming kernels (grids of
 * 2^14 pixels or more) keep all num_modes waves of a position in flight, so a workspace sized
 * for fewer waves than num_modes falls back to one workgroup per position.
 */
size_t pty_grid_workspace_bytes(int h, int w, int num_slices, int ckpt_every, long long num_positions, int adjoint);

/*
 * pty_forward / pty_forward_checkpointed on an h x w grid (same reference functions,
 * simulations.py:501-603, 1130-1443).  Arrays as for pty_forward with n x n replaced by h x w;
 * prop_y has h entries (exp(-i pi lambda dz qy^2) / h), prop_x has w (with / w).  patterns
 * and amplitude may each be NULL (not both).  checkpoints (may be NULL): complex64
 * [num_positions * num_modes][(S-1)/ckpt_every][h][w], the wave entering slices k, 2k, ...
 */
int pty_forward_grid(int h, int w, int num_slices, long long num_positions, int num_modes, const void* slab,
                     long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                     const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                     float* patterns, void* amplitude, void* checkpoints, int ckpt_every, void* workspace,
                     size_t workspace_bytes, void* stream);

/*
 * pty_adjoint / pty_adjoint_amplitude on an h x w grid (the VJP of the same functions,
 * src/ptyrodactyl/multislice/multislice_recon.py:406-434, 849-885, 1400-1436, 1915-1951).
 * Exactly one of grad_patterns (float32 [P][h][w]) and grad_amplitude (complex64 [P][M][h][w]),
 * both fftshifted, seeds it.  ACCUMULATES into grad_slab, grad_probe_k (may be NULL) and
 * grad_shifts (float64 [P][2], may be NULL; summed per position in FP64 in a fixed order).
 */
int pty_adjoint_grid(int h, int w, int num_slices, long long num_positions, int num_modes, const void* slab,
                     long long slab_h, long long slab_w, const void* probe_k, const float* mode_weights,
                     const int32_t* origins, const double* shifts_px, const void* prop_y, const void* prop_x,
                     const float* grad_patterns, const void* grad_amplitude, const void* checkpoints, int ckpt_every,
                     void* grad_slab, void* grad_probe_k, double* grad_shifts, void* workspace,
                     size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PTYRODACTYL_AMD_H */
