// Multislice 4D-STEM forward kernels for gfx950.
//
// One "wave" below is one probe mode at one scan position.  For every wave
// the kernels compute, in complex64 with FP32 accumulation,
//   psi_0     = ifft2(fft2(probe) * ramp(shift))
//   phi_s     = psi_s * T_s(window at origin)         s = 0..S-1
//   psi_{s+1} = ifft2(fft2(phi_s) * P)                s < S-1
//   I        += w_m |fftshift(fft2(phi_{S-1}))|^2
// which is the scan body of the reference's _cbed_amplitude_from_slice_provider
// (src/ptyrodactyl/multislice/simulations.py:543-602) applied to the shifted
// probe of shift_beam_fourier (simulations.py:1287-1321) and reduced as in
// apply_distribution (multislice/reduce.py:81-86).
//
// The Fresnel propagator exp(-i pi lambda dz (qx^2 + qy^2)) factors into
// Py(qy) Px(qx), so fft2 -> *P -> ifft2 is applied as a row operator
// (fft_x, *Px, ifft_x) and a column operator (fft_y, *Py, ifft_y).
//
// Two execution strategies:
//  * n <= 128 (32, 64, 128): one workgroup per scan position keeps the whole wave resident
//    in LDS across all slices (128^2 c64 = 128 KiB of the 160 KiB LDS) and
//    streams only the transmission window from HBM; the transmission multiply
//    is fused into the row phase and |.|^2 into the last column phase.
//  * n >= 256: the wave exceeds LDS, so the slices run as passes over a batch
//    of waves held in a device workspace sized to stay in the L2/Infinity-Cache
//    working set.  The passes alternate axis and each applies two slices'
//    operators along its axis, so a slice costs one trip of the wave through
//    memory.
#include <stdlib.h>

#include "common.hpp"
#include "forward.hpp"

namespace pty {

// ---------------------------------------------------------------------------
// LDS-resident kernel (n = 64, 128)
// ---------------------------------------------------------------------------

// The resident wave keeps row y at LDS row pad16(y): one spare row after every 16 spreads the
// column exchange's stride-16 accesses over the banks, and every row/column address is a
// per-thread base plus a compile-time offset (no per-element index arithmetic).
// Column x of the resident wave used as FFT exchange scratch.
struct ColBufResident {
  float2* col;
  int rs;
  __device__ __forceinline__ float2 ld(int e) const { return col[pad16(e) * rs]; }
  __device__ __forceinline__ void st(int e, float2 v) const { col[pad16(e) * rs] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return col[pe * rs]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { col[pe * rs] = v; }
  __device__ __forceinline__ void sync() const { wave_sync(); }
};

template <int N, bool DET>
__global__ __launch_bounds__(N* N / 16) void k_forward_resident(FwdParams a) {
  constexpr int T = N / 16;       // threads per line
  constexpr int NT = N * N / 16;  // threads per workgroup
  constexpr int RS = N + N / 16;  // padded row stride (complex elements)
  constexpr int G = 64 / T;       // lines per wavefront
  extern __shared__ float2 lds[];
  float2* wave = lds;
  float2* tw = wave + (N + N / 16) * RS;
  float2* py = tw + N;
  float2* px = py + N;
  float2* ramp_y = px + N;
  float2* ramp_x = ramp_y + N;
  float2* twf = ramp_x + N;  // modulus-correcting low parts of tw, py, px
  float2* pyf = twf + N;
  float2* pxf = pyf + N;
  double* red = (double*)(pxf + N);  // NT / 64 doubles for the detector sums

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int p = blockIdx.x;
  const WaveCtx g = resolve_wave(a, p);
  fill_twiddles(tw, twf, N, tid, NT);
  stage_factors(py, pyf, g.prop_y, N, tid, NT);
  stage_factors(px, pxf, g.prop_x, N, tid, NT);
  for (int i = tid; i < N; i += NT) {
    ramp_y[i] = shift_ramp(i, N, g.my, g.dy);
    ramp_x[i] = shift_ramp(i, N, g.mx, g.dx);
  }
  const TwTable twt{tw, twf};
  const FixTable pyt{py, pyf}, pxt{px, pxf};
  __syncthreads();

  for (int m = 0; m < a.M; ++m) {
    for (int s = -1; s < a.S; ++s) {
      // Thread coordinates are re-derived every iteration through an opaque
      // copy so the compiler keeps per-thread LDS addresses local to each
      // phase instead of hoisting dozens of them across the loop (spills).
      int tid_ = tid;
      asm volatile("" : "+v"(tid_));
      const int lane_ = tid_ & 63, wid_ = tid_ >> 6;
      if (s >= 0) {  // row phase of slice s
        const int rt = lane_ % T, ry = wid_ * G + lane_ / T;
        const RowBufWave rb{wave + pad16(ry) * RS};
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = rb.row[pad16(rt + T * q)];
        if (a.ckpt && s > 0 && s % a.ckpt_every == 0) {  // checkpoint the wave entering slice s
          const int nc = num_checkpoints(a.S, a.ckpt_every);
          float2* ck = a.ckpt + (((size_t)p * a.M + m) * nc + (s / a.ckpt_every - 1)) * N * N + (size_t)ry * N;
#pragma unroll
          for (int q = 0; q < 16; ++q) ck[rt + T * q] = v[q];
        }
        if (s == 0) line_fft<N, true>(v, rt, rb, twt);  // finish the shifted probe: ifft along x
        int gy = g.oy + ry;
        if (gy >= g.slab_h) gy -= (int)g.slab_h;
        const float2* trow = g.slab + (long long)s * g.stride + (long long)gy * g.slab_w;
        multiply_window_row<N>(v, trow, g.ox, rt, (int)g.slab_w);
        line_fft<N, false>(v, rt, rb, twt);
        if (s < a.S - 1) {
#pragma unroll
          for (int q = 0; q < 16; ++q) v[q] = pxt.mul(v[q], rt + T * q);
          line_fft<N, true>(v, rt, rb, twt);
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 16; ++q) rb.row[pad16(rt + T * q)] = v[q];
        __syncthreads();
      }
      {  // column phase
        const int ct = lane_ / G, cx = wid_ * G + lane_ % G;
        const ColBufResident cb{wave + pad16(cx), RS};
        float2* col = wave + pad16(cx);
        float2 v[16];
        if (s < 0) {  // probe_k column * shift ramp, inverse FFT along y
          const float2 rx = ramp_x[cx];
          const float2* pk = g.probe_k + (size_t)m * N * N + cx;
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int ky = ct + T * q;
            v[q] = cmul(cmul(pk[(size_t)ky * N], ramp_y[ky]), rx);
          }
          line_fft<N, true>(v, ct, cb, twt);
        } else {
#pragma unroll
          for (int q = 0; q < 16; ++q) v[q] = col[pad16(ct + T * q) * RS];
          line_fft<N, false>(v, ct, cb, twt);
          if (s < a.S - 1) {
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = pyt.mul(v[q], ct + T * q);
            line_fft<N, true>(v, ct, cb, twt);
          }
        }
        if (s < a.S - 1) {
          wave_sync();
#pragma unroll
          for (int q = 0; q < 16; ++q) col[pad16(ct + T * q) * RS] = v[q];
        } else {
          const int xs = (cx + N / 2) & (N - 1);
          const float wm = (a.weights ? a.weights[m] : 1.f);
          if (a.patterns) {
            float* out = a.patterns + g.pat_off + xs;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
              float* o = out + (size_t)((ct + T * q + N / 2) & (N - 1)) * N;
              *o = (m == 0 ? 0.f : *o) + wm * cabs2(v[q]);  // same thread owns the pixel for every mode
            }
          }
          if (a.amplitude) {
            float2* amp = a.amplitude + ((size_t)p * a.M + m) * N * N + xs;
#pragma unroll
            for (int q = 0; q < 16; ++q) amp[(size_t)((ct + T * q + N / 2) & (N - 1)) * N] = v[q];
          }
          // virtual detectors: the masked sum is linear in the intensity, so each mode's w |A|^2 is
          // summed and added to the position's signal in mode order
          for (int d = 0; DET && d < a.det_count; ++d) {
            const unsigned char* mk = a.det_masks + (size_t)d * N * N + xs;
            double part = 0.0;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
              const float val = wm * cabs2(v[q]);
              part += mk[(size_t)((ct + T * q + N / 2) & (N - 1)) * N] ? (double)val : 0.0;
            }
            const double total = block_sum_fp64<NT>(part, red);
            if (tid == 0) {
              double* o = a.det_out + (size_t)p * a.det_count + d;
              *o = (m == 0 ? 0.0 : *o) + total;
            }
          }
        }
        __syncthreads();
      }
    }
  }
}

template <int N>
static size_t resident_lds_bytes() {
  return sizeof(float2) * ((size_t)(N + N / 16) * (N + N / 16) + 8 * N) +
         sizeof(double) * (N * N / 16 / 64);
}

// ---------------------------------------------------------------------------
// Streaming passes (n = 256 ... 4096)
// ---------------------------------------------------------------------------
//
// The propagator's row operator Rx = ifft_x Px fft_x and column operator Cy = ifft_y Py fft_y act on different
// axes, so they commute and slice s may apply them in either order: psi_{s+1} = Rx Cy (psi_s T_s) = Cy Rx (psi_s T_s).
// The passes therefore alternate axis, and each pass applies everything along its own axis between two trips of
// the wave through memory: the operator that completes slice s - 1, the transmission T_s, and the first operator
// of slice s.  A slice costs one read and one write of the wave instead of two of each.
//
//   pass 0            probe_k x shift ramp, inverse transform along the first axis
//   pass 1 (s = 0)    inverse transform along this axis (psi_0 complete), x T_0, first operator of slice 0
//   pass s + 1        second operator of slice s - 1 (psi_s complete), checkpoint, x T_s, first operator of slice s
//   last slice        ... x T_{S-1}, forward transform along this axis (no propagation after the last slice)
//   pass S + 1        forward transform along columns, |.|^2 and the fftshift store
//
// The first axis is chosen from the parity of S so that pass S + 1 always runs on columns.

enum : int { PH_INIT = 1, PH_FIRST = 2, PH_START = 4 };

// Window column x of slice s times the 16 points a thread holds of it (rows t + (N / 16) q).  A window whose rows
// do not wrap (uniform per workgroup) reads the slice through a buffer view, one lane offset plus a uniform row
// offset per point; a slice of 4 GiB or more, or a window that wraps, computes each address.
template <int N>
__device__ __forceinline__ void multiply_window_col(float2 (&v)[16], const WaveCtx& g, int s, int x, int t) {
  constexpr int T = N / 16;
  int gx = g.ox + x;
  if (gx >= g.slab_w) gx -= (int)g.slab_w;
  const float2* slice = g.slab + (long long)s * g.stride;
  const long long w = g.slab_w;
  const unsigned long long slice_bytes = (unsigned long long)g.stride * sizeof(float2);
  if (g.oy + N <= g.slab_h && slice_bytes <= 0xFFFFFFFFull) {
    const BufView tv(slice, slice_bytes);
    const unsigned lo = (unsigned)(((long long)(g.oy + t) * w + gx) * sizeof(float2));
    const unsigned qs = (unsigned)(T * w * sizeof(float2));
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], tv.ld2(lo, qs * q));
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      int gy = g.oy + t + T * q;
      if (gy >= g.slab_h) gy -= (int)g.slab_h;
      v[q] = cmul(v[q], slice[(long long)gy * w + gx]);
    }
  }
}

// Row pass: each group of N / 16 threads holds one row of a wave (v[q] = row[t + (N / 16) q]).
template <int N, int F>
__global__ __launch_bounds__(256) void k_row_phase(FwdParams a, float2* waves, int w0, int s) {
  constexpr int T = N / 16;
  constexpr int RPB = 256 / T;  // rows per workgroup
  constexpr int RS = N + N / 16;
  constexpr bool INIT = F & PH_INIT, FIRST = F & PH_FIRST, START = F & PH_START;
  constexpr bool PROP = !INIT && (!FIRST || START);
  __shared__ float2 rows[RPB * RS];
  constexpr bool DF = N <= DF_MAX_LINE;
  __shared__ float2 tw[twiddle_entries<N>()];
  __shared__ float2 px[PROP ? N : 1];
  __shared__ float2 twf[DF ? twiddle_entries<N>() : 1];
  __shared__ float2 pxf[DF && PROP ? N : 1];
  const int tid = threadIdx.x;
  // grid (nb, N / RPB): consecutive blocks -> different waves (XCD round robin)
  const int wl = blockIdx.x;
  const int w = w0 + wl;
  const int p = a.M == 1 ? w : w / a.M;
  const int t = tid % T, lr = tid / T;
  const int y = blockIdx.y * RPB + lr;
  const WaveCtx g = resolve_wave<INIT>(a, p);
  stage_twiddles<N, DF>(tw, twf, tid, 256);
  if constexpr (PROP) stage_factors<DF>(px, pxf, g.prop_x, N, tid, 256);
  const TwTab<DF> twt{tw, twf};
  const FixTab<DF> pxt{px, pxf};
  const typename RowBufFor<(T <= 64)>::type rb{rows + lr * RS};
  __syncthreads();

  float2* wrow = waves + ((size_t)wl * N + y) * N;
  float2 v[16];
  if constexpr (INIT) {  // probe_k row * shift ramp, inverse FFT along x
    const float2 ry = shift_ramp(y, N, g.my, g.dy);
    const float2* pk = g.probe_k + (size_t)(w % a.M) * N * N + (size_t)y * N;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int kx = t + T * q;
      v[q] = cmul(cmul(pk[kx], shift_ramp(kx, N, g.mx, g.dx)), ry);
    }
    line_fft<N, true>(v, t, rb, twt);
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = wrow[t + T * q];
    if constexpr (FIRST) {
      line_fft<N, true>(v, t, rb, twt);  // completes psi_0
    } else {  // second operator of slice s - 1
      line_fft<N, false>(v, t, rb, twt);
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = pxt.mul(v[q], t + T * q);
      line_fft<N, true>(v, t, rb, twt);
      if (a.ckpt && s % a.ckpt_every == 0) {  // checkpoint the wave entering slice s
        const int nc = num_checkpoints(a.S, a.ckpt_every);
        float2* ck = a.ckpt + ((size_t)w * nc + (s / a.ckpt_every - 1)) * N * N + (size_t)y * N;
#pragma unroll
        for (int q = 0; q < 16; ++q) ck[t + T * q] = v[q];
      }
    }
    int gy = g.oy + y;
    if (gy >= g.slab_h) gy -= (int)g.slab_h;
    const float2* trow = g.slab + (long long)s * g.stride + (long long)gy * g.slab_w;
    multiply_window_row<N>(v, trow, g.ox, t, (int)g.slab_w);
    line_fft<N, false>(v, t, rb, twt);
    if constexpr (START) {  // first operator of slice s
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = pxt.mul(v[q], t + T * q);
      line_fft<N, true>(v, t, rb, twt);
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) wrow[t + T * q] = v[q];
}

// Column-tile scratch: element e of column g at pad16(e) * NC + g.
template <int NC>
struct ColBufBlock {
  float2* base;
  __device__ __forceinline__ float2 ld(int e) const { return base[pad16(e) * NC]; }
  __device__ __forceinline__ void st(int e, float2 v) const { base[pad16(e) * NC] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return base[pe * NC]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { base[pe * NC] = v; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};

// Column pass: each workgroup holds NC columns of one wave, thread (gc, t) the points t + (N / 16) q of column gc.
template <int N, int F>
__global__ __launch_bounds__(256) void k_col_phase(FwdParams a, float2* waves, int w0, int nb, int s) {
  constexpr int T = N / 16;
  constexpr int NC = 256 / T;  // columns per workgroup
  constexpr int PS = N + N / 16;
  constexpr bool INIT = F & PH_INIT, FIRST = F & PH_FIRST, START = F & PH_START;
  constexpr bool PROP = !INIT && (!FIRST || START);
  __shared__ float2 buf[PS * NC];
  constexpr bool DF = N <= DF_MAX_LINE;
  __shared__ float2 tw[twiddle_entries<N>()];
  __shared__ float2 py[PROP ? N : 1];
  __shared__ float2 twf[DF ? twiddle_entries<N>() : 1];
  __shared__ float2 pyf[DF && PROP ? N : 1];
  const int tid = threadIdx.x;
  const int gc = tid % NC, t = tid / NC;
  int wl, cblk;
  if constexpr (NC > 2) {  // grid (nb, N / NC), as column_block maps it
    wl = blockIdx.x;
    cblk = blockIdx.y;
  } else {
    column_block<NC>((int)blockIdx.x, nb, N / NC, wl, cblk);
  }
  const int x = cblk * NC + gc;
  const int w = w0 + wl;
  const int p = a.M == 1 ? w : w / a.M;
  const WaveCtx g = resolve_wave<INIT>(a, p);
  stage_twiddles<N, DF>(tw, twf, tid, 256);
  if constexpr (PROP) stage_factors<DF>(py, pyf, g.prop_y, N, tid, 256);
  const TwTab<DF> twt{tw, twf};
  const FixTab<DF> pyt{py, pyf};
  const ColBufBlock<NC> cb{buf + gc};
  __syncthreads();

  // point q of the thread's column sits lo + q * QS bytes into an n x n array
  constexpr unsigned QS = (unsigned)(T * N * sizeof(float2));
  constexpr unsigned long long NN8 = (unsigned long long)N * N * sizeof(float2);
  const unsigned lo = (unsigned)((t * N + x) * sizeof(float2));
  const BufView wv(waves + (size_t)wl * N * N, NN8);
  float2 v[16];
  if constexpr (INIT) {  // probe_k column * shift ramp, inverse FFT along y
    const float2 rx = shift_ramp(x, N, g.mx, g.dx);
    const BufView pk(g.probe_k + (size_t)(w % a.M) * N * N, NN8);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = cmul(cmul(pk.ld2(lo, QS * q), shift_ramp(t + T * q, N, g.my, g.dy)), rx);
    line_fft<N, true>(v, t, cb, twt);
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = wv.ld2(lo, QS * q);
    if constexpr (FIRST) {
      line_fft<N, true>(v, t, cb, twt);  // completes psi_0
    } else {  // second operator of slice s - 1
      line_fft<N, false>(v, t, cb, twt);
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = pyt.mul(v[q], t + T * q);
      line_fft<N, true>(v, t, cb, twt);
      if (a.ckpt && s % a.ckpt_every == 0) {  // checkpoint the wave entering slice s
        const int nc = num_checkpoints(a.S, a.ckpt_every);
        const BufView ck(a.ckpt + ((size_t)w * nc + (s / a.ckpt_every - 1)) * N * N, NN8);
#pragma unroll
        for (int q = 0; q < 16; ++q) ck.st2(v[q], lo, QS * q);
      }
    }
    multiply_window_col<N>(v, g, s, x, t);
    line_fft<N, false>(v, t, cb, twt);
    if constexpr (START) {  // first operator of slice s
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = pyt.mul(v[q], t + T * q);
      line_fft<N, true>(v, t, cb, twt);
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) wv.st2(v[q], lo, QS * q);
}

// Last pass: the forward transform along columns, |.|^2 summed over the modes of a position with their weights,
// and the fftshifted store.  One workgroup runs every mode of its position, so each pixel has one writer.
template <int N, bool DET>
__global__ __launch_bounds__(256) void k_col_final(FwdParams a, float2* waves, int w0, int np) {
  constexpr int T = N / 16;
  constexpr int NC = 256 / T;
  constexpr int PS = N + N / 16;
  __shared__ float2 buf[PS * NC];
  constexpr bool DF = N <= DF_MAX_LINE;
  __shared__ float2 tw[twiddle_entries<N>()];
  __shared__ float2 twf[DF ? twiddle_entries<N>() : 1];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const int gc = tid % NC, t = tid / NC;
  int pl, cblk;  // position inside the batch, column block
  if constexpr (NC > 2) {
    pl = blockIdx.x;
    cblk = blockIdx.y;
  } else {
    column_block<NC>((int)blockIdx.x, np, N / NC, pl, cblk);
  }
  const int x = cblk * NC + gc;
  const int p = a.M == 1 ? w0 + pl : w0 / a.M + pl;
  const WaveCtx g = resolve_wave<false>(a, p);
  stage_twiddles<N, DF>(tw, twf, tid, 256);
  const TwTab<DF> twt{tw, twf};
  const ColBufBlock<NC> cb{buf + gc};
  __syncthreads();

  constexpr unsigned QS = (unsigned)(T * N * sizeof(float2));
  constexpr unsigned long long NN8 = (unsigned long long)N * N * sizeof(float2);
  const unsigned lo = (unsigned)((t * N + x) * sizeof(float2));
  const int xs = (x + N / 2) & (N - 1);
  // fftshifted row of point q: t + (N / 16) ((q + 8) mod 16)
  const unsigned lo_s = (unsigned)((t * N + xs) * sizeof(float2));
  float acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.f;
  for (int m = 0; m < a.M; ++m) {
    const BufView wv(waves + ((size_t)pl * a.M + m) * N * N, NN8);
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = wv.ld2(lo, QS * q);
    line_fft<N, false>(v, t, cb, twt);
    const float wm = (a.weights ? a.weights[m] : 1.f);
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] += wm * cabs2(v[q]);
    if (a.amplitude) {
      const BufView amp(a.amplitude + ((size_t)p * a.M + m) * N * N, NN8);
#pragma unroll
      for (int q = 0; q < 16; ++q) amp.st2(v[q], lo_s, QS * ((q + 8) & 15));
    }
  }
  if (a.patterns) {
    const BufView out(a.patterns + g.pat_off, NN8 / 2);
#pragma unroll
    for (int q = 0; q < 16; ++q) out.st1(acc[q], lo_s / 2, (QS / 2) * ((q + 8) & 15));
  }
  // virtual detectors: this tile's masked sums go to row 0, columns cblk * NC + d of the position's
  // first wave in the workspace (only this workgroup reads those columns; its reads are complete
  // after the reduction's barrier) and k_det_finalize adds the tiles in order
  for (int d = 0; DET && d < a.det_count; ++d) {
    const unsigned char* mk = a.det_masks + (size_t)d * N * N + xs;
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) part += mk[(size_t)((t + T * q + N / 2) & (N - 1)) * N] ? (double)acc[q] : 0.0;
    const double total = block_sum_fp64<256>(part, red);
    if (tid == 0) {
      double* slot = (double*)(waves + (size_t)pl * a.M * N * N + cblk * NC + d);
      *slot = total;
    }
  }
}

// Sum of the N / NC tile partials of every (position, detector) of a streaming batch, in tile order.
template <int N>
__global__ __launch_bounds__(256) void k_det_finalize(FwdParams a, const float2* waves, long long p0, int np) {
  constexpr int NC = 256 / (N / 16);
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= np * a.det_count) return;
  const int pl = idx / a.det_count, d = idx % a.det_count;
  const double* row = (const double*)(waves + (size_t)pl * a.M * N * N);
  double s = 0.0;
  for (int cb = 0; cb < N / NC; ++cb) s += row[cb * NC + d];
  a.det_out[(p0 + pl) * a.det_count + d] = s;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

template <int N, bool DET>
static hipError_t launch_resident_variant(const FwdParams& a, hipStream_t st) {
  const size_t lds = resident_lds_bytes<N>();
  static bool attr_set[64] = {false};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (dev >= 0 && dev < 64 && !attr_set[dev]) {
    hipError_t e = hipFuncSetAttribute((const void*)k_forward_resident<N, DET>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr_set[dev] = true;
  }
  hipLaunchKernelGGL((k_forward_resident<N, DET>), dim3(a.P), dim3(N * N / 16), lds, st, a);
  return hipGetLastError();
}

// The detector epilogue is a separate instantiation so the plain forward keeps its register budget.
template <int N>
static hipError_t launch_resident(const FwdParams& a, hipStream_t st) {
  return a.det_count > 0 ? launch_resident_variant<N, true>(a, st) : launch_resident_variant<N, false>(a, st);
}

// Grid of the forward column passes over `units` waves or positions: (units, N / NC) while a workgroup
// holds more than two columns, else the one-dimensional XCD-aware order of column_block.
template <int N>
static dim3 column_grid(int units) {
  constexpr int NC = 256 / (N / 16);
  return NC > 2 ? dim3(units, N / NC) : dim3(units * (N / NC));
}

// Pass of slice s on rows or columns (flags PH_FIRST when s = 0, PH_START while s < S - 1).
template <int N>
static void launch_phase(bool col, bool first, bool start, const FwdParams& a, float2* ws, int w0, int nb, int s,
                         hipStream_t st) {
  constexpr int RPB = 256 / (N / 16);
  const dim3 rg(nb, N / RPB), cg = column_grid<N>(nb), b(256);
  if (col) {
    if (first && start) hipLaunchKernelGGL((k_col_phase<N, PH_FIRST | PH_START>), cg, b, 0, st, a, ws, w0, nb, s);
    else if (first) hipLaunchKernelGGL((k_col_phase<N, PH_FIRST>), cg, b, 0, st, a, ws, w0, nb, s);
    else if (start) hipLaunchKernelGGL((k_col_phase<N, PH_START>), cg, b, 0, st, a, ws, w0, nb, s);
    else hipLaunchKernelGGL((k_col_phase<N, 0>), cg, b, 0, st, a, ws, w0, nb, s);
  } else {
    if (first && start) hipLaunchKernelGGL((k_row_phase<N, PH_FIRST | PH_START>), rg, b, 0, st, a, ws, w0, s);
    else if (first) hipLaunchKernelGGL((k_row_phase<N, PH_FIRST>), rg, b, 0, st, a, ws, w0, s);
    else if (start) hipLaunchKernelGGL((k_row_phase<N, PH_START>), rg, b, 0, st, a, ws, w0, s);
    else hipLaunchKernelGGL((k_row_phase<N, 0>), rg, b, 0, st, a, ws, w0, s);
  }
}

template <int N>
static hipError_t launch_streaming_batch(const FwdParams& a, float2* ws, long long p0, int np, hipStream_t st) {
  constexpr int RPB = 256 / (N / 16);
  const int nb = np * a.M;
  const int w0 = (int)(p0 * a.M);
  // first axis from the parity of S: pass S + 1, the final column transform, then falls on columns
  bool col = (a.S % 2) == 1;
  if (col)
    hipLaunchKernelGGL((k_col_phase<N, PH_INIT>), column_grid<N>(nb), dim3(256), 0, st, a, ws, w0, nb, 0);
  else
    hipLaunchKernelGGL((k_row_phase<N, PH_INIT>), dim3(nb, N / RPB), dim3(256), 0, st, a, ws, w0, 0);
  for (int s = 0; s < a.S; ++s) {
    col = !col;
    launch_phase<N>(col, s == 0, s < a.S - 1, a, ws, w0, nb, s, st);
  }
  if (a.det_count > 0) {
    hipLaunchKernelGGL((k_col_final<N, true>), column_grid<N>(np), dim3(256), 0, st, a, ws, w0, np);
    hipLaunchKernelGGL(k_det_finalize<N>, dim3((np * a.det_count + 255) / 256), dim3(256), 0, st, a, ws, p0, np);
  } else {
    hipLaunchKernelGGL((k_col_final<N, false>), column_grid<N>(np), dim3(256), 0, st, a, ws, w0, np);
  }
  return hipGetLastError();
}

// Extra launch lanes of the streaming forward: library streams per device (default priority) that fork
// from the caller's stream and join it again, so that several batches' passes overlap on the device.
constexpr int MAX_LANES = 4;

struct Lanes {
  hipStream_t stream[MAX_LANES] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[MAX_LANES] = {};
};

// Three lanes of 128 waves at n = 256 measured fastest (256^2 x 64 slices, 4096 positions: 87.2 ms against
// 92.4 ms for one stream of 512 waves on the same box; 2 lanes 87.0-89.6, 4 lanes 97.7-102 ms): the tail of each
// pass overlaps the next batch's passes.  Other sizes keep one lane.  PTY_FORWARD_LANES overrides.
static int forward_lanes(int n) {
  const char* v = getenv("PTY_FORWARD_LANES");
  const int k = v ? atoi(v) : (n == 256 ? 3 : 1);
  return k < 1 ? 1 : (k > MAX_LANES ? MAX_LANES : k);
}

static hipError_t lanes_for_device(Lanes** out) {
  static Lanes lanes[64];
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  Lanes& l = lanes[dev];
  if (!l.fork) {
    e = hipEventCreateWithFlags(&l.fork, hipEventDisableTiming);
    for (int i = 1; i < MAX_LANES && e == hipSuccess; ++i) {
      e = hipStreamCreateWithFlags(&l.stream[i], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&l.join[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) return e;
  }
  *out = &l;
  return hipSuccess;
}

template <int N>
static hipError_t launch_streaming(const FwdParams& a, float2* ws, size_t ws_bytes, hipStream_t st) {
  constexpr int T = N / 16;
  constexpr int NC = 256 / T;
  // a detector's tile partial goes to one of the workgroup's own NC columns
  if (a.det_count > NC) return hipErrorInvalidValue;
  const size_t wave_bytes = (size_t)N * N * sizeof(float2);
  const long long max_waves = (long long)(ws_bytes / wave_bytes);
  int lanes = forward_lanes(N);
  if (a.max_lanes > 0 && lanes > a.max_lanes) lanes = a.max_lanes;
  if (lanes > 1) {  // a stream under graph capture keeps its launches on itself
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) lanes = 1;
  }
  while (lanes > 1 && (max_waves / lanes / a.M < 16 || a.P < 16 * lanes)) --lanes;
  long long pb = (max_waves / lanes) / a.M;
  if (pb < 1) return hipErrorInvalidValue;
  if (pb > a.P) pb = a.P;
  if (pb >= 8) pb -= pb % 8;  // keep a wave on one XCD in both passes
  Lanes* ln = nullptr;
  if (lanes > 1) {
    hipError_t e = lanes_for_device(&ln);
    if (e == hipSuccess) e = hipEventRecord(ln->fork, st);
    for (int i = 1; i < lanes && e == hipSuccess; ++i) e = hipStreamWaitEvent(ln->stream[i], ln->fork, 0);
    if (e != hipSuccess) return e;
  }
  int batch = 0;
  for (long long p0 = 0; p0 < a.P; p0 += pb, ++batch) {
    const int np = (int)((a.P - p0) < pb ? (a.P - p0) : pb);
    const int l = batch % lanes;
    hipError_t e = launch_streaming_batch<N>(a, ws + (size_t)l * pb * a.M * N * N, p0, np, l ? ln->stream[l] : st);
    if (e != hipSuccess) return e;
  }
  for (int i = 1; i < lanes; ++i) {
    hipError_t e = hipEventRecord(ln->join[i], ln->stream[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, ln->join[i], 0);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

size_t forward_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves) {
  if (n <= 128) return 0;
  long long waves = num_positions * (long long)num_modes;
  if (max_batch_waves > 0 && waves > max_batch_waves) {
    waves = (max_batch_waves / num_modes) * num_modes;
    if (waves < num_modes) waves = num_modes;
  }
  return (size_t)waves * n * n * sizeof(float2);
}

hipError_t launch_forward(const FwdParams& a, void* ws, size_t ws_bytes, hipStream_t st) {
  switch (a.n) {
    case 32:
      return launch_resident<32>(a, st);
    case 64:
      return launch_resident<64>(a, st);
    case 128:
      return launch_resident<128>(a, st);
    case 256:
      return launch_streaming<256>(a, (float2*)ws, ws_bytes, st);
    case 512:
      return launch_streaming<512>(a, (float2*)ws, ws_bytes, st);
    case 1024:
      return launch_streaming<1024>(a, (float2*)ws, ws_bytes, st);
    case 2048:
      return launch_streaming<2048>(a, (float2*)ws, ws_bytes, st);
    case 4096:
      return launch_streaming<4096>(a, (float2*)ws, ws_bytes, st);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace pty
