// Forward-mode (Jacobian-vector product) multislice kernels for gfx950.
//
// Every wave (one probe mode at one scan position) carries its primal value and its tangent
// through the slices, in complex64 with FP32 accumulation:
//   psi_0  = ifft2(pk R)                       dpsi_0  = ifft2(dpk R + pk dR),  dR = R (-2 pi i k.ds / n)
//   phi_s  = psi_s T_s                         dphi_s  = dpsi_s T_s + psi_s dT_s
//   psi_s+1 = ifft2(fft2(phi_s) P)             dpsi_s+1 = ifft2(fft2(dphi_s) P)          s < S-1
//   A = fftshift(fft2(phi_S-1))                dA = fftshift(fft2(dphi_S-1))
//   I = sum_m w_m |A_m|^2                      dI = sum_m w_m 2 Re(conj(A_m) dA_m)
// which is the tangent of the scan body of the reference's _cbed_amplitude_from_slice_provider
// (src/ptyrodactyl/multislice/simulations.py:543-602) on the shifted probe of shift_beam_fourier
// (simulations.py:1287-1321); the reference's Gauss-Newton / Levenberg-Marquardt solvers take it
// through jax.jvp (src/ptyrodactyl/jacobian/operators.py:30-185).
//
// Both waves of a pair live in a device workspace ([wave][primal, tangent][n][n]) and every slice is
// one row pass and one column pass over the batch, as in the streaming forward.  Each line goes
// through the same transforms, multiplies and table corrections as in the forward kernels of the
// same size, so the primal patterns equal pty_forward's to rounding.
#include "common.hpp"
#include "jvp.hpp"

namespace pty {

template <int N>
struct JvpShape {
  static constexpr int T = N / 16;                          // threads per line
  static constexpr int LINES = (256 / T < N) ? 256 / T : N;  // lines per workgroup
  static constexpr int THREADS = LINES * T;
};

enum : int { JROW_FIRST = 1, JROW_LAST = 2 };
enum : int { JCOL_INIT = 0, JCOL_MID = 1, JCOL_FINAL = 2 };

template <int N, int FLAGS>
__global__ __launch_bounds__(JvpShape<N>::THREADS) void k_jvp_rows(JvpParams a, float2* waves, int nb, int s) {
  using SH = JvpShape<N>;
  constexpr int T = SH::T, RPB = SH::LINES, RS = N + N / 16, NT = SH::THREADS;
  constexpr bool DF = N <= DF_MAX_LINE;
  __shared__ float2 rows[RPB * RS];
  __shared__ float2 tw[twiddle_entries<N>()];
  __shared__ float2 px[N];
  __shared__ float2 twf[DF ? twiddle_entries<N>() : 1];
  __shared__ float2 pxf[DF ? N : 1];
  const int tid = threadIdx.x;
  const int wl = blockIdx.x % nb, rblk = blockIdx.x / nb;
  const long long w = a.w0 + wl;
  const long long p = w / a.f.M;
  const int t = tid % T, lr = tid / T;
  const int y = rblk * RPB + lr;
  const WaveCtx g = resolve_wave(a.f, p);
  stage_twiddles<N, DF>(tw, twf, tid, NT);
  if (!(FLAGS & JROW_LAST)) stage_factors<DF>(px, pxf, g.prop_x, N, tid, NT);
  const TwTab<DF> twt{tw, twf};
  const FixTab<DF> pxt{px, pxf};
  const RowBufWave rb{rows + lr * RS};
  __syncthreads();

  float2* prow = waves + ((size_t)wl * 2 * N + y) * N;  // primal row; the tangent wave follows
  float2* trow = prow + (size_t)N * N;
  float2 v[16], dv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    v[q] = prow[t + T * q];
    dv[q] = trow[t + T * q];
  }
  if (FLAGS & JROW_FIRST) {  // finish both incident waves: ifft along x
    line_fft<N, true>(v, t, rb, twt);
    line_fft<N, true>(dv, t, rb, twt);
  }
  int gy = g.oy + y;
  if (gy >= g.slab_h) gy -= (int)g.slab_h;
  const long long off = (long long)s * g.stride + (long long)gy * g.slab_w;
  const float2* srow = g.slab + off;
  const float2* dsrow = a.d_slab ? a.d_slab + off : nullptr;
  const int sw = (int)g.slab_w, x0 = g.ox + t;
  auto points = [&](auto col) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int x = col(q);
      const float2 tv = srow[x];
      dv[q] = cmul(dv[q], tv);
      if (dsrow) dv[q] = cadd(dv[q], cmul(v[q], dsrow[x]));
      v[q] = cmul(v[q], tv);
    }
  };
  if (g.ox + N <= sw) {  // the window row does not wrap: constant offsets from one base
    points([&](int q) { return x0 + T * q; });
  } else {
    points([&](int q) {
      const int x = x0 + T * q;
      return x >= sw ? x - sw : x;
    });
  }
  line_fft<N, false>(v, t, rb, twt);
  line_fft<N, false>(dv, t, rb, twt);
  if (!(FLAGS & JROW_LAST)) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      v[q] = pxt.mul(v[q], t + T * q);
      dv[q] = pxt.mul(dv[q], t + T * q);
    }
    line_fft<N, true>(v, t, rb, twt);
    line_fft<N, true>(dv, t, rb, twt);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    prow[t + T * q] = v[q];
    trow[t + T * q] = dv[q];
  }
}

template <int NC>
struct JvpColBuf {
  float2* base;
  __device__ __forceinline__ float2 ld(int e) const { return base[pad16(e) * NC]; }
  __device__ __forceinline__ void st(int e, float2 v) const { base[pad16(e) * NC] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return base[pe * NC]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { base[pe * NC] = v; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};

// Signed frequency index (fftfreq order, Nyquist negative) of bin k.
__device__ __forceinline__ int signed_bin(int k, int n) { return k < n / 2 ? k : k - n; }

template <int N, int MODE>
__global__ __launch_bounds__(JvpShape<N>::THREADS) void k_jvp_cols(JvpParams a, float2* waves, int nb) {
  using SH = JvpShape<N>;
  constexpr int T = SH::T, NC = SH::LINES, PS = N + N / 16, NT = SH::THREADS;
  constexpr bool DF = N <= DF_MAX_LINE;
  __shared__ float2 buf[PS * NC];
  __shared__ float2 tw[twiddle_entries<N>()];
  __shared__ float2 py[N];
  __shared__ float2 twf[DF ? twiddle_entries<N>() : 1];
  __shared__ float2 pyf[DF ? N : 1];
  const int tid = threadIdx.x;
  const int gc = tid % NC, t = tid / NC;
  int unit, cblk;  // wave (INIT / MID) or position (FINAL) inside the batch, column block
  column_block<NC>((int)blockIdx.x, nb, N / NC, unit, cblk);
  const int x = cblk * NC + gc;
  const long long pos = MODE == JCOL_FINAL ? a.w0 / a.f.M + unit : (a.w0 + unit) / a.f.M;
  const WaveCtx g = resolve_wave(a.f, pos);
  stage_twiddles<N, DF>(tw, twf, tid, NT);
  if (MODE == JCOL_MID) stage_factors<DF>(py, pyf, g.prop_y, N, tid, NT);
  const TwTab<DF> twt{tw, twf};
  const FixTab<DF> pyt{py, pyf};
  const JvpColBuf<NC> cb{buf + gc};
  __syncthreads();

  if constexpr (MODE == JCOL_INIT) {
    const int m = (int)((a.w0 + unit) % a.f.M);
    const float2 rx = shift_ramp(x, N, g.mx, g.dx);
    const float2* pk = g.probe_k + (size_t)m * N * N + x;
    const float2* dpk = a.d_probe_k ? a.d_probe_k + (size_t)m * N * N + x : nullptr;
    // d(ramp)/d(shift) = ramp * (-2 pi i k / n) per axis
    const float dsy = a.d_shifts ? (float)a.d_shifts[2 * pos] : 0.f;
    const float dsx = a.d_shifts ? (float)a.d_shifts[2 * pos + 1] : 0.f;
    const float kxs = (float)signed_bin(x, N) * dsx;
    float2 v[16], dv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int ky = t + T * q;
      const float2 ry = shift_ramp(ky, N, g.my, g.dy);
      v[q] = cmul(cmul(pk[(size_t)ky * N], ry), rx);
      dv[q] = dpk ? cmul(cmul(dpk[(size_t)ky * N], ry), rx) : make_float2(0.f, 0.f);
      if (a.d_shifts) {
        const float c = 6.283185307179586f * ((float)signed_bin(ky, N) * dsy + kxs) / (float)N;
        dv[q] = cadd(dv[q], cscale(mul_mi(v[q]), c));
      }
    }
    line_fft<N, true>(v, t, cb, twt);
    line_fft<N, true>(dv, t, cb, twt);
    float2* pcol = waves + (size_t)unit * 2 * N * N + x;
    float2* tcol = pcol + (size_t)N * N;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      pcol[(size_t)(t + T * q) * N] = v[q];
      tcol[(size_t)(t + T * q) * N] = dv[q];
    }
  } else if constexpr (MODE == JCOL_MID) {
    float2* pcol = waves + (size_t)unit * 2 * N * N + x;
    float2* tcol = pcol + (size_t)N * N;
    float2 v[16], dv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      v[q] = pcol[(size_t)(t + T * q) * N];
      dv[q] = tcol[(size_t)(t + T * q) * N];
    }
    line_fft<N, false>(v, t, cb, twt);
    line_fft<N, false>(dv, t, cb, twt);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      v[q] = pyt.mul(v[q], t + T * q);
      dv[q] = pyt.mul(dv[q], t + T * q);
    }
    line_fft<N, true>(v, t, cb, twt);
    line_fft<N, true>(dv, t, cb, twt);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      pcol[(size_t)(t + T * q) * N] = v[q];
      tcol[(size_t)(t + T * q) * N] = dv[q];
    }
  } else {
    const int pl = unit;  // position inside the batch
    float acc[16], dacc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = dacc[q] = 0.f;
    const int xs = (x + N / 2) & (N - 1);
    for (int m = 0; m < a.f.M; ++m) {
      const float2* pcol = waves + ((size_t)pl * a.f.M + m) * 2 * N * N + x;
      const float2* tcol = pcol + (size_t)N * N;
      float2 v[16], dv[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        v[q] = pcol[(size_t)(t + T * q) * N];
        dv[q] = tcol[(size_t)(t + T * q) * N];
      }
      line_fft<N, false>(v, t, cb, twt);
      line_fft<N, false>(dv, t, cb, twt);
      const float wm = a.f.weights ? a.f.weights[m] : 1.f;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        acc[q] += wm * cabs2(v[q]);
        dacc[q] += 2.f * wm * (v[q].x * dv[q].x + v[q].y * dv[q].y);
      }
      const size_t abase = ((size_t)pos * a.f.M + m) * N * N + xs;
      if (a.f.amplitude) {
#pragma unroll
        for (int q = 0; q < 16; ++q) a.f.amplitude[abase + (size_t)((t + T * q + N / 2) & (N - 1)) * N] = v[q];
      }
      if (a.d_amplitude) {
#pragma unroll
        for (int q = 0; q < 16; ++q) a.d_amplitude[abase + (size_t)((t + T * q + N / 2) & (N - 1)) * N] = dv[q];
      }
    }
    const size_t pbase = (size_t)g.pat_off + xs;
    if (a.f.patterns) {
#pragma unroll
      for (int q = 0; q < 16; ++q) a.f.patterns[pbase + (size_t)((t + T * q + N / 2) & (N - 1)) * N] = acc[q];
    }
    if (a.d_patterns) {
#pragma unroll
      for (int q = 0; q < 16; ++q) a.d_patterns[pbase + (size_t)((t + T * q + N / 2) & (N - 1)) * N] = dacc[q];
    }
  }
}

template <int N>
static hipError_t launch_jvp_n(JvpParams a, float2* ws, size_t ws_bytes, hipStream_t st) {
  using SH = JvpShape<N>;
  constexpr int RPB = SH::LINES, NC = SH::LINES;
  const size_t pair_bytes = 2 * (size_t)N * N * sizeof(float2);
  long long pb = (long long)(ws_bytes / pair_bytes) / a.f.M;  // positions per batch
  if (pb < 1) return hipErrorInvalidValue;
  if (pb > a.f.P) pb = a.f.P;
  if (pb >= 8) pb -= pb % 8;
  for (long long p0 = 0; p0 < a.f.P; p0 += pb) {
    const int np = (int)((a.f.P - p0) < pb ? (a.f.P - p0) : pb);
    const int nb = np * a.f.M;
    a.w0 = p0 * a.f.M;
    const dim3 rows(nb * (N / RPB)), cols(nb * (N / NC)), fin(np * (N / NC)), blk(SH::THREADS);
    hipLaunchKernelGGL((k_jvp_cols<N, JCOL_INIT>), cols, blk, 0, st, a, ws, nb);
    for (int s = 0; s < a.f.S; ++s) {
      const bool first = s == 0, last = s == a.f.S - 1;
      if (first && last)
        hipLaunchKernelGGL((k_jvp_rows<N, JROW_FIRST | JROW_LAST>), rows, blk, 0, st, a, ws, nb, s);
      else if (first)
        hipLaunchKernelGGL((k_jvp_rows<N, JROW_FIRST>), rows, blk, 0, st, a, ws, nb, s);
      else if (last)
        hipLaunchKernelGGL((k_jvp_rows<N, JROW_LAST>), rows, blk, 0, st, a, ws, nb, s);
      else
        hipLaunchKernelGGL((k_jvp_rows<N, 0>), rows, blk, 0, st, a, ws, nb, s);
      if (!last) hipLaunchKernelGGL((k_jvp_cols<N, JCOL_MID>), cols, blk, 0, st, a, ws, nb);
    }
    hipLaunchKernelGGL((k_jvp_cols<N, JCOL_FINAL>), fin, blk, 0, st, a, ws, np);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

size_t jvp_workspace_bytes(int n, long long num_positions, int num_modes, long long max_batch_waves) {
  long long waves = num_positions * (long long)num_modes;
  if (max_batch_waves > 0 && waves > max_batch_waves) {
    waves = (max_batch_waves / num_modes) * num_modes;
    if (waves < num_modes) waves = num_modes;
  }
  return (size_t)waves * 2 * n * n * sizeof(float2);
}

hipError_t launch_jvp(const JvpParams& a, void* ws, size_t ws_bytes, hipStream_t st) {
  switch (a.f.n) {
    case 32:
      return launch_jvp_n<32>(a, (float2*)ws, ws_bytes, st);
    case 64:
      return launch_jvp_n<64>(a, (float2*)ws, ws_bytes, st);
    case 128:
      return launch_jvp_n<128>(a, (float2*)ws, ws_bytes, st);
    case 256:
      return launch_jvp_n<256>(a, (float2*)ws, ws_bytes, st);
    case 512:
      return launch_jvp_n<512>(a, (float2*)ws, ws_bytes, st);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace pty
