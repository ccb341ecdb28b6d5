// Streaming reverse-mode multislice (n = 256 ... 4096) for gfx950.
//
// The adjoint recurrence is that of adjoint.hip (the reverse of the scan the
// reference differentiates with jax.value_and_grad,
// src/ptyrodactyl/multislice/multislice_recon.py:1915-1951 through
// simulations.py:543-602):
//   g phi_{S-1} = F^H(2 w g A),  gT_s += conj(psi_s) g phi_s,
//   g psi_s = conj(T_s) g phi_s,  g phi_{s-1} = Prop^H(g psi_s),
//   g probe_k += conj(R) F(g psi_0),
// over k-slice segments recomputed from the forward's checkpoints.
#include <type_traits>

#include "adjoint.hpp"
#include "common.hpp"
#include "passes.hpp"

namespace pty {

// ---------------------------------------------------------------------------
// Streaming adjoint (n = 256 ... 4096): passes over a batch of waves in global
// scratch.  As in the streaming forward, the row and column halves of the
// propagator (and of its adjoint) commute, so the passes alternate axis and a
// pass applies everything along its axis between two trips of the wave through
// memory.  A reverse pass of slice s completes Prop^H of slice s along its axis
// (g phi_s is then whole), adds the contribution g phi_s conj(psi_s), applies
// conj(T_s) and starts Prop^H of slice s - 1; the segment recompute runs the
// forward passes and writes each psi_s as it completes.
// ---------------------------------------------------------------------------

struct WaveBuf {  // per-wave buffers: base + wl * stride
  float2* base;
  long long stride;
  __device__ __forceinline__ float2* at(int wl) const { return base + (size_t)wl * stride; }
};

// What a pass does before, at and after the slice's pointwise step.
enum : int { A_PRE_NONE = 0, A_PRE_PROBE = 1, A_PRE_IFFT = 2, A_PRE_FOP = 3, A_PRE_AOP = 4, A_PRE_SEED = 5 };
enum : int { A_MID_NONE = 0, A_MID_T = 1, A_MID_ADJ = 2 };
enum : int { A_POST_NONE = 0, A_POST_FOP = 1, A_POST_AOP = 2, A_POST_FFT = 3, A_POST_PROBE = 4 };

template <int N, bool COL, int PRE, int MID, int POST>
__global__ __launch_bounds__(256) void k_adj_pass(AdjParams a, WaveBuf src, WaveBuf dst, WaveBuf psi_out,
                                                  WaveBuf psi_in, long long w0, int nb, int s, int seg_idx) {
  constexpr int T = N / 16, LPB = 256 / T;  // lines per workgroup
  constexpr int RS = N + N / 16;
  constexpr bool DF = N <= DF_MAX_LINE;
  constexpr bool PROP = PRE == A_PRE_FOP || PRE == A_PRE_AOP || POST == A_POST_FOP || POST == A_POST_AOP;
  constexpr bool SHIFTS = PRE == A_PRE_PROBE || POST == A_POST_PROBE;
  __shared__ float2 buf[LPB * RS];
  __shared__ float2 tw[twiddle_entries<N>()];
  __shared__ float2 pf[PROP ? N : 1];
  __shared__ float2 twf[DF ? twiddle_entries<N>() : 1];
  __shared__ float2 pff[DF && PROP ? N : 1];
  const int tid = threadIdx.x;
  int wl, lblk, t, lin;
  if constexpr (!COL) {  // grid (nb, N / LPB)
    wl = blockIdx.x;
    lblk = blockIdx.y;
    t = tid % T;
    lin = tid / T;
  } else {
    if constexpr (LPB > 2) {
      wl = blockIdx.x;
      lblk = blockIdx.y;
    } else {
      column_block<LPB>((int)blockIdx.x, nb, N / LPB, wl, lblk);
    }
    t = tid / LPB;
    lin = tid % LPB;
  }
  const LineGeo<N, COL> L(t, lblk * LPB + lin);
  const long long w = w0 + wl;
  const int M = a.f.M;
  const int p = M == 1 ? (int)w : (int)(w / M), m = M == 1 ? 0 : (int)(w % M);
  const WaveCtx g = resolve_wave<SHIFTS>(a.f, p);
  stage_twiddles<N, DF>(tw, twf, tid, 256);
  if constexpr (PROP) stage_factors<DF>(pf, pff, COL ? g.prop_y : g.prop_x, N, tid, 256);
  const TwTab<DF> twt{tw, twf};
  const FixTab<DF> pt{pf, pff};
  using RowBuf = typename RowBufFor<(T <= 64)>::type;
  using Buf = typename std::conditional<COL, ColBufBlk<LPB>, RowBuf>::type;
  const Buf lb{COL ? buf + lin : buf + lin * RS};
  __syncthreads();

  float2 v[16];
  // ---- before the slice's pointwise step
  if constexpr (PRE == A_PRE_PROBE) {  // probe_k x shift ramp, inverse FFT along this axis
    const BufView pk(g.probe_k + (size_t)m * N * N, (unsigned long long)N * N * sizeof(float2));
    const float2 r_line = COL ? shift_ramp(L.li, N, g.mx, g.dx) : shift_ramp(L.li, N, g.my, g.dy);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int kq = t + (N / 16) * q;
      const float2 r_pt = COL ? shift_ramp(kq, N, g.my, g.dy) : shift_ramp(kq, N, g.mx, g.dx);
      v[q] = cmul(cmul(pk.ld2(L.lo, L.QS * q), r_pt), r_line);
    }
    line_fft<N, true>(v, t, lb, twt);
  } else {
    L.load(v, src.at(wl));
    if constexpr (PRE == A_PRE_IFFT) line_fft<N, true>(v, t, lb, twt);
    if constexpr (PRE == A_PRE_FOP || PRE == A_PRE_AOP) {
      line_fft<N, false>(v, t, lb, twt);
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = PRE == A_PRE_FOP ? pt.mul(v[q], t + T * q) : pt.mulc(v[q], t + T * q);
      line_fft<N, true>(v, t, lb, twt);
    }
    if constexpr (PRE == A_PRE_SEED) {  // fft, g Phi = 2 w g A (or dL/dA as given), inverse fft
      line_fft<N, false>(v, t, lb, twt);
      if (a.grad_amp) {
        const float2* ga = a.grad_amp + ((size_t)p * M + m) * N * N;
#pragma unroll
        for (int q = 0; q < 16; ++q)
          v[q] = ga[(size_t)((L.ky(q) + N / 2) & (N - 1)) * N + ((L.kx(q) + N / 2) & (N - 1))];
      } else {
        const float two_w = 2.0f * (a.f.weights ? a.f.weights[m] : 1.f);
        const float* gi = a.grad_i + g.pat_off;
#pragma unroll
        for (int q = 0; q < 16; ++q)
          v[q] = cscale(v[q], two_w * gi[(size_t)((L.ky(q) + N / 2) & (N - 1)) * N + ((L.kx(q) + N / 2) & (N - 1))]);
      }
      line_fft<N, true>(v, t, lb, twt);
    }
  }
  if (psi_out.base) L.store(v, psi_out.at(wl));  // psi_s is whole here (segment recompute)
  // ---- the slice's pointwise step
  if constexpr (MID == A_MID_T) {
    float2 tv[16];
    load_window<N, COL>(tv, g.slab, g, s, L);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], tv[q]);
  } else if constexpr (MID == A_MID_ADJ) {
    // two groups of eight points: each group issues its transmission and psi loads before using them, and
    // fewer values stay live than with all sixteen at once (128 VGPRs, four waves per SIMD)
    float2 tv[16];
    load_window<N, COL>(tv, g.slab, g, s, L);
    const BufView pv(psi_in.at(wl), (unsigned long long)N * N * sizeof(float2));
    const BufView gv(a.gt_contrib ? a.gt_contrib + ((size_t)wl * a.contrib_slices + seg_idx) * N * N : a.grad_slab,
                     (unsigned long long)N * N * sizeof(float2));
    float2* gs = a.grad_slab + g.grad_off + (long long)s * g.stride;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float2 ps[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) ps[u] = pv.ld2(L.lo, L.QS * (8 * h + u));
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = 8 * h + u;
        const float2 gt = cmulc(v[q], ps[u]);  // g phi conj(psi)
        if (a.gt_contrib) {  // window coordinates, gathered in a fixed order by k_grad_reduce
          gv.st2(gt, L.lo, L.QS * q);
        } else {
          const long long o = L.slab_offset(g, q);
          unsafeAtomicAdd(&gs[o].x, gt.x);
          unsafeAtomicAdd(&gs[o].y, gt.y);
        }
        v[q] = cmulc(v[q], tv[q]);  // conj(T) g phi
      }
    }
  }
  // ---- after it
  if constexpr (POST == A_POST_FOP || POST == A_POST_AOP) {
    line_fft<N, false>(v, t, lb, twt);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = POST == A_POST_FOP ? pt.mul(v[q], t + T * q) : pt.mulc(v[q], t + T * q);
    line_fft<N, true>(v, t, lb, twt);
  }
  if constexpr (POST == A_POST_FFT || POST == A_POST_PROBE) line_fft<N, false>(v, t, lb, twt);
  if constexpr (POST == A_POST_PROBE) {  // g probe_k += conj(R) F(g psi_0); its per-wave part also feeds dL/dshift
    if (a.grad_probe_k || a.gp_contrib) {
      const float2 r_line = COL ? shift_ramp(L.li, N, g.mx, g.dx) : shift_ramp(L.li, N, g.my, g.dy);
      float2 val[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int kq = t + T * q;
        const float2 r_pt = COL ? shift_ramp(kq, N, g.my, g.dy) : shift_ramp(kq, N, g.mx, g.dx);
        val[q] = cmulc(v[q], COL ? cmul(r_pt, r_line) : cmul(r_line, r_pt));
      }
      if (a.gp_contrib) {
        L.store(val, a.gp_contrib + (size_t)(w - a.w_begin) * N * N);  // gathered by k_probe_reduce
      } else {
        float2* gp = a.grad_probe_k + g.gprobe_off + (size_t)m * N * N;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          float2* o = gp + (size_t)L.ky(q) * N + L.kx(q);
          unsafeAtomicAdd(&o->x, val[q].x);
          unsafeAtomicAdd(&o->y, val[q].y);
        }
      }
    }
  } else {
    if (dst.base) L.store(v, dst.at(wl));
  }
}

// ---------------------------------------------------------------------------
// Host launcher
// ---------------------------------------------------------------------------

template <int N, bool COL, int PRE, int MID, int POST>
static void adj_pass_axis(const AdjParams& a, WaveBuf src, WaveBuf dst, WaveBuf po, WaveBuf pi, long long w0, int nb,
                          int s, int seg, hipStream_t st) {
  constexpr int LPB = 256 / (N / 16);
  const dim3 grid = (!COL || LPB > 2) ? dim3(nb, N / LPB) : dim3(nb * (N / LPB));
  hipLaunchKernelGGL((k_adj_pass<N, COL, PRE, MID, POST>), grid, dim3(256), 0, st, a, src, dst, po, pi, w0, nb, s,
                     seg);
}

// One pass on columns (col) or rows.
template <int N, int PRE, int MID, int POST>
static void adj_pass(bool col, const AdjParams& a, WaveBuf src, WaveBuf dst, WaveBuf po, WaveBuf pi, long long w0,
                     int nb, int s, int seg, hipStream_t st) {
  if (col) adj_pass_axis<N, true, PRE, MID, POST>(a, src, dst, po, pi, w0, nb, s, seg, st);
  else adj_pass_axis<N, false, PRE, MID, POST>(a, src, dst, po, pi, w0, nb, s, seg, st);
}

template <int N>
static hipError_t launch_adj_streaming_n(const AdjParams& a0, void* ws, size_t ws_bytes, hipStream_t st) {
  const long long NN = (long long)N * N;
  const int S = a0.f.S, k = a0.k, nc = num_checkpoints(S, k), C = nc + 1;
  const long long waves = a0.f.P * a0.f.M;
  // gathered (deterministic) when there are no descriptors and the workspace holds the contributions
  // of one k-slice segment and the probe: k + 1 more waves per wave in flight
  const bool gathered = !a0.f.desc && ws_bytes >= (size_t)(2 * k + 3) * NN * sizeof(float2);
  long long nbmax = (long long)(ws_bytes / ((size_t)(gathered ? 2 * k + 3 : k + 2) * NN * sizeof(float2)));
  if (nbmax > waves) nbmax = waves;
  if (nbmax < 1 || (a0.grad_shifts && !gathered)) return hipErrorInvalidValue;
  if (nbmax >= 8) nbmax -= nbmax % 8;
  float2* base = (float2*)ws;
  const WaveBuf none{nullptr, 0};
  for (long long w0 = 0; w0 < waves; w0 += nbmax) {
    const int nb = (int)((waves - w0) < nbmax ? (waves - w0) : nbmax);
    const WaveBuf seg{base, (long long)k * NN};
    const WaveBuf X{base + (size_t)nb * k * NN, NN};
    const WaveBuf Gb{base + (size_t)nb * (k + 1) * NN, NN};
    AdjParams a = a0;
    a.w_begin = w0;
    a.w_end = w0 + nb;
    a.contrib_slices = k;
    a.gt_contrib = gathered ? base + (size_t)nb * (k + 2) * NN : nullptr;
    a.gp_contrib = gathered && (a0.grad_probe_k || a0.grad_shifts) ? base + (size_t)nb * (2 * k + 2) * NN : nullptr;
    bool rcol = false;  // axis of the reverse pass of the next slice (rows for the last slice)
    for (int c = C - 1; c >= 0; --c) {
      const int s0 = c * k, s1 = min(S, s0 + k);
      // psi_s of the segment: slot s - s0 of seg, written by the pass that completes it; psi_{s0} of a segment
      // with c > 0 is read from its checkpoint in place
      const WaveBuf ck = c > 0 ? WaveBuf{const_cast<float2*>(a.ckpt) + ((size_t)w0 * nc + (c - 1)) * NN,
                                         (long long)nc * NN}
                               : none;
      auto slot = [&](int s) { return WaveBuf{seg.base + (size_t)(s - s0) * NN, seg.stride}; };
      auto psi = [&](int s) { return (s == s0 && c > 0) ? ck : slot(s); };
      // segment recompute: forward passes alternating axis from columns
      bool col = true;
      if (c == 0) {
        adj_pass<N, A_PRE_PROBE, A_MID_NONE, A_POST_NONE>(col, a, none, X, none, none, w0, nb, 0, 0, st);
        col = !col;
        if (s1 > 1)
          adj_pass<N, A_PRE_IFFT, A_MID_T, A_POST_FOP>(col, a, X, X, slot(0), none, w0, nb, 0, 0, st);
        else
          adj_pass<N, A_PRE_IFFT, A_MID_NONE, A_POST_NONE>(col, a, X, none, slot(0), none, w0, nb, 0, 0, st);
        col = !col;
      } else if (s1 - s0 > 1) {
        adj_pass<N, A_PRE_NONE, A_MID_T, A_POST_FOP>(col, a, ck, X, none, none, w0, nb, s0, 0, st);
        col = !col;
      }
      for (int s = s0 + 1; s < s1; ++s) {
        if (s < s1 - 1)
          adj_pass<N, A_PRE_FOP, A_MID_T, A_POST_FOP>(col, a, X, X, slot(s), none, w0, nb, s, 0, st);
        else
          adj_pass<N, A_PRE_FOP, A_MID_NONE, A_POST_NONE>(col, a, X, none, slot(s), none, w0, nb, s, 0, st);
        col = !col;
      }
      int s = s1 - 1;
      if (c == C - 1) {  // A = F(psi_{S-1} T_{S-1}), the seed, and g phi_{S-1} = F^H(g Phi)
        adj_pass<N, A_PRE_NONE, A_MID_T, A_POST_FFT>(rcol, a, psi(S - 1), X, none, none, w0, nb, S - 1, 0, st);
        adj_pass<N, A_PRE_SEED, A_MID_NONE, A_POST_NONE>(!rcol, a, X, Gb, none, none, w0, nb, S - 1, 0, st);
        if (S > 1)
          adj_pass<N, A_PRE_IFFT, A_MID_ADJ, A_POST_AOP>(rcol, a, Gb, Gb, none, psi(S - 1), w0, nb, S - 1,
                                                          S - 1 - s0, st);
        else
          adj_pass<N, A_PRE_IFFT, A_MID_ADJ, A_POST_FFT>(rcol, a, Gb, Gb, none, psi(0), w0, nb, 0, 0, st);
        rcol = !rcol;
        s = S - 2;
      }
      for (; s >= s0; --s) {
        if (s > 0)
          adj_pass<N, A_PRE_AOP, A_MID_ADJ, A_POST_AOP>(rcol, a, Gb, Gb, none, psi(s), w0, nb, s, s - s0, st);
        else
          adj_pass<N, A_PRE_AOP, A_MID_ADJ, A_POST_FFT>(rcol, a, Gb, Gb, none, psi(0), w0, nb, 0, 0, st);
        rcol = !rcol;
      }
      if (gathered) {  // the segment's slices
        hipError_t e = launch_adj_reduces(a, w0, nb, s0, s1 - s0, false, st);
        if (e != hipSuccess) return e;
      }
    }
    // F(g psi_0) along the second axis and the probe gradient conj(R) F(g psi_0)
    adj_pass<N, A_PRE_NONE, A_MID_NONE, A_POST_PROBE>(rcol, a, Gb, none, none, none, w0, nb, 0, 0, st);
    if (gathered && (a0.grad_probe_k || a0.grad_shifts)) {
      hipError_t e = launch_adj_reduces(a, w0, nb, 0, 0, true, st);
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_adj_streaming(const AdjParams& a, void* ws, size_t ws_bytes, hipStream_t st) {
  switch (a.f.n) {
    case 256:
      return launch_adj_streaming_n<256>(a, ws, ws_bytes, st);
    case 512:
      return launch_adj_streaming_n<512>(a, ws, ws_bytes, st);
    case 1024:
      return launch_adj_streaming_n<1024>(a, ws, ws_bytes, st);
    case 2048:
      return launch_adj_streaming_n<2048>(a, ws, ws_bytes, st);
    case 4096:
      return launch_adj_streaming_n<4096>(a, ws, ws_bytes, st);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace pty
