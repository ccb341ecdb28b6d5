// Reverse-mode (adjoint) multislice kernels for gfx950.
//
// Forward per wave (see forward.hip):  psi_0 = F^H(probe_k . R),
// phi_s = psi_s T_s,  psi_{s+1} = Prop(phi_s),  A = F(phi_{S-1}),
// I = sum_m w_m |A_m|^2 (fftshifted).  Given g = dL/dI the adjoint is
//   gPhi      = 2 w_m g . A
//   gphi_S-1  = F^H gPhi
//   gT_s     += conj(psi_s) gphi_s          gpsi_s = conj(T_s) gphi_s
//   gphi_s-1  = Prop^H(gpsi_s)              (propagator conjugated)
//   gprobe_k += conj(R) F(gpsi_0)
// with gradients in the dL/dRe + i dL/dIm convention.  This is the reverse
// of the scan the reference differentiates with jax.value_and_grad
// (src/ptyrodactyl/multislice/multislice_recon.py:1915-1951 through
// simulations.py:543-602); the reference keeps every slice's residuals,
// while here the wave entering every k-th slice was checkpointed by the
// forward and each k-slice segment is recomputed before it is reversed.
#include "adjoint.hpp"
#include "common.hpp"

namespace pty {

enum : int { PRE_NONE = 0, PRE_IFFT = 1, PRE_ADJPROP = 2 };
enum : int { POST_NONE = 0, POST_PROP = 1, POST_FFT = 2 };
enum : int { PW_NONE = 0, PW_FWD = 1, PW_ADJ = 2 };
enum : int { C_PROP = 0, C_ADJPROP = 1, C_IFFT_PROBE = 2, C_FINAL_ADJ = 3, C_PROBE_GRAD = 4 };

// Row work shared by both strategies: v holds row y of one wave in the
// strided distribution (thread t owns x = t + T q).
template <int N, int PRE, int POST, class Buf, class TW, class FX>
__device__ __forceinline__ void row_ops(float2 (&v)[16], int t, int y, int s, int pw, const AdjParams& a,
                                        const WaveCtx& g, float2* seg_row, bool seg_store, float2* gt_row,
                                        const Buf& rb, const TW& twt, const FX& px) {
  constexpr int T = N / 16;
  if constexpr (PRE == PRE_IFFT) line_fft<N, true>(v, t, rb, twt);
  if constexpr (PRE == PRE_ADJPROP) {
    line_fft<N, false>(v, t, rb, twt);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = px.mulc(v[q], t + T * q);
    line_fft<N, true>(v, t, rb, twt);
  }
  if (seg_store) {
#pragma unroll
    for (int q = 0; q < 16; ++q) seg_row[t + T * q] = v[q];
  }
  if (pw != PW_NONE) {
    int gy = g.oy + y;
    if (gy >= g.slab_h) gy -= (int)g.slab_h;
    const long long off = (long long)s * g.stride + (long long)gy * g.slab_w;
    const float2* trow = g.slab + off;
    const int sw = (int)g.slab_w;
    // slab column of point q: a window row that does not wrap (uniform per wave) addresses its points at
    // constant offsets from one base; the wrapped case computes each index
    auto points = [&](auto col) {
      // groups of GP points: each group issues all of its loads before using any of them, so the loads
      // overlap instead of waiting one by one under register pressure
      constexpr int GP = N <= 128 ? 4 : 8;
#pragma unroll
      for (int grp = 0; grp < 16 / GP; ++grp) {
        float2 tv[GP], sv[GP];
#pragma unroll
        for (int u = 0; u < GP; ++u) {
          const int q = GP * grp + u;
          tv[u] = trow[col(q)];
          if (pw == PW_ADJ) sv[u] = seg_row[t + T * q];
        }
        asm volatile("" ::: "memory");
        if (pw == PW_FWD) {
#pragma unroll
          for (int u = 0; u < GP; ++u) v[GP * grp + u] = cmul(v[GP * grp + u], tv[u]);
        } else {
          float2* grow = a.grad_slab + g.grad_off + off;
#pragma unroll
          for (int u = 0; u < GP; ++u) {
            const int q = GP * grp + u;
            const float2 gt = cmulc(v[q], sv[u]);  // gphi * conj(psi)
            if (gt_row) {
              gt_row[t + T * q] = gt;  // window coordinates, gathered by k_grad_reduce
            } else {
              const int x = col(q);
              unsafeAtomicAdd(&grow[x].x, gt.x);
              unsafeAtomicAdd(&grow[x].y, gt.y);
            }
            v[q] = cmulc(v[q], tv[u]);  // conj(T) * gphi
          }
        }
      }
    };
    const int x0 = g.ox + t;
    if (g.ox + N <= sw) {
      points([&](int q) { return x0 + T * q; });
    } else {
      points([&](int q) {
        const int x = x0 + T * q;
        return x >= sw ? x - sw : x;
      });
    }
  }
  if constexpr (POST == POST_PROP) {
    line_fft<N, false>(v, t, rb, twt);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = px.mul(v[q], t + T * q);
    line_fft<N, true>(v, t, rb, twt);
  }
  if constexpr (POST == POST_FFT) line_fft<N, false>(v, t, rb, twt);
}

// Column work: v holds column x (thread t owns ky = t + T q).  Returns false
// when the result was consumed (probe gradient) and must not be stored.
template <int N, int OP, class Buf, class TW, class FX>
__device__ __forceinline__ bool col_ops(float2 (&v)[16], int t, int x, long long w, const AdjParams& a,
                                        const WaveCtx& g, const Buf& cb, const TW& twt, const FX& py,
                                        const float2* ramp_y, const float2* ramp_x) {
  constexpr int T = N / 16;
  const int p = (int)(w / a.f.M), m = (int)(w % a.f.M);
  if constexpr (OP == C_IFFT_PROBE) {
    const float2 rx = ramp_x ? ramp_x[x] : shift_ramp(x, N, g.mx, g.dx);
    const float2* pk = g.probe_k + (size_t)m * N * N + x;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int ky = t + T * q;
      const float2 ry = ramp_y ? ramp_y[ky] : shift_ramp(ky, N, g.my, g.dy);
      v[q] = cmul(cmul(pk[(size_t)ky * N], ry), rx);
    }
    line_fft<N, true>(v, t, cb, twt);
    return true;
  } else {
    line_fft<N, false>(v, t, cb, twt);
    if constexpr (OP == C_PROP || OP == C_ADJPROP) {
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = OP == C_PROP ? py.mul(v[q], t + T * q) : py.mulc(v[q], t + T * q);
      line_fft<N, true>(v, t, cb, twt);
      return true;
    } else if constexpr (OP == C_FINAL_ADJ) {
      if (a.grad_amp) {  // amplitude seed: gPhi = dL/dA as given
        const float2* ga = a.grad_amp + ((size_t)p * a.f.M + m) * N * N + ((x + N / 2) & (N - 1));
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = ga[(size_t)((t + T * q + N / 2) & (N - 1)) * N];
      } else {  // intensity seed: gPhi = 2 w g A
        const float two_w = 2.0f * (a.f.weights ? a.f.weights[m] : 1.f);
        const float* gi = a.grad_i + g.pat_off + ((x + N / 2) & (N - 1));
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = cscale(v[q], two_w * gi[(size_t)((t + T * q + N / 2) & (N - 1)) * N]);
      }
      line_fft<N, true>(v, t, cb, twt);
      return true;
    } else {  // C_PROBE_GRAD (its per-wave contributions also feed the position gradient)
      if (a.grad_probe_k || a.gp_contrib) {
        const float2 rx = ramp_x ? ramp_x[x] : shift_ramp(x, N, g.mx, g.dx);
        float2* gp = a.grad_probe_k ? a.grad_probe_k + g.gprobe_off + (size_t)m * N * N + x : nullptr;
        float2* gc = a.gp_contrib ? a.gp_contrib + (size_t)(w - a.w_begin) * N * N + x : nullptr;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int ky = t + T * q;
          const float2 ry = ramp_y ? ramp_y[ky] : shift_ramp(ky, N, g.my, g.dy);
          const float2 val = cmulc(v[q], cmul(ry, rx));
          if (gc) {
            gc[(size_t)ky * N] = val;  // gathered by k_probe_reduce
          } else {
            unsafeAtomicAdd(&gp[(size_t)ky * N].x, val.x);
            unsafeAtomicAdd(&gp[(size_t)ky * N].y, val.y);
          }
        }
      }
      return false;
    }
  }
}

// ---------------------------------------------------------------------------
// LDS-resident adjoint (n = 32, 64, 128): one workgroup per wave, looping
// over waves so that the global scratch is bounded by the grid size.
// ---------------------------------------------------------------------------

// The resident wave keeps row y at LDS row pad16(y) (one spare row per 16): column-exchange accesses
// spread over the banks and every address is a per-thread base plus a compile-time offset.
struct ColBufRes {
  float2* col;
  int rs;
  __device__ __forceinline__ float2 ld(int e) const { return col[pad16(e) * rs]; }
  __device__ __forceinline__ void st(int e, float2 v) const { col[pad16(e) * rs] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return col[pe * rs]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { col[pe * rs] = v; }
  __device__ __forceinline__ void sync() const { wave_sync(); }
};

template <int N>
struct ResShared {
  float2* wave;
  TwTable tw;
  FixTable py;
  FixTable px;
  const float2* ramp_y;
  const float2* ramp_x;
};

__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <int N, int PRE, int POST>
__device__ __forceinline__ void res_row(const ResShared<N>& sh, const AdjParams& a, const WaveCtx& g, int s, int pw,
                                        float2* seg, bool seg_store, float2* gt_wave = nullptr) {
  constexpr int T = N / 16, RS = N + N / 16, G = 64 / T;
  const int tid = opaque((int)threadIdx.x);
  const int lane = tid & 63, wid = tid >> 6;
  const int rt = lane % T, ry = wid * G + lane / T;
  const RowBufWave rb{sh.wave + pad16(ry) * RS};
  const TwTable twt = sh.tw;
  float2 v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = rb.row[pad16(rt + T * q)];
  float2* gt_row = gt_wave ? gt_wave + ((size_t)s * N + ry) * N : nullptr;
  row_ops<N, PRE, POST>(v, rt, ry, s, pw, a, g, seg ? seg + (size_t)ry * N : nullptr, seg_store, gt_row, rb, twt,
                        sh.px);
  wave_sync();
#pragma unroll
  for (int q = 0; q < 16; ++q) rb.row[pad16(rt + T * q)] = v[q];
  __syncthreads();
}

template <int N, int OP>
__device__ __forceinline__ void res_col(const ResShared<N>& sh, const AdjParams& a, const WaveCtx& g, long long w) {
  constexpr int T = N / 16, RS = N + N / 16, G = 64 / T;
  const int tid = opaque((int)threadIdx.x);
  const int lane = tid & 63, wid = tid >> 6;
  const int ct = lane / G, cx = wid * G + lane % G;
  float2* col = sh.wave + pad16(cx);
  const ColBufRes cb{col, RS};
  const TwTable twt = sh.tw;
  float2 v[16];
  if constexpr (OP != C_IFFT_PROBE) {
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = col[pad16(ct + T * q) * RS];
  }
  if (col_ops<N, OP>(v, ct, cx, w, a, g, cb, twt, sh.py, sh.ramp_y, sh.ramp_x)) {
    wave_sync();
#pragma unroll
    for (int q = 0; q < 16; ++q) col[pad16(ct + T * q) * RS] = v[q];
  }
  __syncthreads();
}

template <int N>
__device__ __forceinline__ void res_copy(const ResShared<N>& sh, float2* gmem, bool to_lds) {
  constexpr int T = N / 16, RS = N + N / 16, G = 64 / T;
  const int tid = opaque((int)threadIdx.x);
  const int lane = tid & 63, wid = tid >> 6;
  const int rt = lane % T, ry = wid * G + lane / T;
  float2* row = sh.wave + pad16(ry) * RS;
  float2* grow = gmem + (size_t)ry * N;
  float2 v[16];
  if (to_lds) {
    // all sixteen loads are issued before the first LDS write, so the copy costs one memory latency
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = grow[rt + T * q];
    asm volatile("" ::: "memory");
#pragma unroll
    for (int q = 0; q < 16; ++q) row[pad16(rt + T * q)] = v[q];
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = row[pad16(rt + T * q)];
#pragma unroll
    for (int q = 0; q < 16; ++q) grow[rt + T * q] = v[q];
  }
  __syncthreads();
}

template <int N>
__global__ __launch_bounds__(N* N / 16) void k_adjoint_resident(AdjParams a, float2* scratch) {
  constexpr int NT = N * N / 16, RS = N + N / 16;
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
  const int tid = threadIdx.x;
  fill_twiddles(tw, twf, N, tid, NT);
  const ResShared<N> sh{wave, TwTable{tw, twf}, FixTable{py, pyf}, FixTable{px, pxf}, ramp_y, ramp_x};
  const int S = a.f.S, k = a.k;
  const int nc = num_checkpoints(S, k), C = nc + 1;
  float2* seg = scratch + (size_t)blockIdx.x * (k + 1) * N * N;
  float2* park = seg + (size_t)k * N * N;
  for (long long w = a.w_begin + blockIdx.x; w < a.w_end; w += gridDim.x) {
    const int p = (int)(w / a.f.M);
    const WaveCtx g = resolve_wave(a.f, p);
    float2* gtw = a.gt_contrib ? a.gt_contrib + (size_t)(w - a.w_begin) * a.contrib_slices * N * N : nullptr;
    __syncthreads();
    stage_factors(py, pyf, g.prop_y, N, tid, NT);
    stage_factors(px, pxf, g.prop_x, N, tid, NT);
    for (int i = tid; i < N; i += NT) {
      ramp_y[i] = shift_ramp(i, N, g.my, g.dy);
      ramp_x[i] = shift_ramp(i, N, g.mx, g.dx);
    }
    __syncthreads();
    for (int c = C - 1; c >= 0; --c) {
      const int s0 = c * k, s1 = min(S, s0 + k);
      if (c < C - 1) res_copy<N>(sh, park, false);
      // recompute psi_{s0} .. psi_{s1-1} into seg
      if (c == 0)
        res_col<N, C_IFFT_PROBE>(sh, a, g, w);
      else
        res_copy<N>(sh, const_cast<float2*>(a.ckpt) + ((size_t)w * nc + (c - 1)) * N * N, true);
      for (int s = s0; s < s1; ++s) {
        float2* sg = seg + (size_t)(s - s0) * N * N;
        const bool more = s < s1 - 1;
        if (s == 0) {
          if (more)
            res_row<N, PRE_IFFT, POST_PROP>(sh, a, g, s, PW_FWD, sg, true);
          else
            res_row<N, PRE_IFFT, POST_NONE>(sh, a, g, s, PW_NONE, sg, true);
        } else {
          if (more)
            res_row<N, PRE_NONE, POST_PROP>(sh, a, g, s, PW_FWD, sg, true);
          else
            res_row<N, PRE_NONE, POST_NONE>(sh, a, g, s, PW_NONE, sg, true);
        }
        if (more) res_col<N, C_PROP>(sh, a, g, w);
      }
      int s = s1 - 1;
      if (c == C - 1) {  // exit plane: gphi_{S-1} = F^H (2 w g A)
        res_row<N, PRE_NONE, POST_FFT>(sh, a, g, S - 1, PW_FWD, nullptr, false);
        res_col<N, C_FINAL_ADJ>(sh, a, g, w);
        res_row<N, PRE_IFFT, POST_NONE>(sh, a, g, S - 1, PW_ADJ, seg + (size_t)(S - 1 - s0) * N * N, false, gtw);
        s = S - 2;
      } else {
        res_copy<N>(sh, park, true);
      }
      for (; s >= s0; --s) {
        res_col<N, C_ADJPROP>(sh, a, g, w);
        res_row<N, PRE_ADJPROP, POST_NONE>(sh, a, g, s, PW_ADJ, seg + (size_t)(s - s0) * N * N, false, gtw);
      }
    }
    res_row<N, PRE_NONE, POST_FFT>(sh, a, g, 0, PW_NONE, nullptr, false);
    res_col<N, C_PROBE_GRAD>(sh, a, g, w);
  }
}

template <int N>
static size_t resident_adj_lds_bytes() {
  return sizeof(float2) * ((size_t)(N + N / 16) * (N + N / 16) + 8 * N);
}

// ---------------------------------------------------------------------------
// Deterministic gathers of the per-wave contributions
// ---------------------------------------------------------------------------

constexpr int RED_TILE = 32;

// Offset of index i in the periodic frame that starts at o (0 <= i, o < m).
__device__ __forceinline__ int frame_offset(int i, int o, int m) {
  const int d = i - o;
  return d < 0 ? d + m : d;
}

// One workgroup per 32 x 32 tile of slab slice s_first + blockIdx.z.  Waves w0 .. w0 + cnt - 1 are
// visited in order; every pixel adds the contribution of each window that covers it, so the sum
// has a fixed order and the gradient is bitwise reproducible.  Contributions of wave w0 + i for
// buffer slice blockIdx.z sit at gt_contrib[(i * contrib_slices + blockIdx.z) * n * n].
template <int N, int TH>
__global__ __launch_bounds__(256) void k_grad_reduce(AdjParams a, long long w0, int cnt, int s_first) {
  constexpr int RPT = TH / 8;  // rows per thread: 8 rows per sweep of the 256 threads
  constexpr int HU = 16 / RPT;  // hits loaded together: sixteen loads in flight per thread
  __shared__ int3 hits[256];   // (wave index in the chunk, oy, ox) of the windows meeting this tile
  __shared__ int wave_hits[4];
  const int sb = blockIdx.z;
  const int H = (int)a.f.slab_h, W = (int)a.f.slab_w;
  const int Y0 = blockIdx.y * TH, X0 = blockIdx.x * RED_TILE;
  const int ylen = min(TH, H - Y0), xlen = min(RED_TILE, W - X0);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int tx = tid % RED_TILE, ty = tid / RED_TILE;
  const int X = X0 + tx;
  float2 acc[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) acc[r] = make_float2(0.f, 0.f);
  for (int base = 0; base < cnt; base += 256) {
    // 1. each thread tests one window against the tile; the hits are compacted in wave order
    const int i = base + tid;
    int oy = 0, ox = 0;
    bool hit = false;
    if (i < cnt) {
      const long long p = (w0 + i) / a.f.M;
      oy = wrap_index(a.f.origins[2 * p], H);
      ox = wrap_index(a.f.origins[2 * p + 1], W);
      // the tile meets the periodic window [o, o + N) if its first row / column lies inside it or the
      // tile runs past the window's start
      const int ey = frame_offset(Y0, oy, H), ex = frame_offset(X0, ox, W);
      hit = (ey < N || ey + ylen > H) && (ex < N || ex + xlen > W);
    }
    const unsigned long long mask = __ballot(hit);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();  // the previous chunk's list has been consumed
    if (lane == 0) wave_hits[wid] = __popcll(mask);
    __syncthreads();
    int offset = 0, total = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      offset += v < wid ? wave_hits[v] : 0;
      total += wave_hits[v];
    }
    if (hit) hits[offset + before] = make_int3(i, oy, ox);
    __syncthreads();
    // 2. ascending wave order over the hits: a fixed summation order.  HU hits at a time: all their
    // loads are issued first (clamped to in-window addresses, always in bounds) and the values are
    // then added through selects, so the loads stay unconditional and overlap.
    if (tx < xlen) {
      for (int h0 = 0; h0 < total; h0 += HU) {
        float2 val[HU][RPT];
        bool use[HU][RPT];
#pragma unroll
        for (int u = 0; u < HU; ++u) {
          const int3 e = hits[min(h0 + u, total - 1)];
          const int dx = frame_offset(X, e.z, W);
          const float2* cw = a.gt_contrib + ((size_t)e.x * a.contrib_slices + sb) * N * N + min(dx, N - 1);
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            const int yl = ty + 8 * r;
            const int dy = frame_offset(Y0 + yl, e.y, H);
            val[u][r] = cw[min(dy, N - 1) * N];
            use[u][r] = h0 + u < total && yl < ylen && dy < N && dx < N;
          }
        }
#pragma unroll
        for (int u = 0; u < HU; ++u)
#pragma unroll
          for (int r = 0; r < RPT; ++r)
            acc[r] = cadd(acc[r], use[u][r] ? val[u][r] : make_float2(0.f, 0.f));
      }
    }
  }
  if (tx >= xlen) return;
  float2* gs = a.grad_slab + (size_t)(s_first + sb) * H * W;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int yl = ty + 8 * r;
    if (yl < ylen) {
      const size_t o = (size_t)(Y0 + yl) * W + X;
      gs[o] = cadd(gs[o], acc[r]);
    }
  }
}

// Probe gradient: element e of mode m sums gp_contrib[i][e] over the waves w0 + i of mode m.  Eight
// interleaved partial sums keep eight loads in flight; they are combined in a fixed order.
template <int N>
__global__ __launch_bounds__(256) void k_probe_reduce(AdjParams a, long long w0, int cnt) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int M = a.f.M;
  if (idx >= (long long)M * N * N) return;
  const int m = (int)(idx / (N * N));
  const int e = (int)(idx % (N * N));
  const int first = (int)(((m - w0 % M) % M + M) % M);  // first i with (w0 + i) % M == m
  const float2* src = a.gp_contrib + e;
  const size_t stride = (size_t)M * N * N;
  float2 part[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) part[u] = make_float2(0.f, 0.f);
  const int count = first < cnt ? (cnt - first + M - 1) / M : 0;  // waves of mode m in the chunk
  const float2* s0 = src + (size_t)first * N * N;
  int j = 0;
  for (; j + 8 <= count; j += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) part[u] = cadd(part[u], s0[(size_t)(j + u) * stride]);
  }
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if (j + u < count) part[u] = cadd(part[u], s0[(size_t)(j + u) * stride]);
  float2 acc = part[0];
#pragma unroll
  for (int u = 1; u < 8; ++u) acc = cadd(acc, part[u]);
  a.grad_probe_k[idx] = cadd(a.grad_probe_k[idx], acc);
}

// Position gradient.  With u = probe_k R, R(k) = exp(-2 pi i k.s / n) / n^2 and the stored per-wave
// gp = conj(R) F(g psi_0), dL/ds_y = Re sum_k conj(gp) probe_k (-2 pi i k_y / n)
//                                  = (2 pi / n) sum_k k_y Im(conj(gp) probe_k),
// k in fftfreq order (Nyquist negative); likewise for x.  One workgroup per position of the chunk,
// summing its modes' waves in the chunk in a fixed order.
template <int N>
__global__ __launch_bounds__(256) void k_pos_grad(AdjParams a, long long w0, int cnt) {
  __shared__ double red[4];
  const int M = a.f.M;
  const long long p = w0 / M + blockIdx.x;
  double sy = 0.0, sx = 0.0;
  for (int m = 0; m < M; ++m) {
    const long long w = p * M + m;
    if (w < w0 || w >= w0 + cnt) continue;  // uniform over the workgroup
    const float2* gp = a.gp_contrib + (size_t)(w - w0) * N * N;
    const float2* pk = a.f.probe_k + (size_t)m * N * N;
    for (int e = threadIdx.x; e < N * N; e += 256) {
      const int ky = e / N, kx = e % N;
      const float2 v = cmulc(pk[e], gp[e]);  // probe_k conj(gp)
      const double im = (double)v.y;
      sy += (double)(ky < N / 2 ? ky : ky - N) * im;
      sx += (double)(kx < N / 2 ? kx : kx - N) * im;
    }
  }
  const double ty = block_sum_fp64<256>(sy, red);
  const double tx = block_sum_fp64<256>(sx, red);
  if (threadIdx.x == 0) {
    const double scale = 2.0 * 3.14159265358979323846 / N;
    a.grad_shifts[2 * p] += scale * ty;
    a.grad_shifts[2 * p + 1] += scale * tx;
  }
}

template <int N>
static void launch_reduces(const AdjParams& a, long long w0, int cnt, int s_first, int s_count, bool probe,
                           hipStream_t st) {
  if (s_count > 0) {
    const unsigned gx = (unsigned)((a.f.slab_w + RED_TILE - 1) / RED_TILE);
    const unsigned tiles32 = gx * (unsigned)((a.f.slab_h + 31) / 32) * (unsigned)s_count;
    if (tiles32 >= 2048) {  // enough 32 x 32 tiles to fill the GPU
      const dim3 grid(gx, (unsigned)((a.f.slab_h + 31) / 32), (unsigned)s_count);
      hipLaunchKernelGGL((k_grad_reduce<N, 32>), grid, dim3(256), 0, st, a, w0, cnt, s_first);
    } else {  // small slabs: 32 x 8 tiles, four times the workgroups
      const dim3 grid(gx, (unsigned)((a.f.slab_h + 7) / 8), (unsigned)s_count);
      hipLaunchKernelGGL((k_grad_reduce<N, 8>), grid, dim3(256), 0, st, a, w0, cnt, s_first);
    }
  }
  if (probe && a.grad_probe_k) {
    const long long total = (long long)a.f.M * N * N;
    hipLaunchKernelGGL(k_probe_reduce<N>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a, w0, cnt);
  }
  if (probe && a.grad_shifts) {
    const long long positions = (w0 + cnt - 1) / a.f.M - w0 / a.f.M + 1;
    hipLaunchKernelGGL(k_pos_grad<N>, dim3((unsigned)positions), dim3(256), 0, st, a, w0, cnt);
  }
}

hipError_t launch_adj_reduces(const AdjParams& a, long long w0, int cnt, int s_first, int s_count, bool probe,
                              hipStream_t st) {
  switch (a.f.n) {
    case 256: launch_reduces<256>(a, w0, cnt, s_first, s_count, probe, st); break;
    case 512: launch_reduces<512>(a, w0, cnt, s_first, s_count, probe, st); break;
    case 1024: launch_reduces<1024>(a, w0, cnt, s_first, s_count, probe, st); break;
    case 2048: launch_reduces<2048>(a, w0, cnt, s_first, s_count, probe, st); break;
    case 4096: launch_reduces<4096>(a, w0, cnt, s_first, s_count, probe, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static int device_cus() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
  }
  return cus > 0 ? cus : 256;
}

static long long resident_slots(int n) {
  const int per_cu = n >= 128 ? 1 : (n == 64 ? 4 : 8);
  return (long long)device_cus() * per_cu;
}

// Waves per streaming batch: 512 of 256^2, 128 of 512^2, then 16 / 4 / 1 of 1024^2 / 2048^2 / 4096^2
// (about 1.4 GiB of segment, work and contribution buffers at k = 4).
long long adjoint_streaming_batch(int n) {
  if (n <= 256) return 512;
  if (n == 512) return 128;
  const long long q = n / 1024;  // 1, 2, 4
  return 16 / (q * q);
}

// Budget for the contribution buffer of the resident path: 4 GiB of the 288 GB of HBM.
constexpr size_t CONTRIB_BUDGET = (size_t)4 << 30;

static long long balanced_chunk(long long waves, long long max_chunk) {
  const long long chunks = (waves + max_chunk - 1) / max_chunk;
  return (waves + chunks - 1) / chunks;
}

size_t adjoint_min_workspace_bytes(int n, int k) {
  const size_t wave_bytes = (size_t)n * n * sizeof(float2);
  return (size_t)(n <= 128 ? k + 1 : k + 2) * wave_bytes;
}

size_t adjoint_workspace_bytes(int n, int S, int k, long long waves) {
  const size_t wave_bytes = (size_t)n * n * sizeof(float2);
  if (n <= 128) {
    long long slots = resident_slots(n);
    if (slots > waves) slots = waves;
    const size_t per_wave = (size_t)(S + 1) * wave_bytes;  // slab contributions of every slice + probe
    long long chunk = (long long)(CONTRIB_BUDGET / per_wave);
    if (chunk < slots) chunk = slots;
    if (chunk > waves) chunk = waves;
    return (size_t)slots * (k + 1) * wave_bytes + (size_t)chunk * per_wave;
  }
  long long nb = adjoint_streaming_batch(n);
  if (nb > waves) nb = waves;
  return (size_t)nb * (2 * k + 3) * wave_bytes;  // seg[k], X, G, k slices of contributions, probe
}

template <int N>
static hipError_t launch_adj_resident(const AdjParams& a0, void* ws, size_t ws_bytes, hipStream_t st) {
  const size_t lds = resident_adj_lds_bytes<N>();
  static bool attr_set[64] = {false};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (dev >= 0 && dev < 64 && !attr_set[dev]) {
    hipError_t e = hipFuncSetAttribute((const void*)k_adjoint_resident<N>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr_set[dev] = true;
  }
  const long long waves = a0.f.P * a0.f.M;
  const size_t NN = (size_t)N * N;
  const size_t per_slot = (size_t)(a0.k + 1) * NN * sizeof(float2);
  const size_t per_wave = (size_t)(a0.f.S + 1) * NN * sizeof(float2);
  long long slots = resident_slots(N);
  if (slots > waves) slots = waves;
  AdjParams a = a0;
  a.gt_contrib = nullptr;
  a.gp_contrib = nullptr;
  a.contrib_slices = a0.f.S;
  long long chunk = waves;
  bool gathered = false;
  if (!a0.f.desc && ws_bytes >= slots * per_slot + (size_t)slots * per_wave) {
    gathered = true;
    chunk = balanced_chunk(waves, (long long)((ws_bytes - slots * per_slot) / per_wave));
  } else {
    const long long fit = (long long)(ws_bytes / per_slot);
    if (slots > fit) slots = fit;
  }
  if (slots < 1 || (a0.grad_shifts && !gathered)) return hipErrorInvalidValue;
  float2* scratch = (float2*)ws;
  float2* contrib = scratch + (size_t)slots * (a0.k + 1) * NN;
  for (long long c0 = 0; c0 < waves; c0 += chunk) {
    const long long cnt = waves - c0 < chunk ? waves - c0 : chunk;
    a.w_begin = c0;
    a.w_end = c0 + cnt;
    if (gathered) {
      a.gt_contrib = contrib;
      a.gp_contrib = a0.grad_probe_k || a0.grad_shifts ? contrib + (size_t)chunk * a0.f.S * NN : nullptr;
    }
    const long long grid = slots < cnt ? slots : cnt;
    hipLaunchKernelGGL(k_adjoint_resident<N>, dim3((unsigned)grid), dim3(N * N / 16), lds, st, a, scratch);
    if (gathered) launch_reduces<N>(a, c0, (int)cnt, 0, a0.f.S, a0.grad_probe_k || a0.grad_shifts, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_adjoint(const AdjParams& a, void* ws, size_t ws_bytes, hipStream_t st) {
  switch (a.f.n) {
    case 32:
      return launch_adj_resident<32>(a, ws, ws_bytes, st);
    case 64:
      return launch_adj_resident<64>(a, ws, ws_bytes, st);
    case 128:
      return launch_adj_resident<128>(a, ws, ws_bytes, st);
    case 256:
    case 512:
    case 1024:
    case 2048:
    case 4096:
      return launch_adj_streaming(a, ws, ws_bytes, st);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace pty
