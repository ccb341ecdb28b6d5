// Persistent, XCD-local streaming forward for windows that exceed LDS (n = 256).
//
// The per-slice row and column passes of the streaming forward exchange the
// whole wave through memory twice per slice.  With one launch per pass that
// traffic leaves the L2 (each XCD has its own 4 MB L2, kernel boundaries
// write it back), so the path runs at HBM speed.  This kernel keeps small
// groups of waves inside one XCD's L2 for the whole slice loop:
//
//  * every workgroup reads its XCD id (HW_REG_XCC_ID) and takes tasks from
//    that XCD's ticket counter only; the waves of a group are touched only by
//    workgroups of that XCD, so they live in that XCD's L2;
//  * two groups per XCD are interleaved phase by phase (init column pass,
//    row pass s, column pass s, ..., final column pass), so a workgroup that
//    finishes a task of one group moves on to the other group while the
//    first group's phase completes;
//  * a task of phase p waits until the group's completed-task counter shows
//    phases 0..p-1 done.  Producers wait for their stores (s_waitcnt 0) before
//    counting; consumers read the wave with device-scope loads, which bypass
//    the CU's L1 and read the XCD's L2;
//  * tasks are claimed dynamically, so no co-residency assumption is needed;
//    every wait is bounded and reports an error instead of hanging.
//
// Arithmetic is identical to k_rows / k_cols in forward.hip (same FFT code,
// same operation order), so results are bitwise equal to the per-pass path.
#include "common.hpp"
#include "forward.hpp"
#include "xl2.hpp"

namespace pty {

namespace {

constexpr int XL2_THREADS = 256;
constexpr int XL2_XCDS = 8;
constexpr int XL2_MAX_GROUP = 16;
constexpr int XL2_LINE = 16;  // ints per 64-byte line
// per XCD: ticket counter, then one completion counter per (slot, wave of the group), each on its own line
constexpr int XL2_STATE_STRIDE = XL2_LINE * (1 + 2 * XL2_MAX_GROUP);
constexpr int XL2_SPIN_LIMIT = 1 << 22;  // about a second of polling; legitimate waits last microseconds

__device__ __forceinline__ float2 load_l2(const float2* p) {
  const unsigned long long v =
      __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  float2 r;
  __builtin_memcpy(&r, &v, sizeof(r));
  return r;
}

__device__ __forceinline__ int xcc_id() {
  // HW_REG_XCC_ID (hwreg 20), bits [3:0]
  return __builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & 0xF;
}

template <int NC>
struct ColBufX {
  float2* base;
  __device__ __forceinline__ float2 ld(int e) const { return base[pad16(e) * NC]; }
  __device__ __forceinline__ void st(int e, float2 v) const { base[pad16(e) * NC] = v; }
  __device__ __forceinline__ float2 ldp(int pe) const { return base[pe * NC]; }
  __device__ __forceinline__ void stp(int pe, float2 v) const { base[pe * NC] = v; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};

}  // namespace

template <int N>
__global__ __launch_bounds__(XL2_THREADS, 2) void k_forward_xl2(FwdParams a, float2* ws, int* state, int g,
                                                             long long waves) {
  constexpr int T = N / 16;
  constexpr int RPB = XL2_THREADS / T;  // rows per task
  constexpr int NC = XL2_THREADS / T;   // columns per task
  constexpr int RS = N + N / 16;
  constexpr int BLK = N / RPB;          // tasks per wave and phase
  __shared__ float2 scratch[(RPB * RS > (N + N / 16) * NC) ? RPB * RS : (N + N / 16) * NC];
  __shared__ float2 tw[twiddle_entries<N>()];
  __shared__ float2 pvx[N];
  __shared__ float2 pvy[N];
  constexpr bool DF = N <= DF_MAX_LINE;  // the same tables as the per-pass kernels of this size
  __shared__ float2 twf[DF ? twiddle_entries<N>() : 1];
  __shared__ float2 pvxf[DF ? N : 1];
  __shared__ float2 pvyf[DF ? N : 1];
  __shared__ long long s_ticket;
  __shared__ int s_abort;

  const int tid = threadIdx.x;
  const int x = xcc_id() & (XL2_XCDS - 1);
  int* st = state + XL2_LINE + x * XL2_STATE_STRIDE;
  int* ticket = st;
  int* done = st + XL2_LINE;  // done[(slot * XL2_MAX_GROUP + wl) * XL2_LINE]: completed tasks of one wave
  int* err = state;
  stage_twiddles<N, DF>(tw, twf, tid, XL2_THREADS);
  // without position descriptors every wave shares one propagator
  stage_factors<DF>(pvx, pvxf, a.prop_x, N, tid, XL2_THREADS);
  stage_factors<DF>(pvy, pvyf, a.prop_y, N, tid, XL2_THREADS);
  const TwTab<DF> twt{tw, twf};
  const FixTab<DF> pxt{pvx, pvxf}, pyt{pvy, pvyf};

  const int S = a.S;
  const long long groups = (waves + g - 1) / g;
  const long long local_groups = groups > x ? (groups - x + XL2_XCDS - 1) / XL2_XCDS : 0;
  const long long pairs = (local_groups + 1) / 2;
  const int nph = 2 * S + 1;
  const int tpp = g * BLK;
  const long long tickets = pairs * 2LL * nph * tpp;
  float2* slot_base[2] = {ws + ((size_t)(x * 2 + 0) * g) * N * N, ws + ((size_t)(x * 2 + 1) * g) * N * N};

  for (;;) {
    __syncthreads();
    if (tid == 0) {
      s_ticket = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_abort = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const long long t = s_ticket;
    if (t >= tickets || s_abort) break;
    const long long k = t / (2LL * nph * tpp);
    const int r = (int)(t % (2LL * nph * tpp));
    const int phase = r / (2 * tpp);
    const int slot = (r / tpp) & 1;
    const int task = r % tpp;
    const long long lg = 2 * k + slot;  // local group index on this XCD
    if (lg >= local_groups) continue;    // odd tail: slot 1 of the last pair is empty
    const long long gi = x + XL2_XCDS * lg;
    const int wl = task / BLK, blk = task % BLK;
    const long long w = gi * g + wl;
    const bool active = w < waves;

    // wait until phases 0..phase-1 of this wave (and every earlier wave in its place) are complete
    int* wdone = done + (slot * XL2_MAX_GROUP + wl) * XL2_LINE;
    if (tid == 0) {
      const int target = (int)((k * nph + phase) * BLK);
      int spins = 0;
      while (__hip_atomic_load(wdone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        ++spins;
        if ((spins & 1023) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          s_abort = 1;
          break;
        }
        if (spins > XL2_SPIN_LIMIT) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          s_abort = 1;
          break;
        }
      }
    }
    __syncthreads();
    if (s_abort) break;

    if (active) {
      // an opaque copy of the thread index per task keeps the compiler from hoisting
      // per-thread index and twiddle arithmetic of all six FFT sites out of the loop
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));
      const WaveCtx c = resolve_wave(a, w);
      float2* wave = slot_base[slot] + (size_t)wl * N * N;
      if (phase % 2 == 1) {  // row pass of slice s
        const int s = (phase - 1) / 2;
        const bool first = s == 0, last = s == S - 1;
        const int t16 = tid % T, lr = tid / T, y = blk * RPB + lr;
        const RowBufWave rb{scratch + lr * RS};
        float2* wrow = wave + (size_t)y * N;
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = load_l2(wrow + t16 + T * q);
        if (first) line_fft<N, true>(v, t16, rb, twt);
        int gy = c.oy + y;
        if (gy >= c.slab_h) gy -= (int)c.slab_h;
        const float2* trow = c.slab + (long long)s * c.stride + (long long)gy * c.slab_w;
        multiply_window_row<N>(v, trow, c.ox, t16, (int)c.slab_w);
        line_fft<N, false>(v, t16, rb, twt);
        if (!last) {
#pragma unroll
          for (int q = 0; q < 16; ++q) v[q] = pxt.mul(v[q], t16 + T * q);
          line_fft<N, true>(v, t16, rb, twt);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) wrow[t16 + T * q] = v[q];
      } else {  // column passes: init (phase 0), propagate (slice s), final (phase 2S)
        const int gc = tid % NC, t16 = tid / NC;
        const int xcol = blk * NC + gc;
        const ColBufX<NC> cb{scratch + gc};
        const bool init = phase == 0, fin = phase == 2 * S;
        float2 v[16];
        float2* wcol = wave + xcol;
        if (init) {
          const float2 rx = shift_ramp(xcol, N, c.mx, c.dx);
          const float2* pk = c.probe_k + xcol;
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int ky = t16 + T * q;
            v[q] = cmul(cmul(pk[(size_t)ky * N], shift_ramp(ky, N, c.my, c.dy)), rx);
          }
          line_fft<N, true>(v, t16, cb, twt);
#pragma unroll
          for (int q = 0; q < 16; ++q) wcol[(size_t)(t16 + T * q) * N] = v[q];
        } else {
#pragma unroll
          for (int q = 0; q < 16; ++q) v[q] = load_l2(wcol + (size_t)(t16 + T * q) * N);
          line_fft<N, false>(v, t16, cb, twt);
          if (!fin) {
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = pyt.mul(v[q], t16 + T * q);
            line_fft<N, true>(v, t16, cb, twt);
#pragma unroll
            for (int q = 0; q < 16; ++q) wcol[(size_t)(t16 + T * q) * N] = v[q];
          } else {
            const float wm = a.weights ? a.weights[0] : 1.f;
            const int xs = (xcol + N / 2) & (N - 1);
            if (a.patterns) {
              float* out = a.patterns + c.pat_off;
#pragma unroll
              for (int q = 0; q < 16; ++q) out[(size_t)((t16 + T * q + N / 2) & (N - 1)) * N + xs] = wm * cabs2(v[q]);
            }
          }
        }
      }
    }
    // publish: every thread's stores are acknowledged by the L2 before the task is counted
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(wdone, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void k_xl2_reset(int* state) {
  for (int i = threadIdx.x; i < XL2_LINE + XL2_XCDS * XL2_STATE_STRIDE; i += blockDim.x) state[i] = 0;
}

// Every XCD must have completed all of its tasks: a launch whose workgroups did not reach
// some XCD (a partitioned device, an unexpected XCC numbering) leaves work undone.  Sets the
// error flag to 2 in that case.
__global__ void k_xl2_check(int* state, int S, int blk, int g, long long waves) {
  const int x = threadIdx.x / XL2_MAX_GROUP, wl = threadIdx.x % XL2_MAX_GROUP;
  if (x >= XL2_XCDS || wl >= g) return;
  const long long groups = (waves + g - 1) / g;
  const long long local_groups = groups > x ? (groups - x + XL2_XCDS - 1) / XL2_XCDS : 0;
  const long long per_group = (long long)(2 * S + 1) * blk;
  const long long want0 = ((local_groups + 1) / 2) * per_group, want1 = (local_groups / 2) * per_group;
  const int* done = state + XL2_LINE + x * XL2_STATE_STRIDE + XL2_LINE;
  if (done[wl * XL2_LINE] != want0 || done[(XL2_MAX_GROUP + wl) * XL2_LINE] != want1)
    atomicOr(state, 2);
}

size_t xl2_state_bytes() { return (size_t)(XL2_LINE + XL2_XCDS * XL2_STATE_STRIDE) * sizeof(int); }

size_t xl2_workspace_bytes(int n, int g) { return (size_t)XL2_XCDS * 2 * g * n * n * sizeof(float2); }

hipError_t launch_forward_xl2(const FwdParams& a, void* ws, size_t ws_bytes, int* state, int g, int grid,
                              hipStream_t st) {
  if (a.n != 256 || a.M != 1 || g < 1 || g > XL2_MAX_GROUP || a.desc) return hipErrorInvalidValue;
  if (ws_bytes < xl2_workspace_bytes(a.n, g)) return hipErrorInvalidValue;
  {  // ticket and completion counters are 32-bit
    const long long groups = (a.P * (long long)a.M + g - 1) / g;
    const long long per_xcd = (groups + XL2_XCDS - 1) / XL2_XCDS + 1;
    if (per_xcd * (2LL * a.S + 1) * g * (256 / (XL2_THREADS / 16)) + grid >= (1LL << 31)) return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(k_xl2_reset, dim3(1), dim3(256), 0, st, state);
  const long long waves = a.P * (long long)a.M;
  hipLaunchKernelGGL(k_forward_xl2<256>, dim3(grid), dim3(XL2_THREADS), 0, st, a, (float2*)ws, state, g, waves);
  constexpr int RPB = XL2_THREADS / (256 / 16);
  hipLaunchKernelGGL(k_xl2_check, dim3(1), dim3(XL2_XCDS * XL2_MAX_GROUP), 0, st, state, a.S, 256 / RPB, g, waves);
  return hipGetLastError();
}

}  // namespace pty
