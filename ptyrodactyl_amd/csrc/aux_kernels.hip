// Auxiliary gfx950 kernels around the multislice: transmission set-up,
// FP64 normal equations for Gauss-Newton / Levenberg-Marquardt, and a small
// FP64 damped SPD solve.
#include <math.h>

#include "aux_kernels.hpp"
#include "common.hpp"

namespace pty {

// T = exp(i sigma V) from a float64 potential, stored complex64.  The phase
// is formed and reduced in float64 (``transmission_func``,
// src/ptyrodactyl/multislice/simulations.py:105-159).
__global__ __launch_bounds__(256) void k_transmission(const double* pot, double sigma, float2* out, long long count) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    double s, c;
    sincos(sigma * pot[i], &s, &c);
    out[i] = make_float2((float)c, (float)s);
  }
}

hipError_t launch_transmission(const double* pot, double sigma, float2* out, long long count, hipStream_t st) {
  long long blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_transmission, dim3((unsigned)blocks), dim3(256), 0, st, pot, sigma, out, count);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Square-root-intensity loss of the gradient workloads, in one pass:
//   L = sum_i (sqrt(I_i + eps) - sqrt(M_i))^2   (FP64)
//   dL/dI_i = (sqrt(I_i + eps) - sqrt(M_i)) / sqrt(I_i + eps)   (float32)
// A fixed grid of SQRT_LOSS_BLOCKS workgroups writes one FP64 partial each (fixed-order tree inside the
// workgroup), and one workgroup adds the partials in order, so the loss is bitwise reproducible.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_sqrt_loss(const float* pat, const float* meas, float eps, long long count,
                                                   float* grad, double* partials) {
  __shared__ double red[4];
  double acc = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (long long)SQRT_LOSS_BLOCKS * 256) {
    const float root = sqrtf(pat[i] + eps);
    const float diff = root - sqrtf(meas[i]);
    if (grad) grad[i] = diff / root;
    acc += (double)diff * (double)diff;
  }
  const double total = block_sum_fp64<256>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* partials, int n, double* out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
  const double total = block_sum_fp64<256>(acc, red);
  if (threadIdx.x == 0) *out = total;
}

hipError_t launch_sqrt_loss(const float* pat, const float* meas, float eps, long long count, float* grad,
                            double* loss, double* partials, hipStream_t st) {
  hipLaunchKernelGGL(k_sqrt_loss, dim3(SQRT_LOSS_BLOCKS), dim3(256), 0, st, pat, meas, eps, count, grad, partials);
  if (loss) hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partials, SQRT_LOSS_BLOCKS, loss);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Normal equations.  For every detector pixel of every position:
//   r   = |A|^2 - data                      (A = amp[p][0])
//   J_j = 2 Re(conj(A) dA_j)                (dA_j = amp[p][1 + j])
// and the kernel accumulates JtJ (Q x Q), Jtr (Q) and sum r^2 in FP64.
// Each workgroup reduces tiles of 256 pixels through LDS into per-product
// FP64 accumulators and writes one partial; a second kernel sums the
// partials in a fixed order (deterministic).
// ---------------------------------------------------------------------------

constexpr int NE_TILE = 256;

__global__ __launch_bounds__(256) void k_normal_partials(const float2* amp, const float* data, long long positions,
                                                         int pix, int Q, double* partials) {
  extern __shared__ double lj[];  // [NE_TILE][Q + 1]: J_0..J_{Q-1}, r
  const int tid = threadIdx.x;
  const int stride = Q + 1;
  const int nprod = Q * (Q + 1) / 2 + Q + 1;
  // product index -> (a, b) with b == Q meaning r
  int pa = -1, pb = -1;
  if (tid < nprod) {
    int t = tid, a = 0;
    while (a <= Q && t >= (Q + 1 - a)) {
      t -= (Q + 1 - a);
      ++a;
    }
    pa = a;
    pb = a + t;
  }
  double acc = 0.0;
  const long long total = positions * (long long)pix;
  const long long tiles = (total + NE_TILE - 1) / NE_TILE;
  for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const long long idx = tile * NE_TILE + tid;
    double* row = lj + tid * stride;
    if (idx < total) {
      const long long p = idx / pix;
      const int q = (int)(idx % pix);
      const float2* ap = amp + (size_t)p * (Q + 1) * pix + q;
      const float2 a0 = ap[0];
      for (int j = 0; j < Q; ++j) {
        const float2 d = ap[(size_t)(j + 1) * pix];
        row[j] = 2.0 * ((double)a0.x * (double)d.x + (double)a0.y * (double)d.y);
      }
      row[Q] = ((double)a0.x * a0.x + (double)a0.y * a0.y) - (double)data[idx];
    } else {
      for (int j = 0; j <= Q; ++j) row[j] = 0.0;
    }
    __syncthreads();
    if (pa >= 0) {
      for (int i = 0; i < NE_TILE; ++i) acc += lj[i * stride + pa] * lj[i * stride + pb];
    }
    __syncthreads();
  }
  if (tid < nprod) partials[(size_t)blockIdx.x * nprod + tid] = acc;
}

__global__ __launch_bounds__(256) void k_normal_reduce(const double* partials, int blocks, int Q, double* jtj,
                                                       double* jtr, double* rss) {
  const int nprod = Q * (Q + 1) / 2 + Q + 1;
  for (int t = threadIdx.x; t < nprod; t += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += partials[(size_t)b * nprod + t];
    int tt = t, a = 0;
    while (a <= Q && tt >= (Q + 1 - a)) {
      tt -= (Q + 1 - a);
      ++a;
    }
    const int b2 = a + tt;
    if (a == Q) {
      *rss = s;
    } else if (b2 == Q) {
      jtr[a] = s;
    } else {
      jtj[a * Q + b2] = s;
      jtj[b2 * Q + a] = s;
    }
  }
}

int normal_partial_blocks(long long positions, int pix) {
  const long long tiles = (positions * (long long)pix + NE_TILE - 1) / NE_TILE;
  return (int)(tiles < 1024 ? tiles : 1024);
}

hipError_t launch_normal_equations(const float2* amp, const float* data, long long positions, int pix, int Q,
                                   double* partials, double* jtj, double* jtr, double* rss, hipStream_t st) {
  const int blocks = normal_partial_blocks(positions, pix);
  const size_t lds = (size_t)NE_TILE * (Q + 1) * sizeof(double);
  hipLaunchKernelGGL(k_normal_partials, dim3(blocks), dim3(256), lds, st, amp, data, positions, pix, Q, partials);
  hipLaunchKernelGGL(k_normal_reduce, dim3(1), dim3(256), 0, st, partials, blocks, Q, jtj, jtr, rss);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Damped SPD solve in FP64: (A + damping I) x = b by Cholesky, plus the
// model reduction <x, b> - 1/2 <x, A x> that the LM gain ratio needs.
// One workgroup; Q <= 64.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(64) void k_damped_solve(const double* A, const double* b, double damping, int Q,
                                                    double* x, double* model_reduction, int* status) {
  __shared__ double L[64 * 64];
  __shared__ double y[64];
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * Q; i += blockDim.x) {
    const int r = i / Q, c = i % Q;
    L[r * 64 + c] = A[i] + (r == c ? damping : 0.0);
  }
  __syncthreads();
  if (tid == 0) {
    int ok = 1;
    for (int j = 0; j < Q; ++j) {
      double d = L[j * 64 + j];
      for (int k = 0; k < j; ++k) d -= L[j * 64 + k] * L[j * 64 + k];
      if (!(d > 0.0)) {
        ok = 0;
        d = 1e-300;
      }
      const double ljj = sqrt(d);
      L[j * 64 + j] = ljj;
      for (int i = j + 1; i < Q; ++i) {
        double s = L[i * 64 + j];
        for (int k = 0; k < j; ++k) s -= L[i * 64 + k] * L[j * 64 + k];
        L[i * 64 + j] = s / ljj;
      }
    }
    for (int i = 0; i < Q; ++i) {  // L y = b
      double s = b[i];
      for (int k = 0; k < i; ++k) s -= L[i * 64 + k] * y[k];
      y[i] = s / L[i * 64 + i];
    }
    for (int i = Q - 1; i >= 0; --i) {  // L^T x = y
      double s = y[i];
      for (int k = i + 1; k < Q; ++k) s -= L[k * 64 + i] * x[k];
      x[i] = s / L[i * 64 + i];
    }
    double sb = 0.0, sas = 0.0;
    for (int i = 0; i < Q; ++i) {
      sb += x[i] * b[i];
      double as = 0.0;
      for (int k = 0; k < Q; ++k) as += A[i * Q + k] * x[k];
      sas += x[i] * as;
    }
    *model_reduction = sb - 0.5 * sas;
    *status = ok;
  }
}

hipError_t launch_damped_solve(const double* A, const double* b, double damping, int Q, double* x,
                               double* model_reduction, int* status, hipStream_t st) {
  hipLaunchKernelGGL(k_damped_solve, dim3(1), dim3(64), 0, st, A, b, damping, Q, x, model_reduction, status);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Virtual annular detector: masked sum of each stored pattern
// ---------------------------------------------------------------------------

// One workgroup per pattern.  Each thread sums its pixels (stride 256) in FP64 with the 0/1 mask,
// then a fixed-order tree in LDS combines the threads: the result is bitwise reproducible.
__global__ __launch_bounds__(256) void k_annular(const float* patterns, long long pixels, const unsigned char* mask,
                                                 double* out) {
  __shared__ double red[256];
  const float* pat = patterns + (long long)blockIdx.x * pixels;
  double acc = 0.0;
  for (long long i = threadIdx.x; i < pixels; i += 256) {
    const float v = pat[i];  // unconditional load; the mask selects
    acc += mask[i] ? (double)v : 0.0;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

hipError_t launch_annular(const float* patterns, long long count, long long pixels, const unsigned char* mask,
                          double* out, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (count > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_annular, dim3((unsigned)count), dim3(256), 0, st, patterns, pixels, mask, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Counter-based normal variates with JAX's key layout (the random modes of
// decompose_beam_to_modes, simulations.py:1446-1536).  Threefry-2x32 with 20
// rounds (Salmon et al., SC'11, Random123) keyed by (k0, k1); element i of the
// flat C-order output uses the counter (i >> 32, i & 0xffffffff), the 64-bit
// word out0 << 32 | out1, the uniform bitcast((bits >> 12) | bits(1.0)) - 1
// scaled to [lo, 1) with lo = nextafter(-1, 0), and sqrt(2) erfinv(u).
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned rotl32(unsigned x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ void threefry2x32(unsigned k0, unsigned k1, unsigned& x0, unsigned& x1) {
  const unsigned ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  const int rot[2][4] = {{13, 15, 26, 6}, {17, 29, 16, 24}};
  x0 += ks[0];
  x1 += ks[1];
#pragma unroll
  for (int block = 0; block < 5; ++block) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      x0 += x1;
      x1 = rotl32(x1, rot[block & 1][j]);
      x1 ^= x0;
    }
    x0 += ks[(block + 1) % 3];
    x1 += ks[(block + 2) % 3] + (unsigned)(block + 1);
  }
}

__global__ __launch_bounds__(256) void k_random_normal(unsigned k0, unsigned k1, long long count, double* out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  unsigned x0 = (unsigned)((unsigned long long)i >> 32), x1 = (unsigned)((unsigned long long)i & 0xffffffffull);
  threefry2x32(k0, k1, x0, x1);
  const unsigned long long bits = ((unsigned long long)x0 << 32) | (unsigned long long)x1;
  const double f = __longlong_as_double((long long)((bits >> 12) | 0x3FF0000000000000ull)) - 1.0;
  const double lo = -1.0 + 0x1.0p-53;  // nextafter(-1, 0)
  double u = f * (1.0 - lo) + lo;
  u = u > lo ? u : lo;
  out[i] = 1.4142135623730951 * erfinv(u);
}

hipError_t launch_random_normal(unsigned k0, unsigned k1, long long count, double* out, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_random_normal, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, k0, k1, count, out);
  return hipGetLastError();
}

}  // namespace pty
