// Independent-atom potential builders on gfx950, all in FP64:
//
//  * k_projected_radii — the projected potential of one element at given
//    radii (src/ptyrodactyl/multislice/form_factors.py:323-439);
//  * k_atom_grid — every element's potential on the (height, width) grid,
//    evaluated on the supersampled grid and block-averaged in the same
//    thread (single_atom_potential, atom_potentials.py:588-681);
//  * k_slice_structure — the Fourier-space slices
//      F_s(k) = sum_t Khat_t(k) sum_{atoms a of element t in slice s}
//               exp(2 pi i (kx sx_a + ky sy_a)),
//    which is the reference's per-atom Fourier shift
//    (_process_all_slices, atom_potentials.py:826-993) summed by linearity,
//    so one inverse FFT per slice replaces one FFT pair per atom.
//
// K0 and K1 follow the reference's evaluation: the polynomial
// approximations of Abramowitz & Stegun 9.8.5-9.8.8 (small branch x <= 2
// with I0, I1 by their power series; large branch x > 2).
#include <math.h>

#include "common.hpp"
#include "potential.hpp"

namespace pty {

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kMottBethe = 47.87801;           // V A^2 (Lobato normalisation)
constexpr double kKirklandA0Ek = 0.5292 * 14.4;   // legacy Kirkland normalisation

__device__ __forceinline__ double bessel_i0_series(double x) {
  const double q = 0.25 * x * x;
  double term = 1.0, total = 1.0;
#pragma unroll
  for (int k = 1; k < 26; ++k) {
    term *= q / (double)(k * k);
    total += term;
  }
  return total;
}

__device__ __forceinline__ double bessel_i1_series(double x) {
  const double q = 0.25 * x * x;
  double term = 0.5 * x, total = term;
#pragma unroll
  for (int k = 1; k < 26; ++k) {
    term *= q / (double)(k * (k + 1));
    total += term;
  }
  return total;
}

// A&S 9.8.5 / 9.8.6.
__device__ double bessel_k0(double x) {
  if (x <= 2.0) {
    const double q = 0.25 * x * x;
    const double poly = -0.57721566 + q * (0.42278420 + q * (0.23069756 + q * (0.03488590 + q * (0.00262698 +
                        q * (0.00010750 + q * 0.00000740)))));
    return -log(0.5 * x) * bessel_i0_series(x) + poly;
  }
  const double y = 2.0 / x;
  const double poly = 1.25331414 + y * (-0.07832358 + y * (0.02189568 + y * (-0.01062446 + y * (0.00587872 +
                      y * (-0.00251540 + y * 0.00053208)))));
  return exp(-x) / sqrt(x) * poly;
}

// A&S 9.8.7 / 9.8.8.
__device__ double bessel_k1(double x) {
  if (x <= 2.0) {
    const double q = 0.25 * x * x;
    const double poly = 1.0 + q * (0.15443144 + q * (-0.67278579 + q * (-0.18156897 + q * (-0.01919402 +
                        q * (-0.00110404 + q * -0.00004686)))));
    return log(0.5 * x) * bessel_i1_series(x) + poly / x;
  }
  const double y = 2.0 / x;
  const double poly = 1.25331414 + y * (0.23498619 + y * (-0.03655620 + y * (0.01504268 + y * (-0.00780353 +
                      y * (0.00325614 + y * -0.00068245)))));
  return exp(-x) / sqrt(x) * poly;
}

// Projected potential (V A) at radius r of one element's coefficient row.
__device__ __forceinline__ double projected_potential(int kind, const double* row, double r) {
  r = fmax(r, 1e-10);
  if (kind == PARAM_LOBATO) {
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const double a = row[2 * i], b = row[2 * i + 1];
      const double sb = sqrt(b);
      const double arg = 2.0 * kPi * r / sb;
      sum += a / b * bessel_k0(arg) + a * kPi * r / (b * sb) * bessel_k1(arg);
    }
    return 2.0 * kPi * kMottBethe * sum;
  }
  double lor[3], gau[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    lor[i] = row[2 * i] * bessel_k0(2.0 * kPi * r * sqrt(row[2 * i + 1]));
    const double c = row[6 + 2 * i], d = row[7 + 2 * i];
    gau[i] = c / d * exp(-(kPi * kPi) * r * r / d);
  }
  return 4.0 * kPi * kPi * kKirklandA0Ek * ((lor[0] + lor[1]) + lor[2]) +
         2.0 * kPi * kPi * kKirklandA0Ek * ((gau[0] + gau[1]) + gau[2]);
}

// dV/dr of projected_potential: the Bessel identities K0'(x) = -K1(x) and (x K1(x))' = -x K0(x) applied to the
// same A&S evaluations.  Below the radius floor the potential does not depend on r.
__device__ __forceinline__ double projected_potential_dr(int kind, const double* row, double r) {
  if (r < 1e-10) return 0.0;
  if (kind == PARAM_LOBATO) {
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const double a = row[2 * i], b = row[2 * i + 1];
      const double sb = sqrt(b);
      const double arg = 2.0 * kPi * r / sb;
      // d/dr [a/b K0(arg)] = -a/b (2 pi / sqrt b) K1(arg);  d/dr [a pi r / b^1.5 K1(arg)] = -a pi / b^1.5 arg K0(arg)
      sum -= a / b * (2.0 * kPi / sb) * bessel_k1(arg) + a * kPi / (b * sb) * arg * bessel_k0(arg);
    }
    return 2.0 * kPi * kMottBethe * sum;
  }
  double lor = 0.0, gau = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double sb = sqrt(row[2 * i + 1]);
    lor -= row[2 * i] * 2.0 * kPi * sb * bessel_k1(2.0 * kPi * r * sb);
    const double c = row[6 + 2 * i], d = row[7 + 2 * i];
    gau -= c / d * exp(-(kPi * kPi) * r * r / d) * (2.0 * kPi * kPi * r / d);
  }
  return 4.0 * kPi * kPi * kKirklandA0Ek * lor + 2.0 * kPi * kPi * kKirklandA0Ek * gau;
}

__global__ __launch_bounds__(256) void k_projected_radii(int kind, const double* __restrict__ row, long long count,
                                                         const double* __restrict__ r, double* __restrict__ out) {
  double p[PARAM_STRIDE];
#pragma unroll
  for (int i = 0; i < PARAM_STRIDE; ++i) p[i] = row[i];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = projected_potential(kind, p, r[i]);
}

// out[t][i][j] = mean over the supersampled block of pixel (i, j) of the
// potential of element t centred at centers[t] (x, y in A; origin at the
// supersampled index (height*ss/2, width*ss/2)).
__global__ __launch_bounds__(256) void k_atom_grid(int kind, const double* __restrict__ rows, double pixel_size,
                                                   int height, int width, int ss, const double* __restrict__ centers,
                                                   double* __restrict__ out) {
  const int t = blockIdx.y;
  double p[PARAM_STRIDE];
#pragma unroll
  for (int i = 0; i < PARAM_STRIDE; ++i) p[i] = rows[(long long)t * PARAM_STRIDE + i];
  const double cx = centers ? centers[2 * t] : 0.0;
  const double cy = centers ? centers[2 * t + 1] : 0.0;
  const double step = pixel_size / ss;
  const int gh2 = (height * ss) / 2, gw2 = (width * ss) / 2;
  const long long pix = (long long)height * width;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < pix;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / width), j = (int)(e % width);
    double acc = 0.0;
    for (int a = 0; a < ss; ++a) {
      const double dy = (double)(i * ss + a - gh2) * step - cy;
      for (int b = 0; b < ss; ++b) {
        const double dx = (double)(j * ss + b - gw2) * step - cx;
        acc += projected_potential(kind, p, sqrt(dx * dx + dy * dy + 1e-10));
      }
    }
    out[(long long)t * pix + e] = acc / (double)(ss * ss);
  }
}

// Centre gradient of k_atom_grid: block b of element t sums, over its pixels (grid stride) and their
// supersamples, grad_out[t][i][j] * dV/dr * dr/dc with dr/dc = -(x - c) / r, and writes the (x, y) pair to
// partial[t][b] after a fixed-order tree reduction; k_atom_center_finish adds the blocks in order.
constexpr int CG_BLOCKS = 512;

__global__ __launch_bounds__(256) void k_atom_center_grad(int kind, const double* __restrict__ rows,
                                                          double pixel_size, int height, int width, int ss,
                                                          const double* __restrict__ centers,
                                                          const double* __restrict__ grad_out,
                                                          double2* __restrict__ partial) {
  __shared__ double2 red[256];
  const int t = blockIdx.y;
  double p[PARAM_STRIDE];
#pragma unroll
  for (int i = 0; i < PARAM_STRIDE; ++i) p[i] = rows[(long long)t * PARAM_STRIDE + i];
  const double cx = centers ? centers[2 * t] : 0.0;
  const double cy = centers ? centers[2 * t + 1] : 0.0;
  const double step = pixel_size / ss;
  const int gh2 = (height * ss) / 2, gw2 = (width * ss) / 2;
  const long long pix = (long long)height * width;
  const double* g = grad_out + (long long)t * pix;
  double gx = 0.0, gy = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < pix;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / width), j = (int)(e % width);
    const double ge = g[e];
    double sx = 0.0, sy = 0.0;
    for (int a = 0; a < ss; ++a) {
      const double dy = (double)(i * ss + a - gh2) * step - cy;
      for (int b = 0; b < ss; ++b) {
        const double dx = (double)(j * ss + b - gw2) * step - cx;
        const double r = sqrt(dx * dx + dy * dy + 1e-10);
        const double w = projected_potential_dr(kind, p, r) / r;
        sx -= w * dx;
        sy -= w * dy;
      }
    }
    gx = fma(ge, sx, gx);
    gy = fma(ge, sy, gy);
  }
  red[threadIdx.x] = make_double2(gx, gy);
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      const double2 o = red[threadIdx.x + h];
      red[threadIdx.x].x += o.x;
      red[threadIdx.x].y += o.y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(long long)t * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_atom_center_finish(int num_types, int blocks, double scale,
                                                            const double2* __restrict__ partial,
                                                            double* __restrict__ grad_centers) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num_types) return;
  double sx = 0.0, sy = 0.0;
  for (int b = 0; b < blocks; ++b) {
    sx += partial[(long long)t * blocks + b].x;
    sy += partial[(long long)t * blocks + b].y;
  }
  grad_centers[2 * t] = sx * scale;
  grad_centers[2 * t + 1] = sy * scale;
}

// ---------------------------------------------------------------------------
// Structure-factor slices.  A workgroup owns a 64 x 128 tile of one slice's
// Fourier grid; thread (tr, tc) (tc = lane % 16, tr = 4 wave + lane / 16)
// holds the 4 x 8 pixels rows tr + 16 i, columns tc + 16 j.  For each
// element the slice's atoms stream through LDS in chunks of 16: the phase
// factors u_a(ky) = exp(2 pi i ky sy_a) (64 rows) and v_a(kx) =
// exp(2 pi i kx sx_a) (128 columns) are generated in runs of 16 consecutive
// frequencies, each run from one sincospi and a multiplicative step
// exp(2 pi i s / n) (a fresh sincospi where fftfreq wraps to negative
// frequencies); then every thread accumulates sum_a u_a(ky_i) v_a(kx_j) for
// its 32 pixels, 12 LDS reads per 128 FMAs.  The element's sum is weighted
// by Khat_t and written (first element) or added (later elements) into the
// slice in HBM by the thread that owns the pixels.  Frequencies follow
// numpy's fftfreq.
// ---------------------------------------------------------------------------

constexpr int SF_ROWS = 64;
constexpr int SF_COLS = 128;
constexpr int SF_RPT = 4;
constexpr int SF_CPT = 8;
constexpr int SF_CHUNK = 16;
constexpr int SF_RUN = 16;
constexpr int SF_RUNS_PER_ATOM = (SF_ROWS + SF_COLS) / SF_RUN;  // 4 row runs + 8 column runs

__device__ __forceinline__ double2 phase_factor(int g, int n, double shift) {
  const int k = (g < (n + 1) / 2) ? g : g - n;
  double t = (double)k * shift / (double)n;
  t -= rint(t);
  double s, c;
  sincospi(2.0 * t, &s, &c);
  return make_double2(c, s);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// Phase factors of atoms c0 .. c0 + n - 1 for one 64 x 128 tile: u[a][row] = exp(2 pi i ky sy_a) and
// v[a][col] = exp(2 pi i kx sx_a), runs of 16 frequencies per thread (see above).
__device__ __forceinline__ void stage_phases(double2 (*u_lds)[SF_ROWS], double2 (*v_lds)[SF_COLS],
                                             const double* __restrict__ shifts, long long c0, int n, int height,
                                             int width, int y0, int x0) {
  if ((int)threadIdx.x < n * SF_RUNS_PER_ATOM) {
    const int a = threadIdx.x / SF_RUNS_PER_ATOM, run = threadIdx.x % SF_RUNS_PER_ATOM;
    const double* sh = shifts + 2 * (c0 + a);
    const bool is_row = run < SF_ROWS / SF_RUN;
    const int dim = is_row ? height : width;
    const double shift = is_row ? sh[1] : sh[0];
    const int q0 = is_row ? run * SF_RUN : (run - SF_ROWS / SF_RUN) * SF_RUN;
    const int g0 = (is_row ? y0 : x0) + q0;
    double2* dst = is_row ? &u_lds[a][q0] : &v_lds[a][q0];
    double ts = shift / (double)dim;
    ts -= rint(ts);
    double ss, cs;
    sincospi(2.0 * ts, &ss, &cs);
    const double2 step = make_double2(cs, ss);
    const int wrap = (dim + 1) / 2;
    double2 val = phase_factor(g0, dim, shift);
    dst[0] = val;
    for (int q = 1; q < SF_RUN; ++q) {
      val = (g0 + q == wrap) ? phase_factor(g0 + q, dim, shift) : cmul(val, step);
      dst[q] = val;
    }
  }
}

__global__ __launch_bounds__(256) void k_slice_structure(int num_types, int height, int width, long long num_atoms,
                                                         const double* __restrict__ shifts,
                                                         const long long* __restrict__ segments,
                                                         const double2* __restrict__ kernel_k,
                                                         double2* __restrict__ out_k) {
  __shared__ double2 u_lds[SF_CHUNK][SF_ROWS];
  __shared__ double2 v_lds[SF_CHUNK][SF_COLS];
  const int s = blockIdx.z;
  const int x0 = blockIdx.x * SF_COLS, y0 = blockIdx.y * SF_ROWS;
  const int lane = threadIdx.x & 63;
  const int tc = lane & 15, tr = (threadIdx.x >> 6) * 4 + (lane >> 4);
  const long long plane = (long long)height * width;
  double2* out_s = out_k + (long long)s * plane;
  bool written = false;

  for (int t = 0; t < num_types; ++t) {
    long long a0 = segments[(long long)s * num_types + t];
    long long a1 = segments[(long long)s * num_types + t + 1];
    a0 = a0 < 0 ? 0 : (a0 > num_atoms ? num_atoms : a0);
    a1 = a1 < 0 ? 0 : (a1 > num_atoms ? num_atoms : a1);
    if (a1 <= a0) continue;
    double2 acc[SF_RPT][SF_CPT];
#pragma unroll
    for (int i = 0; i < SF_RPT; ++i)
#pragma unroll
      for (int j = 0; j < SF_CPT; ++j) acc[i][j] = make_double2(0.0, 0.0);
    for (long long c0 = a0; c0 < a1; c0 += SF_CHUNK) {
      const int n = (int)((a1 - c0) < SF_CHUNK ? (a1 - c0) : SF_CHUNK);
      __syncthreads();
      stage_phases(u_lds, v_lds, shifts, c0, n, height, width, y0, x0);
      __syncthreads();
#pragma unroll 2
      for (int a = 0; a < n; ++a) {
        double2 u[SF_RPT], v[SF_CPT];
#pragma unroll
        for (int i = 0; i < SF_RPT; ++i) u[i] = u_lds[a][tr + 16 * i];
#pragma unroll
        for (int j = 0; j < SF_CPT; ++j) v[j] = v_lds[a][tc + 16 * j];
#pragma unroll
        for (int i = 0; i < SF_RPT; ++i)
#pragma unroll
          for (int j = 0; j < SF_CPT; ++j) {
            acc[i][j].x = fma(u[i].x, v[j].x, fma(-u[i].y, v[j].y, acc[i][j].x));
            acc[i][j].y = fma(u[i].x, v[j].y, fma(u[i].y, v[j].x, acc[i][j].y));
          }
      }
    }
    const double2* kt = kernel_k + (long long)t * plane;
#pragma unroll
    for (int i = 0; i < SF_RPT; ++i) {
      const int row = y0 + tr + 16 * i;
      if (row >= height) continue;
#pragma unroll
      for (int j = 0; j < SF_CPT; ++j) {
        const int col = x0 + tc + 16 * j;
        if (col >= width) continue;
        const long long e = (long long)row * width + col;
        double2 f = cmul(kt[e], acc[i][j]);
        if (written) {
          const double2 o = out_s[e];
          f.x += o.x;
          f.y += o.y;
        }
        out_s[e] = f;
      }
    }
    written = true;
  }
  if (!written) {
#pragma unroll
    for (int i = 0; i < SF_RPT; ++i) {
      const int row = y0 + tr + 16 * i;
      if (row >= height) continue;
#pragma unroll
      for (int j = 0; j < SF_CPT; ++j) {
        const int col = x0 + tc + 16 * j;
        if (col < width) out_s[(long long)row * width + col] = make_double2(0.0, 0.0);
      }
    }
  }
}

// Shift gradient of the structure-factor slices.  With A_s = ifft2 of dL/dV_s (zero outside the cropped
// region) and W_st = 2 pi i Khat_t A_s, atom a of element t in slice s has
//   dL/dsx_a = Re sum_k W_st(k) kx exp(2 pi i (kx sx_a + ky sy_a)),   dL/dsy_a likewise with ky,
// kx, ky in cycles per pixel (numpy's fftfreq).  A workgroup owns a 64 x 128 tile as k_slice_structure does; per
// element it holds W for its 32 pixels, stages the phases of 16 atoms at a time, sums its pixels for each atom,
// reduces the workgroup in a fixed order and writes the tile's pair to partial[tile][atom];
// k_shift_grad_finish adds the tiles of each atom in order.
__global__ __launch_bounds__(256) void k_slice_shift_grad(int num_types, int height, int width, long long num_atoms,
                                                          const double* __restrict__ shifts,
                                                          const long long* __restrict__ segments,
                                                          const double2* __restrict__ kernel_k,
                                                          const double2* __restrict__ adj_k,
                                                          double2* __restrict__ partial) {
  __shared__ double2 u_lds[SF_CHUNK][SF_ROWS];
  __shared__ double2 v_lds[SF_CHUNK][SF_COLS];
  __shared__ double2 wsum[4][SF_CHUNK];
  const int s = blockIdx.z;
  const int x0 = blockIdx.x * SF_COLS, y0 = blockIdx.y * SF_ROWS;
  const long long tile = (long long)blockIdx.y * gridDim.x + blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int tc = lane & 15, tr = wid * 4 + (lane >> 4);
  const long long plane = (long long)height * width;
  const double2* as = adj_k + (long long)s * plane;
  double kx[SF_CPT], ky[SF_RPT];
#pragma unroll
  for (int j = 0; j < SF_CPT; ++j) {
    const int col = x0 + tc + 16 * j;
    kx[j] = (double)(col < (width + 1) / 2 ? col : col - width) / (double)width;
  }
#pragma unroll
  for (int i = 0; i < SF_RPT; ++i) {
    const int row = y0 + tr + 16 * i;
    ky[i] = (double)(row < (height + 1) / 2 ? row : row - height) / (double)height;
  }
  for (int t = 0; t < num_types; ++t) {
    long long a0 = segments[(long long)s * num_types + t];
    long long a1 = segments[(long long)s * num_types + t + 1];
    a0 = a0 < 0 ? 0 : (a0 > num_atoms ? num_atoms : a0);
    a1 = a1 < 0 ? 0 : (a1 > num_atoms ? num_atoms : a1);
    if (a1 <= a0) continue;
    const double2* kt = kernel_k + (long long)t * plane;
    double2 wk[SF_RPT][SF_CPT];  // 2 pi i Khat_t A_s, zero outside the grid
#pragma unroll
    for (int i = 0; i < SF_RPT; ++i) {
      const int row = y0 + tr + 16 * i;
#pragma unroll
      for (int j = 0; j < SF_CPT; ++j) {
        const int col = x0 + tc + 16 * j;
        double2 w = make_double2(0.0, 0.0);
        if (row < height && col < width) {
          const long long e = (long long)row * width + col;
          const double2 z = cmul(kt[e], as[e]);
          w = make_double2(-2.0 * kPi * z.y, 2.0 * kPi * z.x);
        }
        wk[i][j] = w;
      }
    }
    for (long long c0 = a0; c0 < a1; c0 += SF_CHUNK) {
      const int n = (int)((a1 - c0) < SF_CHUNK ? (a1 - c0) : SF_CHUNK);
      __syncthreads();
      stage_phases(u_lds, v_lds, shifts, c0, n, height, width, y0, x0);
      __syncthreads();
      for (int a = 0; a < n; ++a) {
        double gx = 0.0, gy = 0.0;
#pragma unroll
        for (int i = 0; i < SF_RPT; ++i) {
          double2 sum = make_double2(0.0, 0.0), sumx = make_double2(0.0, 0.0);
#pragma unroll
          for (int j = 0; j < SF_CPT; ++j) {
            const double2 c = cmul(wk[i][j], v_lds[a][tc + 16 * j]);
            sum.x += c.x;
            sum.y += c.y;
            sumx.x = fma(c.x, kx[j], sumx.x);
            sumx.y = fma(c.y, kx[j], sumx.y);
          }
          const double2 u = u_lds[a][tr + 16 * i];
          gx += u.x * sumx.x - u.y * sumx.y;
          gy += ky[i] * (u.x * sum.x - u.y * sum.y);
        }
        // fixed-order reduction: butterfly within the wave, then the four waves in order
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          gx += __shfl_xor(gx, off, 64);
          gy += __shfl_xor(gy, off, 64);
        }
        if (lane == 0) wsum[wid][a] = make_double2(gx, gy);
      }
      __syncthreads();
      if ((int)threadIdx.x < n) {
        const int a = threadIdx.x;
        double2 r = wsum[0][a];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
          r.x += wsum[w][a].x;
          r.y += wsum[w][a].y;
        }
        partial[tile * num_atoms + c0 + a] = r;
      }
    }
  }
}

// Atoms [segments[0], segments[num_segments]) (clamped) are the atoms of the slices just reduced.
__global__ __launch_bounds__(256) void k_shift_grad_finish(long long num_atoms, const long long* __restrict__ segments,
                                                           int num_segments, int tiles,
                                                           const double2* __restrict__ partial,
                                                           double* __restrict__ grad_shifts) {
  long long a_begin = segments[0], a_end = segments[num_segments];
  a_begin = a_begin < 0 ? 0 : (a_begin > num_atoms ? num_atoms : a_begin);
  a_end = a_end < 0 ? 0 : (a_end > num_atoms ? num_atoms : a_end);
  for (long long a = a_begin + blockIdx.x * (long long)blockDim.x + threadIdx.x; a < a_end;
       a += (long long)gridDim.x * blockDim.x) {
    double gx = 0.0, gy = 0.0;
    for (int t = 0; t < tiles; ++t) {
      const double2 p = partial[(long long)t * num_atoms + a];
      gx += p.x;
      gy += p.y;
    }
    grad_shifts[2 * a] = gx;
    grad_shifts[2 * a + 1] = gy;
  }
}

long long grid_blocks(long long work) {
  long long b = (work + 255) / 256;
  if (b > 65536) b = 65536;
  return b < 1 ? 1 : b;
}

}  // namespace

hipError_t launch_projected_radii(int kind, const double* row, long long count, const double* r, double* out,
                                  hipStream_t st) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_projected_radii, dim3((unsigned)grid_blocks(count)), dim3(256), 0, st, kind, row, count, r,
                     out);
  return hipGetLastError();
}

hipError_t launch_atom_grid(int kind, int num_types, const double* rows, double pixel_size, int height, int width,
                            int supersampling, const double* centers, double* out, hipStream_t st) {
  const long long pix = (long long)height * width;
  if (num_types == 0 || pix == 0) return hipSuccess;
  long long bx = grid_blocks(pix);
  if (bx > 4096) bx = 4096;
  hipLaunchKernelGGL(k_atom_grid, dim3((unsigned)bx, (unsigned)num_types), dim3(256), 0, st, kind, rows,
                     pixel_size, height, width, supersampling, centers, out);
  return hipGetLastError();
}

size_t atom_center_grad_workspace_bytes(int num_types) {
  return (size_t)(num_types > 0 ? num_types : 0) * CG_BLOCKS * sizeof(double2);
}

hipError_t launch_atom_center_grad(int kind, int num_types, const double* rows, double pixel_size, int height,
                                   int width, int supersampling, const double* centers, const double* grad_out,
                                   double* grad_centers, void* ws, hipStream_t st) {
  const long long pix = (long long)height * width;
  if (num_types == 0) return hipSuccess;
  long long bx = grid_blocks(pix);
  if (bx > CG_BLOCKS) bx = CG_BLOCKS;
  double2* partial = (double2*)ws;
  hipLaunchKernelGGL(k_atom_center_grad, dim3((unsigned)bx, (unsigned)num_types), dim3(256), 0, st, kind, rows,
                     pixel_size, height, width, supersampling, centers, grad_out, partial);
  hipLaunchKernelGGL(k_atom_center_finish, dim3((unsigned)((num_types + 255) / 256)), dim3(256), 0, st, num_types,
                     (int)bx, 1.0 / ((double)supersampling * supersampling), partial, grad_centers);
  return hipGetLastError();
}

hipError_t launch_slice_structure(int num_slices, int num_types, int height, int width, long long num_atoms,
                                  const double* shifts, const long long* segments, const double2* kernel_k,
                                  double2* out_k, hipStream_t st) {
  if (num_slices == 0 || height == 0 || width == 0) return hipSuccess;
  const dim3 grid((unsigned)((width + SF_COLS - 1) / SF_COLS), (unsigned)((height + SF_ROWS - 1) / SF_ROWS),
                  (unsigned)num_slices);
  hipLaunchKernelGGL(k_slice_structure, grid, dim3(256), 0, st, num_types, height, width, num_atoms, shifts,
                     segments, kernel_k, out_k);
  return hipGetLastError();
}

size_t slice_shift_grad_workspace_bytes(int height, int width, long long num_atoms) {
  const long long tiles = (long long)((width + SF_COLS - 1) / SF_COLS) * ((height + SF_ROWS - 1) / SF_ROWS);
  return (size_t)(tiles * (num_atoms > 0 ? num_atoms : 0)) * sizeof(double2);
}

hipError_t launch_slice_shift_grad(int num_slices, int num_types, int height, int width, long long num_atoms,
                                   const double* shifts, const long long* segments, const double2* kernel_k,
                                   const double2* adj_k, double* grad_shifts, void* ws, hipStream_t st) {
  if (num_slices == 0 || num_atoms == 0) return hipSuccess;
  const dim3 grid((unsigned)((width + SF_COLS - 1) / SF_COLS), (unsigned)((height + SF_ROWS - 1) / SF_ROWS),
                  (unsigned)num_slices);
  double2* partial = (double2*)ws;
  hipLaunchKernelGGL(k_slice_shift_grad, grid, dim3(256), 0, st, num_types, height, width, num_atoms, shifts,
                     segments, kernel_k, adj_k, partial);
  const int tiles = (int)(grid.x * grid.y);
  hipLaunchKernelGGL(k_shift_grad_finish, dim3((unsigned)grid_blocks(num_atoms)), dim3(256), 0, st, num_atoms,
                     segments, num_slices * num_types, tiles, (const double2*)partial, grad_shifts);
  return hipGetLastError();
}

}  // namespace pty
