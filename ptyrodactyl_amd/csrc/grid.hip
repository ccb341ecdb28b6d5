// Multislice forward and adjoint on grids of any size h x w (1 to 4096 per
// axis): the shapes the power-of-two kernels do not take, such as odd,
// prime or non-square probe windows.
//
// One workgroup owns one scan position at a time (a grid-stride loop over
// positions; workgroups never communicate).  The wave lives in the
// workgroup's slice of the device workspace.  Each pass moves a group of rows
// or columns through LDS, runs a mixed-radix Stockham transform on them there
// and writes them back, with the pointwise products of the multislice (shift
// ramp, transmission, propagator, |A|^2, and the adjoint's conjugate
// products) applied in LDS between the transforms.  Per slice the wave makes
// one column pass (inverse y transform, times T_s, forward y transform) and
// one row pass (forward x transform, times the propagator, inverse x
// transform), the same factorisation as the power-of-two streaming kernels.
//
// Stage of radix r after the radices so far multiply to ns (Stockham
// autosort):
//   out[e] = sum_q in[j + q L / r] * w^(q (lo + k ns)),  w = exp(-+2 pi i / (ns r)),
//   with lo = e mod ns, k = (e / ns) mod r, j = (e / (ns r)) ns + lo.
// Radices 2 to 5 run one butterfly (all r outputs of one j) per thread with
// constant-coefficient r-point DFTs; larger prime radices one output per
// thread.  Twiddles come from an exact table exp(-2 pi i e / L) built in
// FP64, so no error accumulates across stages beyond the FP32 sums.
#include <math.h>

#include <stdlib.h>

#include <algorithm>

#include "common.hpp"
#include "grid.hpp"

namespace pty {
namespace {

constexpr int GT = 256;             // threads per workgroup
constexpr int GROUP_ELEMS = 2048;   // elements per LDS line group
constexpr int MAX_BLOCKS = 2048;    // workgroups per launch
constexpr size_t WS_BUDGET = size_t(2) << 30;
constexpr long long DEFAULT_STREAM_MIN_PIXELS = 1LL << 14;

struct Axis {
  int L;            // transform length
  int ls;           // line stride in LDS (odd for even L, to spread banks)
  int nst;          // stages
  const int* radix; // LDS
  const float2* tw; // LDS, exp(-2 pi i e / L)
};

__device__ __forceinline__ float2 gmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ float2 gmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 gscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// q = x / d, r = x mod d for 0 <= x < 2^24 and d >= 1: the float estimate x * (1/d) is within one
// of the quotient, and the remainder test corrects it.
__device__ __forceinline__ void divmod(int x, int d, float rd, int& q, int& r) {
  q = (int)((float)x * rd);
  r = x - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
}

// (major, minor) of o = major * mod + minor, stepped by the workgroup size without divisions.
struct Walk {
  int major, minor, dmaj, dmin, mod;
  __device__ Walk(int o, int m) : mod(m) {
    const float rm = 1.0f / (float)m;
    divmod(o, m, rm, major, minor);
    divmod(GT, m, rm, dmaj, dmin);
  }
  __device__ __forceinline__ void step() {
    major += dmaj;
    minor += dmin;
    if (minor >= mod) {
      minor -= mod;
      ++major;
    }
  }
};

// fftfreq bin of index kidx (Nyquist negative for even L, as numpy)
__device__ __forceinline__ int signed_bin(int kidx, int L) { return kidx < (L + 1) / 2 ? kidx : kidx - L; }

// Index of bin kidx in the fftshifted layout.
__device__ __forceinline__ int shifted_index(int kidx, int L) {
  const int i = kidx + L / 2;
  return i >= L ? i - L : i;
}

// exp(-2 pi i k s / L) for the shift s = m + d pixels, with k m reduced modulo L first so that
// large integer shifts lose no precision.
__device__ __forceinline__ float2 ramp1(int kidx, int L, int m, float d) {
  const int k = signed_bin(kidx, L);
  long long km = ((long long)k * m) % L;
  if (km < 0) km += L;
  const float t = (float)km / (float)L + (float)k * d / (float)L;
  float s, c;
  sincospif(-2.0f * t, &s, &c);
  return make_float2(c, s);
}

// cos and sin of 2 pi j k / R for j, k = 1 .. (R - 1) / 2, R odd, evaluated in double at compile time (integer
// reduction of j k to [0, R / 2], then the Taylor series of cx_sincos_octant, accurate to 1e-16 up to pi).
template <int R>
struct OddDftConsts {
  static constexpr int H = (R - 1) / 2;
  float c[H][H], s[H][H];  // [k - 1][j - 1]
  constexpr OddDftConsts() : c{}, s{} {
    for (int k = 1; k <= H; ++k)
      for (int j = 1; j <= H; ++j) {
        long long m = (long long)j * k % R;
        double sign = 1.0;
        if (2 * m > R) {
          m = R - m;
          sign = -1.0;
        }
        double sn = 0.0, cs = 0.0;
        cx_sincos_octant(m, R, sn, cs);
        c[k - 1][j - 1] = (float)cs;
        s[k - 1][j - 1] = (float)(sign * sn);
      }
  }
};

// R-point DFT for an odd R, by the symmetric pairs s_j = x_j + x_{R-j}, d_j = x_j - x_{R-j} (j = 1 .. H):
//   y_0 = x_0 + sum_j s_j,  y_k, y_{R-k} = t_k -+ i u_k  (forward; the inverse swaps the signs)
//   t_k = x_0 + sum_j cos(2 pi j k / R) s_j,  u_k = sum_j sin(2 pi j k / R) d_j.
// 2 H^2 complex multiply-adds instead of the (R - 1)^2 complex products of the direct sum.
template <int R, bool INV>
__device__ __forceinline__ void small_dft_odd(float2* v) {
  constexpr int H = (R - 1) / 2;
  constexpr OddDftConsts<R> K{};
  const float2 x0 = v[0];
  float2 sp[H], dm[H];
  float2 y0 = x0;
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    sp[j - 1] = make_float2(v[j].x + v[R - j].x, v[j].y + v[R - j].y);
    dm[j - 1] = make_float2(v[j].x - v[R - j].x, v[j].y - v[R - j].y);
    y0 = make_float2(y0.x + sp[j - 1].x, y0.y + sp[j - 1].y);
  }
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    float2 t = x0;
    float2 u = make_float2(K.s[k - 1][0] * dm[0].x, K.s[k - 1][0] * dm[0].y);
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      t = make_float2(fmaf(K.c[k - 1][j - 1], sp[j - 1].x, t.x), fmaf(K.c[k - 1][j - 1], sp[j - 1].y, t.y));
      if (j > 1) u = make_float2(fmaf(K.s[k - 1][j - 1], dm[j - 1].x, u.x), fmaf(K.s[k - 1][j - 1], dm[j - 1].y, u.y));
    }
    // -i u = (u.y, -u.x)
    const float2 iu = INV ? make_float2(-u.y, u.x) : make_float2(u.y, -u.x);
    v[k] = make_float2(t.x + iu.x, t.y + iu.y);
    v[R - k] = make_float2(t.x - iu.x, t.y - iu.y);
  }
  v[0] = y0;
}

// In-register DFT of R = 2, 3, 4, 5, 7, 11 or 13 points, y_k = sum_q x_q exp(-+2 pi i q k / R) (INV: +).
template <int R, bool INV>
__device__ __forceinline__ void small_dft(float2* v) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    // forward: y1 = d02 - i d13, y3 = d02 + i d13 (inverse swaps them)
    const float2 mi = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
    v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
    v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
    v[1] = make_float2(d02.x + mi.x, d02.y + mi.y);
    v[3] = make_float2(d02.x - mi.x, d02.y - mi.y);
  } else if constexpr (R == 3) {
    constexpr float S3 = 0.86602540378443865f;  // sin(2 pi / 3)
    const float2 s12 = make_float2(v[1].x + v[2].x, v[1].y + v[2].y);
    const float2 d12 = make_float2(v[1].x - v[2].x, v[1].y - v[2].y);
    const float2 t = make_float2(v[0].x - 0.5f * s12.x, v[0].y - 0.5f * s12.y);
    // forward: u = -i sin(2 pi/3) (x1 - x2)
    const float2 u = INV ? make_float2(-S3 * d12.y, S3 * d12.x) : make_float2(S3 * d12.y, -S3 * d12.x);
    v[0] = make_float2(v[0].x + s12.x, v[0].y + s12.y);
    v[1] = make_float2(t.x + u.x, t.y + u.y);
    v[2] = make_float2(t.x - u.x, t.y - u.y);
  } else if constexpr (R == 7 || R == 11 || R == 13) {
    small_dft_odd<R, INV>(v);
  } else {
    static_assert(R == 5, "radix");
    // symmetric form: with s1 = x1 + x4, d1 = x1 - x4, s2 = x2 + x3, d2 = x2 - x3,
    //   y0 = x0 + s1 + s2,  y1,4 = t1 -+ i u1,  y2,3 = t2 -+ i u2  (forward; the inverse swaps the signs)
    //   t1 = x0 + c1 s1 + c2 s2,  t2 = x0 + c2 s1 + c1 s2,  u1 = n1 d1 + n2 d2,  u2 = n2 d1 - n1 d2
    // with c1, c2 = cos(2 pi / 5), cos(4 pi / 5) and n1, n2 = sin(2 pi / 5), sin(4 pi / 5)
    constexpr float C1 = 0.30901699437494745f, C2 = -0.80901699437494745f;
    constexpr float N1 = 0.95105651629515357f, N2 = 0.58778525229247314f;
    const float2 x0 = v[0];
    const float2 s1 = make_float2(v[1].x + v[4].x, v[1].y + v[4].y);
    const float2 d1 = make_float2(v[1].x - v[4].x, v[1].y - v[4].y);
    const float2 s2 = make_float2(v[2].x + v[3].x, v[2].y + v[3].y);
    const float2 d2 = make_float2(v[2].x - v[3].x, v[2].y - v[3].y);
    const float2 t1 = make_float2(fmaf(C2, s2.x, fmaf(C1, s1.x, x0.x)), fmaf(C2, s2.y, fmaf(C1, s1.y, x0.y)));
    const float2 t2 = make_float2(fmaf(C1, s2.x, fmaf(C2, s1.x, x0.x)), fmaf(C1, s2.y, fmaf(C2, s1.y, x0.y)));
    const float2 u1 = make_float2(fmaf(N2, d2.x, N1 * d1.x), fmaf(N2, d2.y, N1 * d1.y));
    const float2 u2 = make_float2(fmaf(-N1, d2.x, N2 * d1.x), fmaf(-N1, d2.y, N2 * d1.y));
    // -i u = (u.y, -u.x)
    const float2 iu1 = INV ? make_float2(-u1.y, u1.x) : make_float2(u1.y, -u1.x);
    const float2 iu2 = INV ? make_float2(-u2.y, u2.x) : make_float2(u2.y, -u2.x);
    v[0] = make_float2(x0.x + s1.x + s2.x, x0.y + s1.y + s2.y);
    v[1] = make_float2(t1.x + iu1.x, t1.y + iu1.y);
    v[4] = make_float2(t1.x - iu1.x, t1.y - iu1.y);
    v[2] = make_float2(t2.x + iu2.x, t2.y + iu2.y);
    v[3] = make_float2(t2.x - iu2.x, t2.y - iu2.y);
  }
}

// One Stockham stage of radix R in {2, 3, 4, 5, 7, 11, 13}: a thread takes a butterfly (line, j), j < L / R,
// twiddles its R inputs by w^(q lo) (w = exp(-+2 pi i / (ns R)), lo = j mod ns), runs the R-point
// DFT and writes outputs (j / ns) ns R + lo + k ns.
template <int R, bool INV>
__device__ void butterfly_stage(const float2* a, float2* b, int nl, int L, int ls, int ns, const float2* tw) {
  const int nb = L / R;
  const int tstep = L / (ns * R);
  const int total = nl * nb;
  const float rns = 1.0f / (float)ns;
  Walk wk(threadIdx.x, nb);
  for (int t = threadIdx.x; t < total; t += GT, wk.step()) {
    const int line = wk.major;
    const int j = wk.minor;
    int jb, lo;
    divmod(j, ns, rns, jb, lo);
    const float2* src = a + line * ls + j;
    float2 v[R];
    v[0] = src[0];  // the twiddle of input 0 is 1
    const int tq = lo * tstep;
#pragma unroll
    for (int q = 1; q < R; ++q) {
      const float2 x = src[q * nb];
      const float2 w = tw[q * tq];
      const float wi = INV ? -w.y : w.y;
      v[q] = make_float2(x.x * w.x - x.y * wi, x.x * wi + x.y * w.x);
    }
    small_dft<R, INV>(v);
    float2* dst = b + line * ls + jb * ns * R + lo;
#pragma unroll
    for (int k = 0; k < R; ++k) dst[k * ns] = v[k];
  }
}

// Stage of any radix r, one output per thread:
//   out[e] = sum_q in[j + q L / r] w^(q (lo + k ns)), lo = e mod ns, k = (e / ns) mod r,
//   j = (e / (ns r)) ns + lo.
__device__ void generic_stage(const float2* a, float2* b, int nl, int L, int ls, int ns, int r, const float2* tw,
                              bool inv) {
  const int mr = ns * r;
  const int tstep = L / mr;
  const int lr = L / r;
  const int total = nl * L;
  const float rns = 1.0f / (float)ns, rr = 1.0f / (float)r;
  Walk wk(threadIdx.x, L);
  for (int o = threadIdx.x; o < total; o += GT, wk.step()) {
    const int line = wk.major;
    const int e = wk.minor;
    int hi, lo, jr, k;
    divmod(e, ns, rns, hi, lo);
    divmod(hi, r, rr, jr, k);
    const int j = jr * ns + lo;
    const int step = lo + k * ns;
    const float2* src = a + line * ls + j;
    float re = src[0].x, im = src[0].y;  // the q = 0 term has the factor 1
    int ex = step;
    for (int q = 1; q < r; ++q) {
      const float2 x = src[q * lr];
      const float2 t = tw[ex * tstep];
      const float ti = inv ? -t.y : t.y;
      re += x.x * t.x - x.y * ti;
      im += x.x * ti + x.y * t.x;
      ex += step;
      if (ex >= mr) ex -= mr;
    }
    b[line * ls + e] = make_float2(re, im);
  }
}

// Transforms `nl` lines held at stride ax.ls in `a` (ping-pong with `b`); returns the buffer that
// holds the result.  Ends with a barrier.
template <bool INV, bool WIDE>
__device__ float2* fft_lines_dir(float2* a, float2* b, int nl, const Axis& ax) {
  const int L = ax.L, ls = ax.ls;
  int ns = 1;
  for (int st = 0; st < ax.nst; ++st) {
    const int r = ax.radix[st];
    switch (r) {
      case 2: butterfly_stage<2, INV>(a, b, nl, L, ls, ns, ax.tw); break;
      case 3: butterfly_stage<3, INV>(a, b, nl, L, ls, ns, ax.tw); break;
      case 4: butterfly_stage<4, INV>(a, b, nl, L, ls, ns, ax.tw); break;
      case 5: butterfly_stage<5, INV>(a, b, nl, L, ls, ns, ax.tw); break;
      default:
        if constexpr (WIDE) {
          if (r == 7) {
            butterfly_stage<7, INV>(a, b, nl, L, ls, ns, ax.tw);
            break;
          }
          if (r == 11) {
            butterfly_stage<11, INV>(a, b, nl, L, ls, ns, ax.tw);
            break;
          }
          if (r == 13) {
            butterfly_stage<13, INV>(a, b, nl, L, ls, ns, ax.tw);
            break;
          }
        }
        generic_stage(a, b, nl, L, ls, ns, r, ax.tw, INV);
        break;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    ns *= r;
  }
  return a;
}

template <bool WIDE>
__device__ float2* fft_lines(float2* a, float2* b, int nl, const Axis& ax, bool inv) {
  return inv ? fft_lines_dir<true, WIDE>(a, b, nl, ax) : fft_lines_dir<false, WIDE>(a, b, nl, ax);
}

enum { NONE = 0, FWD = 1, INV = 2 };

// One pass over lines [l_begin, l_end) of one axis.  COLS: a line is a column of the h x w wave (the
// element index is the row), and global accesses put consecutive threads on consecutive columns;
// otherwise a line is a row.  load(line, e) -> value; transform fft1; mid(line, e, value) -> value
// (if has_mid); transform fft2; store(line, e, value).
template <bool COLS, bool WIDE, class Load, class Mid, class Store>
__device__ void line_pass(int l_begin, int l_end, int g_max, const Axis& ax, float2* bufa, float2* bufb, int fft1,
                          bool has_mid, int fft2, Load load, Mid mid, Store store) {
  const int L = ax.L, ls = ax.ls;
  for (int l0 = l_begin; l0 < l_end; l0 += g_max) {
    const int g = min(g_max, l_end - l0);
    const int total = g * L;
    {
      Walk wk(threadIdx.x, COLS ? g : L);
      for (int o = threadIdx.x; o < total; o += GT, wk.step()) {
        const int line = COLS ? wk.minor : wk.major;
        const int e = COLS ? wk.major : wk.minor;
        bufa[line * ls + e] = load(l0 + line, e);
      }
    }
    __syncthreads();
    float2* cur = bufa;
    float2* oth = bufb;
    if (fft1 != NONE) {
      cur = fft_lines<WIDE>(cur, oth, g, ax, fft1 == INV);
      oth = cur == bufa ? bufb : bufa;
    }
    if (has_mid) {
      Walk wk(threadIdx.x, COLS ? g : L);
      for (int o = threadIdx.x; o < total; o += GT, wk.step()) {
        const int line = COLS ? wk.minor : wk.major;
        const int e = COLS ? wk.major : wk.minor;
        cur[line * ls + e] = mid(l0 + line, e, cur[line * ls + e]);
      }
      __syncthreads();
    }
    if (fft2 != NONE) cur = fft_lines<WIDE>(cur, oth, g, ax, fft2 == INV);
    {
      Walk wk(threadIdx.x, COLS ? g : L);
      for (int o = threadIdx.x; o < total; o += GT, wk.step()) {
        const int line = COLS ? wk.minor : wk.major;
        const int e = COLS ? wk.major : wk.minor;
        store(l0 + line, e, cur[line * ls + e]);
      }
    }
    __syncthreads();
  }
}

struct Setup {
  Axis ay, ax;
  float2* bufa;
  float2* bufb;
};

// LDS layout: twiddles along y [h], along x [w], then two ping-pong buffers of gcap elements.
__device__ Setup setup(const GridParams& a, int* s_rad, float2* lds) {
  Setup u;
  float2* twy = lds;
  float2* twx = lds + a.h;
  u.bufa = twx + a.w;
  u.bufb = u.bufa + a.gcap;
  for (int e = threadIdx.x; e < a.h; e += GT) {
    double s, c;
    sincospi(-2.0 * (double)e / (double)a.h, &s, &c);
    twy[e] = make_float2((float)c, (float)s);
  }
  for (int e = threadIdx.x; e < a.w; e += GT) {
    double s, c;
    sincospi(-2.0 * (double)e / (double)a.w, &s, &c);
    twx[e] = make_float2((float)c, (float)s);
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < GRID_MAX_STAGES; ++i) {
      s_rad[i] = a.ry[i];
      s_rad[GRID_MAX_STAGES + i] = a.rx[i];
    }
  }
  __syncthreads();
  u.ay = Axis{a.h, a.lsy, a.ny, s_rad, twy};
  u.ax = Axis{a.w, a.lsx, a.nx, s_rad + GRID_MAX_STAGES, twx};
  return u;
}

struct Ident {
  __device__ float2 operator()(int, int, float2 v) const { return v; }
};

struct Position {
  int oy, ox;  // window origin reduced into the slab
  int my, mx;  // integer shift modulo h, w
  float dy, dx;
};

__device__ Position position(const GridParams& a, long long p) {
  Position q;
  q.oy = wrap_index(a.origins[2 * p], a.slab_h);
  q.ox = wrap_index(a.origins[2 * p + 1], a.slab_w);
  split_shift_px(a.shifts[2 * p], a.h, q.my, q.dy);
  split_shift_px(a.shifts[2 * p + 1], a.w, q.mx, q.dx);
  return q;
}

// Slab element of window pixel (y, x).
__device__ __forceinline__ long long slab_index(const GridParams& a, const Position& q, int y, int x) {
  long long ry = q.oy + y;
  if (ry >= a.slab_h) ry -= a.slab_h;
  long long rx = q.ox + x;
  if (rx >= a.slab_w) rx -= a.slab_w;
  return ry * a.slab_w + rx;
}

// Pieces of the forward shared by the per-position and the streaming kernels; each works on lines
// [l0, l1) of one wave wb (h x w complex64, (ky, x) layout between passes).

// Incident wave of mode m: rows of probe_k times the shift ramp (with the 1/(h w) of the inverse
// transform), inverse transform along x.
template <bool WIDE>
__device__ void incident_rows(const GridParams& a, const Axis& ax, float2* ba, float2* bb, int l0, int l1,
                              const Position& q, int m, float2* wb) {
  const int h = a.h, w = a.w;
  const float inv_hw = 1.0f / ((float)h * (float)w);
  const float2* pk = a.probe_k + (long long)m * h * w;
  line_pass<false, WIDE>(
      l0, l1, a.gx, ax, ba, bb, INV, false, NONE,
      [&](int ky, int kx) {
        const float2 r = gmul(ramp1(ky, h, q.my, q.dy), ramp1(kx, w, q.mx, q.dx));
        return gmul(pk[(long long)ky * w + kx], gscale(r, inv_hw));
      },
      Ident{}, [&](int ky, int x, float2 v) { wb[(long long)ky * w + x] = v; });
}

// Rows: forward along x, times Py Px (or its conjugate), inverse along x.
template <bool CONJ, bool WIDE>
__device__ void propagate_rows(const GridParams& a, const Axis& ax, float2* ba, float2* bb, int l0, int l1,
                               float2* wb) {
  const int w = a.w;
  line_pass<false, WIDE>(
      l0, l1, a.gx, ax, ba, bb, FWD, true, INV, [&](int ky, int x) { return wb[(long long)ky * w + x]; },
      [&](int ky, int kx, float2 v) {
        const float2 f = gmul(a.prop_y[ky], a.prop_x[kx]);
        return CONJ ? gmulc(f, v) : gmul(v, f);
      },
      [&](int ky, int x, float2 v) { wb[(long long)ky * w + x] = v; });
}

// Columns of slice s: to real space along y, checkpoint (if ck), times T_s, back along y.
template <bool WIDE>
__device__ void slice_cols(const GridParams& a, const Axis& ay, float2* ba, float2* bb, int l0, int l1,
                           const Position& q, int s, float2* wb, float2* ck) {
  const int w = a.w;
  const float2* ts = a.slab + (long long)s * a.slab_h * a.slab_w;
  line_pass<true, WIDE>(
      l0, l1, a.gy, ay, ba, bb, INV, true, FWD, [&](int x, int ky) { return wb[(long long)ky * w + x]; },
      [&](int x, int y, float2 v) {
        if (ck) ck[(long long)y * w + x] = v;
        return gmul(v, ts[slab_index(a, q, y, x)]);
      },
      [&](int x, int ky, float2 v) { wb[(long long)ky * w + x] = v; });
}

// Rows after the last slice: forward along x gives the detector amplitude A(ky, kx) of mode m at
// position p; writes it fftshifted and adds w_m |A|^2 to the pattern (stores it for m = 0).
template <bool WIDE>
__device__ void final_rows(const GridParams& a, const Axis& ax, float2* ba, float2* bb, int l0, int l1,
                           long long p, int m, float2* wb) {
  const int h = a.h, w = a.w;
  const long long hw = (long long)h * w;
  const float wm = a.weights[m];
  float* pat = a.patterns ? a.patterns + p * hw : nullptr;
  float2* amp = a.amplitude ? a.amplitude + (p * a.M + m) * hw : nullptr;
  line_pass<false, WIDE>(
      l0, l1, a.gx, ax, ba, bb, FWD, false, NONE, [&](int ky, int x) { return wb[(long long)ky * w + x]; }, Ident{},
      [&](int ky, int kx, float2 v) {
        const long long o = (long long)shifted_index(ky, h) * w + shifted_index(kx, w);
        if (amp) amp[o] = v;
        if (pat) {
          const float inten = wm * (v.x * v.x + v.y * v.y);
          pat[o] = m == 0 ? inten : pat[o] + inten;
        }
      });
}

__device__ __forceinline__ float2* checkpoint_of(const GridParams& a, long long p, int m, int s) {
  if (!a.ckpt || s == 0 || s % a.ckpt_every != 0) return nullptr;
  const int nc = num_checkpoints(a.S, a.ckpt_every);
  return a.ckpt + ((p * a.M + m) * nc + (s / a.ckpt_every - 1)) * (long long)a.h * a.w;
}

template <bool WIDE>
__global__ __launch_bounds__(GT) void k_grid_forward(GridParams a, float2* ws) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[2 * GRID_MAX_STAGES];
  const Setup u = setup(a, s_rad, lds);
  const int h = a.h, w = a.w;
  const long long hw = (long long)h * w;
  float2* wb = ws + (long long)blockIdx.x * hw;
  // With a pattern output a workgroup runs all modes of its position (one writer per pattern pixel);
  // amplitudes alone make every (position, mode) wave an independent task.
  const bool per_wave = a.patterns == nullptr;
  const long long tasks = per_wave ? a.P * a.M : a.P;
  for (long long t = blockIdx.x; t < tasks; t += gridDim.x) {
    const long long p = per_wave ? t / a.M : t;
    const int m0 = per_wave ? (int)(t % a.M) : 0;
    const int m1 = per_wave ? m0 + 1 : a.M;
    const Position q = position(a, p);
    for (int m = m0; m < m1; ++m) {
      incident_rows<WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, q, m, wb);
      for (int s = 0; s < a.S; ++s) {
        slice_cols<WIDE>(a, u.ay, u.bufa, u.bufb, 0, w, q, s, wb, checkpoint_of(a, p, m, s));
        if (s + 1 < a.S)
          propagate_rows<false, WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, wb);
        else
          final_rows<WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, p, m, wb);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Streaming forward for large grids: every pass is one launch over (wave, line group) workgroups, so
// that a few positions still fill the device.  The twiddle tables are built once in the workspace.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(GT) void k_grid_twiddles(int h, int w, float2* tw) {
  const int e = blockIdx.x * GT + threadIdx.x;
  if (e >= h + w) return;
  const int L = e < h ? h : w;
  const int i = e < h ? e : e - h;
  double sn, cs;
  sincospi(-2.0 * (double)i / (double)L, &sn, &cs);
  tw[e] = make_float2((float)cs, (float)sn);
}

// Axis of one streaming pass: its twiddle table copied into LDS, then the two line buffers.
template <bool Y>
__device__ Axis stream_axis(const GridParams& a, const float2* twg, float2* lds, int* s_rad, float2*& ba,
                            float2*& bb) {
  const int L = Y ? a.h : a.w;
  const int ls = Y ? a.lsy : a.lsx;
  const int g = Y ? a.gy : a.gx;
  for (int e = threadIdx.x; e < L; e += GT) lds[e] = twg[(Y ? 0 : a.h) + e];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < GRID_MAX_STAGES; ++i) s_rad[i] = Y ? a.ry[i] : a.rx[i];
  }
  __syncthreads();
  ba = lds + L;
  bb = ba + g * ls;
  return Axis{L, ls, Y ? a.ny : a.nx, s_rad, lds};
}

// Blocks b = group * nw + wave; wave wl lives at waves + wl * stride.
template <bool WIDE>
__global__ __launch_bounds__(GT) void k_gs_incident(GridParams a, const float2* twg, float2* waves, long long stride,
                                                     long long p0, int nw) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  float2 *ba, *bb;
  const Axis ax = stream_axis<false>(a, twg, lds, s_rad, ba, bb);
  const int wl = blockIdx.x % nw, l0 = (blockIdx.x / nw) * a.gx;
  const long long p = p0 + wl / a.M;
  incident_rows<WIDE>(a, ax, ba, bb, l0, min(a.h, l0 + a.gx), position(a, p), wl % a.M, waves + (long long)wl * stride);
}

template <bool WIDE>
__global__ __launch_bounds__(GT) void k_gs_cols(GridParams a, const float2* twg, float2* waves, long long p0, int nw,
                                                 int s) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  float2 *ba, *bb;
  const Axis ay = stream_axis<true>(a, twg, lds, s_rad, ba, bb);
  const int wl = blockIdx.x % nw, l0 = (blockIdx.x / nw) * a.gy;
  const long long p = p0 + wl / a.M;
  const int m = wl % a.M;
  slice_cols<WIDE>(a, ay, ba, bb, l0, min(a.w, l0 + a.gy), position(a, p), s, waves + (long long)wl * a.h * a.w,
             checkpoint_of(a, p, m, s));
}

template <bool CONJ, bool WIDE>
__global__ __launch_bounds__(GT) void k_gs_rows(GridParams a, const float2* twg, float2* waves, long long stride,
                                                 int nw) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  float2 *ba, *bb;
  const Axis ax = stream_axis<false>(a, twg, lds, s_rad, ba, bb);
  const int wl = blockIdx.x % nw, l0 = (blockIdx.x / nw) * a.gx;
  propagate_rows<CONJ, WIDE>(a, ax, ba, bb, l0, min(a.h, l0 + a.gx), waves + (long long)wl * stride);
}

// Blocks b = group * np + position; each block runs the modes of its position in order, so the pattern
// sum over modes has one writer per pixel.
template <bool WIDE>
__global__ __launch_bounds__(GT) void k_gs_final(GridParams a, const float2* twg, float2* waves, long long p0,
                                                  int np) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  float2 *ba, *bb;
  const Axis ax = stream_axis<false>(a, twg, lds, s_rad, ba, bb);
  const int pl = blockIdx.x % np, l0 = (blockIdx.x / np) * a.gx;
  for (int m = 0; m < a.M; ++m)
    final_rows<WIDE>(a, ax, ba, bb, l0, min(a.h, l0 + a.gx), p0 + pl, m,
               waves + ((long long)pl * a.M + m) * a.h * a.w);
}

// Pieces of the adjoint shared by the per-position and the streaming kernels.

// Recompute pass over columns of slice s: the wave entering slice s, from a checkpoint (real space) or
// from wb ((ky, x) layout, inverse along y), is stored in sg; if `advance`, times T_s and forward along y
// into wb.
template <bool WIDE>
__device__ void recompute_cols(const GridParams& a, const Axis& ay, float2* ba, float2* bb, int l0, int l1,
                               const Position& q, int s, const float2* src, bool from_ckpt, bool advance, float2* sg,
                               float2* wb) {
  const int w = a.w;
  const float2* ts = a.slab + (long long)s * a.slab_h * a.slab_w;
  line_pass<true, WIDE>(
      l0, l1, a.gy, ay, ba, bb, from_ckpt ? NONE : INV, true, advance ? FWD : NONE,
      [&](int x, int e) { return src[(long long)e * w + x]; },
      [&](int x, int y, float2 v) {
        sg[(long long)y * w + x] = v;
        return advance ? gmul(v, ts[slab_index(a, q, y, x)]) : v;
      },
      [&](int x, int ky, float2 v) {
        if (advance) wb[(long long)ky * w + x] = v;
      });
}

// Rows after the last slice: detector amplitude, seed dL/dA (from dL/dI or given), adjoint of the forward
// x transform into gb.
template <bool WIDE>
__device__ void seed_rows(const GridParams& a, const Axis& ax, float2* ba, float2* bb, int l0, int l1, long long p,
                          int m, const float2* wb, float2* gb) {
  const int h = a.h, w = a.w;
  const long long hw = (long long)h * w;
  const float wm = a.weights[m];
  const float* gpat = a.grad_patterns ? a.grad_patterns + p * hw : nullptr;
  const float2* gamp = a.grad_amplitude ? a.grad_amplitude + (p * a.M + m) * hw : nullptr;
  line_pass<false, WIDE>(
      l0, l1, a.gx, ax, ba, bb, FWD, true, INV, [&](int ky, int x) { return wb[(long long)ky * w + x]; },
      [&](int ky, int kx, float2 amp) {
        const long long o = (long long)shifted_index(ky, h) * w + shifted_index(kx, w);
        if (gamp) return gamp[o];
        return gscale(amp, 2.0f * wm * gpat[o]);
      },
      [&](int ky, int x, float2 v) { gb[(long long)ky * w + x] = v; });
}

// Back through slice s over columns: dL/dphi (inverse along y), the contribution conj(psi_s) dL/dphi to dL/dT_s
// (stored at window pixel (y, x) of `contrib` for a gather, or added to the slab gradient with float32 atomics
// when contrib is null), dL/dpsi_s = conj(T_s) dL/dphi, forward along y into gb.
template <bool WIDE>
__device__ void back_cols(const GridParams& a, const Axis& ay, float2* ba, float2* bb, int l0, int l1,
                          const Position& q, int s, const float2* sg, float2* gb, float2* contrib) {
  const int w = a.w;
  const long long sstride = a.slab_h * a.slab_w;
  const float2* ts = a.slab + (long long)s * sstride;
  float2* gs = a.grad_slab + (long long)s * sstride;
  line_pass<true, WIDE>(
      l0, l1, a.gy, ay, ba, bb, INV, true, FWD, [&](int x, int ky) { return gb[(long long)ky * w + x]; },
      [&](int x, int y, float2 g) {
        const long long si = slab_index(a, q, y, x);
        const float2 cg = gmulc(sg[(long long)y * w + x], g);
        if (contrib) {
          contrib[(long long)y * w + x] = cg;
        } else {
          unsafeAtomicAdd(&gs[si].x, cg.x);
          unsafeAtomicAdd(&gs[si].y, cg.y);
        }
        return gmulc(ts[si], g);
      },
      [&](int x, int ky, float2 v) { gb[(long long)ky * w + x] = v; });
}

// Rows of dL/dpsi_0, forward along x: dL/d(probe_k) += conj(R) dL/du (float32 atomics) and this thread's
// share of dL/dshift = Re sum conj(dL/du) du/dshift, u = probe_k R.
template <bool WIDE>
__device__ void probe_rows(const GridParams& a, const Axis& ax, float2* ba, float2* bb, int l0, int l1,
                           const Position& q, int m, const float2* gb, double& gsy, double& gsx) {
  const int h = a.h, w = a.w;
  const long long hw = (long long)h * w;
  const float inv_hw = 1.0f / ((float)h * (float)w);
  const double two_pi = 6.283185307179586;
  const float2* pk = a.probe_k + (long long)m * hw;
  float2* gp = a.grad_probe_k ? a.grad_probe_k + (long long)m * hw : nullptr;
  const bool want_shift = a.grad_shifts != nullptr;
  line_pass<false, WIDE>(
      l0, l1, a.gx, ax, ba, bb, FWD, true, NONE, [&](int ky, int x) { return gb[(long long)ky * w + x]; },
      [&](int ky, int kx, float2 gu) {
        const float2 r = gscale(gmul(ramp1(ky, h, q.my, q.dy), ramp1(kx, w, q.mx, q.dx)), inv_hw);
        const long long o = (long long)ky * w + kx;
        if (gp) {
          const float2 cg = gmulc(r, gu);
          unsafeAtomicAdd(&gp[o].x, cg.x);
          unsafeAtomicAdd(&gp[o].y, cg.y);
        }
        if (want_shift) {
          // du/ds_y = u (-2 pi i ky / h): Re(conj(gu) u (-2 pi i f)) = 2 pi f Im(conj(gu) u)
          const float2 z = gmulc(gu, gmul(pk[o], r));
          gsy += two_pi * (double)signed_bin(ky, h) / (double)h * (double)z.y;
          gsx += two_pi * (double)signed_bin(kx, w) / (double)w * (double)z.y;
        }
        return gu;
      },
      [&](int, int, float2) {});
}

__device__ __forceinline__ const float2* checkpoint_in(const GridParams& a, long long p, int m, int c) {
  const int nc = num_checkpoints(a.S, a.ckpt_every);
  return a.ckpt + ((p * a.M + m) * nc + (c - 1)) * (long long)a.h * a.w;
}

template <bool WIDE>
__global__ __launch_bounds__(GT) void k_grid_adjoint(GridParams a, float2* ws) {
  __shared__ int s_rad[2 * GRID_MAX_STAGES];
  __shared__ double s_red[GT / 64];
  extern __shared__ float2 lds[];
  const Setup u = setup(a, s_rad, lds);
  const int h = a.h, w = a.w;
  const long long hw = (long long)h * w;
  const int k = a.ckpt_every;
  const int nc = num_checkpoints(a.S, k);
  float2* wb = ws + (long long)blockIdx.x * (k + 2) * hw;  // forward wave
  float2* gb = wb + hw;                                      // gradient wave
  float2* seg = gb + hw;                                     // waves entering the slices of one segment
  for (long long p = blockIdx.x; p < a.P; p += gridDim.x) {
    const Position q = position(a, p);
    double gsy = 0.0, gsx = 0.0;
    for (int m = 0; m < a.M; ++m) {
      for (int c = nc; c >= 0; --c) {
        const int s0 = c * k;
        const int s1 = min(a.S, s0 + k);
        // recompute the waves entering slices s0 .. s1-1 (and the detector seed in the last segment)
        if (c == 0) incident_rows<WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, q, m, wb);
        for (int s = s0; s < s1; ++s) {
          const bool from_ckpt = c > 0 && s == s0;
          const bool advance = s + 1 < s1 || s == a.S - 1;
          recompute_cols<WIDE>(a, u.ay, u.bufa, u.bufb, 0, w, q, s, from_ckpt ? checkpoint_in(a, p, m, c) : wb,
                               from_ckpt, advance, seg + (long long)(s - s0) * hw, wb);
          if (s + 1 < s1)
            propagate_rows<false, WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, wb);
          else if (s == a.S - 1)
            seed_rows<WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, p, m, wb, gb);
        }
        // back through slices s1-1 .. s0
        for (int s = s1 - 1; s >= s0; --s) {
          back_cols<WIDE>(a, u.ay, u.bufa, u.bufb, 0, w, q, s, seg + (long long)(s - s0) * hw, gb, nullptr);
          if (s > 0)
            propagate_rows<true, WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, gb);
          else
            probe_rows<WIDE>(a, u.ax, u.bufa, u.bufb, 0, h, q, m, gb, gsy, gsx);
        }
      }
    }
    if (a.grad_shifts) {
      const double ty = block_sum_fp64<GT>(gsy, s_red);
      const double tx = block_sum_fp64<GT>(gsx, s_red);
      if (threadIdx.x == 0) {
        a.grad_shifts[2 * p] += ty;
        a.grad_shifts[2 * p + 1] += tx;
      }
      __syncthreads();
    }
  }
}

// Streaming adjoint: per wave the workspace holds (k + 2) h w elements (forward wave, gradient wave, k
// segment waves); every pass is one launch over (wave, line group) workgroups.  Blocks b = group * nw + wave.

template <bool WIDE>
__global__ __launch_bounds__(GT) void k_gsa_recompute(GridParams a, const float2* twg, float2* base, long long p0,
                                                       int nw, int s, int c) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  float2 *ba, *bb;
  const Axis ay = stream_axis<true>(a, twg, lds, s_rad, ba, bb);
  const long long hw = (long long)a.h * a.w;
  const int k = a.ckpt_every;
  const int wl = blockIdx.x % nw, l0 = (blockIdx.x / nw) * a.gy;
  const long long p = p0 + wl / a.M;
  const int m = wl % a.M;
  const int s0 = c * k, s1 = min(a.S, s0 + k);
  const bool from_ckpt = c > 0 && s == s0;
  const bool advance = s + 1 < s1 || s == a.S - 1;
  float2* wb = base + (long long)wl * (k + 2) * hw;
  recompute_cols<WIDE>(a, ay, ba, bb, l0, min(a.w, l0 + a.gy), position(a, p), s,
                 from_ckpt ? checkpoint_in(a, p, m, c) : wb, from_ckpt, advance, wb + (2 + s - s0) * hw, wb);
}

template <bool WIDE>
__global__ __launch_bounds__(GT) void k_gsa_seed(GridParams a, const float2* twg, float2* base, long long p0,
                                                  int nw) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  float2 *ba, *bb;
  const Axis ax = stream_axis<false>(a, twg, lds, s_rad, ba, bb);
  const long long hw = (long long)a.h * a.w;
  const int wl = blockIdx.x % nw, l0 = (blockIdx.x / nw) * a.gx;
  float2* wb = base + (long long)wl * (a.ckpt_every + 2) * hw;
  seed_rows<WIDE>(a, ax, ba, bb, l0, min(a.h, l0 + a.gx), p0 + wl / a.M, wl % a.M, wb, wb + hw);
}

template <bool WIDE>
__global__ __launch_bounds__(GT) void k_gsa_back(GridParams a, const float2* twg, float2* base, long long p0, int nw,
                                                  int s, int s0, float2* contrib) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  float2 *ba, *bb;
  const Axis ay = stream_axis<true>(a, twg, lds, s_rad, ba, bb);
  const long long hw = (long long)a.h * a.w;
  const int wl = blockIdx.x % nw, l0 = (blockIdx.x / nw) * a.gy;
  float2* wb = base + (long long)wl * (a.ckpt_every + 2) * hw;
  back_cols<WIDE>(a, ay, ba, bb, l0, min(a.w, l0 + a.gy), position(a, p0 + wl / a.M), s, wb + (2 + s - s0) * hw,
                  wb + hw, contrib + (long long)wl * hw);
}

// dL/dT_s of one batch: every slab pixel adds the contributions (window coordinates, contrib[wl][y][x]) of the
// waves whose windows cover it, in ascending wave order, so the sum has a fixed order and is bitwise
// reproducible.  One workgroup per 32 x TH tile of the slab: the waves meeting the tile are compacted in order
// with a ballot, then each thread sums its RPT pixels over them with HU windows' loads in flight (clamped to
// in-window addresses, always in bounds, and selected afterwards).
constexpr int GATHER_TILE = 32;

__device__ __forceinline__ int frame_offset(int i, int o, int m) {
  const int d = i - o;
  return d < 0 ? d + m : d;
}

template <int TH>
__global__ __launch_bounds__(GT) void k_gsa_gather(GridParams a, const float2* contrib, long long p0, int nw, int s) {
  constexpr int RPT = TH / 8;   // rows per thread: 8 rows per sweep of the 256 threads
  constexpr int HU = 16 / RPT;  // windows loaded together: sixteen loads in flight per thread
  __shared__ int3 hits[GT];     // (wave in the batch, oy, ox) of the windows meeting this tile
  __shared__ int wave_hits[GT / 64];
  const int H = (int)a.slab_h, W = (int)a.slab_w, h = a.h, w = a.w;
  const long long hw = (long long)h * w;
  const int Y0 = blockIdx.y * TH, X0 = blockIdx.x * GATHER_TILE;
  const int ylen = min(TH, H - Y0), xlen = min(GATHER_TILE, W - X0);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int tx = tid % GATHER_TILE, ty = tid / GATHER_TILE;
  const int X = X0 + tx;
  float2 acc[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) acc[r] = make_float2(0.f, 0.f);
  for (int base = 0; base < nw; base += GT) {
    const int i = base + tid;
    int oy = 0, ox = 0;
    bool hit = false;
    if (i < nw) {
      const long long p = p0 + i / a.M;
      oy = wrap_index(a.origins[2 * p], a.slab_h);
      ox = wrap_index(a.origins[2 * p + 1], a.slab_w);
      // the tile meets the periodic window [o, o + extent) if its first row / column lies inside it or the
      // tile runs past the window's start
      const int ey = frame_offset(Y0, oy, H), ex = frame_offset(X0, ox, W);
      hit = (ey < h || ey + ylen > H) && (ex < w || ex + xlen > W);
    }
    const unsigned long long mask = __ballot(hit);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();  // the previous chunk's list has been consumed
    if (lane == 0) wave_hits[wid] = __popcll(mask);
    __syncthreads();
    int offset = 0, total = 0;
#pragma unroll
    for (int v = 0; v < GT / 64; ++v) {
      offset += v < wid ? wave_hits[v] : 0;
      total += wave_hits[v];
    }
    if (hit) hits[offset + before] = make_int3(i, oy, ox);
    __syncthreads();
    if (tx < xlen) {
      for (int h0 = 0; h0 < total; h0 += HU) {
        float2 val[HU][RPT];
        bool use[HU][RPT];
#pragma unroll
        for (int u = 0; u < HU; ++u) {
          const int3 e = hits[min(h0 + u, total - 1)];
          const int dx = frame_offset(X, e.z, W);
          const float2* cw = contrib + (long long)e.x * hw + min(dx, w - 1);
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            const int yl = ty + 8 * r;
            const int dy = frame_offset(Y0 + yl, e.y, H);
            val[u][r] = cw[(long long)min(dy, h - 1) * w];
            use[u][r] = h0 + u < total && yl < ylen && dy < h && dx < w;
          }
        }
#pragma unroll
        for (int u = 0; u < HU; ++u)
#pragma unroll
          for (int r = 0; r < RPT; ++r)
            if (use[u][r]) acc[r] = make_float2(acc[r].x + val[u][r].x, acc[r].y + val[u][r].y);
      }
    }
  }
  if (tx >= xlen) return;
  float2* gs = a.grad_slab + (long long)s * H * W;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int yl = ty + 8 * r;
    if (yl < ylen) {
      const long long o = (long long)(Y0 + yl) * W + X;
      gs[o] = make_float2(gs[o].x + acc[r].x, gs[o].y + acc[r].y);
    }
  }
}

// Probe gradient rows; with shift gradients each block writes its FP64 partial sums to
// partials[2 b], partials[2 b + 1].
template <bool WIDE>
__global__ __launch_bounds__(GT) void k_gsa_probe(GridParams a, const float2* twg, float2* base, long long p0, int nw,
                                                   double* partials) {
  extern __shared__ float2 lds[];
  __shared__ int s_rad[GRID_MAX_STAGES];
  __shared__ double s_red[GT / 64];
  float2 *ba, *bb;
  const Axis ax = stream_axis<false>(a, twg, lds, s_rad, ba, bb);
  const long long hw = (long long)a.h * a.w;
  const int wl = blockIdx.x % nw, l0 = (blockIdx.x / nw) * a.gx;
  float2* wb = base + (long long)wl * (a.ckpt_every + 2) * hw;
  double gsy = 0.0, gsx = 0.0;
  probe_rows<WIDE>(a, ax, ba, bb, l0, min(a.h, l0 + a.gx), position(a, p0 + wl / a.M), wl % a.M, wb + hw, gsy, gsx);
  if (a.grad_shifts) {
    const double ty = block_sum_fp64<GT>(gsy, s_red);
    const double tx = block_sum_fp64<GT>(gsx, s_red);
    if (threadIdx.x == 0) {
      partials[2 * (long long)blockIdx.x] = ty;
      partials[2 * (long long)blockIdx.x + 1] = tx;
    }
  }
}

// dL/dshift of each position of the batch: its partials summed over row groups and modes in a fixed order.
__global__ __launch_bounds__(GT) void k_gsa_shift_reduce(GridParams a, const double* partials, long long p0, int np,
                                                          int row_groups) {
  const int pl = blockIdx.x * GT + threadIdx.x;
  if (pl >= np) return;
  const int nw = np * a.M;
  double sy = 0.0, sx = 0.0;
  for (int g = 0; g < row_groups; ++g)
    for (int m = 0; m < a.M; ++m) {
      const long long b = (long long)g * nw + (long long)pl * a.M + m;
      sy += partials[2 * b];
      sx += partials[2 * b + 1];
    }
  a.grad_shifts[2 * (p0 + pl)] += sy;
  a.grad_shifts[2 * (p0 + pl) + 1] += sx;
}

int factor_edge(int L, int* r) {
  int n = 0, v = L;
  while (v % 4 == 0) {
    r[n++] = 4;
    v /= 4;
  }
  for (int f = 2; v > 1; ++f)
    while (v % f == 0) {
      r[n++] = f;
      v /= f;
    }
  return n;
}

// Whether an edge has a radix-7, -11 or -13 stage.  The per-position kernels come in two builds: with the
// symmetric butterflies of those radices (which take more registers: three waves per SIMD instead of four) and
// without them (those stages then run as generic one-output-per-thread stages).  The streaming kernels always
// have them: their LDS use holds them at four waves per SIMD either way.
bool wide_radices(const GridParams& a) {
  for (int i = 0; i < GRID_MAX_STAGES; ++i)
    for (int r : {a.ry[i], a.rx[i]})
      if (r == 7 || r == 11 || r == 13) return true;
  return false;
}

int lds_stride(int L) { return L % 2 == 0 ? L + 1 : L; }

size_t prepare(GridParams& a) {
  for (int i = 0; i < GRID_MAX_STAGES; ++i) a.ry[i] = a.rx[i] = 1;
  a.ny = factor_edge(a.h, a.ry);
  a.nx = factor_edge(a.w, a.rx);
  a.lsy = lds_stride(a.h);
  a.lsx = lds_stride(a.w);
  a.gy = std::max(1, std::min(GROUP_ELEMS / a.h, a.w));  // columns per group (line length h)
  a.gx = std::max(1, std::min(GROUP_ELEMS / a.w, a.h));  // rows per group (line length w)
  a.gcap = std::max(a.gy * a.lsy, a.gx * a.lsx);
  return (size_t)(a.h + a.w + 2 * a.gcap) * sizeof(float2);
}

int clamp_ckpt(int k, int S) { return std::max(1, std::min(k, S)); }

hipError_t set_lds(const void* fn, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// Twiddle tables of the streaming forward at the start of its workspace, 256-byte aligned.
size_t twiddle_bytes(int h, int w) { return (((size_t)(h + w) * sizeof(float2)) + 255) & ~(size_t)255; }

// Grids of at least this many pixels run the streaming forward (PTY_GRID_STREAM_MIN_PIXELS overrides).
long long stream_min_pixels() {
  const char* v = getenv("PTY_GRID_STREAM_MIN_PIXELS");
  return v ? atoll(v) : DEFAULT_STREAM_MIN_PIXELS;
}

long long blocks_for(long long P, size_t per_block, size_t ws_bytes) {
  const long long fit = per_block ? (long long)(ws_bytes / per_block) : 0;
  return std::min(std::min(P, (long long)MAX_BLOCKS), fit);
}

// Streaming forward: one launch per pass over (wave, line group) workgroups, batches of bpos positions.
template <bool WIDE>
hipError_t stream_forward(const GridParams& a, float2* tw, float2* waves, long long bpos, hipStream_t st) {
  const long long hw = (long long)a.h * a.w;
  const size_t lds_rows = (size_t)(a.w + 2 * a.gx * a.lsx) * sizeof(float2);
  const size_t lds_cols = (size_t)(a.h + 2 * a.gy * a.lsy) * sizeof(float2);
  hipError_t e = set_lds((const void*)k_gs_incident<WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gs_rows<false, WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gs_final<WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gs_cols<WIDE>, lds_cols);
  if (e != hipSuccess) return e;
  const long long row_groups = (a.h + a.gx - 1) / a.gx;
  const long long col_groups = (a.w + a.gy - 1) / a.gy;
  hipLaunchKernelGGL(k_grid_twiddles, dim3((unsigned)((a.h + a.w + GT - 1) / GT)), dim3(GT), 0, st, a.h, a.w, tw);
  for (long long p0 = 0; p0 < a.P; p0 += bpos) {
    const int np = (int)std::min(bpos, a.P - p0);
    const int nw = np * a.M;
    hipLaunchKernelGGL(k_gs_incident<WIDE>, dim3((unsigned)(nw * row_groups)), dim3(GT), lds_rows, st, a, tw, waves,
                       hw, p0, nw);
    for (int s = 0; s < a.S; ++s) {
      hipLaunchKernelGGL(k_gs_cols<WIDE>, dim3((unsigned)(nw * col_groups)), dim3(GT), lds_cols, st, a, tw, waves, p0,
                         nw, s);
      if (s + 1 < a.S)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gs_rows<false, WIDE>), dim3((unsigned)(nw * row_groups)), dim3(GT),
                           lds_rows, st, a, tw, waves, hw, nw);
      else
        hipLaunchKernelGGL(k_gs_final<WIDE>, dim3((unsigned)(np * row_groups)), dim3(GT), lds_rows, st, a, tw, waves,
                           p0, np);
    }
  }
  return hipGetLastError();
}

// Streaming adjoint: per batch of bpos positions, the segment recompute and the reverse passes, one launch per
// pass over (wave, line group) workgroups.  Each reverse column pass stores its dL/dT contributions per wave in
// `contrib` (one wave of the batch each), and a gather adds them to the slab gradient in wave order.
template <bool WIDE>
hipError_t stream_adjoint(const GridParams& a, float2* tw, double* partials, float2* base, float2* contrib,
                          long long bpos, hipStream_t st) {
  const long long hw = (long long)a.h * a.w;
  const int k = a.ckpt_every;
  const long long row_groups = (a.h + a.gx - 1) / a.gx;
  const long long col_groups = (a.w + a.gy - 1) / a.gy;
  const long long stride = (long long)(k + 2) * hw;
  const size_t lds_rows = (size_t)(a.w + 2 * a.gx * a.lsx) * sizeof(float2);
  const size_t lds_cols = (size_t)(a.h + 2 * a.gy * a.lsy) * sizeof(float2);
  hipError_t e = set_lds((const void*)k_gs_incident<WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gs_rows<false, WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gs_rows<true, WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gsa_seed<WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gsa_probe<WIDE>, lds_rows);
  if (e == hipSuccess) e = set_lds((const void*)k_gsa_recompute<WIDE>, lds_cols);
  if (e == hipSuccess) e = set_lds((const void*)k_gsa_back<WIDE>, lds_cols);
  if (e != hipSuccess) return e;
  const int nc = num_checkpoints(a.S, k);
  // gather tiles: 32 x 32 when there are enough of them to fill the GPU, else 32 x 8
  const unsigned gtx = (unsigned)((a.slab_w + GATHER_TILE - 1) / GATHER_TILE);
  const bool tall = (long long)gtx * ((a.slab_h + 31) / 32) >= 1024;
  const dim3 gather_grid(gtx, (unsigned)(tall ? (a.slab_h + 31) / 32 : (a.slab_h + 7) / 8));
  hipLaunchKernelGGL(k_grid_twiddles, dim3((unsigned)((a.h + a.w + GT - 1) / GT)), dim3(GT), 0, st, a.h, a.w, tw);
  for (long long p0 = 0; p0 < a.P; p0 += bpos) {
    const int np = (int)std::min(bpos, a.P - p0);
    const int nw = np * a.M;
    const dim3 rows((unsigned)(nw * row_groups)), cols((unsigned)(nw * col_groups));
    for (int c = nc; c >= 0; --c) {
      const int s0 = c * k, s1 = std::min(a.S, s0 + k);
      if (c == 0) hipLaunchKernelGGL(k_gs_incident<WIDE>, rows, dim3(GT), lds_rows, st, a, tw, base, stride, p0, nw);
      for (int s = s0; s < s1; ++s) {
        hipLaunchKernelGGL(k_gsa_recompute<WIDE>, cols, dim3(GT), lds_cols, st, a, tw, base, p0, nw, s, c);
        if (s + 1 < s1)
          hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gs_rows<false, WIDE>), rows, dim3(GT), lds_rows, st, a, tw, base,
                             stride, nw);
        else if (s == a.S - 1)
          hipLaunchKernelGGL(k_gsa_seed<WIDE>, rows, dim3(GT), lds_rows, st, a, tw, base, p0, nw);
      }
      for (int s = s1 - 1; s >= s0; --s) {
        hipLaunchKernelGGL(k_gsa_back<WIDE>, cols, dim3(GT), lds_cols, st, a, tw, base, p0, nw, s, s0, contrib);
        if (tall)
          hipLaunchKernelGGL(k_gsa_gather<32>, gather_grid, dim3(GT), 0, st, a, (const float2*)contrib, p0, nw, s);
        else
          hipLaunchKernelGGL(k_gsa_gather<8>, gather_grid, dim3(GT), 0, st, a, (const float2*)contrib, p0, nw, s);
        if (s > 0)
          hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gs_rows<true, WIDE>), rows, dim3(GT), lds_rows, st, a, tw, base + hw,
                             stride, nw);
        else
          hipLaunchKernelGGL(k_gsa_probe<WIDE>, rows, dim3(GT), lds_rows, st, a, tw, base, p0, nw, partials);
      }
    }
    if (a.grad_shifts)
      hipLaunchKernelGGL(k_gsa_shift_reduce, dim3((unsigned)((np + GT - 1) / GT)), dim3(GT), 0, st, a,
                         (const double*)partials, p0, np, (int)row_groups);
  }
  return hipGetLastError();
}

}  // namespace

size_t grid_block_bytes(int h, int w, int ckpt_every, bool adjoint) {
  const size_t wave = (size_t)h * (size_t)w * sizeof(float2);
  return adjoint ? wave * (size_t)(std::max(1, ckpt_every) + 2) : wave;
}

size_t grid_workspace_bytes(int h, int w, int ckpt_every, long long num_positions, bool adjoint) {
  // the streaming adjoint keeps one more wave per position: its dL/dT contributions for the gather
  const size_t per = grid_block_bytes(h, w, ckpt_every, adjoint) + (adjoint ? (size_t)h * w * sizeof(float2) : 0);
  long long b = std::min(num_positions, (long long)MAX_BLOCKS);
  b = std::min(b, std::max(1LL, (long long)(WS_BUDGET / per)));
  b = std::max(1LL, b);
  if (!adjoint) return per * (size_t)b + twiddle_bytes(h, w);
  // the streaming adjoint also keeps the twiddle tables and (row groups) x 2 FP64 shift partials per position
  const int gx = std::max(1, std::min(GROUP_ELEMS / w, h));
  const size_t row_groups = (size_t)((h + gx - 1) / gx);
  return (per + row_groups * 16) * (size_t)b + twiddle_bytes(h, w) + 256;
}

hipError_t launch_grid_forward(GridParams a, void* ws, size_t ws_bytes, hipStream_t st) {
  a.ckpt_every = a.ckpt ? clamp_ckpt(a.ckpt_every, a.S) : 1;
  const size_t lds = prepare(a);
  const long long hw = (long long)a.h * a.w;
  const size_t twb = twiddle_bytes(a.h, a.w);
  const size_t per_pos = (size_t)a.M * (size_t)hw * sizeof(float2);
  if (hw >= stream_min_pixels() && ws_bytes >= twb + per_pos) {
    // streaming: one launch per pass over (wave, line group) workgroups
    float2* tw = (float2*)ws;
    float2* waves = (float2*)((char*)ws + twb);
    const long long bpos = std::min(a.P, (long long)((ws_bytes - twb) / per_pos));
    return wide_radices(a) ? stream_forward<true>(a, tw, waves, bpos, st)
                           : stream_forward<false>(a, tw, waves, bpos, st);
  }
  const long long tasks = a.patterns ? a.P : a.P * a.M;
  const long long nb = blocks_for(tasks, grid_block_bytes(a.h, a.w, 1, false), ws_bytes);
  if (nb < 1) return hipErrorInvalidValue;
  const void* kern = wide_radices(a) ? (const void*)k_grid_forward<true> : (const void*)k_grid_forward<false>;
  hipError_t e = set_lds(kern, lds);
  if (e != hipSuccess) return e;
  if (wide_radices(a))
    hipLaunchKernelGGL(k_grid_forward<true>, dim3((unsigned)nb), dim3(GT), lds, st, a, (float2*)ws);
  else
    hipLaunchKernelGGL(k_grid_forward<false>, dim3((unsigned)nb), dim3(GT), lds, st, a, (float2*)ws);
  return hipGetLastError();
}

hipError_t launch_grid_adjoint(GridParams a, void* ws, size_t ws_bytes, hipStream_t st) {
  a.ckpt_every = clamp_ckpt(a.ckpt_every, a.S);
  const size_t lds = prepare(a);
  const int k = a.ckpt_every;
  const long long hw = (long long)a.h * a.w;
  const long long row_groups = (a.h + a.gx - 1) / a.gx;
  const size_t twb = twiddle_bytes(a.h, a.w);
  // per position: M waves of (k + 3) h w elements ((k + 2) waves and the dL/dT contributions) and M x row_groups
  // pairs of shift partial sums
  const size_t per_pos = (size_t)a.M * ((size_t)(k + 3) * (size_t)hw * sizeof(float2) + (size_t)row_groups * 16);
  if (hw >= stream_min_pixels() && ws_bytes >= twb + 256 + per_pos) {
    const long long bpos = std::min(a.P, (long long)((ws_bytes - twb - 256) / per_pos));
    float2* tw = (float2*)ws;
    double* partials = (double*)((char*)ws + twb);
    const size_t waves_off = (twb + (size_t)bpos * a.M * row_groups * 16 + 255) & ~(size_t)255;
    float2* base = (float2*)((char*)ws + waves_off);
    float2* contrib = base + bpos * a.M * (k + 2) * hw;
    return wide_radices(a) ? stream_adjoint<true>(a, tw, partials, base, contrib, bpos, st)
                           : stream_adjoint<false>(a, tw, partials, base, contrib, bpos, st);
  }
  const long long nb = blocks_for(a.P, grid_block_bytes(a.h, a.w, k, true), ws_bytes);
  if (nb < 1) return hipErrorInvalidValue;
  const void* kern = wide_radices(a) ? (const void*)k_grid_adjoint<true> : (const void*)k_grid_adjoint<false>;
  hipError_t e = set_lds(kern, lds);
  if (e != hipSuccess) return e;
  if (wide_radices(a))
    hipLaunchKernelGGL(k_grid_adjoint<true>, dim3((unsigned)nb), dim3(GT), lds, st, a, (float2*)ws);
  else
    hipLaunchKernelGGL(k_grid_adjoint<false>, dim3((unsigned)nb), dim3(GT), lds, st, a, (float2*)ws);
  return hipGetLastError();
}

}  // namespace pty
