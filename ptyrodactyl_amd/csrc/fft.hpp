// Register/LDS FFT building blocks for gfx950 (64-wide wavefronts).
//
// Data distribution used throughout: a group of T threads transforms one
// line of L = T*E complex64 points; thread t holds v[q] = x[t + T*q] for
// q in [0, E).  The forward and inverse transforms leave the output in the
// same distribution (v[q] = X[t + T*q]), so a pointwise multiply between a
// forward and an inverse transform needs no data exchange.  Internal passes
// are Stockham radix-R passes that exchange through an LDS scratch owned by
// the group; every stage is written for this build (no vendor FFT library).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define PTY_HD __host__ __device__ __forceinline__

namespace pty {

PTY_HD float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
PTY_HD float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
PTY_HD float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
PTY_HD float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
PTY_HD float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Modulus correction of a complex64 factor whose exact value has modulus r (r2 = r * r): the real d
// with |z (1 + d)| = r to about 1e-15.  The tables below keep z * d as a second float2 (a double-float
// pair with z) and multiply with cmul_df, so the correction survives the float32 rounding of the
// product (scaling an already rounded product by 1 + d would round back to the same value).  Each float32 component carries a rounding error of up to half
// an ulp, so |z| misses r by up to ~6e-8; the factors of a unitary operator (the FFT twiddles, the
// Fresnel propagator) are applied the same way to every position and slice, so that modulus error
// is systematic and grows with the number of transforms, while a phase error of the same size is
// harmless.  A factor whose modulus differs from r by more than rounding explains is not a rounded
// unit-modulus value and is used as given (d = 0).
__device__ __forceinline__ float modulus_fix(float2 z, double r2) {
  const double m2 = (double)z.x * (double)z.x + (double)z.y * (double)z.y;
  if (!(m2 > 0.0)) return 0.f;
  const double d = 0.5 * (r2 - m2) / m2;
  return fabs(d) < 2.5e-7 ? (float)d : 0.f;
}
PTY_HD float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
PTY_HD float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }
// a * (-i) and a * (+i)
PTY_HD float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }
PTY_HD float2 mul_pi(float2 a) { return make_float2(-a.y, a.x); }

// Complex multiply by a factor given as double-float parts (hi + lo); the
// small terms are accumulated first so the product carries the factor's
// full precision (the modulus-corrected tables below).
PTY_HD float2 cmul_df(float2 z, float crh, float crl, float cih, float cil) {
  float re = fmaf(z.x, crl, -z.y * cil);
  re = fmaf(-z.y, cih, re);
  re = fmaf(z.x, crh, re);
  float im = fmaf(z.x, cil, z.y * crl);
  im = fmaf(z.y, crh, im);
  im = fmaf(z.x, cih, im);
  return make_float2(re, im);
}

// Lines up to this length (the LDS-resident kernels) carry the modulus-correcting low parts of their
// twiddle and propagator tables (two more float2 tables in LDS, two more FMAs per multiply).  The
// streaming kernels (n >= 256) keep plain float32 tables: there the extra LDS lowers occupancy from
// four to three workgroups per CU and the extra work costs 15 % of the 256^2 x 64-slice forward,
// while their patterns already stay within 3.2e-6 of the complex128 oracle.
constexpr int DF_MAX_LINE = 128;

// z * exp(i theta) as three shears (lifting steps; A. W. Paeth, "A fast algorithm for general raster
// rotation", Graphics Interface 1986; the lifting form of FFT twiddles in S. Oraintara, Y.-J. Chen and
// T. Q. Nguyen, "Integer fast Fourier transform", IEEE Trans. Signal Process. 50(3), 2002):
//   x1 = x - t y,  y' = y + s x1,  x' = x1 - t y',  t = tan(theta / 2), s = sin(theta).
// Each step is one FMA, and the map has determinant exactly 1 for any rounded t and s: rounding the constants
// to float32 leaves no systematic error in the modulus, only a phase error of the constants' size.  A complex
// multiply by rounded (cos, sin) scales every point it touches by the same |W| != 1, an error that adds up
// coherently over the ~100 chained transforms of a thick specimen.
PTY_HD float2 rot_shear(float2 z, float t, float s) {
  const float x1 = fmaf(-t, z.y, z.x);
  const float y = fmaf(s, x1, z.y);
  return make_float2(fmaf(-t, y, x1), y);
}

// Multiply by W16^E = exp(-2 pi i E / 16) (forward) or its conjugate: W16^E = (-i)^k W16^r with E = 4 k + r,
// r in {-1, 0, 1, 2}; the rotation by -r pi / 8 runs as three shears (tan(pi / 16), sin(pi / 8) or
// tan(pi / 8), sin(pi / 4)) and the quarter turn is a swap with a sign change.  Three FMAs against the
// eight of tw16_df; the phase error of the float32 constants (about 3e-8) stays.
template <int E, bool INV>
PTY_HD float2 tw16_shear(float2 z) {
  constexpr int e = ((E % 16) + 16) % 16;
  constexpr int r = (e % 4 == 3) ? -1 : e % 4;
  constexpr int k = ((e - r) / 4) % 4;
  constexpr float T1 = 0.19891236737965800691f, S1 = 0.38268343236508977173f;  // tan(pi/16), sin(pi/8)
  constexpr float T2 = 0.41421356237309504880f, S2 = 0.70710678118654752440f;  // tan(pi/8), sin(pi/4)
  // rotation angle of W16^r: -r pi / 8 (forward), +r pi / 8 (inverse)
  constexpr float sg = (INV ? 1.f : -1.f) * (r < 0 ? -1.f : 1.f);
  float2 w = z;
  if constexpr (r == 1 || r == -1) w = rot_shear(z, sg * T1, sg * S1);
  else if constexpr (r == 2) w = rot_shear(z, sg * T2, sg * S2);
  // (-i)^k forward, (+i)^k inverse
  constexpr int kk = INV ? (4 - k) % 4 : k;
  if constexpr (kk == 0) return w;
  else if constexpr (kk == 1) return mul_mi(w);
  else if constexpr (kk == 2) return make_float2(-w.x, -w.y);
  else return mul_pi(w);
}

// Multiply by W16^E = exp(-2 pi i E / 16) (forward) or its conjugate with double-float constants: exact to
// float32 rounding of the product, in modulus and phase (the LDS-resident kernels, whose gradients carry the
// tightest parity bound).  With float32 constants alone the 256^2 x 64-slice patterns move from 3.2e-6 to
// 1.3e-5 of the complex128 oracle (measured), since their modulus error, about 5e-8, is systematic.
template <int E, bool INV>
PTY_HD float2 tw16_df(float2 z) {
  constexpr int e = ((E % 16) + 16) % 16;
  // cos(pi/8), sin(pi/8), 1/sqrt(2) split into float32 hi + lo
  constexpr float C1H = 0.9238795042037964f, C1L = 2.830748968563057e-08f;
  constexpr float S1H = 0.3826834261417389f, S1L = 6.2233507236442165e-09f;
  constexpr float R2H = 0.7071067690849304f, R2L = 1.2101617485882343e-08f;
  if constexpr (e == 0) {
    return z;
  } else if constexpr (e == 4) {
    return INV ? mul_pi(z) : mul_mi(z);
  } else if constexpr (e == 8) {
    return make_float2(-z.x, -z.y);
  } else if constexpr (e == 12) {
    return INV ? mul_mi(z) : mul_pi(z);
  } else if constexpr (e % 4 == 2) {
    // W = R2 (sr + i si): z W = R2 (sr zr - si zi, sr zi + si zr)
    constexpr float sr = (e == 2 || e == 14) ? 1.f : -1.f;
    constexpr float si0 = (e == 2 || e == 6) ? -1.f : 1.f;
    constexpr float si = INV ? -si0 : si0;
    const float a = sr * z.x - si * z.y;
    const float b = sr * z.y + si * z.x;
    return make_float2(fmaf(a, R2H, a * R2L), fmaf(b, R2H, b * R2L));
  } else {
    // odd e: (cos, -sin)(2 pi e / 16) from the pi/8 pair with signs
    constexpr int q = e / 4;           // quadrant
    constexpr bool near = (e % 4) == 1;  // angle q*pi/2 + pi/8 (else + 3pi/8)
    // base (c, s) for the angle within the quadrant: exp(-i theta)
    constexpr float bch = near ? C1H : S1H, bcl = near ? C1L : S1L;
    constexpr float bsh = near ? S1H : C1H, bsl = near ? S1L : C1L;
    // rotate by quadrant: exp(-i (q pi/2 + theta)) = (-i)^q (c - i s)
    float crh, crl, cih, cil;
    if constexpr (q == 0) { crh = bch; crl = bcl; cih = -bsh; cil = -bsl; }
    else if constexpr (q == 1) { crh = -bsh; crl = -bsl; cih = -bch; cil = -bcl; }
    else if constexpr (q == 2) { crh = -bch; crl = -bcl; cih = bsh; cil = bsl; }
    else { crh = bsh; crl = bsl; cih = bch; cil = bcl; }
    if constexpr (INV) { cih = -cih; cil = -cil; }
    return cmul_df(z, crh, crl, cih, cil);
  }
}

// W16^E by the double-float constants (EXACT) or by three shears.
template <int E, bool INV, bool EXACT>
PTY_HD float2 tw16(float2 z) {
  if constexpr (EXACT) return tw16_df<E, INV>(z);
  else return tw16_shear<E, INV>(z);
}

template <bool INV>
PTY_HD void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
PTY_HD void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
  t3 = INV ? mul_pi(t3) : mul_mi(t3);
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// Natural-order in-register DFT of size R (R in {2, 4, 8, 16}); EXACT selects the double-float W16 constants.
template <int R, bool INV, bool EXACT>
struct Dft;

template <bool INV, bool EXACT>
struct Dft<2, INV, EXACT> {
  PTY_HD static void run(float2* a) { dft2<INV>(a[0], a[1]); }
};

template <bool INV, bool EXACT>
struct Dft<4, INV, EXACT> {
  PTY_HD static void run(float2* a) { dft4<INV>(a[0], a[1], a[2], a[3]); }
};

// 8 = 2 x 4 (decimation in time): X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k].
template <bool INV, bool EXACT>
struct Dft<8, INV, EXACT> {
  PTY_HD static void run(float2* a) {
    float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    o1 = tw16<2, INV, EXACT>(o1);
    o2 = tw16<4, INV, EXACT>(o2);
    o3 = tw16<6, INV, EXACT>(o3);
    a[0] = cadd(e0, o0);
    a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1);
    a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2);
    a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3);
    a[7] = csub(e3, o3);
  }
};

// 16 = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2.
template <bool INV, bool EXACT>
struct Dft<16, INV, EXACT> {
  PTY_HD static void run(float2* a) {
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4<INV>(a[n2], a[4 + n2], a[8 + n2], a[12 + n2]);
    // a[4*k1 + n2] now holds the inner DFT; apply W16^(n2*k1)
    a[5] = tw16<1, INV, EXACT>(a[5]);
    a[6] = tw16<2, INV, EXACT>(a[6]);
    a[7] = tw16<3, INV, EXACT>(a[7]);
    a[9] = tw16<2, INV, EXACT>(a[9]);
    a[10] = tw16<4, INV, EXACT>(a[10]);
    a[11] = tw16<6, INV, EXACT>(a[11]);
    a[13] = tw16<3, INV, EXACT>(a[13]);
    a[14] = tw16<6, INV, EXACT>(a[14]);
    a[15] = tw16<9, INV, EXACT>(a[15]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<INV>(a[4 * k1], a[4 * k1 + 1], a[4 * k1 + 2], a[4 * k1 + 3]);
    // a[4*k1 + k2] = X[k1 + 4*k2]: transpose the 4x4 index grid into natural order
    float2 t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = a[i];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = t[4 * k1 + k2];
  }
};

// Radix schedule for a line of length L with E = 16 points per thread.
template <int L>
struct Plan;
template <>
struct Plan<32> {
  static constexpr int NP = 2;
  static constexpr int R[3] = {16, 2, 1};
};
template <>
struct Plan<64> {
  static constexpr int NP = 2;
  static constexpr int R[3] = {16, 4, 1};
};
template <>
struct Plan<128> {
  static constexpr int NP = 2;
  static constexpr int R[3] = {16, 8, 1};
};
template <>
struct Plan<256> {
  static constexpr int NP = 2;
  static constexpr int R[3] = {16, 16, 1};
};
template <>
struct Plan<512> {
  static constexpr int NP = 3;
  static constexpr int R[3] = {16, 16, 2};
};
template <>
struct Plan<1024> {
  static constexpr int NP = 3;
  static constexpr int R[3] = {16, 16, 4};
};
template <>
struct Plan<2048> {
  static constexpr int NP = 3;
  static constexpr int R[3] = {16, 16, 8};
};
template <>
struct Plan<4096> {
  static constexpr int NP = 3;
  static constexpr int R[3] = {16, 16, 16};
};

// Twiddle table of a line of length L: every pass after the first, with NS points before it and radix R,
// multiplies point (j, r) (j < NS, r < R) by W_L^(j r L / (NS R)), W_L^m = exp(-2 pi i m / L).  The table
// holds these factors pass by pass, at twiddle_slot<L, NS>() + r NS + j: the pass after the first 16-point
// one (NS = 16) from 0, the third pass of lines of 512 and more (NS = 256) from 256.  A thread's factors of
// one pass then sit at constant offsets from one address.
template <int L>
constexpr int twiddle_entries() {
  return L <= 256 ? L : 256 + L;
}
template <int L, int NS>
constexpr int twiddle_slot() {
  return NS == 16 ? 0 : 256;
}

// One Stockham pass.  NS = product of the radices of earlier passes.
// Buf provides ldp(pe), stp(pe, v) and sync() for the group's scratch, where pe = pad16(e) is the padded
// index of element e (one pad slot after every 16 elements); TW::apply<INV, SLOT, NS>(x, j, r) returns
// x * the table factor of point (j, r) (x * its conjugate when INV).  Every element a thread touches in
// one pass sits at a constant padded distance from its first one (the strides below are multiples of 16,
// or the run stays inside one 16-element block), so each access is one address plus an immediate offset.
template <int L, int E, int R, int NS, bool INV, bool FIRST, bool LAST, class Buf, class TW>
__device__ __forceinline__ void stockham_pass(float2 (&v)[E], int t, Buf& buf, const TW& tw) {
  constexpr int T = L / E;
  constexpr int NB = E / R;
  static_assert(FIRST || (L / R) % 16 == 0, "load stride must be a multiple of 16");
  static_assert(LAST || NS % 16 == 0 || (NS == 1 && R == 16), "store run must stay padded-contiguous");
  constexpr int LD_STEP = (L / R) + (L / R) / 16;           // padded distance of e and e + L / R
  constexpr int ST_STEP = NS % 16 == 0 ? NS + NS / 16 : NS;  // padded distance of idx and idx + NS
  float2 x[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + T * b;
    const int pj = j + (j >> 4);
#pragma unroll
    for (int r = 0; r < R; ++r) x[b][r] = FIRST ? v[b + NB * r] : buf.ldp(pj + LD_STEP * r);
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + T * b;
    if constexpr (!FIRST) {
      // opaque twiddle index: the forward and inverse transforms of a line read the same entries, and
      // reusing them across the pointwise multiply between the two would keep 45 values live
      int jm = j % NS;
      asm volatile("" : "+v"(jm));
#pragma unroll
      for (int r = 1; r < R; ++r) x[b][r] = tw.template apply<INV, twiddle_slot<L, NS>(), NS>(x[b][r], jm, r);
    }
    Dft<R, INV, TW::EXACT>::run(x[b]);
  }
  if constexpr (LAST) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + NB * r] = x[b][r];
  } else {
    buf.sync();  // every read of the previous contents is done
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + T * b;
      const int idx = (j / NS) * NS * R + (j % NS);
      const int pidx = idx + (idx >> 4);
#pragma unroll
      for (int r = 0; r < R; ++r) buf.stp(pidx + ST_STEP * r, x[b][r]);
    }
    buf.sync();  // writes visible to the group
  }
}

// Full line FFT (forward when INV == false, unnormalised inverse otherwise).
template <int L, bool INV, class Buf, class TW>
__device__ __forceinline__ void line_fft(float2 (&v)[16], int t, Buf& buf, const TW& tw) {
  using P = Plan<L>;
  if constexpr (P::NP == 2) {
    stockham_pass<L, 16, P::R[0], 1, INV, true, false>(v, t, buf, tw);
    stockham_pass<L, 16, P::R[1], P::R[0], INV, false, true>(v, t, buf, tw);
  } else {
    stockham_pass<L, 16, P::R[0], 1, INV, true, false>(v, t, buf, tw);
    stockham_pass<L, 16, P::R[1], P::R[0], INV, false, false>(v, t, buf, tw);
    stockham_pass<L, 16, P::R[2], P::R[0] * P::R[1], INV, false, true>(v, t, buf, tw);
  }
}

// z * (hi + lo) and z * conj(hi + lo) for a double-float complex factor.
PTY_HD float2 cmul_pair(float2 z, float2 hi, float2 lo) { return cmul_df(z, hi.x, lo.x, hi.y, lo.y); }
PTY_HD float2 cmulc_pair(float2 z, float2 hi, float2 lo) { return cmul_df(z, hi.x, lo.x, -hi.y, -lo.y); }

// Twiddle multiply from the LDS tables (float32 factors laid out as twiddle_slot describes) and, when
// DF, their modulus-correcting low parts.
template <bool DF>
struct TwTab {
  // the double-float tables come with the double-float W16 constants, plain float32 tables with the shears
  static constexpr bool EXACT = DF;
  const float2* tab;
  const float2* lo;
  template <bool INV, int SLOT, int NS>
  __device__ __forceinline__ float2 apply(float2 x, int j, int r) const {
    const float2* t = tab + SLOT + j;
    if constexpr (DF) {
      const float2* l = lo + SLOT + j;
      return INV ? cmulc_pair(x, t[r * NS], l[r * NS]) : cmul_pair(x, t[r * NS], l[r * NS]);
    } else {
      return INV ? cmulc(x, t[r * NS]) : cmul(x, t[r * NS]);
    }
  }
};
using TwTable = TwTab<true>;

// Fill the twiddle tables of a line of length L (twiddle_entries<L>() entries each) and (DF) the low
// parts that make each rounded entry unit-modulus.
template <bool DF = true>
__device__ __forceinline__ void fill_twiddles(float2* tab, float2* lo, int L, int tid, int nthreads) {
  const int entries = L <= 256 ? L : 256 + L;
  for (int e = tid; e < entries; e += nthreads) {
    int m;  // exponent of W_L, below L
    if (L <= 256) {
      m = (e % 16) * (e / 16);  // two passes: NS = 16, R = L / 16
    } else if (e < 256) {
      m = (e % 16) * (e / 16) * (L / 256);  // second of three passes: NS = 16, R = 16
    } else {
      m = ((e - 256) % 256) * ((e - 256) / 256);  // third pass: NS = 256
    }
    float s, c;
    sincospif(-2.0f * (float)m / (float)L, &s, &c);
    const float2 w = make_float2(c, s);
    tab[e] = w;
    if constexpr (DF) {
      const float d = modulus_fix(w, 1.0);
      lo[e] = make_float2(w.x * d, w.y * d);
    }
  }
}

// Twiddle tables of the streaming line lengths, built at compile time: entry e holds exp(-2 pi i m / L) with m as
// fill_twiddles maps it, evaluated in double by octant reduction and a Taylor series (|angle| <= pi / 4, 14 terms:
// truncation below 1e-20) and rounded to float32.  The streaming passes copy these into LDS instead of evaluating
// one sincospif per entry in every workgroup.
struct TwPair {
  float x, y;
};

constexpr double kTwoPi = 6.283185307179586476925286766559;

// sin and cos of 2 pi r / L for 0 <= r <= L / 8 (Taylor series in double).
constexpr void cx_sincos_octant(long long r, long long L, double& s, double& c) {
  const double x = kTwoPi * (double)r / (double)L;
  double term_s = x, term_c = 1.0, sum_s = x, sum_c = 1.0;
  const double x2 = x * x;
  for (int k = 1; k <= 14; ++k) {
    term_s *= -x2 / ((double)(2 * k) * (double)(2 * k + 1));
    term_c *= -x2 / ((double)(2 * k - 1) * (double)(2 * k));
    sum_s += term_s;
    sum_c += term_c;
  }
  s = sum_s;
  c = sum_c;
}

// (cos, -sin)(2 pi m / L) for L a multiple of 8, by exact integer reduction to the first octant.
constexpr TwPair cx_twiddle(long long m, long long L) {
  m %= L;
  const long long q = m / (L / 4), rq = m % (L / 4);  // quadrant and offset in it
  double s = 0.0, c = 0.0;
  if (rq <= L / 8) {
    cx_sincos_octant(rq, L, s, c);
  } else {  // angle = pi / 2 - 2 pi (L / 4 - rq) / L
    double s2 = 0.0, c2 = 0.0;
    cx_sincos_octant(L / 4 - rq, L, s2, c2);
    s = c2;
    c = s2;
  }
  double cs = c, sn = s;  // rotate by q quarter turns
  for (long long i = 0; i < q; ++i) {
    const double t = cs;
    cs = -sn;
    sn = t;
  }
  return TwPair{(float)cs, (float)-sn};
}

template <int L>
struct TwiddleData {
  TwPair e[twiddle_entries<L>()];
  constexpr TwiddleData() : e{} {
    for (int i = 0; i < twiddle_entries<L>(); ++i) {
      long long m;
      if (L <= 256) m = (long long)(i % 16) * (i / 16);
      else if (i < 256) m = (long long)(i % 16) * (i / 16) * (L / 256);
      else m = (long long)((i - 256) % 256) * ((i - 256) / 256);
      e[i] = cx_twiddle(m, L);
    }
  }
};

template <int L>
__device__ const TwiddleData<L> kTwiddles{};

// Stage the twiddle table of a line of length L in LDS (and, when DF, the modulus-correcting low parts):
// streaming lengths copy the compile-time table, the LDS-resident ones evaluate and correct each entry.
template <int L, bool DF>
__device__ __forceinline__ void stage_twiddles(float2* tab, float2* lo, int tid, int nthreads) {
  if constexpr (DF || L < 256) {
    fill_twiddles<DF>(tab, lo, L, tid, nthreads);
  } else {
    for (int e = tid; e < twiddle_entries<L>(); e += nthreads) {
      const TwPair w = kTwiddles<L>.e[e];
      tab[e] = make_float2(w.x, w.y);
    }
  }
}

// A complex64 factor table of modulus r (the separable Fresnel propagator, r = 1/n) staged in LDS
// with (DF) the modulus-correcting low parts of its entries.
template <bool DF>
struct FixTab {
  const float2* z;
  const float2* lo;
  __device__ __forceinline__ float2 mul(float2 v, int i) const {
    if constexpr (DF) return cmul_pair(v, z[i], lo[i]);
    else return cmul(v, z[i]);
  }
  __device__ __forceinline__ float2 mulc(float2 v, int i) const {
    if constexpr (DF) return cmulc_pair(v, z[i], lo[i]);
    else return cmulc(v, z[i]);
  }
};
using FixTable = FixTab<true>;

// Copy n propagator entries (modulus 1/n each) from global memory into LDS with (DF) their low parts.
template <bool DF = true>
__device__ __forceinline__ void stage_factors(float2* z, float2* lo, const float2* src, int n, int tid,
                                              int nthreads) {
  const double r2 = 1.0 / ((double)n * (double)n);
  for (int i = tid; i < n; i += nthreads) {
    const float2 v = src[i];
    z[i] = v;
    if constexpr (DF) {
      const float d = modulus_fix(v, r2);
      lo[i] = make_float2(v.x * d, v.y * d);
    }
  }
}

// Wavefront-local synchronisation for groups that never span two waves.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace pty
