// In-place Optax update rules for the reconstructors' parameters.
//
// The reference builds its optimizers with optax 0.2.3
// (src/ptyrodactyl/multislice/multislice_recon.py:70-83):
//   adam(lr)                                   b1 0.9, b2 0.999, eps 1e-8, eps_root 0
//   adagrad(lr, initial_accumulator_value=0, eps=1e-8)
//   rmsprop(lr, decay=0.9, eps=1e-8, initial_scale=0)
// and hands them conj(jax.grad), which is torch's dL/dRe + i dL/dIm
// (multislice_recon.py:166-210).  One thread updates one (real or complex)
// element: parameter, first moment and second moment are read once and
// written once, so a step costs (3 + 3) words of HBM traffic per element
// (adam) and the launch is bound by HBM bandwidth.
//
// Per element, with g the incoming gradient and |g|^2 = Re(conj(g) g):
//   adam     mu = (1-b1) g + b1 mu;  nu = (1-b2) |g|^2 + b2 nu
//            u  = (mu / (1 - b1^t)) / (sqrt(nu / (1 - b2^t)) + eps)
//   adagrad  nu = |g|^2 + nu;  u = (nu > 0 ? 1 / sqrt(nu + eps) : 0) g
//   rmsprop  nu = (1-d) |g|^2 + d nu;  u = g / sqrt(nu + eps)
//   p = p + (-lr) u
// Scalar coefficients are formed in double on the host and rounded to the
// parameter precision, as Optax's weakly typed Python scalars are.
#include <hip/hip_runtime.h>

#include "optim.hpp"

namespace pty {
namespace {

template <typename T, int RULE, bool CPLX>
__global__ __launch_bounds__(256) void k_optax(long long count, T* __restrict__ p, const T* __restrict__ g,
                                                T* __restrict__ mu, T* __restrict__ nu, OptaxStep s) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  constexpr int W = CPLX ? 2 : 1;
  const T gr = g[W * i];
  const T gi = CPLX ? g[W * i + 1] : T(0);
  const T g2 = CPLX ? gr * gr + gi * gi : gr * gr;
  T dr, di = T(0);
  if constexpr (RULE == OPTAX_ADAM) {
    const T mr = (T)s.one_minus_b1 * gr + (T)s.b1 * mu[W * i];
    mu[W * i] = mr;
    T mi = T(0);
    if constexpr (CPLX) {
      mi = (T)s.one_minus_b1 * gi + (T)s.b1 * mu[W * i + 1];
      mu[W * i + 1] = mi;
    }
    const T v = (T)s.one_minus_b2 * g2 + (T)s.b2 * nu[i];
    nu[i] = v;
    const T den = sqrt(v / (T)s.bc2) + (T)s.eps;
    dr = (mr / (T)s.bc1) / den;
    if constexpr (CPLX) di = (mi / (T)s.bc1) / den;
  } else if constexpr (RULE == OPTAX_ADAGRAD) {
    const T t = g2 + nu[i];
    nu[i] = t;
    const T inv = t > T(0) ? T(1) / sqrt(t + (T)s.eps) : T(0);
    dr = inv * gr;
    if constexpr (CPLX) di = inv * gi;
  } else {  // OPTAX_RMSPROP
    const T v = (T)s.one_minus_b1 * g2 + (T)s.b1 * nu[i];
    nu[i] = v;
    const T inv = T(1) / sqrt(v + (T)s.eps);
    dr = inv * gr;
    if constexpr (CPLX) di = inv * gi;
  }
  const T step = (T)s.neg_lr;
  p[W * i] = p[W * i] + step * dr;
  if constexpr (CPLX) p[W * i + 1] = p[W * i + 1] + step * di;
}

template <typename T, int RULE>
hipError_t launch_rule(bool cplx, long long count, void* p, const void* g, void* mu, void* nu, const OptaxStep& s,
                       hipStream_t st) {
  const dim3 grid((unsigned)((count + 255) / 256)), blk(256);
  if (cplx)
    hipLaunchKernelGGL((k_optax<T, RULE, true>), grid, blk, 0, st, count, (T*)p, (const T*)g, (T*)mu, (T*)nu, s);
  else
    hipLaunchKernelGGL((k_optax<T, RULE, false>), grid, blk, 0, st, count, (T*)p, (const T*)g, (T*)mu, (T*)nu, s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_typed(int rule, bool cplx, long long count, void* p, const void* g, void* mu, void* nu,
                        const OptaxStep& s, hipStream_t st) {
  switch (rule) {
    case OPTAX_ADAM:
      return launch_rule<T, OPTAX_ADAM>(cplx, count, p, g, mu, nu, s, st);
    case OPTAX_ADAGRAD:
      return launch_rule<T, OPTAX_ADAGRAD>(cplx, count, p, g, mu, nu, s, st);
    case OPTAX_RMSPROP:
      return launch_rule<T, OPTAX_RMSPROP>(cplx, count, p, g, mu, nu, s, st);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_optax(int rule, bool fp64, bool cplx, long long count, void* p, const void* g, void* mu, void* nu,
                        const OptaxStep& s, hipStream_t st) {
  if (count == 0) return hipSuccess;
  return fp64 ? launch_typed<double>(rule, cplx, count, p, g, mu, nu, s, st)
              : launch_typed<float>(rule, cplx, count, p, g, mu, nu, s, st);
}

}  // namespace pty
