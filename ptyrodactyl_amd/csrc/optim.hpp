// Host-side declaration of the in-place Optax update kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace pty {

enum OptaxRule { OPTAX_ADAM = 0, OPTAX_ADAGRAD = 1, OPTAX_RMSPROP = 2 };

// Scalars of one step, formed in double on the host.  adam: b1, b2 and the bias corrections
// bc1 = 1 - b1^t, bc2 = 1 - b2^t; rmsprop: b1 holds the decay; adagrad uses eps only.
struct OptaxStep {
  double b1, one_minus_b1, b2, one_minus_b2, bc1, bc2, eps, neg_lr;
};

hipError_t launch_optax(int rule, bool fp64, bool cplx, long long count, void* p, const void* g, void* mu, void* nu,
                        const OptaxStep& s, hipStream_t st);

}  // namespace pty
