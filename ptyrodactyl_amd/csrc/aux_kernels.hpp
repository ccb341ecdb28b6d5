// Host-side declarations of the auxiliary kernels (transmission set-up,
// FP64 normal equations, damped SPD solve).
#pragma once
#include <hip/hip_runtime.h>

namespace pty {

hipError_t launch_transmission(const double* pot, double sigma, float2* out, long long count, hipStream_t st);
constexpr int SQRT_LOSS_BLOCKS = 1024;  // workgroups (and FP64 partials) of the square-root-intensity loss
hipError_t launch_sqrt_loss(const float* pat, const float* meas, float eps, long long count, float* grad,
                            double* loss, double* partials, hipStream_t st);
int normal_partial_blocks(long long positions, int pix);
hipError_t launch_normal_equations(const float2* amp, const float* data, long long positions, int pix, int Q,
                                   double* partials, double* jtj, double* jtr, double* rss, hipStream_t st);
hipError_t launch_damped_solve(const double* A, const double* b, double damping, int Q, double* x,
                               double* model_reduction, int* status, hipStream_t st);
hipError_t launch_random_normal(unsigned k0, unsigned k1, long long count, double* out, hipStream_t st);
hipError_t launch_annular(const float* patterns, long long count, long long pixels, const unsigned char* mask,
                          double* out, hipStream_t st);

}  // namespace pty
