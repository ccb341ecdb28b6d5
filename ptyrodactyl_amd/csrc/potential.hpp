// Host-side declarations of the independent-atom potential kernels (FP64).
#pragma once
#include <hip/hip_runtime.h>

namespace pty {

// Parameterisations of the projected atomic potential.
constexpr int PARAM_LOBATO = 0;    // row: a1, b1, ..., a5, b5 (10 values)
constexpr int PARAM_KIRKLAND = 1;  // row: a1, b1, a2, b2, a3, b3, c1, d1, c2, d2, c3, d3 (12 values)
constexpr int PARAM_STRIDE = 12;   // rows are passed padded to 12 doubles

hipError_t launch_projected_radii(int kind, const double* row, long long count, const double* r, double* out,
                                  hipStream_t st);
hipError_t launch_atom_grid(int kind, int num_types, const double* rows, double pixel_size, int height, int width,
                            int supersampling, const double* centers, double* out, hipStream_t st);
size_t atom_center_grad_workspace_bytes(int num_types);
hipError_t launch_atom_center_grad(int kind, int num_types, const double* rows, double pixel_size, int height,
                                   int width, int supersampling, const double* centers, const double* grad_out,
                                   double* grad_centers, void* ws, hipStream_t st);
size_t slice_shift_grad_workspace_bytes(int height, int width, long long num_atoms);
hipError_t launch_slice_shift_grad(int num_slices, int num_types, int height, int width, long long num_atoms,
                                   const double* shifts, const long long* segments, const double2* kernel_k,
                                   const double2* adj_k, double* grad_shifts, void* ws, hipStream_t st);
hipError_t launch_slice_structure(int num_slices, int num_types, int height, int width, long long num_atoms,
                                  const double* shifts, const long long* segments, const double2* kernel_k,
                                  double2* out_k, hipStream_t st);

}  // namespace pty
