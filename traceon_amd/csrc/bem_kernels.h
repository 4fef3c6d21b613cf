// Host-visible launchers of the BEM kernels (bem.hip), the LU (lu.hip) and the tracer (trace.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace tb {

// Per target element (matrix row): collocation point at alpha = 0, normal there,
// flux factor 2(K-1)/(K+1) and the row kind (0 potential, 1 flux, 2 unknown type).
struct RowInfo {
  double r0, z0, nr, nz, fac;
  int kind, pad;
};

hipError_t launch_elements(const double* lines, const uint8_t* types, const double* values, int N, double* jac, double* pos,
                           double* src, RowInfo* rows, int* bad_type, hipStream_t st);
hipError_t launch_src_from_buffers(const double* jac, const double* pos, int N, double* src, hipStream_t st);
hipError_t launch_fill(double* M, size_t ld, float* M32, size_t ld32, const RowInfo* rows, const double* src, int N, int row0,
                       int row1, int accumulate, hipStream_t st);
hipError_t launch_self(double* M, size_t ld, float* M32, size_t ld32, const double* lines, const RowInfo* rows, int N, int row0,
                       int row1, double* diag_out, hipStream_t st);
hipError_t launch_sources(const double* q, const double* jac, const double* pos, int N, int nq, int pstride, int zcol, double* sw,
                          double* sr, double* sz, hipStream_t st);
hipError_t launch_field_points(const double* pts, int M, const double* sw, const double* sr, const double* sz, int Ns, double* pot,
                               double* field, int accumulate, hipStream_t st);
hipError_t launch_current_points(const double* pts, int M, const double* sw, const double* sr, const double* sz, int Ns,
                                 double* field, int accumulate, hipStream_t st);
hipError_t launch_current_axial_potential(const double* zq, int M, const double* sw, const double* sr, const double* sz, int Ns,
                                          double* out, hipStream_t st);
hipError_t launch_axial_derivs(double* out, const double* zq, int Nz, const double* sw, const double* sr, const double* sz, int Ns,
                               int current, hipStream_t st);
hipError_t launch_axial_coeffs(const double* d, const double* zq, int Nz, double* C, double* scratch, hipStream_t st);
hipError_t launch_axial_points(const double* pts, int M, const double* zq, int Nz, const double* C, double* pot, double* field,
                               hipStream_t st);
hipError_t launch_jac3d(const double* tri, int N, double* jac, double* pos, hipStream_t st);
hipError_t launch_charge(const double* lines, const double* q, int N, double* out, hipStream_t st);
hipError_t launch_scalar(int op, const double* in, int n, double* out, hipStream_t st);
hipError_t launch_line_eval(int op, const double* in, int n, double* out, hipStream_t st);
hipError_t launch_self_integrand(int kind, const double* in, int n, double* out, hipStream_t st);
hipError_t launch_tri_eval(int op, const double* in, int n, double* out, hipStream_t st);
hipError_t launch_plane(const double* p0, const double* nrm, const double* S, int Np, int* found, double* out, hipStream_t st);

}  // namespace tb
