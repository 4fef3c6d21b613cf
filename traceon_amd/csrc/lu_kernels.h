// Host-visible drivers of the dense LU and solves (lu.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace tb {

constexpr int LU_NB = 128;   // panel width of the blocked factorisation

// gate (device int, may be null): the launch does nothing when *gate <= gp (see lu.hip);
// vmax (may be null): max |C| of each workgroup into vmax[0 .. gemm_tiles(M, N)) (bit patterns of doubles)
template <typename T>
hipError_t gemm(int M, int N, int K, T alpha, const T* A, size_t lda, const T* B, size_t ldb, T beta, T* C, size_t ldc,
                hipStream_t st, const int* gate = nullptr, int gp = 0, unsigned long long* vmax = nullptr);
size_t gemm_tiles(int M, int N);   // workgroups of one gemm launch (the slots vmax needs)
template <typename T> size_t lu_workspace_bytes(int N);
// Launch count, summed HIP-event duration, algorithmic FLOP (2 M N K per product) and
// algorithmic HBM bytes (A and B read once, C read and written) of the MFMA GEMM launches
// inside one factorisation.
struct KernelTiming {
  long long launches;
  double ms;
  double flops;
  double bytes;
};
// In-place LU of the row-major N x N matrix A (leading dimension lda) with partial pivoting.
// ipiv (device, N ints) receives 0-based pivot rows; info_dev (device int, zeroed by the caller)
// receives 1 + the first zero pivot column, or stays 0.
// gemm_timing (may be null): HIP events around every GEMM launch on stream st.
template <typename T>
hipError_t lu_factor(T* A, size_t lda, int N, int* ipiv, void* ws, int* info_dev, hipStream_t st, int panel_groups,
                     KernelTiming* gemm_timing = nullptr);
// X (nrhs x N, row per right-hand side, nrhs <= 64) <- A^{-1} X using the factors and the
// workspace of the same factorisation (block inverses, permutation)
template <typename T>
hipError_t lu_solve(const T* LU, size_t lda, int N, const int* ipiv, double* X, int nrhs, const void* ws, hipStream_t st);
hipError_t residual(const double* A, size_t lda, int N, const double* B, const double* X, double* R, int nrhs,
                    unsigned long long* rmax, hipStream_t st);
hipError_t axpy(double* X, const double* D, size_t n, hipStream_t st);
// W.err of the last lu_solve with this workspace (1: a wait of its triangular solves passed the
// spin limit and the solution is NaN) into the device int *status
template <typename T>
hipError_t lu_solve_status(const void* ws, int N, int* status, hipStream_t st);
// *out = bit pattern of max |x[i]| (a nonnegative double; NaN gives the largest pattern)
hipError_t absmax(const double* x, size_t n, unsigned long long* out, hipStream_t st);
hipError_t convert_d2f(const double* S, size_t lds, float* D, size_t ldd, int rows, int cols, hipStream_t st);

}  // namespace tb
