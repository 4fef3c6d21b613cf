// BEM kernels for gfx950: element buffers, dense FP64 matrix fill, log-singular
// diagonal, field sums at point batches, axial derivatives and their interpolation
// coefficients.  Each kernel names the reference function it replaces.
//
// Layout in HBM (all FP64 unless stated):
//   lines    (N,4,3)       reference layout, read once by k_elements
//   jac,pos  (N,16),(N,16,2) reference layout (EffectivePointCharges), kept for the API
//   src      3 x (16,N)    SoA (r, z, w) per Gauss point, k-major: lane j of a wave reads
//                          consecutive j -> fully coalesced in the fill
//   rows     (N) RowInfo   per target element: (r0, z0, nr, nz, flux factor, kind)
//   matrix   (N,ld)        row-major, one 64-wide wave = 64 consecutive columns
#include "radial_math.h"
#include "bem_kernels.h"

namespace tb {

// ------------------------------------------------------------------------------
// Element preparation: fill_jacobian_buffer_radial (radial.c:224-246) plus the
// per-row data of fill_matrix_radial (radial.c:265-292) and the SoA source copy.
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_elements(const double* __restrict__ lines, const uint8_t* __restrict__ types,
                                                  const double* __restrict__ values, int N, double* __restrict__ jac,
                                                  double* __restrict__ pos, double* __restrict__ src, RowInfo* __restrict__ rows,
                                                  int* __restrict__ bad_type) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * NQ) return;
  int i = t / NQ, k = t % NQ;
  Line4 L = load_line4(lines + 12 * (size_t)i);
  double r, z, j;
  line4_point(L, GL_X[k], r, z, j);
  double w = GL_W[k] * j;
  if (jac) jac[(size_t)i * NQ + k] = w;
  if (pos) { pos[(size_t)i * 2 * NQ + 2 * k] = r; pos[(size_t)i * 2 * NQ + 2 * k + 1] = z; }
  if (src) {
    src[(size_t)k * N + i] = r;
    src[(size_t)(NQ + k) * N + i] = z;
    src[(size_t)(2 * NQ + k) * N + i] = w;
  }
  if (rows && k == 0) {
    RowInfo R;
    double j0;
    line4_point(L, 0.0, R.r0, R.z0, j0);
    line4_normal(L, 0.0, R.nr, R.nz);
    int ty = types ? types[i] : VOLTAGE_FIXED;
    if (ty == VOLTAGE_FIXED || ty == VOLTAGE_FUN || ty == MAGNETOSTATIC_POT) { R.kind = 0; R.fac = 0.0; }
    else if (ty == DIELECTRIC || ty == MAGNETIZABLE) { R.kind = 1; R.fac = flux_factor(values[i]); }
    else { R.kind = 2; R.fac = 0.0; if (bad_type) atomicOr(bad_type, 1); }
    rows[i] = R;
  }
}

// SoA sources from caller-provided (N,16) jacobian and (N,16,2) position buffers
__global__ __launch_bounds__(256) void k_src_from_buffers(const double* __restrict__ jac, const double* __restrict__ pos, int N,
                                                          double* __restrict__ src) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * NQ) return;
  int i = t / NQ, k = t % NQ;
  src[(size_t)k * N + i] = pos[2 * (size_t)t];
  src[(size_t)(NQ + k) * N + i] = pos[2 * (size_t)t + 1];
  src[(size_t)(2 * NQ + k) * N + i] = jac[t];
}

// ------------------------------------------------------------------------------
// Dense fill, fill_matrix_radial (radial.c:249-310).
// Block = 256 consecutive columns x ROWS rows; the rows' data are wave-uniform
// (scalar loads), the source Gauss points are per-lane coalesced loads.  Each
// entry is the 16-point sum over the source element, written once (or added,
// reproducing the reference's accumulate semantics).  Optionally also writes an
// FP32 copy of the same entries for the mixed-precision factorisation.
// ------------------------------------------------------------------------------
template <int ROWS>
__global__ __launch_bounds__(256) void k_fill(double* __restrict__ M, size_t ld, float* __restrict__ M32, size_t ld32,
                                              const RowInfo* __restrict__ rows, const double* __restrict__ src, int N,
                                              int row0, int row1, int accumulate) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int ib = row0 + blockIdx.y * ROWS;
  const int jj = j < N ? j : N - 1;
  const double* sr = src;
  const double* sz = src + (size_t)NQ * N;
  const double* sw = src + (size_t)2 * NQ * N;
  double acc[ROWS];
#pragma unroll
  for (int t = 0; t < ROWS; t++) acc[t] = 0.0;

  // a block of potential rows only (every row of a Dirichlet geometry) runs a loop without
  // the per-row branches, whose joins cost register copies of the accumulators in every pair
  bool pot_only = ib + ROWS - 1 <= row1;
#pragma unroll
  for (int t = 0; t < ROWS; t++) pot_only = pot_only && rows[min(ib + t, row1)].kind == 0;
  if (pot_only) {
    double r0[ROWS], z0[ROWS];
#pragma unroll
    for (int t = 0; t < ROWS; t++) { r0[t] = rows[ib + t].r0; z0[t] = rows[ib + t].z0; }
#pragma unroll 2
    for (int k = 0; k < NQ; k++) {
      const double r = sr[(size_t)k * N + jj], z = sz[(size_t)k * N + jj], w = sw[(size_t)k * N + jj];
#pragma unroll
      for (int t = 0; t < ROWS; t++) acc[t] = fma(w, ring_potential(r0[t], r - r0[t], z - z0[t]), acc[t]);
    }
  } else {
#pragma unroll 2
    for (int k = 0; k < NQ; k++) {
      const double r = sr[(size_t)k * N + jj], z = sz[(size_t)k * N + jj], w = sw[(size_t)k * N + jj];
#pragma unroll
      for (int t = 0; t < ROWS; t++) {
        const int i = ib + t;
        if (i <= row1) {
          const RowInfo R = rows[i];
          const double dr = r - R.r0, dz = z - R.z0;
          if (R.kind == 0) {
            acc[t] = fma(w, ring_potential(R.r0, dr, dz), acc[t]);
          } else if (R.kind == 1) {
            double gr, gz;
            ring_grad(R.r0, dr, dz, gr, gz);
            // field_dot_normal_radial: factor * (n . E), E = -grad
            acc[t] = fma(w, -R.fac * (R.nr * gr + R.nz * gz), acc[t]);
          } else {
            acc[t] = __builtin_nan("");
          }
        }
      }
    }
  }
  if (j >= N) return;
#pragma unroll
  for (int t = 0; t < ROWS; t++) {
    const int i = ib + t;
    if (i <= row1) {
      double v = acc[t];
      double* dst = M + (size_t)i * ld + j;
      if (accumulate) v += *dst;
      *dst = v;
      if (M32) M32[(size_t)i * ld32 + j] = (float)v;
    }
  }
}

// ------------------------------------------------------------------------------
// Singular diagonal: self_potential_radial / self_field_dot_normal_radial
// (radial.c:182-222) integrated over alpha in [-1,1] as solver.py:239-247 does with
// QUADPACK.  Here: split at the singular point alpha = 0, geometric grading towards it
// (ratio SIGMA, NLEV levels, the innermost panel [0, SIGMA^(NLEV-1)] below 1e-13) and
// the 16-point Gauss rule on every panel.  A log singularity at distance SIGMA/(1-SIGMA)
// panel lengths leaves a per-panel relative error below ~1e-17, far inside QUADPACK's
// 1e-9 request.  One 64-lane wave per element, shuffles for the reduction.
// ------------------------------------------------------------------------------
constexpr double SIGMA = 0.3;
constexpr int NLEV = 26;
constexpr int NSELF = 2 * NLEV * NQ;   // 832 nodes per element

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_self(double* __restrict__ M, size_t ld, float* __restrict__ M32, size_t ld32,
                                              const double* __restrict__ lines, const RowInfo* __restrict__ rows, int N,
                                              int row0, int row1, double* __restrict__ diag_out) {
  const int lane = threadIdx.x & 63;
  const int i = row0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i > row1) return;
  const Line4 L = load_line4(lines + 12 * (size_t)i);
  const RowInfo R = rows[i];
  double acc = 0.0;
  for (int n = lane; n < NSELF; n += 64) {
    const int side = n / (NLEV * NQ);
    const int m = (n % (NLEV * NQ)) / NQ;
    const int q = n % NQ;
    const double b = pow(SIGMA, (double)m);
    const double a = (m == NLEV - 1) ? 0.0 : b * SIGMA;
    const double half = 0.5 * (b - a);
    double alpha = fma(half, GL_X[q], 0.5 * (a + b));
    if (side) alpha = -alpha;
    double dr, dz, j;
    line4_delta(L, alpha, dr, dz, j);
    double f;
    if (R.kind == 0) {
      f = j * ring_potential(R.r0, dr, dz);
    } else {
      double gr, gz;
      ring_grad(R.r0, dr, dz, gr, gz);
      f = j * -R.fac * (R.nr * gr + R.nz * gz);
    }
    acc = fma(half * GL_W[q], f, acc);
  }
  acc = wave_sum(acc);
  if (lane == 0) {
    double v = (R.kind == 0) ? acc : (R.kind == 1 ? acc - 1.0 : __builtin_nan(""));
    if (M) M[(size_t)i * ld + i] = v;
    if (M32) M32[(size_t)i * ld32 + i] = (float)v;
    if (diag_out) diag_out[i - row0] = acc;   // the integral itself
  }
}

// ------------------------------------------------------------------------------
// Sources for the field sums: w = q * jac, flattened over (element, Gauss point).
// (the products radial.c:116, 159-160 form inside their loops)
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sources(const double* __restrict__ q, const double* __restrict__ jac,
                                                 const double* __restrict__ pos, int N, int nq, int pstride, int zcol,
                                                 double* __restrict__ sw, double* __restrict__ sr, double* __restrict__ sz) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N * nq) return;
  int i = s / nq;
  sw[s] = q[i] * jac[s];
  sr[s] = pos[(size_t)s * pstride];
  sz[s] = pos[(size_t)s * pstride + zcol];
}

// ------------------------------------------------------------------------------
// Field sums at a batch of points: potential_radial / field_radial (radial.c:106-173).
// One lane per point; the loop over sources is wave-uniform, so the source triple
// comes in through scalar loads and every lane shares it.  FP64 VALU bound.
// ------------------------------------------------------------------------------
template <bool POT, bool FIELD>
__global__ __launch_bounds__(256) void k_field_points(const double* __restrict__ pts, int M, const double* __restrict__ sw,
                                                      const double* __restrict__ sr, const double* __restrict__ sz, int Ns,
                                                      double* __restrict__ pot, double* __restrict__ field, int accumulate) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int tt = t < M ? t : M - 1;
  const double x = pts[3 * (size_t)tt], y = pts[3 * (size_t)tt + 1], z0 = pts[3 * (size_t)tt + 2];
  const double r0 = sqrt(x * x + y * y);
  double P = 0.0, Er = 0.0, Ez = 0.0;
#pragma unroll 2
  for (int s = 0; s < Ns; s++) {
    const double w = sw[s], dr = sr[s] - r0, dz = sz[s] - z0;
    if (POT && FIELD) {
      double ph, gr, gz;
      ring_all(r0, dr, dz, ph, gr, gz);
      P = fma(w, ph, P); Er = fma(-w, gr, Er); Ez = fma(-w, gz, Ez);
    } else if (POT) {
      P = fma(w, ring_potential(r0, dr, dz), P);
    } else {
      double gr, gz;
      ring_grad(r0, dr, dz, gr, gz);
      Er = fma(-w, gr, Er); Ez = fma(-w, gz, Ez);
    }
  }
  if (t >= M) return;
  if (POT) pot[t] = accumulate ? pot[t] + P : P;
  if (FIELD) {
    double fx = 0.0, fy = 0.0;
    if (r0 >= AXIS_EPS) { fx = x / r0 * Er; fy = y / r0 * Er; }
    double* f = field + 3 * (size_t)t;
    if (accumulate) { f[0] += fx; f[1] += fy; f[2] += Ez; }
    else { f[0] = fx; f[1] = fy; f[2] = Ez; }
  }
}

// The same sums for a few points: PPW points per workgroup, whose 256 threads take the
// sources s = tid, tid + 256, ... and meet in a fixed-order reduction (wave butterflies, then
// the four waves in order).  With fewer points than the GPU has lanes, one lane per point
// would leave most of the chip idle while each lane walks every source.
template <int PPW, int NV>
__device__ __forceinline__ void block_sums(double (&v)[PPW][NV], double (&red)[4][PPW][NV]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < PPW; i++)
#pragma unroll
    for (int k = 0; k < NV; k++) v[i][k] = wave_sum(v[i][k]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < PPW; i++)
#pragma unroll
      for (int k = 0; k < NV; k++) red[wv][i][k] = v[i][k];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PPW; i++)
#pragma unroll
    for (int k = 0; k < NV; k++) v[i][k] = ((red[0][i][k] + red[1][i][k]) + red[2][i][k]) + red[3][i][k];
}

template <bool POT, bool FIELD, int PPW>
__global__ __launch_bounds__(256) void k_field_points_wg(const double* __restrict__ pts, int M, const double* __restrict__ sw,
                                                         const double* __restrict__ sr, const double* __restrict__ sz, int Ns,
                                                         double* __restrict__ pot, double* __restrict__ field, int accumulate) {
  __shared__ double red[4][PPW][3];
  const int p0 = blockIdx.x * PPW;
  double x[PPW], y[PPW], z0[PPW], r0[PPW], v[PPW][3];
#pragma unroll
  for (int i = 0; i < PPW; i++) {
    const int t = min(p0 + i, M - 1);
    x[i] = pts[3 * (size_t)t]; y[i] = pts[3 * (size_t)t + 1]; z0[i] = pts[3 * (size_t)t + 2];
    r0[i] = sqrt(x[i] * x[i] + y[i] * y[i]);
    v[i][0] = v[i][1] = v[i][2] = 0.0;
  }
  for (int s = threadIdx.x; s < Ns; s += 256) {
    const double w = sw[s], rs = sr[s], zz = sz[s];
#pragma unroll
    for (int i = 0; i < PPW; i++) {
      const double dr = rs - r0[i], dz = zz - z0[i];
      if (POT && FIELD) {
        double ph, gr, gz;
        ring_all(r0[i], dr, dz, ph, gr, gz);
        v[i][0] = fma(w, ph, v[i][0]); v[i][1] = fma(-w, gr, v[i][1]); v[i][2] = fma(-w, gz, v[i][2]);
      } else if (POT) {
        v[i][0] = fma(w, ring_potential(r0[i], dr, dz), v[i][0]);
      } else {
        double gr, gz;
        ring_grad(r0[i], dr, dz, gr, gz);
        v[i][1] = fma(-w, gr, v[i][1]); v[i][2] = fma(-w, gz, v[i][2]);
      }
    }
  }
  block_sums<PPW, 3>(v, red);
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int i = 0; i < PPW; i++) {
    const int t = p0 + i;
    if (t >= M) break;
    if (POT) pot[t] = accumulate ? pot[t] + v[i][0] : v[i][0];
    if (FIELD) {
      double fx = 0.0, fy = 0.0;
      if (r0[i] >= AXIS_EPS) { fx = x[i] / r0[i] * v[i][1]; fy = y[i] / r0[i] * v[i][1]; }
      double* f = field + 3 * (size_t)t;
      if (accumulate) { f[0] += fx; f[1] += fy; f[2] += v[i][2]; }
      else { f[0] = fx; f[1] = fy; f[2] = v[i][2]; }
    }
  }
}

template <int PPW>
__global__ __launch_bounds__(256) void k_current_points_wg(const double* __restrict__ pts, int M, const double* __restrict__ sw,
                                                           const double* __restrict__ sr, const double* __restrict__ sz, int Ns,
                                                           double* __restrict__ field, int accumulate) {
  __shared__ double red[4][PPW][2];
  const int p0 = blockIdx.x * PPW;
  double x[PPW], y[PPW], z0[PPW], r[PPW], v[PPW][2];
#pragma unroll
  for (int i = 0; i < PPW; i++) {
    const int t = min(p0 + i, M - 1);
    x[i] = pts[3 * (size_t)t]; y[i] = pts[3 * (size_t)t + 1]; z0[i] = pts[3 * (size_t)t + 2];
    r[i] = sqrt(x[i] * x[i] + y[i] * y[i]);
    v[i][0] = v[i][1] = 0.0;
  }
  for (int s = threadIdx.x; s < Ns; s += 256) {
    const double w = sw[s], rs = sr[s], zz = sz[s];
#pragma unroll
    for (int i = 0; i < PPW; i++) {
      double hr, hz;
      current_ring_field(r[i], z0[i], rs, zz, hr, hz);
      v[i][0] = fma(w, hr, v[i][0]); v[i][1] = fma(w, hz, v[i][1]);
    }
  }
  block_sums<PPW, 2>(v, red);
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int i = 0; i < PPW; i++) {
    const int t = p0 + i;
    if (t >= M) break;
    double fx = 0.0, fy = 0.0;
    if (r[i] >= AXIS_EPS) { fx = x[i] / r[i] * v[i][0]; fy = y[i] / r[i] * v[i][0]; }
    double* f = field + 3 * (size_t)t;
    if (accumulate) { f[0] += fx; f[1] += fy; f[2] += v[i][1]; }
    else { f[0] = fx; f[1] = fy; f[2] = v[i][1]; }
  }
}

__global__ __launch_bounds__(256) void k_current_axial_potential_wg(const double* __restrict__ zq, int M, const double* __restrict__ sw,
                                                                    const double* __restrict__ sr, const double* __restrict__ sz,
                                                                    int Ns, double* __restrict__ out) {
  __shared__ double red[4][1][1];
  const double z0 = zq[blockIdx.x];
  double v[1][1] = {{0.0}};
  for (int s = threadIdx.x; s < Ns; s += 256) v[0][0] = fma(sw[s], current_ring_axial_potential(z0, sr[s], sz[s]), v[0][0]);
  block_sums<1, 1>(v, red);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0][0];
}

// current_field_radial (radial.c:56-83) at a batch of points; sources (w, r, z)
__global__ __launch_bounds__(256) void k_current_points(const double* __restrict__ pts, int M, const double* __restrict__ sw,
                                                        const double* __restrict__ sr, const double* __restrict__ sz, int Ns,
                                                        double* __restrict__ field, int accumulate) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int tt = t < M ? t : M - 1;
  const double x = pts[3 * (size_t)tt], y = pts[3 * (size_t)tt + 1], z0 = pts[3 * (size_t)tt + 2];
  const double r = sqrt(x * x + y * y);
  double Hr = 0.0, Hz = 0.0;
  for (int s = 0; s < Ns; s++) {
    double hr, hz;
    current_ring_field(r, z0, sr[s], sz[s], hr, hz);
    Hr = fma(sw[s], hr, Hr); Hz = fma(sw[s], hz, Hz);
  }
  if (t >= M) return;
  double fx = 0.0, fy = 0.0;
  if (r >= AXIS_EPS) { fx = x / r * Hr; fy = y / r * Hr; }
  double* f = field + 3 * (size_t)t;
  if (accumulate) { f[0] += fx; f[1] += fy; f[2] += Hz; }
  else { f[0] = fx; f[1] = fy; f[2] = Hz; }
}

// current_potential_axial (radial.c:39-53) at a batch of z values
__global__ __launch_bounds__(256) void k_current_axial_potential(const double* __restrict__ zq, int M, const double* __restrict__ sw,
                                                                 const double* __restrict__ sr, const double* __restrict__ sz,
                                                                 int Ns, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M) return;
  const double z0 = zq[t];
  double P = 0.0;
  for (int s = 0; s < Ns; s++) P = fma(sw[s], current_ring_axial_potential(z0, sr[s], sz[s]), P);
  out[t] = P;
}

// ------------------------------------------------------------------------------
// Axial derivatives: axial_derivatives_radial (radial.c:19-36) and
// current_axial_derivatives_radial (radial.c:86-103), accumulating into (Nz, 9).
// One block per z sample, threads stride over the sources, block reduction.
// ------------------------------------------------------------------------------
template <bool CURRENT>
__global__ __launch_bounds__(256) void k_axial_derivs(double* __restrict__ out, const double* __restrict__ zq, int Nz,
                                                      const double* __restrict__ sw, const double* __restrict__ sr,
                                                      const double* __restrict__ sz, int Ns) {
  __shared__ double red[4][NDERIV];
  const int iz = blockIdx.x;
  const double z0 = zq[iz];
  double acc[NDERIV];
#pragma unroll
  for (int l = 0; l < NDERIV; l++) acc[l] = 0.0;
  for (int s = threadIdx.x; s < Ns; s += blockDim.x) {
    double d[NDERIV];
    const double r = sr[s];
    if (CURRENT) {
      current_ring_axial_derivs(z0, r, sz[s], d);
#pragma unroll
      for (int l = 0; l < NDERIV; l++) acc[l] = fma(sw[s], d[l], acc[l]);
    } else {
      ring_axial_derivs(z0, r, sz[s], d);
      const double w = sw[s] * (0.5 * r);
#pragma unroll
      for (int l = 0; l < NDERIV; l++) acc[l] = fma(w, d[l], acc[l]);
    }
  }
#pragma unroll
  for (int l = 0; l < NDERIV; l++) acc[l] = wave_sum(acc[l]);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int l = 0; l < NDERIV; l++) red[wv][l] = acc[l];
  }
  __syncthreads();
  if (threadIdx.x < NDERIV) {
    double v = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    out[(size_t)iz * NDERIV + threadIdx.x] += v;
  }
}

// ------------------------------------------------------------------------------
// Interpolation coefficients (field.py:906-937), layout (Nz-1, 9, 6) highest power first.
// Orders 0..6: quintic Hermite on each interval matching value, first and second
// derivative (the samples of orders l, l+1, l+2) at both ends -- the polynomial
// scipy's BPoly.from_derivatives builds, written in closed form.
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_quintic(const double* __restrict__ d, const double* __restrict__ zq, int Nz,
                                                 double* __restrict__ C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (Nz - 1) * 7) return;
  const int i = t / 7, l = t % 7;
  const double h = zq[i + 1] - zq[i];
  const double y0 = d[(size_t)i * NDERIV + l], y1 = d[(size_t)(i + 1) * NDERIV + l];
  const double p0 = d[(size_t)i * NDERIV + l + 1], p1 = d[(size_t)(i + 1) * NDERIV + l + 1];
  const double s0 = d[(size_t)i * NDERIV + l + 2], s1 = d[(size_t)(i + 1) * NDERIV + l + 2];
  const double D0 = y1 - y0 - p0 * h - 0.5 * s0 * h * h;   // value mismatch of the quadratic Taylor part
  const double D1 = (p1 - p0 - s0 * h) * h;                 // slope mismatch, scaled by h
  const double D2 = (s1 - s0) * h * h;                      // curvature mismatch, scaled by h^2
  const double h3 = h * h * h;
  double* c = C + ((size_t)i * NDERIV + l) * 6;
  c[0] = (6.0 * D0 - 3.0 * D1 + 0.5 * D2) / (h3 * h * h);
  c[1] = (-15.0 * D0 + 7.0 * D1 - D2) / (h3 * h);
  c[2] = (10.0 * D0 - 4.0 * D1 + 0.5 * D2) / h3;
  c[3] = 0.5 * s0;
  c[4] = p0;
  c[5] = y0;
}

// Orders 7, 8: not-a-knot cubic spline (scipy CubicSpline default) in slots 2..5.
// The knot slopes s solve a tridiagonal system (Thomas algorithm).  Rows 1..n-2:
//   dx_i s_{i-1} + 2 (dx_{i-1} + dx_i) s_i + dx_{i-1} s_{i+1} = 3 (dx_i sl_{i-1} + dx_{i-1} sl_i),
// row 0 and row n-1 the not-a-knot end conditions.  One wave per order walks the
// elimination: its lanes form the rows of a block of 64 with coalesced loads, and the
// wave-uniform recurrence takes them one by one through readlane, the lane of each row
// keeping that row's result for one coalesced store per block.  scratch: 3 * Nz doubles per
// order (cp, dp, then the slopes in place of dp).
__device__ __forceinline__ double uniform_lane(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// row i of the slope system for order l: (sub, diag, sup, rhs)
__device__ __forceinline__ void spline_row(const double* __restrict__ d, const double* __restrict__ zq, int n, int l, int i,
                                           double& sub, double& diag, double& sup, double& rhs) {
  auto y = [&](int k) { return d[(size_t)k * NDERIV + l]; };
  auto dx = [&](int k) { return zq[k + 1] - zq[k]; };
  auto sl = [&](int k) { return (y(k + 1) - y(k)) / dx(k); };
  if (i == 0) {
    const double dd = dx(0) + dx(1);
    sub = 0.0; diag = dx(1); sup = dd;
    rhs = ((dx(0) + 2.0 * dd) * dx(1) * sl(0) + dx(0) * dx(0) * sl(1)) / dd;
  } else if (i == n - 1) {
    const double dd = dx(n - 3) + dx(n - 2);
    sub = dd; diag = dx(n - 3); sup = 0.0;
    rhs = (dx(n - 2) * dx(n - 2) * sl(n - 3) + (2.0 * dd + dx(n - 2)) * dx(n - 3) * sl(n - 2)) / dd;
  } else {
    sub = dx(i); diag = 2.0 * (dx(i - 1) + dx(i)); sup = dx(i - 1);
    rhs = 3.0 * (dx(i) * sl(i - 1) + dx(i - 1) * sl(i));
  }
}

// 64 threads, blockIdx.x = order - 7: forward elimination, then back substitution
__global__ __launch_bounds__(64) void k_spline_slopes(const double* __restrict__ d, const double* __restrict__ zq, int Nz,
                                                      double* __restrict__ scratch) {
  const int l = 7 + blockIdx.x, lane = threadIdx.x, n = Nz;
  double* cp = scratch + (size_t)blockIdx.x * 3 * Nz;
  double* dp = cp + Nz;
  double cprev = 0.0, dprev = 0.0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    double sub = 0.0, diag = 1.0, sup = 0.0, rhs = 0.0;
    if (i0 + lane < n) spline_row(d, zq, n, l, i0 + lane, sub, diag, sup, rhs);
    double mc = 0.0, md = 0.0;
    const int cnt = min(64, n - i0);
#pragma unroll 8
    for (int t = 0; t < cnt; t++) {
      const double a = uniform_lane(sub, t), b = uniform_lane(diag, t), c = uniform_lane(sup, t), r = uniform_lane(rhs, t);
      const double den = b - a * cprev;
      cprev = c / den;
      dprev = (r - a * dprev) / den;
      if (lane == t) { mc = cprev; md = dprev; }
    }
    if (i0 + lane < n) { cp[i0 + lane] = mc; dp[i0 + lane] = md; }
  }
  // back substitution from the last row: s_{n-1} = dp_{n-1}, s_i = dp_i - cp_i s_{i+1}
  double snext = 0.0;
  for (int i1 = n; i1 > 0; i1 -= 64) {
    const int i0 = max(0, i1 - 64), cnt = i1 - i0;
    double c = 0.0, dv = 0.0;
    if (lane < cnt) { c = cp[i0 + lane]; dv = dp[i0 + lane]; }
    double ms = 0.0;
#pragma unroll 8
    for (int t = cnt - 1; t >= 0; t--) {
      const double ct = uniform_lane(c, t), dt = uniform_lane(dv, t);
      snext = (i0 + t == n - 1) ? dt : dt - ct * snext;
      if (lane == t) ms = snext;
    }
    if (lane < cnt) dp[i0 + lane] = ms;   // the slopes replace dp
  }
}

// slots 2..5 of orders 7, 8 on every interval from the slopes
__global__ __launch_bounds__(256) void k_spline_coeffs(const double* __restrict__ d, const double* __restrict__ zq, int Nz,
                                                       double* __restrict__ C, const double* __restrict__ scratch) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * (Nz - 1)) return;
  const int o = t / (Nz - 1), i = t % (Nz - 1), l = 7 + o;
  const double* s = scratch + (size_t)o * 3 * Nz + Nz;
  const double h = zq[i + 1] - zq[i], y0 = d[(size_t)i * NDERIV + l], m = (d[(size_t)(i + 1) * NDERIV + l] - y0) / h;
  const double q = (s[i] + s[i + 1] - 2.0 * m) / h;
  double* c = C + ((size_t)i * NDERIV + l) * 6;
  c[0] = 0.0; c[1] = 0.0;
  c[2] = q / h;
  c[3] = (m - s[i]) / h - q;
  c[4] = s[i];
  c[5] = y0;
}

// potential_radial_derivs / field_radial_derivs (radial.c:313-384) at a batch of points
__global__ __launch_bounds__(256) void k_axial_points(const double* __restrict__ pts, int M, const double* __restrict__ zq, int Nz,
                                                      const double* __restrict__ C, double* __restrict__ pot,
                                                      double* __restrict__ field) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M) return;
  const double x = pts[3 * (size_t)t], y = pts[3 * (size_t)t + 1], z = pts[3 * (size_t)t + 2];
  if (pot) pot[t] = axial_potential(zq, Nz, C, x, y, z);
  if (field) axial_field(zq, Nz, C, x, y, z, field + 3 * (size_t)t);
}

// fill_jacobian_buffer_3d (three_d.c:34-62): 12-point rule on each triangle
__global__ __launch_bounds__(256) void k_jac3d(const double* __restrict__ tri, int N, double* __restrict__ jac,
                                               double* __restrict__ pos) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * NQ3) return;
  const int i = t / NQ3, k = t % NQ3;
  const double* T = tri + 9 * (size_t)i;
  double e1[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]};
  double e2[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]};
  double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2], c2 = e1[0] * e2[1] - e1[1] * e2[0];
  double area = 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  jac[t] = 2.0 * TRI_W[k] * area;
  for (int c = 0; c < 3; c++) pos[3 * (size_t)t + c] = T[c] + TRI_B1[k] * e1[c] + TRI_B2[k] * e2[c];
}

// charge_radial (radial.c:124-148): total charge of each element at unit-free density q
__global__ __launch_bounds__(256) void k_charge(const double* __restrict__ lines, const double* __restrict__ q, int N,
                                                double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  Line4 L = load_line4(lines + 12 * (size_t)i);
  double s = 0.0;
  for (int k = 0; k < NQ; k++) {
    double r, z, j;
    line4_point(L, GL_X[k], r, z, j);
    s += 2.0 * PI * r * GL_W[k] * j * q[i];
  }
  out[i] = s;
}

// Elementwise helpers behind the reference's scalar entry points.
__global__ void k_scalar(int op, const double* __restrict__ in, int n, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double* a = in + 4 * (size_t)t;
  double* o = out + 4 * (size_t)t;
  switch (op) {
    case 0: o[0] = ellipkm1_(a[0]); break;
    case 1: o[0] = ellipem1_(a[0]); break;
    case 2: o[0] = ellipk_(a[0]); break;
    case 3: o[0] = ellipe_(a[0]); break;
    case 4: o[0] = ring_potential(a[0], a[2], a[3]); break;
    case 5: { double gr, gz; ring_grad(a[0], a[2], a[3], gr, gz); o[0] = gr; break; }
    case 6: { double gr, gz; ring_grad(a[0], a[2], a[3], gr, gz); o[0] = gz; break; }
    case 7: o[0] = flux_factor(a[0]); break;
    case 8: o[0] = current_ring_axial_potential(a[0], a[1], a[2]); break;
    case 9: { double hr, hz; current_ring_field(a[0], a[1], a[2], a[3], hr, hz); o[0] = hr; o[1] = hz; break; }
    case 10: {   // normal_2d (utilities_2d.c:18-30): tangent p1->p2 rotated to (ty, -tx)
      double tx = a[2] - a[0], ty = a[3] - a[1];
      double len = sqrt(tx * tx + ty * ty);
      o[0] = ty / len; o[1] = -tx / len;
      break;
    }
    default: o[0] = __builtin_nan("");
  }
}

// Element geometry at one parameter value: position_and_jacobian_radial (op 0),
// delta_position_and_jacobian_radial (op 1), higher_order_normal_radial (op 2).
// in: per item (alpha, line[12]) ; out: per item (r, z, jac)
__global__ void k_line_eval(int op, const double* __restrict__ in, int n, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double* a = in + 13 * (size_t)t;
  Line4 L = load_line4(a + 1);
  double* o = out + 3 * (size_t)t;
  if (op == 0) line4_point(L, a[0], o[0], o[1], o[2]);
  else if (op == 1) line4_delta(L, a[0], o[0], o[1], o[2]);
  else { line4_normal(L, a[0], o[0], o[1]); o[2] = 0.0; }
}

// The integrands of the singular diagonal at a batch of parameter values:
// self_potential_radial (kind 0) and self_field_dot_normal_radial (kind 1), radial.c:182-222.
// The target is the element's own point at alpha = 0 (its normal there for kind 1), the
// source the offset of the point at alpha.  in per item (alpha, K, line[12]); out: one value.
__global__ void k_self_integrand(int kind, const double* __restrict__ in, int n, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double* a = in + 14 * (size_t)t;
  const Line4 L = load_line4(a + 2);
  double r0, z0, j0, dr, dz, j;
  line4_point(L, 0.0, r0, z0, j0);
  line4_delta(L, a[0], dr, dz, j);
  if (kind == 0) {
    out[t] = j * ring_potential(r0, dr, dz);
  } else {
    double nr, nz, gr, gz;
    line4_normal(L, 0.0, nr, nz);
    ring_grad(r0, dr, dz, gr, gz);
    out[t] = j * -flux_factor(a[1]) * (nr * gr + nz * gz);   // factor * (n . E), E = -grad
  }
}

// Flat-triangle helpers of the 3D utilities (utilities_3d.c:95-134, three_d.c:310-328):
// op 0 normal_3d (unit normal of (v1 - v0) x (v2 - v0)), op 1 position_and_jacobian_3d
// (v0 + alpha (v1 - v0) + beta (v2 - v0), jacobian = twice the area), op 2 the area.
// in per item (alpha, beta, tri[9]); out per item 4 values.
__global__ void k_tri_eval(int op, const double* __restrict__ in, int n, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double* a = in + 11 * (size_t)t;
  const double* T = a + 2;
  const double e1[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]};
  const double e2[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]};
  const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  double* o = out + 4 * (size_t)t;
  if (op == 0) {
    for (int k = 0; k < 3; k++) o[k] = c[k] / len;
    o[3] = 0.0;
  } else if (op == 1) {
    for (int k = 0; k < 3; k++) o[k] = T[k] + a[0] * e1[k] + a[1] * e2[k];
    o[3] = len;   // 2 * area
  } else {
    o[0] = 0.5 * len;
    o[1] = o[2] = o[3] = 0.0;
  }
}

// plane_intersection (tracing.c:184-223): the last index i whose signed distance
// differs in sign from i+1 (zero counts as negative), found in parallel with an atomic max.
__global__ void k_plane_scan(const double* __restrict__ p0, const double* __restrict__ nrm, const double* __restrict__ S, int Np,
                             int* __restrict__ found) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= Np) return;
  const double nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  const double* a = S + 6 * (size_t)i;
  const double* b = S + 6 * (size_t)(i + 1);
  double ka = (nrm[0] * (p0[0] - a[0]) + nrm[1] * (p0[1] - a[1]) + nrm[2] * (p0[2] - a[2])) / nn;
  double kb = (nrm[0] * (p0[0] - b[0]) + nrm[1] * (p0[1] - b[1]) + nrm[2] * (p0[2] - b[2])) / nn;
  if ((ka > 0) != (kb > 0)) atomicMax(found, i);
}

__global__ void k_plane_interp(const double* __restrict__ p0, const double* __restrict__ nrm, const double* __restrict__ S,
                               const int* __restrict__ found, double* __restrict__ out) {
  const int i = *found;
  if (threadIdx.x != 0) return;
  if (i < 0) { for (int c = 0; c < 6; c++) out[c] = __builtin_nan(""); return; }
  const double nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  const double* a = S + 6 * (size_t)i;
  const double* b = S + 6 * (size_t)(i + 1);
  double ka = (nrm[0] * (p0[0] - a[0]) + nrm[1] * (p0[1] - a[1]) + nrm[2] * (p0[2] - a[2])) / nn;
  double kb = (nrm[0] * (p0[0] - b[0]) + nrm[1] * (p0[1] - b[1]) + nrm[2] * (p0[2] - b[2])) / nn;
  double diff = ka - kb;
  double f = -kb / diff, fp = ka / diff;
  for (int c = 0; c < 6; c++) out[c] = fp * b[c] + f * a[c];
}

// ------------------------------------------------------------------------------
// Host launchers
// ------------------------------------------------------------------------------
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_elements(const double* lines, const uint8_t* types, const double* values, int N, double* jac, double* pos,
                           double* src, RowInfo* rows, int* bad_type, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_elements, dim3(cdiv((size_t)N * NQ, 256)), dim3(256), 0, st, lines, types, values, N, jac, pos, src, rows,
                     bad_type);
  return hipGetLastError();
}

hipError_t launch_src_from_buffers(const double* jac, const double* pos, int N, double* src, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_src_from_buffers, dim3(cdiv((size_t)N * NQ, 256)), dim3(256), 0, st, jac, pos, N, src);
  return hipGetLastError();
}

constexpr int FILL_ROWS = 4;
hipError_t launch_fill(double* M, size_t ld, float* M32, size_t ld32, const RowInfo* rows, const double* src, int N, int row0,
                       int row1, int accumulate, hipStream_t st) {
  if (N <= 0 || row1 < row0) return hipSuccess;
  int nrows = row1 - row0 + 1;
  // grid.y is limited to 65535: tile the row range
  const int max_y = 65535;
  for (int r = row0; r <= row1; r += max_y * FILL_ROWS) {
    int r1 = r + max_y * FILL_ROWS - 1;
    if (r1 > row1) r1 = row1;
    dim3 grid(cdiv(N, 256), cdiv(r1 - r + 1, FILL_ROWS));
    hipLaunchKernelGGL(k_fill<FILL_ROWS>, grid, dim3(256), 0, st, M, ld, M32, ld32, rows, src, N, r, r1, accumulate);
  }
  (void)nrows;
  return hipGetLastError();
}

hipError_t launch_self(double* M, size_t ld, float* M32, size_t ld32, const double* lines, const RowInfo* rows, int N, int row0,
                       int row1, double* diag_out, hipStream_t st) {
  if (N <= 0 || row1 < row0) return hipSuccess;
  hipLaunchKernelGGL(k_self, dim3(cdiv(row1 - row0 + 1, 4)), dim3(256), 0, st, M, ld, M32, ld32, lines, rows, N, row0, row1,
                     diag_out);
  return hipGetLastError();
}

hipError_t launch_sources(const double* q, const double* jac, const double* pos, int N, int nq, int pstride, int zcol, double* sw,
                          double* sr, double* sz, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sources, dim3(cdiv((size_t)N * nq, 256)), dim3(256), 0, st, q, jac, pos, N, nq, pstride, zcol, sw, sr, sz);
  return hipGetLastError();
}

// Below this many points the sums run a workgroup per point (or four), above it a lane per
// point (the lane kernel fills the GPU from about 2^18 points on: 1024 workgroups, 4 waves per SIMD).
constexpr int FEW_POINTS = 1 << 17;

hipError_t launch_field_points(const double* pts, int M, const double* sw, const double* sr, const double* sz, int Ns, double* pot,
                               double* field, int accumulate, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  if (M < FEW_POINTS) {
    const bool four = M >= 4096;
    const dim3 g(four ? cdiv(M, 4) : (unsigned)M), b(256);
    if (pot && field) {
      if (four) hipLaunchKernelGGL((k_field_points_wg<true, true, 4>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
      else hipLaunchKernelGGL((k_field_points_wg<true, true, 1>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
    } else if (pot) {
      if (four) hipLaunchKernelGGL((k_field_points_wg<true, false, 4>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
      else hipLaunchKernelGGL((k_field_points_wg<true, false, 1>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
    } else if (field) {
      if (four) hipLaunchKernelGGL((k_field_points_wg<false, true, 4>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
      else hipLaunchKernelGGL((k_field_points_wg<false, true, 1>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
    }
    return hipGetLastError();
  }
  dim3 g(cdiv(M, 256)), b(256);
  if (pot && field) hipLaunchKernelGGL((k_field_points<true, true>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
  else if (pot) hipLaunchKernelGGL((k_field_points<true, false>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
  else if (field) hipLaunchKernelGGL((k_field_points<false, true>), g, b, 0, st, pts, M, sw, sr, sz, Ns, pot, field, accumulate);
  return hipGetLastError();
}

hipError_t launch_current_points(const double* pts, int M, const double* sw, const double* sr, const double* sz, int Ns,
                                 double* field, int accumulate, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  if (M < FEW_POINTS) {
    if (M >= 4096)
      hipLaunchKernelGGL(k_current_points_wg<4>, dim3(cdiv(M, 4)), dim3(256), 0, st, pts, M, sw, sr, sz, Ns, field, accumulate);
    else
      hipLaunchKernelGGL(k_current_points_wg<1>, dim3(M), dim3(256), 0, st, pts, M, sw, sr, sz, Ns, field, accumulate);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_current_points, dim3(cdiv(M, 256)), dim3(256), 0, st, pts, M, sw, sr, sz, Ns, field, accumulate);
  return hipGetLastError();
}

hipError_t launch_current_axial_potential(const double* zq, int M, const double* sw, const double* sr, const double* sz, int Ns,
                                          double* out, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  if (M < FEW_POINTS) {
    hipLaunchKernelGGL(k_current_axial_potential_wg, dim3(M), dim3(256), 0, st, zq, M, sw, sr, sz, Ns, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_current_axial_potential, dim3(cdiv(M, 256)), dim3(256), 0, st, zq, M, sw, sr, sz, Ns, out);
  return hipGetLastError();
}

hipError_t launch_axial_derivs(double* out, const double* zq, int Nz, const double* sw, const double* sr, const double* sz, int Ns,
                               int current, hipStream_t st) {
  if (Nz <= 0 || Ns <= 0) return hipSuccess;
  if (current) hipLaunchKernelGGL(k_axial_derivs<true>, dim3(Nz), dim3(256), 0, st, out, zq, Nz, sw, sr, sz, Ns);
  else hipLaunchKernelGGL(k_axial_derivs<false>, dim3(Nz), dim3(256), 0, st, out, zq, Nz, sw, sr, sz, Ns);
  return hipGetLastError();
}

hipError_t launch_axial_coeffs(const double* d, const double* zq, int Nz, double* C, double* scratch, hipStream_t st) {
  if (Nz < 4) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_quintic, dim3(cdiv((size_t)(Nz - 1) * 7, 256)), dim3(256), 0, st, d, zq, Nz, C);
  hipLaunchKernelGGL(k_spline_slopes, dim3(2), dim3(64), 0, st, d, zq, Nz, scratch);
  hipLaunchKernelGGL(k_spline_coeffs, dim3(cdiv((size_t)(Nz - 1) * 2, 256)), dim3(256), 0, st, d, zq, Nz, C, (const double*)scratch);
  return hipGetLastError();
}

hipError_t launch_axial_points(const double* pts, int M, const double* zq, int Nz, const double* C, double* pot, double* field,
                               hipStream_t st) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axial_points, dim3(cdiv(M, 256)), dim3(256), 0, st, pts, M, zq, Nz, C, pot, field);
  return hipGetLastError();
}

hipError_t launch_jac3d(const double* tri, int N, double* jac, double* pos, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jac3d, dim3(cdiv((size_t)N * NQ3, 256)), dim3(256), 0, st, tri, N, jac, pos);
  return hipGetLastError();
}

hipError_t launch_charge(const double* lines, const double* q, int N, double* out, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_charge, dim3(cdiv(N, 256)), dim3(256), 0, st, lines, q, N, out);
  return hipGetLastError();
}

hipError_t launch_scalar(int op, const double* in, int n, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scalar, dim3(cdiv(n, 256)), dim3(256), 0, st, op, in, n, out);
  return hipGetLastError();
}

hipError_t launch_line_eval(int op, const double* in, int n, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_line_eval, dim3(cdiv(n, 256)), dim3(256), 0, st, op, in, n, out);
  return hipGetLastError();
}

hipError_t launch_self_integrand(int kind, const double* in, int n, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_self_integrand, dim3(cdiv(n, 256)), dim3(256), 0, st, kind, in, n, out);
  return hipGetLastError();
}

hipError_t launch_tri_eval(int op, const double* in, int n, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tri_eval, dim3(cdiv(n, 256)), dim3(256), 0, st, op, in, n, out);
  return hipGetLastError();
}

hipError_t launch_plane(const double* p0, const double* nrm, const double* S, int Np, int* found, double* out, hipStream_t st) {
  if (Np >= 2) hipLaunchKernelGGL(k_plane_scan, dim3(cdiv(Np - 1, 256)), dim3(256), 0, st, p0, nrm, S, Np, found);
  hipLaunchKernelGGL(k_plane_interp, dim3(1), dim3(64), 0, st, p0, nrm, S, found, out);
  return hipGetLastError();
}

}  // namespace tb
