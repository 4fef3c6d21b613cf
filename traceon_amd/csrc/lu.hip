// Dense LU with partial pivoting and solves, for the BEM charge system
// (replaces numpy.linalg.solve -> LAPACK dgesv at traceon/solver.py:173).
//
// Row-major storage (the fill's layout).  Blocked right-looking factorisation, panel
// width NB.  Per panel at column k0:
//   1. k_panel   : cooperative kernel, one workgroup per CU.  The m x NB panel is split
//                  by rows over the workgroups and held in LDS for the whole panel;
//                  each column costs ONE grid barrier: every workgroup publishes its
//                  local pivot candidate together with that row's panel values, and the
//                  owner of the next diagonal row publishes that row, so after the
//                  barrier everyone can swap and update without a second exchange.
//   2. k_laswp   : the panel's row interchanges applied to the other columns.
//   3. k_diag_block : inverses of the unit-lower and upper triangles of the NB x NB
//                  diagonal block (they also serve the triangular solves).
//   4. gemm      : U12 = inv(L11) * A12  (into a workspace), A22 -= L21 * U12 (MFMA),
//                  U12 copied back.
// Before step 1 each panel first tries the unpivoted fast path described below.
// Pivot choice is LAPACK's (largest |a|, first index on ties).
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <stdint.h>
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <vector>
#include "lu_kernels.h"

namespace cg = cooperative_groups;

namespace tb {

template <typename T> __device__ __forceinline__ T tabs(T x) { return x < T(0) ? -x : x; }

// Gate of the speculative factorisation: the kernels of panel p run only while no panel
// q <= p has failed the no-pivot test (*gate holds the first failed panel, or nblk).
// A null gate always runs.
__device__ __forceinline__ bool gated_off(const int* gate, int p) { return gate != nullptr && *gate <= p; }

// ---------------------------------------------------------------------------------
// Panel factorisation
// ---------------------------------------------------------------------------------
template <typename T, int NB>
__global__ __launch_bounds__(256) void k_panel(T* __restrict__ A, size_t lda, int N, int k0, int* __restrict__ ipiv,
                                               double* __restrict__ wsd, int* __restrict__ wsi, T* __restrict__ wsr,
                                               int* __restrict__ info) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* P = reinterpret_cast<T*>(smem_raw);          // local panel rows, R x NB
  __shared__ T prow[NB], drow[NB];
  __shared__ double red_v[256];
  __shared__ int red_i[256], red_s[256];
  __shared__ int s_p, s_slot;

  cg::grid_group grid = cg::this_grid();
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int m = N - k0;
  const int R = (m + G - 1) / G;
  const int rbeg = k0 + g * R;
  const int rend = min(rbeg + R, N);
  const int nloc = rend > rbeg ? rend - rbeg : 0;
  const int nbw = min(NB, N - k0);   // panel width (last panel may be narrower)

  volatile double* cand_val = wsd;          // [2][G]
  volatile int* cand_idx = wsi;             // [2][G]
  volatile T* cand_row = wsr;               // [2][G][NB]
  volatile T* diag_row = wsr + 2 * (size_t)G * NB;   // [2][NB]

  for (int e = tid; e < nloc * NB; e += 256) {
    int r = e / NB, c = e % NB;
    P[e] = c < nbw ? A[(size_t)(rbeg + r) * lda + k0 + c] : T(0);
  }
  __syncthreads();

  // publish the local candidate for column j among local rows with global index >= first
  auto publish = [&](int j, int first, int buf) {
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int r = tid; r < nloc; r += 256) {
      int gr = rbeg + r;
      if (gr < first) continue;
      double v = (double)tabs(P[r * NB + j]);
      if (v > best || (v == best && gr < bi)) { best = v; bi = gr; }
    }
    red_v[tid] = best; red_i[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) {
        double v2 = red_v[tid + s]; int i2 = red_i[tid + s];
        if (v2 > red_v[tid] || (v2 == red_v[tid] && i2 < red_i[tid])) { red_v[tid] = v2; red_i[tid] = i2; }
      }
      __syncthreads();
    }
    int bi0 = red_i[0];
    if (tid == 0) { cand_val[buf * G + g] = red_v[0]; cand_idx[buf * G + g] = bi0; }
    if (bi0 != 0x7fffffff && red_v[0] >= 0.0) {
      for (int c = tid; c < NB; c += 256) cand_row[((size_t)buf * G + g) * NB + c] = P[(bi0 - rbeg) * NB + c];
    }
    if (first >= rbeg && first < rend) {
      for (int c = tid; c < NB; c += 256) diag_row[(size_t)buf * NB + c] = P[(first - rbeg) * NB + c];
    }
    __threadfence();
    __syncthreads();
  };

  publish(0, k0, 0);
  grid.sync();

  for (int j = 0; j < nbw; j++) {
    const int buf = j & 1;
    const int d = k0 + j;
    // global pivot: largest candidate, smallest row on ties
    {
      double v = -1.0; int i = 0x7fffffff, s = -1;
      if (tid < G) { v = cand_val[buf * G + tid]; i = cand_idx[buf * G + tid]; s = tid; }
      red_v[tid] = v; red_i[tid] = i; red_s[tid] = s;
      __syncthreads();
      for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) {
          double v2 = red_v[tid + st]; int i2 = red_i[tid + st];
          if (v2 > red_v[tid] || (v2 == red_v[tid] && i2 < red_i[tid])) {
            red_v[tid] = v2; red_i[tid] = i2; red_s[tid] = red_s[tid + st];
          }
        }
        __syncthreads();
      }
      if (tid == 0) { s_p = red_i[0]; s_slot = red_s[0]; }
      __syncthreads();
    }
    // No candidate (every remaining entry of the column is NaN): keep the row in place
    // (no swap) and report the column as singular; the index stays inside the matrix.
    const bool none = (s_slot < 0) || (s_p < d) || (s_p >= N);
    const int p = none ? d : s_p, slot = s_slot;
    for (int c = tid; c < NB; c += 256) {
      drow[c] = diag_row[(size_t)buf * NB + c];
      prow[c] = none ? drow[c] : cand_row[((size_t)buf * G + slot) * NB + c];
    }
    __syncthreads();
    if (g == 0 && tid == 0) {
      ipiv[d] = p;
      if ((none || !(tabs(prow[j]) > T(0))) && *info == 0) *info = d + 1;
    }
    __syncthreads();
    if (p != d && p >= rbeg && p < rend)
      for (int c = tid; c < NB; c += 256) P[(p - rbeg) * NB + c] = drow[c];
    if (d >= rbeg && d < rend)
      for (int c = tid; c < NB; c += 256) P[(d - rbeg) * NB + c] = prow[c];
    __syncthreads();
    // scale the sub-diagonal part of column j, then rank-1 update of the columns right of j
    const T inv = T(1) / prow[j];
    const int first = max(d + 1, rbeg);
    const int nr = rend > first ? rend - first : 0;
    for (int r = tid; r < nr; r += 256) P[(first - rbeg + r) * NB + j] *= inv;
    __syncthreads();
    const int ncol = nbw - j - 1;
    if (ncol > 0) {
      for (int e = tid; e < nr * ncol; e += 256) {
        int r = first - rbeg + e / ncol, c = j + 1 + e % ncol;
        P[r * NB + c] -= P[r * NB + j] * prow[c];
      }
      __syncthreads();
      publish(j + 1, d + 1, buf ^ 1);
      grid.sync();
    }
  }
  __syncthreads();
  for (int e = tid; e < nloc * NB; e += 256) {
    int r = e / NB, c = e % NB;
    if (c < nbw) A[(size_t)(rbeg + r) * lda + k0 + c] = P[e];
  }
}

// Row interchanges of one panel applied to the columns outside it (LAPACK dlaswp):
// one thread per column, swaps in panel order.
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(T* __restrict__ A, size_t lda, int N, int k0, int nb, const int* __restrict__ ipiv) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int ncols = N - nb;   // columns outside [k0, k0+nb)
  if (c >= ncols) return;
  int col = c < k0 ? c : c + nb;
  for (int j = 0; j < nb; j++) {
    int d = k0 + j, p = ipiv[d];
    if (p != d && p >= 0 && p < N) {
      T a = A[(size_t)d * lda + col], b = A[(size_t)p * lda + col];
      A[(size_t)d * lda + col] = b;
      A[(size_t)p * lda + col] = a;
    }
  }
}

// ---------------------------------------------------------------------------------
// No-pivot fast path of one panel.  Partial pivoting leaves every row in place exactly
// when each multiplier satisfies |l| <= 1 (the diagonal entry is then the largest of its
// column, ties going to the diagonal as in LAPACK).  So the panel is first factored
// without interchanges (the diagonal block by k_diag_block, the rows below it as
// L21 = A21 * inv(U11) on MFMA); if every computed |l| <= 1 the result is the partially
// pivoted factorisation (ipiv = identity, up to the rounding of the multipliers) and no
// grid-wide pivot search is needed.  Otherwise the panel is restored from its copy and
// factored by k_panel.  Pure potential (Dirichlet) BEM rows never need a swap; rows of
// flux type (dielectric / magnetizable) usually do.
// ---------------------------------------------------------------------------------

// Diagonal block of a panel in one workgroup of 256 threads (4 waves) with the NB x NB block
// in LDS, row-major with a padded row stride.  Outside nb x nb the block is padded with the
// identity.
//   factor = 1: unpivoted LU of the block, right-looking in 16-column sub-panels s:
//     1. the sub-panel (rows 16s..NB-1 of its 16 columns) is factored in registers, one row
//        per lane: every wave holds the sub-panel's 16 diagonal rows in lanes 0-15 and up to
//        48 of the rows below them in lanes 16-63, and eliminates them redundantly, the
//        pivot row reaching its lanes by readlane, so the 16 steps need no barrier.
//        Multipliers are formed with the reciprocal of the pivot, as LAPACK's dgetf2 scales
//        its column; max |l| -> lmax;
//     2. the block row right of it is solved with the sub-panel's unit-lower 16 x 16
//        diagonal block, one thread per column (column-oriented, so the 15 dependent steps
//        each carry up to 15 independent FMAs);
//     3. the trailing block is updated with the rank-16 product in 4 x 4 register tiles.
//   Then, always: Y = inv(L11) (unit lower) and Z = inv(U11), in place (the strictly lower
//   part of the block becomes Y's, the upper part Z's).  The 16 x 16 diagonal blocks of both
//   are inverted by 16-lane groups, one column per lane; then three doubling levels
//   b = 16, 32, 64 form the off-diagonal blocks of the 2b x 2b inverses,
//     Y21 = -Y22 (L21 Y11),   Z12 = -Z11 (U12 Z22),
//   each product in 4 x 4 register tiles (16-byte LDS reads over runs of 4 k; only the 4 x 4
//   block on a triangle's diagonal is masked) and written back over the block it replaces.
//   Y and Z are written row-major NB x NB with zero padding.
// Work: 2/3 NB^3 FMAs for the LU and as many for the inverses, about 40 barriers in all.
template <typename T>
constexpr int diag_row_stride() { return LU_NB + 16 / (int)sizeof(T); }
template <typename T>
constexpr size_t diag_lds_bytes() { return sizeof(T) * LU_NB * diag_row_stride<T>(); }

__device__ __forceinline__ float lane_bcast(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// reciprocal of a pivot: FP32 by the hardware reciprocal (1 ulp; the FP32 factors only
// precondition the FP64 refinement), FP64 correctly rounded
__device__ __forceinline__ float pivot_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double pivot_rcp(double x) { return 1.0 / x; }

// four consecutive elements of an LDS row (16-byte aligned start)
template <typename T>
__device__ __forceinline__ void lds_load4(const T* p, T (&v)[4]) {
  if constexpr (sizeof(T) == 4) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v x = *reinterpret_cast<const f4v*>(p);
    v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
  } else {
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v x = *reinterpret_cast<const d2v*>(p), y = *reinterpret_cast<const d2v*>(p + 2);
    v[0] = x[0]; v[1] = x[1]; v[2] = y[0]; v[3] = y[1];
  }
}
// acc[i][j] += sum over k in [k0, k1) of S(r + i, ka + k) * S(kb + k, c + j), k1 - k0 a multiple of 4
template <typename T, int DS>
__device__ __forceinline__ void tile_mac(T (&acc)[4][4], const T* S, int r, int ka, int kb, int c, int k0, int k1) {
#pragma unroll 2
  for (int k = k0; k < k1; k += 4) {
    T a[4][4], q[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++) lds_load4(&S[(r + i) * DS + ka + k], a[i]);
#pragma unroll
    for (int kk = 0; kk < 4; kk++) lds_load4(&S[(kb + k + kk) * DS + c], q[kk]);
#pragma unroll
    for (int kk = 0; kk < 4; kk++)
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i][kk], q[kk][j], acc[i][j]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_diag_block(T* __restrict__ A, size_t lda, int k0, int nb, int factor,
                                                   int* __restrict__ ipiv, T* __restrict__ Y, T* __restrict__ Z,
                                                   unsigned long long* __restrict__ lmax, const int* gate, int gp,
                                                   T* __restrict__ backup) {
  constexpr int NB = LU_NB, DS = diag_row_stride<T>(), PER = NB * NB / 256, CH = 16, VEC = 16 / (int)sizeof(T), PERV = PER / VEC;
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  if (gated_off(gate, gp)) return;
  // this kernel is on the factorisation's critical path and shares its CU with trailing-update
  // workgroups: its waves take issue priority over theirs
  __builtin_amdgcn_s_setprio(3);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);   // element (r, c) at r * DS + c
  __shared__ T udinv[NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // the block: 16-byte row segments when A allows them (all in flight at once), else elements
  const bool vec = (lda % VEC) == 0 && (k0 % VEC) == 0 && (((uintptr_t)A) % 16) == 0 && nb == NB;
  if (vec) {
    vec_t x[PERV];
#pragma unroll
    for (int i = 0; i < PERV; i++) {
      const int e = tid + 256 * i, r = e / (NB / VEC), c = (e % (NB / VEC)) * VEC;
      x[i] = *reinterpret_cast<const vec_t*>(A + (size_t)(k0 + r) * lda + k0 + c);
    }
#pragma unroll
    for (int i = 0; i < PERV; i++) {
      const int e = tid + 256 * i, r = e / (NB / VEC), c = (e % (NB / VEC)) * VEC;
      *reinterpret_cast<vec_t*>(&S[r * DS + c]) = x[i];
      if (backup) *reinterpret_cast<vec_t*>(&backup[r * NB + c]) = x[i];   // for a restore when the panel fails
    }
  } else {
#pragma unroll 1
    for (int i0 = 0; i0 < PER; i0 += CH) {
      T x[CH];
#pragma unroll
      for (int i = 0; i < CH; i++) {
        const int e = tid + 256 * (i0 + i), r = e / NB, c = e % NB;
        x[i] = (r < nb && c < nb) ? A[(size_t)(k0 + r) * lda + k0 + c] : (r == c ? T(1) : T(0));
      }
#pragma unroll
      for (int i = 0; i < CH; i++) {
        const int e = tid + 256 * (i0 + i), r = e / NB, c = e % NB;
        S[r * DS + c] = x[i];
        if (backup) backup[e] = x[i];
      }
    }
  }
  __syncthreads();

  if (factor) {
    T m = T(0);   // max |l|
    for (int p0 = 0; p0 < NB; p0 += 16) {
      {
        const int row = lane < 16 ? p0 + lane : p0 + 16 + 48 * wave + (lane - 16);
        const bool own = row < NB && (lane >= 16 || wave == 0);   // stored back by this lane
        T x[16];
#pragma unroll
        for (int j = 0; j < 16; j++) x[j] = row < NB ? S[row * DS + p0 + j] : T(0);
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const T rp = pivot_rcp(lane_bcast(x[j], j));   // row p0 + j sits in lane j
          T u[16];
#pragma unroll
          for (int k = j + 1; k < 16; k++) u[k] = lane_bcast(x[k], j);
          if (lane > j) {
            const T l = x[j] * rp;
            x[j] = l;
#pragma unroll
            for (int k = j + 1; k < 16; k++) x[k] = fma(-l, u[k], x[k]);
            const T al = row < nb ? tabs(l) : T(0);
            m = (al > m || al != al) ? al : m;
          }
        }
        __syncthreads();   // every wave has read the sub-panel before any lane writes it
        if (own) {
#pragma unroll
          for (int j = 0; j < 16; j++) S[row * DS + p0 + j] = x[j];
        }
      }
      __syncthreads();
      const int nc = NB - p0 - 16;   // columns right of the sub-panel (and rows below it)
      if (tid < nc) {                // U12 = inv(L_ss) A12, one column per thread
        const int c = p0 + 16 + tid;
        T x[16];
#pragma unroll
        for (int i = 0; i < 16; i++) x[i] = S[(p0 + i) * DS + c];
#pragma unroll
        for (int k = 0; k < 15; k++) {
#pragma unroll
          for (int i = k + 1; i < 16; i++) x[i] = fma(-S[(p0 + i) * DS + p0 + k], x[k], x[i]);
        }
#pragma unroll
        for (int i = 0; i < 16; i++) S[(p0 + i) * DS + c] = x[i];
      }
      __syncthreads();
      const int mt = nc / 4;         // trailing block -= L21 U12, 4 x 4 tiles
      for (int t = tid; t < mt * mt; t += 256) {
        const int r = p0 + 16 + 4 * (t / mt), c = p0 + 16 + 4 * (t % mt);
        T acc[4][4] = {};
        tile_mac<T, DS>(acc, S, r, p0, p0, c, 0, 16);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) S[(r + i) * DS + c + j] -= acc[i][j];
      }
      __syncthreads();
    }
    double md = (double)m;
    for (int o = 32; o >= 1; o >>= 1) {
      const double x = __shfl_xor(md, o, 64);
      md = (x > md || x != x) ? x : md;
    }
    if (lane == 0) {
      if (md != md) md = 1e300;
      if (md > 0.0) atomicMax(lmax, (unsigned long long)__double_as_longlong(md));
    }
    if (vec) {
#pragma unroll
      for (int i = 0; i < PERV; i++) {
        const int e = tid + 256 * i, r = e / (NB / VEC), c = (e % (NB / VEC)) * VEC;
        *reinterpret_cast<vec_t*>(A + (size_t)(k0 + r) * lda + k0 + c) = *reinterpret_cast<const vec_t*>(&S[r * DS + c]);
      }
    } else {
#pragma unroll 1
      for (int i0 = 0; i0 < PER; i0 += CH) {
        T x[CH];
#pragma unroll
        for (int i = 0; i < CH; i++) {
          const int e = tid + 256 * (i0 + i);
          x[i] = S[(e / NB) * DS + e % NB];
        }
#pragma unroll
        for (int i = 0; i < CH; i++) {
          const int e = tid + 256 * (i0 + i), r = e / NB, c = e % NB;
          if (r < nb && c < nb) A[(size_t)(k0 + r) * lda + k0 + c] = x[i];
        }
      }
    }
    if (tid < nb) ipiv[k0 + tid] = k0 + tid;
    __syncthreads();   // the block is read out before the inverses overwrite it
  }

  if (tid < NB) udinv[tid] = T(1) / S[tid * DS + tid];
  __syncthreads();
  {  // 16 x 16 diagonal blocks: lanes 0-31 invert L of blocks 2w, 2w+1, lanes 32-63 U of the same
    const int g = lane >> 4, cl = lane & 15, o = 16 * (2 * wave + (g & 1));
    T x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = (i == cl) ? T(1) : T(0);
    if (g < 2) {
#pragma unroll
      for (int k = 0; k < 15; k++) {
#pragma unroll
        for (int i = k + 1; i < 16; i++) x[i] = fma(-S[(o + i) * DS + o + k], x[k], x[i]);
      }
#pragma unroll
      for (int i = 0; i < 16; i++)
        if (i > cl) S[(o + i) * DS + o + cl] = x[i];
    } else {
#pragma unroll
      for (int k = 15; k >= 0; k--) {
        x[k] *= udinv[o + k];
#pragma unroll
        for (int i = 0; i < k; i++) x[i] = fma(-S[(o + i) * DS + o + k], x[k], x[i]);
      }
#pragma unroll
      for (int i = 0; i < 16; i++)
        if (i <= cl) S[(o + i) * DS + o + cl] = x[i];
    }
  }
  __syncthreads();
  // doubling: for each pair of b x b diagonal blocks at o = 2b p,
  //   L: T = L21 Y11 -> L21's place, then Y21 = -Y22 T;  U: T = U12 Z22 -> U12's place, then Z12 = -Z11 T
  // (Y is unit lower: stored strictly below the diagonal; Z upper: stored on and above it).
  // A triangular factor is masked only in the 4 x 4 block on its diagonal.
  for (int b = 16; b < NB; b *= 2) {
    const int tb = b / 4, ntile = tb * tb, npair = NB / (2 * b), total = 2 * npair * ntile;
#pragma unroll 1
    for (int step = 0; step < 2; step++) {
      T acc[2][4][4] = {};
      int rr[2], cc[2];
#pragma unroll
      for (int n = 0; n < 2; n++) {
        const int it = tid + 256 * n;
        rr[n] = -1;
        cc[n] = 0;
        if (it >= total) continue;
        const int side = it / (npair * ntile), rest = it % (npair * ntile);
        const int o = 2 * b * (rest / ntile), tile = rest % ntile;
        // a wave's lanes share the bound of the k range (tj in step 0, ti in step 1) as far as
        // possible: the other index runs fastest
        const int ti = 4 * (step == 0 ? tile % tb : tile / tb), tj = 4 * (step == 0 ? tile / tb : tile % tb);
        T (&ac)[4][4] = acc[n];
        if (side == 0) {   // lower: rows o+b+ti, columns o+tj
          const int r = o + b + ti, c = o + tj;
          rr[n] = r;
          cc[n] = c;
          if (step == 0) {   // L21 Y11: Y11(o+k, c+j) vanishes for k < tj + j
            tile_mac<T, DS>(ac, S, r, o, o, c, tj + 4, b);
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
#pragma unroll
              for (int j = 0; j < 4; j++) {
                const T y = kk > j ? S[(c + kk) * DS + c + j] : (kk == j ? T(1) : T(0));
#pragma unroll
                for (int i = 0; i < 4; i++) ac[i][j] = fma(S[(r + i) * DS + c + kk], y, ac[i][j]);
              }
          } else {           // Y22 T (negated below): Y22(r+i, o+b+k) vanishes for k > ti + i
            tile_mac<T, DS>(ac, S, r, o + b, o + b, c, 0, ti);
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
#pragma unroll
              for (int i = 0; i < 4; i++) {
                const T y = i > kk ? S[(r + i) * DS + r + kk] : (i == kk ? T(1) : T(0));
#pragma unroll
                for (int j = 0; j < 4; j++) ac[i][j] = fma(y, S[(r + kk) * DS + c + j], ac[i][j]);
              }
          }
        } else {           // upper: rows o+ti, columns o+b+tj
          const int r = o + ti, c = o + b + tj;
          rr[n] = r;
          cc[n] = c;
          if (step == 0) {   // U12 Z22: Z22(o+b+k, c+j) vanishes for k > tj + j
            tile_mac<T, DS>(ac, S, r, o + b, o + b, c, 0, tj);
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
#pragma unroll
              for (int j = 0; j < 4; j++) {
                const T z = kk <= j ? S[(c + kk) * DS + c + j] : T(0);
#pragma unroll
                for (int i = 0; i < 4; i++) ac[i][j] = fma(S[(r + i) * DS + c + kk], z, ac[i][j]);
              }
          } else {           // Z11 T (negated below): Z11(r+i, o+k) vanishes for k < ti + i
            tile_mac<T, DS>(ac, S, r, o, o, c, ti + 4, b);
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
#pragma unroll
              for (int i = 0; i < 4; i++) {
                const T z = i <= kk ? S[(r + i) * DS + r + kk] : T(0);
#pragma unroll
                for (int j = 0; j < 4; j++) ac[i][j] = fma(z, S[(r + kk) * DS + c + j], ac[i][j]);
              }
          }
        }
      }
      __syncthreads();   // every operand read before the results replace them
#pragma unroll
      for (int n = 0; n < 2; n++) {
        if (rr[n] < 0) continue;
        const T sg = step == 0 ? T(1) : T(-1);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) S[(rr[n] + i) * DS + cc[n] + j] = sg * acc[n][i][j];
      }
      __syncthreads();
    }
  }
#pragma unroll 4
  for (int i = 0; i < PERV; i++) {
    const int e = tid + 256 * i, r = e / (NB / VEC), c = (e % (NB / VEC)) * VEC;
    const vec_t x = *reinterpret_cast<const vec_t*>(&S[r * DS + c]);
    vec_t y, z;
#pragma unroll
    for (int v = 0; v < VEC; v++) {
      const bool in = r < nb && c + v < nb;
      y[v] = in ? (r > c + v ? x[v] : (r == c + v ? T(1) : T(0))) : T(0);
      z[v] = in ? (r <= c + v ? x[v] : T(0)) : T(0);
    }
    *reinterpret_cast<vec_t*>(&Y[r * NB + c]) = y;
    *reinterpret_cast<vec_t*>(&Z[r * NB + c]) = z;
  }
}

// After the factorisation: info = 1 + first column whose U diagonal is zero or NaN
// (numpy.linalg.solve reports such a matrix as singular), else unchanged.
template <typename T>
__global__ __launch_bounds__(1024) void k_diag_check(const T* __restrict__ A, size_t lda, int N, int* __restrict__ info) {
  __shared__ int first;
  if (threadIdx.x == 0) first = 0x7fffffff;
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    T d = A[(size_t)i * lda + i];
    if (!(tabs(d) > T(0))) atomicMin(&first, i);
  }
  __syncthreads();
  if (threadIdx.x == 0 && first != 0x7fffffff && (*info == 0 || *info > first + 1)) *info = first + 1;
}

// copy a rows x cols block back into the matrix and track max |value| (bits, nonnegative)
template <typename T>
__global__ __launch_bounds__(256) void k_copy_back_max(const T* __restrict__ S, size_t lds, T* __restrict__ D, size_t ldd, int rows,
                                                       int cols, unsigned long long* __restrict__ vmax, const int* gate, int gp) {
  __shared__ double wm[4];
  if (gated_off(gate, gp)) return;
  const size_t total = (size_t)rows * cols;
  double m = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / cols), c = (int)(e % cols);
    const T v = S[(size_t)r * lds + c];
    D[(size_t)r * ldd + c] = v;
    const double a = (double)tabs(v);
    m = (a > m || a != a) ? a : m;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const double x = __shfl_xor(m, o, 64);
    m = (x > m || x != x) ? x : m;
  }
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int w = 0; w < 4; w++) mm = (wm[w] > mm || wm[w] != wm[w]) ? wm[w] : mm;
    if (mm != mm) mm = 1e300;
    if (mm > 0.0) atomicMax(vmax, (unsigned long long)__double_as_longlong(mm));
  }
}

// ---------------------------------------------------------------------------------
// GEMM on MFMA: C(MxN) = beta*C + alpha*A(MxK)*B(KxN), all row-major.
// Block 256 threads = 2x2 waves, block tile GB x GB (128; 64 for products too small to
// fill the chip with 128-tiles), wave tile GB/2 x GB/2 made of MFMA 16x16 tiles; K staged
// through LDS in chunks of BK.  Instruction per element type:
//   double: v_mfma_f64_16x16x4f64, float: v_mfma_f32_16x16x4f32
// Operand layout of both (16x16x4, one block): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; the four accumulator values v = 0..3 of lane l are D[row][l%16] with
// row = l/16 + 4v for the f64 instruction and row = 4(l/16) + v for the f32 one.
//
// LDS: A as [row][k] with a row stride of 36 words, B as [k][col] with GB + 16 words (in
// units of 4 bytes; 18 and GB + 8 elements for double).  The 16 rows x 4 k an MFMA reads from
// A and the 4 k x 16 cols it reads from B then fall in distinct banks, and both tiles are
// written with one 16-byte store per global vector load.
// ---------------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<double> {
  typedef d4 type;
  static __device__ __forceinline__ d4 mma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MfmaAcc<float> {
  typedef f4 type;
  static __device__ __forceinline__ f4 mma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};


// Schedule of one tile:
//  1. per K chunk: the next chunk's global loads are issued into registers, the MFMAs of the
//     chunk run from LDS (the operand fragments of step s+1 are read before the MFMAs of step
//     s), then the registers are written to LDS;
//  2. write-back through LDS: the accumulators (times alpha) are stored in the MFMA layout,
//     then each thread updates 16-byte row segments of C, 8 in flight, CROWS tile rows per
//     pass, so the read-modify-write of a tile costs a few memory latencies instead of one
//     per element.  The products are summed from zero and C is added once at the end (adding
//     them into C one MFMA at a time rounds each partial sum at the magnitude of C, and an
//     FP32 LU built that way needs more refinement steps).
// NBUF = 1 (FP32, 128-tiles): one LDS buffer and two barriers per chunk, 36 KiB of LDS, so
// three workgroups fit a CU and one's write-back overlaps the others' MFMAs; otherwise two
// buffers and one barrier per chunk.
// Edge tiles: when the operands are 16-byte aligned, K is a multiple of BK, N of the vector
// width and both M and N are at least GB ("fast"), a tile that would cross the edge of C is
// moved back inside it (origin min(bm, M - GB), min(bn, N - GB)) and stores only the rows and
// columns at or beyond its own origin, which the tile before it does not store: every tile
// then takes unguarded vector loads.  Other shapes take element loads with bounds checks.
// TRACK_MAX: also the largest |C| written by each workgroup (NaN counts as large), stored
// as the bit pattern of a nonnegative double in vmax[workgroup] (no atomics: the slots are
// reduced by the consumer).
template <typename T, int BK, bool TRACK_MAX, int GB, bool FAST>
__device__ __forceinline__ void gemm_tile(int M, int N, int K, T alpha, const T* __restrict__ A, size_t lda,
                                          const T* __restrict__ B, size_t ldb, T beta, T* __restrict__ C, size_t ldc,
                                          const int* gate, int gp, unsigned long long* __restrict__ vmax) {
  typedef typename MfmaAcc<T>::type acc_t;
  if (gated_off(gate, gp)) return;
  // the 64-tile products are the per-panel products of the factorisation's critical path,
  // run beside trailing-update workgroups: their waves take issue priority
  if constexpr (GB == 64) __builtin_amdgcn_s_setprio(3);
  constexpr int MT = GB / 32;                             // 16 x 16 MFMA tiles per wave and dimension
  constexpr int VEC = 16 / sizeof(T);                     // elements per 16-byte load
  constexpr int A_LD = (GB * BK) / (256 * VEC);           // vector loads per thread for A
  constexpr int B_LD = (BK * GB) / (256 * VEC);
  constexpr int SA = BK + 16 / sizeof(T);                 // 36 words (float), 18 doubles
  constexpr int SB = GB + 64 / sizeof(T);                 // 144 / 80 words (float), 136 / 72 doubles
  constexpr int KS = BK / 4;                              // MFMA steps per chunk
  constexpr int SC = GB + 4;                              // row stride of the staged C tile
  constexpr int NBUF = (sizeof(T) == 4 && GB == 128) ? 1 : 2;
  constexpr size_t AB_BYTES = sizeof(T) * NBUF * (GB * SA + BK * SB);
  constexpr int CROWS = AB_BYTES >= sizeof(T) * SC * GB ? GB : AB_BYTES >= sizeof(T) * SC * (GB / 2) ? GB / 2 : GB / 4;
  constexpr int SEG = GB / VEC;                           // 16-byte segments per tile row
  constexpr int PER = CROWS * SEG / 256;                  // segments per thread per pass
  static_assert(CROWS * SC * sizeof(T) <= AB_BYTES, "staged C tile exceeds the LDS of the operands");
  static_assert(PER * 256 == CROWS * SEG, "staged rows must split evenly over the threads");
  __shared__ __attribute__((aligned(16))) char smem[AB_BYTES];
  T (*As)[GB][SA] = reinterpret_cast<T (*)[GB][SA]>(smem);
  T (*Bs)[BK][SB] = reinterpret_cast<T (*)[BK][SB]>(smem + sizeof(T) * NBUF * GB * SA);
  T* Cs = reinterpret_cast<T*>(smem);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * (GB / 2), wn = (wave & 1) * (GB / 2);
  const int ci = lane & 15, kq = lane >> 4;
  const int bm = FAST ? min((int)blockIdx.y * GB, M - GB) : (int)blockIdx.y * GB;   // tile origin
  const int bn = FAST ? min((int)blockIdx.x * GB, N - GB) : (int)blockIdx.x * GB;
  const int sm = (int)blockIdx.y * GB - bm, sn = (int)blockIdx.x * GB - bn;         // first row / column this tile stores

  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  vec_t ra[A_LD], rb[B_LD];
  auto fetch = [&](int k0) {
    if constexpr (FAST) {
#pragma unroll
      for (int i = 0; i < A_LD; i++) {
        const int e = (tid + i * 256) * VEC;   // element index in the GB x BK tile, k fastest
        ra[i] = *reinterpret_cast<const vec_t*>(A + (size_t)(bm + e / BK) * lda + k0 + e % BK);
      }
#pragma unroll
      for (int i = 0; i < B_LD; i++) {
        const int e = (tid + i * 256) * VEC;   // element index in the BK x GB tile, n fastest
        rb[i] = *reinterpret_cast<const vec_t*>(B + (size_t)(k0 + e / GB) * ldb + bn + e % GB);
      }
    } else {
#pragma unroll
      for (int i = 0; i < A_LD; i++) {
        const int e = (tid + i * 256) * VEC;
        const int gr = bm + e / BK, gk = k0 + e % BK;
#pragma unroll
        for (int v = 0; v < VEC; v++) ra[i][v] = (gr < M && gk + v < K) ? A[(size_t)gr * lda + gk + v] : T(0);
      }
#pragma unroll
      for (int i = 0; i < B_LD; i++) {
        const int e = (tid + i * 256) * VEC;
        const int gk = k0 + e / GB, gc = bn + e % GB;
#pragma unroll
        for (int v = 0; v < VEC; v++) rb[i][v] = (gk < K && gc + v < N) ? B[(size_t)gk * ldb + gc + v] : T(0);
      }
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_LD; i++) {
      const int e = (tid + i * 256) * VEC;
      *reinterpret_cast<vec_t*>(&As[buf][e / BK][e % BK]) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < B_LD; i++) {
      const int e = (tid + i * 256) * VEC;
      *reinterpret_cast<vec_t*>(&Bs[buf][e / GB][e % GB]) = rb[i];
    }
  };

  const int nk = (K + BK - 1) / BK;
  if (nk > 0) fetch(0);
  acc_t acc[MT][MT];
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < MT; b++) acc[a][b] = acc_t{0, 0, 0, 0};
  if (nk > 0) stash(0);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int cur = NBUF == 1 ? 0 : (kt & 1);
    if (kt + 1 < nk) fetch((kt + 1) * BK);
    T af[2][MT], bf[2][MT];
#pragma unroll
    for (int a = 0; a < MT; a++) af[0][a] = As[cur][wm + a * 16 + ci][kq];
#pragma unroll
    for (int b = 0; b < MT; b++) bf[0][b] = Bs[cur][kq][wn + b * 16 + ci];
#pragma unroll
    for (int s = 0; s < KS; s++) {
      const int p = s & 1;
      if (s + 1 < KS) {
#pragma unroll
        for (int a = 0; a < MT; a++) af[p ^ 1][a] = As[cur][wm + a * 16 + ci][4 * (s + 1) + kq];
#pragma unroll
        for (int b = 0; b < MT; b++) bf[p ^ 1][b] = Bs[cur][4 * (s + 1) + kq][wn + b * 16 + ci];
      }
#pragma unroll
      for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b < MT; b++) acc[a][b] = MfmaAcc<T>::mma(af[p][a], bf[p][b], acc[a][b]);
    }
    if (kt + 1 < nk) {
      if constexpr (NBUF == 1) __syncthreads();   // every wave is done reading the buffer
      stash(NBUF == 1 ? 0 : cur ^ 1);
    }
    __syncthreads();
  }

  // write back through LDS, CROWS tile rows per pass.  The C of a large product (128-tiles, the
  // trailing updates) is read once and written once: it streams past the caches (non-temporal),
  // which leaves L2 to the A and B panels that the other tiles read again
  const bool use_c = beta != T(0);
  constexpr int BATCH = PER < 8 ? PER : 8;
  constexpr bool STREAM_C = GB == 128 && !TRACK_MAX;
  double vm = 0.0;
  auto track = [&](T x) {
    if constexpr (TRACK_MAX) {
      const double a = (double)tabs(x);
      vm = (a > vm || a != a) ? a : vm;
    }
  };
#pragma unroll
  for (int pass = 0; pass < GB / CROWS; pass++) {
    const int r0 = pass * CROWS;
    if (pass > 0) __syncthreads();
#pragma unroll
    for (int a = 0; a < MT; a++) {
      if (wm + a * 16 < r0 || wm + a * 16 >= r0 + CROWS) continue;
#pragma unroll
      for (int b = 0; b < MT; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) Cs[(wm + a * 16 + MfmaAcc<T>::row(lane, v) - r0) * SC + wn + b * 16 + ci] = alpha * acc[a][b][v];
    }
    __syncthreads();
#pragma unroll
    for (int i0 = 0; i0 < PER; i0 += BATCH) {
      vec_t cvv[BATCH];
#pragma unroll
      for (int j = 0; j < BATCH; j++) {
        const int idx = tid + (i0 + j) * 256, row = r0 + idx / SEG, c = (idx % SEG) * VEC;
        if constexpr (FAST) {
          const vec_t* cp = reinterpret_cast<const vec_t*>(C + (size_t)(bm + row) * ldc + bn + c);
          if constexpr (STREAM_C)
            cvv[j] = use_c ? __builtin_nontemporal_load(cp) : vec_t{};
          else
            cvv[j] = use_c ? *cp : vec_t{};
        } else {
#pragma unroll
          for (int v = 0; v < VEC; v++)
            cvv[j][v] = (use_c && bm + row < M && bn + c + v < N) ? C[(size_t)(bm + row) * ldc + bn + c + v] : T(0);
        }
      }
#pragma unroll
      for (int j = 0; j < BATCH; j++) {
        const int idx = tid + (i0 + j) * 256, row = r0 + idx / SEG, c = (idx % SEG) * VEC;
        const vec_t val = *reinterpret_cast<const vec_t*>(&Cs[(row - r0) * SC + c]);
        vec_t out;
#pragma unroll
        for (int v = 0; v < VEC; v++) out[v] = use_c ? fma(beta, cvv[j][v], val[v]) : val[v];
        if constexpr (FAST) {
          if (row >= sm && c >= sn) {
#pragma unroll
            for (int v = 0; v < VEC; v++) track(out[v]);
            vec_t* cp = reinterpret_cast<vec_t*>(C + (size_t)(bm + row) * ldc + bn + c);
            if constexpr (STREAM_C)
              __builtin_nontemporal_store(out, cp);
            else
              *cp = out;
          }
        } else {
#pragma unroll
          for (int v = 0; v < VEC; v++) {
            if (bm + row < M && bn + c + v < N) {
              track(out[v]);
              C[(size_t)(bm + row) * ldc + bn + c + v] = out[v];
            }
          }
        }
      }
    }
  }
  if constexpr (TRACK_MAX) {
    for (int o = 32; o >= 1; o >>= 1) {
      const double x = __shfl_xor(vm, o, 64);
      vm = (x > vm || x != x) ? x : vm;
    }
    __syncthreads();   // Cs is reused for the four wave maxima
    double* wmx = reinterpret_cast<double*>(smem);
    if (lane == 0) wmx[wave] = vm;
    __syncthreads();
    if (tid == 0) {
      double mm = 0.0;
      for (int w = 0; w < 4; w++) mm = (wmx[w] > mm || wmx[w] != wmx[w]) ? wmx[w] : mm;
      if (mm != mm) mm = 1e300;
      vmax[blockIdx.y * gridDim.x + blockIdx.x] = (unsigned long long)__double_as_longlong(mm);
    }
  }
}

template <typename T, int BK, bool TRACK_MAX, int GB, bool FAST>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, T alpha, const T* __restrict__ A, size_t lda,
                                              const T* __restrict__ B, size_t ldb, T beta, T* __restrict__ C, size_t ldc,
                                              const int* gate, int gp, unsigned long long* __restrict__ vmax) {
  gemm_tile<T, BK, TRACK_MAX, GB, FAST>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, gp, vmax);
}
// U12 = inv(L11) A12 on 64-tiles, with the panel's no-pivot test in front (what
// k_panel_check does, without a launch of its own): every workgroup reduces the max |l|
// of the diagonal block (*lmax) and of the L21 product's workgroups (slots); above 1 (or
// NaN) the panel fails: workgroup (0, 0) records it in first_fail and no workgroup writes.
// All workgroups reach the same verdict from the same values.
template <typename T, bool FAST>
__global__ __launch_bounds__(256) void k_gemm_checked(int M, int N, int K, T alpha, const T* __restrict__ A, size_t lda,
                                                      const T* __restrict__ B, size_t ldb, T beta, T* __restrict__ C, size_t ldc,
                                                      int* first_fail, int gp, const unsigned long long* __restrict__ lmax,
                                                      const unsigned long long* __restrict__ slots, int nslot) {
  __shared__ unsigned long long red[256];
  if (gated_off(first_fail, gp)) return;
  unsigned long long m = threadIdx.x == 0 ? *lmax : 0ull;
  for (int i = threadIdx.x; i < nslot; i += 256) m = slots[i] > m ? slots[i] : m;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o && red[threadIdx.x + o] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + o];
    __syncthreads();
  }
  const bool fail = red[0] > 0x3FF0000000000000ull;
  __syncthreads();   // red is done with before the tile reuses nothing of it; keeps the verdict uniform
  if (fail) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0) atomicMin(first_fail, gp);
    return;
  }
  gemm_tile<T, sizeof(T) == 8 ? 16 : 32, false, 64, FAST>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, 0, nullptr);
}

// FP32 128-tiles: with this register budget the compiler keeps the accumulators in
// architectural VGPRs and allocates 144 registers per lane in all, so three workgroups fit a
// SIMD's 512 and, when one of them ends, the diagonal-block kernel of the next panel (one
// wave of 144 registers per SIMD) fits beside the other two
template <bool TRACK_MAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_num_vgpr(72))) void k_gemm_f32(
    int M, int N, int K, float alpha, const float* __restrict__ A, size_t lda, const float* __restrict__ B, size_t ldb, float beta,
    float* __restrict__ C, size_t ldc, const int* gate, int gp, unsigned long long* __restrict__ vmax) {
  gemm_tile<float, 32, TRACK_MAX, 128, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, gp, vmax);
}

// ---------------------------------------------------------------------------------
// FP32 GEMM on BF16 MFMA (the large products of the FP32 factorisation):
// C = beta*C + alpha*A*B with every FP32 operand split into three BF16 parts,
// x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1), round to
// nearest even, 24 significant bits in all), and a*b formed from the six products whose
// parts' orders sum to at most 2 (the three dropped ones are below 2^-24 |a b|).  BF16
// products are exact in FP32 and v_mfma_f32_16x16x32_bf16 accumulates in FP32, so the result
// has the accuracy of an FP32 GEMM (measured: 6e-7 relative against an FP64 product of
// 1024 x 256 x 1536, the FP32 MFMA kernel 8e-7) at 6 x 16 cycles per 16 x 16 x 32 block
// instead of 8 x 32 for v_mfma_f32_16x16x4f32.
// Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16], i = 0..7; the accumulators are those of v_mfma_f32_16x16x4f32.
// Tile 128 x 128, 512 threads (4 x 2 waves of 32 x 64), K chunks of 32.  The split happens
// when a chunk moves from registers to LDS: A as [part][row][k], B as [part][col][k], 64-byte
// rows whose four 16-byte segments are XOR-swizzled by (row / 2) % 4: any 8 consecutive rows
// then put a given 16-byte segment in 8 distinct 16-byte bank groups, so the fragment reads
// (lanes = 16 consecutive rows at one k group) and the B stores (lanes = consecutive columns,
// 16 bytes of 8 consecutive k each) are free of bank conflicts, and an A row's 64 bytes are
// written by 8 consecutive lanes.  48 KiB of LDS: two workgroups (16 waves) per CU, so one
// wave's split and stores overlap another's MFMAs.
// Fast shapes only (see gemm_tile): A rows and C rows 16-byte aligned, K % 32 == 0,
// N % 4 == 0, M, N >= 128; B is read element-wise (its rows need no alignment).
// ---------------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split_bf16x3(float x, __bf16& p0, __bf16& p1, __bf16& p2) {
  p0 = (__bf16)x;
  const float r1 = x - (float)p0;
  p1 = (__bf16)r1;
  p2 = (__bf16)(r1 - (float)p1);
}
// The same split for two values at once: one packed conversion per part (v_cvt_pk_bf16_f32,
// round to nearest even), the parts' values recovered exactly from the packed bits (BF16 is
// the upper half of an FP32).  p0/p1/p2 hold the parts of a in bits 0-15 and of b in 16-31;
// the arithmetic is that of split_bf16x3.
__device__ __forceinline__ void split_pair_bf16x3(float a, float b, unsigned& p0, unsigned& p1, unsigned& p2) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  typedef __bf16 h2v __attribute__((ext_vector_type(2)));
  const h2v v0 = __builtin_convertvector((f2v{a, b}), h2v);
  p0 = __builtin_bit_cast(unsigned, v0);
  const float ra = a - __uint_as_float(p0 << 16), rb = b - __uint_as_float(p0 & 0xffff0000u);
  const h2v v1 = __builtin_convertvector((f2v{ra, rb}), h2v);
  p1 = __builtin_bit_cast(unsigned, v1);
  const float sa = ra - __uint_as_float(p1 << 16), sb = rb - __uint_as_float(p1 & 0xffff0000u);
  const h2v v2 = __builtin_convertvector((f2v{sa, sb}), h2v);
  p2 = __builtin_bit_cast(unsigned, v2);
}
// byte offset of (row, k) in one 128 x 32 BF16 part
__device__ __forceinline__ int x6_off(int row, int k) { return row * 64 + 16 * ((k >> 3) ^ ((row >> 1) & 3)) + 2 * (k & 7); }

__global__ __launch_bounds__(512, 2) void k_gemm_x6(int M, int N, int K, float alpha, const float* __restrict__ A, size_t lda,
                                                    const float* __restrict__ B, size_t ldb, float beta, float* __restrict__ C,
                                                    size_t ldc, const int* gate, int gp, int band) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  constexpr int GB = 128, BK = 32, MA = 2, MB = 4, PART = GB * BK * 2, NT = 512;
  constexpr int SC = GB + 4, CROWS = 64, SEG = GB / 4, PER = CROWS * SEG / NT;
  if (gated_off(gate, gp)) return;
  __shared__ __attribute__((aligned(16))) char smem[6 * PART];   // A parts 0-2, B parts 0-2; later the staged C rows
  char* As = smem;
  char* Bs = smem + 3 * PART;
  float* Cs = reinterpret_cast<float*>(smem);
  static_assert(CROWS * SC * 4 <= 6 * PART, "staged C rows exceed the LDS of the operands");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;   // 8 waves in a 4 x 2 grid of 32 x 64 sub-tiles
  const int li = lane & 15, lq = lane >> 4;
  int ty = blockIdx.y, tx = blockIdx.x;
  if (band > 0) {
    // XCD-aware order: workgroup w runs on XCD w % 8 as its (w / 8)-th; each XCD takes a
    // contiguous range of tiles, walked down bands of `band` tile rows column by column, so
    // that the tiles an XCD holds at once share their A rows and B columns in its L2
    const int gx = gridDim.x, T = gx * gridDim.y, w = ty * gx + tx, x = w & 7, per = T >> 3, rem = T & 7;
    const int t = x * per + min(x, rem) + (w >> 3);
    const int g = t / (band * gx), rows = min(band, (int)gridDim.y - g * band), i = t - g * band * gx;
    ty = g * band + i % rows;
    tx = i / rows;
  }
  const int bm = min(ty * GB, M - GB), bn = min(tx * GB, N - GB);   // tile origin (inside C)
  const int sm = ty * GB - bm, sn = tx * GB - bn;                  // first row / column stored
  f4v acc[MA][MB];
#pragma unroll
  for (int a = 0; a < MA; a++)
#pragma unroll
    for (int b = 0; b < MB; b++) acc[a][b] = f4v{0, 0, 0, 0};
  // per chunk, thread: A elements (tid + 512 i) * 4 .. + 3 of the 128 x 32 tile (k fastest);
  // B column bn + (tid % 128), k = 8 (tid / 128) + 0..7
  f4v ra[2];
  float rb[8];
  const int bc = tid & 127, bkg = tid >> 7;
  // this thread's B column at k = 8 bkg; a chunk's rows follow at a stride of ldb
  const float* const bcol = B + (size_t)(8 * bkg) * ldb + bn + bc;
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int e = (tid + i * NT) * 4;
      ra[i] = *reinterpret_cast<const f4v*>(A + (size_t)(bm + e / BK) * lda + k0 + e % BK);
    }
    const float* q = bcol + (size_t)k0 * ldb;
#pragma unroll
    for (int j = 0; j < 8; j++, q += ldb) rb[j] = *q;
  };
  typedef unsigned u2v __attribute__((ext_vector_type(2)));
  typedef unsigned u4v __attribute__((ext_vector_type(4)));
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int e = (tid + i * NT) * 4;
      unsigned p0[2], p1[2], p2[2];
      split_pair_bf16x3(ra[i][0], ra[i][1], p0[0], p1[0], p2[0]);
      split_pair_bf16x3(ra[i][2], ra[i][3], p0[1], p1[1], p2[1]);
      const int off = x6_off(e / BK, e % BK);
      *reinterpret_cast<u2v*>(As + off) = u2v{p0[0], p0[1]};
      *reinterpret_cast<u2v*>(As + PART + off) = u2v{p1[0], p1[1]};
      *reinterpret_cast<u2v*>(As + 2 * PART + off) = u2v{p2[0], p2[1]};
    }
    {
      unsigned p0[4], p1[4], p2[4];
#pragma unroll
      for (int v = 0; v < 4; v++) split_pair_bf16x3(rb[2 * v], rb[2 * v + 1], p0[v], p1[v], p2[v]);
      const int off = x6_off(bc, 8 * bkg);
      *reinterpret_cast<u4v*>(Bs + off) = u4v{p0[0], p0[1], p0[2], p0[3]};
      *reinterpret_cast<u4v*>(Bs + PART + off) = u4v{p1[0], p1[1], p1[2], p1[3]};
      *reinterpret_cast<u4v*>(Bs + 2 * PART + off) = u4v{p2[0], p2[1], p2[2], p2[3]};
    }
  };
  const int nk = K / BK;
  if (nk > 0) {
    fetch(0);
    stash();
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    if (kt + 1 < nk) fetch((kt + 1) * BK);
    bf16x8_t af[MA][3];
#pragma unroll
    for (int a = 0; a < MA; a++)
#pragma unroll
      for (int p = 0; p < 3; p++) af[a][p] = *reinterpret_cast<const bf16x8_t*>(As + p * PART + x6_off(wm + a * 16 + li, 8 * lq));
#pragma unroll
    for (int b = 0; b < MB; b++) {
      bf16x8_t bq[3];
#pragma unroll
      for (int p = 0; p < 3; p++) bq[p] = *reinterpret_cast<const bf16x8_t*>(Bs + p * PART + x6_off(wn + b * 16 + li, 8 * lq));
#pragma unroll
      for (int a = 0; a < MA; a++) {   // smallest products first
        f4v c = acc[a][b];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][2], bq[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][1], bq[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][0], bq[2], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][1], bq[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][0], bq[1], c, 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][0], bq[0], c, 0, 0, 0);
      }
    }
    if (kt + 1 < nk) {
      __syncthreads();   // every wave is done reading the chunk
      stash();
    }
    __syncthreads();
  }
  // write back through LDS, 64 tile rows per pass, 4 segments of C in flight per thread.  C is
  // read once and written once: it streams past the caches (non-temporal), which leaves L2 to
  // the A and B panels that the other tiles of the launch read again
  const bool use_c = beta != 0.0f;
#pragma unroll
  for (int pass = 0; pass < GB / CROWS; pass++) {
    const int r0 = pass * CROWS;
    if (pass > 0) __syncthreads();
#pragma unroll
    for (int a = 0; a < MA; a++) {
      if (wm + a * 16 < r0 || wm + a * 16 >= r0 + CROWS) continue;
#pragma unroll
      for (int b = 0; b < MB; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) Cs[(wm + a * 16 + 4 * lq + v - r0) * SC + wn + b * 16 + li] = alpha * acc[a][b][v];
    }
    __syncthreads();
    f4v cv[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int idx = tid + j * NT, row = r0 + idx / SEG, c = (idx % SEG) * 4;
      cv[j] = use_c ? __builtin_nontemporal_load(reinterpret_cast<const f4v*>(C + (size_t)(bm + row) * ldc + bn + c)) : f4v{0, 0, 0, 0};
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int idx = tid + j * NT, row = r0 + idx / SEG, c = (idx % SEG) * 4;
      const f4v val = *reinterpret_cast<const f4v*>(&Cs[(row - r0) * SC + c]);
      f4v out;
#pragma unroll
      for (int v = 0; v < 4; v++) out[v] = use_c ? fmaf(beta, cv[j][v], val[v]) : val[v];
      if (row >= sm && c >= sn) __builtin_nontemporal_store(out, reinterpret_cast<f4v*>(C + (size_t)(bm + row) * ldc + bn + c));
    }
  }
}

// copy a rows x cols block (row-major, separate leading dims)
template <typename T>
__global__ __launch_bounds__(256) void k_copy2d(const T* __restrict__ S, size_t lds, T* __restrict__ D, size_t ldd, int rows, int cols,
                                                const int* gate, int gp) {
  if (gated_off(gate, gp)) return;
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)rows * cols) return;
  int r = (int)(e / cols), c = (int)(e % cols);
  D[(size_t)r * ldd + c] = S[(size_t)r * lds + c];
}

// the factors of a panel into the matrix: L21 (rows x nb, ld nb) and U12 (nb x cols, ld cols)
// in one launch
template <typename T>
__global__ __launch_bounds__(256) void k_copy_panel(const T* __restrict__ L, size_t ldl, T* __restrict__ A21, const T* __restrict__ U,
                                                    size_t ldu, T* __restrict__ A12, size_t lda, int rows, int cols, int nb,
                                                    const int* gate, int gp) {
  if (gated_off(gate, gp)) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nl = (size_t)rows * nb;
  if (e < nl) {
    const int r = (int)(e / nb), c = (int)(e % nb);
    A21[(size_t)r * lda + c] = L[(size_t)r * ldl + c];
  } else if (e < nl + (size_t)nb * cols) {
    const size_t f = e - nl;
    const int r = (int)(f / cols), c = (int)(f % cols);
    A12[(size_t)r * lda + c] = U[(size_t)r * ldu + c];
  }
}

__global__ __launch_bounds__(256) void k_d2f(const double* __restrict__ S, size_t lds, float* __restrict__ D, size_t ldd, int rows,
                                             int cols) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)rows * cols) return;
  int r = (int)(e / cols), c = (int)(e % cols);
  D[(size_t)r * ldd + c] = (float)S[(size_t)r * lds + c];
}

// ---------------------------------------------------------------------------------
// Solves with the factors: x = U^{-1} L^{-1} P b for nrhs right-hand sides stored as
// (nrhs, N) rows (the reference's F layout).  Vectors in FP64; factor entries are
// promoted on load (an FP32 factorisation is applied in FP64 arithmetic).
//
// One kernel per triangle.  Workgroup g claims block row b from a counter (in order, so
// every block it waits for belongs to a workgroup that is already running), accumulates
// L[b, j] x_j for the solved blocks j as their completion flags appear, then applies the
// inverse of the diagonal block stored by the factorisation.  Waits are bounded: past the
// spin limit the kernel records an error and finishes.
// ---------------------------------------------------------------------------------

// permutation of the row interchanges: perm[i] = original row that ends at position i
__global__ __launch_bounds__(256) void k_perm(const int* __restrict__ ipiv, int N, int* __restrict__ perm) {
  extern __shared__ int pl[];
  for (int i = threadIdx.x; i < N; i += blockDim.x) pl[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < N; d++) {
      int p = ipiv[d];
      if (p != d && p >= 0 && p < N) { int t = pl[d]; pl[d] = pl[p]; pl[p] = t; }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) perm[i] = pl[i];
}

constexpr long long SPIN_LIMIT = 1ll << 22;   // polls of a flag before a wait gives up (err set)

__global__ void k_set_int(int* p, int v) { *p = v; }

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// max |x| over n values as the bit pattern of a nonnegative double (NaN as the largest
// pattern, so a caller sees a non-finite maximum); *out zeroed by the launcher's first kernel
__global__ __launch_bounds__(256) void k_absmax(const double* __restrict__ x, size_t n, unsigned long long* __restrict__ out) {
  unsigned long long m = 0ull;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double a = fabs(x[i]);
    const unsigned long long b = a != a ? 0x7FFFFFFFFFFFFFFFull : (unsigned long long)__double_as_longlong(a);
    m = b > m ? b : m;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned long long y = __shfl_xor(m, o, 64);
    m = y > m ? y : m;
  }
  if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst) { *dst = *src; }

// panel p of the speculative pass: max |l| of the diagonal block (*lmax) and of the
// workgroups of the L21 product (slots[0 .. nslot)); record p as failed when some |l| > 1
// (or NaN).  Bit patterns of nonnegative doubles order as integers.
__global__ __launch_bounds__(256) void k_panel_check(unsigned long long* __restrict__ lmax,
                                                     const unsigned long long* __restrict__ slots, int nslot,
                                                     int* __restrict__ first_fail, int p) {
  __shared__ unsigned long long red[256];
  if (*first_fail <= p) return;
  unsigned long long m = threadIdx.x == 0 ? *lmax : 0ull;
  for (int i = threadIdx.x; i < nslot; i += 256) m = slots[i] > m ? slots[i] : m;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o && red[threadIdx.x + o] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *lmax = red[0];
    if (red[0] > 0x3FF0000000000000ull) atomicMin(first_fail, p);
  }
}

// Blocked triangular solve with the stored diagonal-block inverses, one workgroup of 512
// threads per block row of 128.  Workgroups take block rows in dependency order from a
// ticket counter (so a waiting workgroup never blocks one that has not started) and walk
// their dependencies in solve order: forward for L, backward for U, in chunks of 1 KB of
// row (256 columns in FP32, 128 in FP64).  Per chunk, wave w loads its rows 16w..16w+15
// with one 16-byte load per lane and row (branch-free: edge lanes load a clamped address
// and discard it), then waits for the solution values of the chunk's columns.
//
// Completion is carried by the values themselves: the output vector is first filled with
// TRSV_PENDING (a signalling-NaN bit pattern that no arithmetic produces), solutions are
// written with relaxed agent-scope atomic stores, and readers poll them with relaxed
// agent-scope atomic loads until no lane of the wave still sees the marker.  Each value is
// its own flag, so no release/acquire fence (an L2 writeback / invalidate on this
// multi-die part) sits on the critical path.  A wait past SPIN_LIMIT polls records the
// error and writes NaN for the block, which propagates to the residual.
// The block's inverse is staged in LDS while the dependencies are summed; the row sums
// are reduced through LDS; x_b = inv(D_b) (b_b - sum) is an LDS matvec with four threads
// per row.
constexpr int TRSV_PAD = LU_NB + 1;   // LDS row stride of the staged inverse
constexpr int TRSV_RED = 16;          // partial sums per row left after two shuffle steps
constexpr int TRSV_THREADS = 512;
constexpr unsigned long long TRSV_PENDING = 0x7FF4A5A5C3C3E1E1ull;
template <typename T>
constexpr size_t trsv_lds_bytes() {
  return sizeof(T) * LU_NB * TRSV_PAD + sizeof(double) * LU_NB * (TRSV_RED + 1) + sizeof(double) * LU_NB;
}

// Coupling blocks of the blocked solves, after the factorisation: for every block row b with
// a neighbour in the solve order, M_b = inv(L_bb) L_{b,b-1} (forward, L) or inv(U_bb) U_{b,b+1}
// (backward, U), NB x NB row-major, zero outside the blocks' extents.  Block b's solution is
// then x_b = inv(D_b) (b_b - sum over the other blocks) - M_b x_neighbour, so the solve of block
// b can finish everything but one matrix-vector product before its neighbour is known.
// One workgroup per block; 8 x 8 output tiles per thread, K staged through LDS in chunks of 32;
// FP64 accumulation, stored as T.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_coupling(const T* __restrict__ LU, size_t lda, int N, const T* __restrict__ Dinv,
                                                       T* __restrict__ Mout) {
  constexpr int NB = LU_NB, KC = 32;
  const int nblk = (N + NB - 1) / NB, b = blockIdx.x;
  const int nbr = LOWER ? b - 1 : b + 1;   // the neighbour block
  if (nbr < 0 || nbr >= nblk) return;
  const int r0 = b * NB, nr = min(NB, N - r0);
  const int c0 = nbr * NB, nc = min(NB, N - c0);
  __shared__ T Ds[NB][KC + 1];   // inv(D_b)[:, k0:k0+KC]
  __shared__ T Bs[KC][NB + 1];   // coupling block rows k0:k0+KC
  const int tid = threadIdx.x, ti = tid / 16, tj = tid % 16;   // rows 8 ti.., columns 8 tj..
  double acc[8][8] = {};
  const T* D = Dinv + (size_t)b * NB * NB;
  for (int k0 = 0; k0 < NB; k0 += KC) {
    for (int e = tid; e < NB * KC; e += 256) {
      const int i = e / KC, k = e % KC;
      Ds[i][k] = D[(size_t)i * NB + k0 + k];
      const int kk = e / NB, j = e % NB;
      Bs[kk][j] = (k0 + kk < nr && j < nc) ? LU[(size_t)(r0 + k0 + kk) * lda + c0 + j] : T(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < KC; k++) {
      double dv[8], bv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) { dv[u] = (double)Ds[8 * ti + u][k]; bv[u] = (double)Bs[k][8 * tj + u]; }
#pragma unroll
      for (int u = 0; u < 8; u++)
#pragma unroll
        for (int v = 0; v < 8; v++) acc[u][v] = fma(dv[u], bv[v], acc[u][v]);
    }
    __syncthreads();
  }
  T* M = Mout + (size_t)b * NB * NB;
#pragma unroll
  for (int u = 0; u < 8; u++)
#pragma unroll
    for (int v = 0; v < 8; v++) {
      const int i = 8 * ti + u, j = 8 * tj + v;
      M[(size_t)i * NB + j] = (i < nr && j < nc) ? (T)acc[u][v] : T(0);
    }
}

// out[0..n) = TRSV_PENDING, *counter = 0 (a kernel, not a memset node, so that the solve
// can be captured in a graph)
__global__ __launch_bounds__(256) void k_trsv_prep(double* __restrict__ out, int n, int* __restrict__ counter) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = __longlong_as_double((long long)TRSV_PENDING);
  if (i == 0) *counter = 0;
}

__device__ __forceinline__ bool trsv_pending(double v) {
  return (unsigned long long)__double_as_longlong(v) == TRSV_PENDING;
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(TRSV_THREADS) void k_trsv_blocked(const T* __restrict__ LU, size_t lda, int N,
                                                               const int* __restrict__ perm, const int* __restrict__ pivoted,
                                                               const double* __restrict__ Bv, double* X,
                                                               const T* __restrict__ Dinv, const T* __restrict__ Mc,
                                                               int* __restrict__ counter, int* __restrict__ err) {
  constexpr int NB = LU_NB;
  constexpr int NW = TRSV_THREADS / 64;   // waves
  constexpr int RPW = NB / NW;            // rows per wave (16)
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int CH = 64 * VEC;            // columns per chunk
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Ds = reinterpret_cast<T*>(smem_raw);                                        // NB x TRSV_PAD
  double* part = reinterpret_cast<double*>(smem_raw + sizeof(T) * NB * TRSV_PAD);   // NB x (TRSV_RED + 1)
  double* vs = part + NB * (TRSV_RED + 1);                                         // NB
  __shared__ int s_b, s_fail;
  __shared__ double s_xn[NW][NB];   // the neighbour block's solution, one copy per wave
  const int nblk = (N + NB - 1) / NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) {
    const int t = atomicAdd(counter, 1);
    s_b = LOWER ? t : nblk - 1 - t;
    s_fail = 0;
  }
  __syncthreads();
  const int b = s_b;
  if (b < 0 || b >= nblk) return;
  const int r0 = b * NB;
  const int nr = min(NB, N - r0);
  // stage inv(D_b) (independent of every other block)
  const T* Db = Dinv + (size_t)b * NB * NB;
  for (int e = tid; e < NB * NB; e += TRSV_THREADS) Ds[(e / NB) * TRSV_PAD + (e % NB)] = Db[e];
  double acc[RPW];
#pragma unroll
  for (int i = 0; i < RPW; i++) acc[i] = 0.0;
  // this wave's rows (clamped into the matrix; rows past N are discarded at the end)
  const T* rowp[RPW];
#pragma unroll
  for (int i = 0; i < RPW; i++) rowp[i] = LU + (size_t)min(r0 + wave * RPW + i, N - 1) * lda;
  const bool vec_ok = (lda % VEC) == 0 && (((uintptr_t)LU) % 16) == 0;
  // the neighbour block in solve order (b - 1 for L, b + 1 for U) enters last, through the
  // coupling block M_b: this thread's share of M_b (row 16w + l%16, columns 32q..32q+31) is
  // loaded now
  const int nbr = LOWER ? b - 1 : b + 1;
  const bool has_nbr = nbr >= 0 && nbr < nblk;
  const int c0n = nbr * NB, ncn = has_nbr ? min(NB, N - c0n) : 0;
  const int mrow = wave * RPW + (lane & 15), mq = lane >> 4;
  vec_t mv[32 / VEC];
  if (has_nbr) {
    const T* mp = Mc + (size_t)b * NB * NB + (size_t)mrow * NB + 32 * mq;
#pragma unroll
    for (int i = 0; i < 32 / VEC; i++) mv[i] = *reinterpret_cast<const vec_t*>(mp + i * VEC);
  }
  // the other dependency columns [lo, hi), chunks in solve order
  const int lo = LOWER ? 0 : (has_nbr ? c0n + ncn : r0 + nr), hi = LOWER ? (has_nbr ? c0n : r0) : N;
  const int nch = (hi - lo + CH - 1) / CH;
  const int top = LOWER ? 0 : nblk * NB;   // U chunks are aligned to the block grid from its end
  bool failed = false;
  for (int k = 0; k < nch; k++) {
    const int cs = LOWER ? top + k * CH : top - (k + 1) * CH;   // nominal chunk start
    const int c = cs + lane * VEC;                               // this lane's first column
    vec_t l[RPW];
    if (vec_ok && cs >= lo && cs + CH <= hi) {
#pragma unroll
      for (int i = 0; i < RPW; i++) l[i] = *reinterpret_cast<const vec_t*>(rowp[i] + c);
    } else {
#pragma unroll
      for (int i = 0; i < RPW; i++)
#pragma unroll
        for (int v = 0; v < VEC; v++) {
          const bool in = c + v >= lo && c + v < hi;
          const T t = rowp[i][in ? c + v : lo];
          l[i][v] = in ? t : T(0);
        }
    }
    double x[VEC];
#pragma unroll
    for (int v = 0; v < VEC; v++) x[v] = 0.0;
    long long spins = 0;
    while (!failed) {   // after one wait has timed out, the rest are skipped
      bool pend = false;
#pragma unroll
      for (int v = 0; v < VEC; v++) {
        if (c + v >= lo && c + v < hi) {
          x[v] = __hip_atomic_load(X + c + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          pend = pend || trsv_pending(x[v]);
        }
      }
      if (!__any(pend)) break;
      if (++spins > SPIN_LIMIT) {
        failed = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (failed) {
#pragma unroll
      for (int v = 0; v < VEC; v++) x[v] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < RPW; i++)
#pragma unroll
      for (int v = 0; v < VEC; v++) acc[i] = fma((double)l[i][v], x[v], acc[i]);
  }
  if (failed && lane == 0) {
    atomicOr(err, 1);
    s_fail = 1;
  }
  // row sums: two shuffle steps leave 16 partials per row, summed from LDS in fixed order
#pragma unroll
  for (int i = 0; i < RPW; i++) {
    double v = acc[i];
    v += __shfl_xor(v, 32, 64);
    v += __shfl_xor(v, 16, 64);
    if (lane < TRSV_RED) part[(wave * RPW + i) * (TRSV_RED + 1) + lane] = v;
  }
  __syncthreads();
  if (tid < NB) {
    double ssum = 0.0;
#pragma unroll
    for (int q = 0; q < TRSV_RED; q++) ssum += part[tid * (TRSV_RED + 1) + q];
    const int row = r0 + tid;
    double bv = 0.0;
    if (row < N) {
      const bool use_perm = LOWER && perm && pivoted && *pivoted;
      bv = Bv[use_perm ? perm[row] : row];
    }
    vs[tid] = row < N ? bv - ssum : 0.0;
  }
  __syncthreads();
  // y = inv(D_b) vs: row 16w + l%16, quarter q = l/16 takes columns [16q, 16q+16) and
  // [64+16q, 80+16q) (the 64 lanes of a step then read 64 distinct banks)
  {
    const int row = mrow, q = mq;
    const T* Dr = Ds + row * TRSV_PAD;
    double xv = 0.0;
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int cc = 0; cc < 16; cc++) {
        const int col = h * 64 + q * 16 + cc;
        xv = fma((double)Dr[col], vs[col], xv);
      }
    xv += __shfl_xor(xv, 16, 64);
    xv += __shfl_xor(xv, 32, 64);
    bool wfail = s_fail != 0;
    if (has_nbr) {
      // x_b = y - M_b x_nbr: each lane polls two of the neighbour's values, the wave shares
      // them through its own LDS row (no workgroup barrier on this path)
      const int ca = c0n + 2 * lane;
      double xa = 0.0, xb = 0.0;
      long long spins = 0;
      while (!wfail) {
        bool pend = false;
        if (ca < c0n + ncn) {
          xa = __hip_atomic_load(X + ca, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          pend = trsv_pending(xa);
        }
        if (ca + 1 < c0n + ncn) {
          xb = __hip_atomic_load(X + ca + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          pend = pend || trsv_pending(xb);
        }
        if (!__any(pend)) break;
        if (++spins > SPIN_LIMIT) {
          wfail = true;
          if (lane == 0) atomicOr(err, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      double* xs = s_xn[wave];
      xs[2 * lane] = ca < c0n + ncn ? xa : 0.0;
      xs[2 * lane + 1] = ca + 1 < c0n + ncn ? xb : 0.0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double z = 0.0;
#pragma unroll
      for (int i = 0; i < 32 / VEC; i++)
#pragma unroll
        for (int v = 0; v < VEC; v++) z = fma((double)mv[i][v], xs[32 * q + i * VEC + v], z);
      z += __shfl_xor(z, 16, 64);
      z += __shfl_xor(z, 32, 64);
      xv -= z;
    }
    if (q == 0 && row < nr) {
      if (wfail) xv = __builtin_nan("");
      __hip_atomic_store(X + r0 + row, xv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// residual r = b - A x with the FP64 matrix, accumulated in double-double: each product
// is split exactly with an FMA (a*x = p + e) and summed with TwoSum, so the residual of
// an already accurate x is itself accurate and refinement converges to the correctly
// rounded solution of the assembled system rather than stalling at cond * eps.
// rmax[r] = bit pattern of max |r| (nonnegative doubles order as integers).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dd_add(double& hi, double& lo, double a, double alo) {
  double s, e;
  two_sum(hi, a, s, e);
  e += lo + alo;
  hi = s + e;
  lo = e - (hi - s);
}

__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ A, size_t lda, int N, const double* __restrict__ B,
                                                  const double* __restrict__ X, double* __restrict__ Rv, int nrhs,
                                                  unsigned long long* __restrict__ rmax) {
  // one wave per row, four independent double-double partial sums per lane; one atomic per
  // workgroup and right-hand side for the maximum.  Rows that start 16-byte aligned are read
  // as pairs of doubles, eight 16-byte loads in flight per lane (8 KB per wave, enough to keep
  // HBM busy with the waves a CU holds); other rows element by element, four in flight.
  __shared__ double wm[4];
  typedef double d2v __attribute__((ext_vector_type(2)));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  const double* a = A + (size_t)(row < N ? row : 0) * lda;
  for (int r = 0; r < nrhs; r++) {
    const double* x = X + (size_t)r * N;
    double av = 0.0;
    if (row < N) {
      double hi[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
      int s = lane;
      if ((((uintptr_t)a | (uintptr_t)x) & 15) == 0) {
        // pair p of the row (elements 2p, 2p+1): lane takes pairs lane, lane + 64, ...
        int pb = 0;
        for (; pb + 8 * 64 <= N / 2; pb += 8 * 64) {
          d2v pa[8], px[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            pa[u] = *reinterpret_cast<const d2v*>(a + 2 * (pb + 64 * u + lane));
            px[u] = *reinterpret_cast<const d2v*>(x + 2 * (pb + 64 * u + lane));
          }
#pragma unroll
          for (int u = 0; u < 8; u++)
#pragma unroll
            for (int v = 0; v < 2; v++) {
              const double pr = pa[u][v] * px[u][v];
              dd_add(hi[(2 * u + v) & 3], lo[(2 * u + v) & 3], pr, fma(pa[u][v], px[u][v], -pr));
            }
        }
        s = 2 * pb + lane;   // the rest element by element
      }
      for (; s + 192 < N; s += 256) {
        double pa[4], px[4];
#pragma unroll
        for (int u = 0; u < 4; u++) { pa[u] = a[s + 64 * u]; px[u] = x[s + 64 * u]; }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const double pr = pa[u] * px[u];
          dd_add(hi[u], lo[u], pr, fma(pa[u], px[u], -pr));
        }
      }
      for (; s < N; s += 64) {
        const double pr = a[s] * x[s];
        dd_add(hi[0], lo[0], pr, fma(a[s], x[s], -pr));
      }
#pragma unroll
      for (int u = 1; u < 4; u++) dd_add(hi[0], lo[0], hi[u], lo[u]);
      double h = hi[0], l = lo[0];
      for (int o = 32; o >= 1; o >>= 1) {
        const double h2 = __shfl_xor(h, o, 64), l2 = __shfl_xor(l, o, 64);
        dd_add(h, l, h2, l2);
      }
      if (lane == 0) {
        // b - (hi + lo)
        double sv, e;
        two_sum(B[(size_t)r * N + row], -h, sv, e);
        const double v = sv + (e - l);
        Rv[(size_t)r * N + row] = v;
        av = v < 0 ? -v : v;
      }
    }
    if (lane == 0) wm[wave] = av;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = wm[0];
      for (int w = 1; w < 4; w++) m = (wm[w] > m || wm[w] != wm[w]) ? wm[w] : m;
      atomicMax(rmax + r, (unsigned long long)__double_as_longlong(m));
    }
    __syncthreads();
  }
}

__global__ void k_axpy(double* __restrict__ X, const double* __restrict__ D, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) X[i] += D[i];
}

// ---------------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------------
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// fewer than two 128 x 128 tiles per CU (the skinny products of a panel): 64 x 64 tiles,
// four times as many workgroups
static inline bool gemm_small(int M, int N) { return (size_t)cdiv(M, 128) * cdiv(N, 128) < 512; }
size_t gemm_tiles(int M, int N) {
  const int t = gemm_small(M, N) ? 64 : 128;
  return (size_t)cdiv(M, t) * cdiv(N, t);
}

// tile rows per band of k_gemm_x6's XCD-aware tile order (0: the plain row-major order).  At
// K = 128 the plain order is faster (C, read and written once, keeps its row locality); from
// K = 256 on the A and B panels (256-512 KiB per tile) make up most of the fetches, and the
// banded order serves them from each XCD's L2.  TRACEON_X6_BAND overrides it for K >= 256.
static int x6_band(int K) {
  static const int env = [] {
    const char* e = getenv("TRACEON_X6_BAND");
    return e && *e ? atoi(e) : -1;
  }();
  if (K < 256) return 0;
  return env >= 0 ? env : 8;
}

template <typename T, int GB, bool TM, bool FAST>
static void launch_gemm(dim3 grid, int M, int N, int K, T alpha, const T* A, size_t lda, const T* B, size_t ldb, T beta, T* C,
                        size_t ldc, hipStream_t st, const int* gate, int gp, unsigned long long* vm) {
  constexpr int BK = sizeof(T) == 8 ? 16 : 32;
  if constexpr (sizeof(T) == 4 && GB == 128 && FAST && !TM)
    hipLaunchKernelGGL(k_gemm_x6, grid, dim3(512), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, gp, x6_band(K));
  else if constexpr (sizeof(T) == 4 && GB == 128 && FAST)
    hipLaunchKernelGGL((k_gemm_f32<TM>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, gp, vm);
  else
    hipLaunchKernelGGL((k_gemm<T, BK, TM, GB, FAST>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, gp, vm);
}

template <typename T, int GB>
static void launch_gemm(int M, int N, int K, T alpha, const T* A, size_t lda, const T* B, size_t ldb, T beta, T* C, size_t ldc,
                        hipStream_t st, const int* gate, int gp, unsigned long long* vm) {
  constexpr int BK = sizeof(T) == 8 ? 16 : 32;
  constexpr size_t VEC = 16 / sizeof(T);
  const bool fast = lda % VEC == 0 && (uintptr_t)A % 16 == 0 && ldb % VEC == 0 && (uintptr_t)B % 16 == 0 && ldc % VEC == 0 &&
                    (uintptr_t)C % 16 == 0 && K % BK == 0 && N % VEC == 0 && M >= GB && N >= GB;
  const dim3 grid(cdiv(N, GB), cdiv(M, GB));
  if (vm) {
    if (fast) launch_gemm<T, GB, true, true>(grid, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st, gate, gp, vm);
    else launch_gemm<T, GB, true, false>(grid, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st, gate, gp, vm);
  } else {
    if (fast) launch_gemm<T, GB, false, true>(grid, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st, gate, gp, vm);
    else launch_gemm<T, GB, false, false>(grid, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st, gate, gp, vm);
  }
}

template <typename T>
hipError_t gemm(int M, int N, int K, T alpha, const T* A, size_t lda, const T* B, size_t ldb, T beta, T* C, size_t ldc,
                hipStream_t st, const int* gate, int gp, unsigned long long* vmax) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (gemm_small(M, N)) launch_gemm<T, 64>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st, gate, gp, vmax);
  else launch_gemm<T, 128>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st, gate, gp, vmax);
  return hipGetLastError();
}
template hipError_t gemm<double>(int, int, int, double, const double*, size_t, const double*, size_t, double, double*, size_t,
                                 hipStream_t, const int*, int, unsigned long long*);

// U12 product with the panel check in front (k_gemm_checked); products of fewer than 512
// 128-tiles only (64-tiles)
template <typename T>
static hipError_t gemm_checked(int M, int N, int K, T alpha, const T* A, size_t lda, const T* B, size_t ldb, T beta, T* C, size_t ldc,
                               hipStream_t st, int* first_fail, int gp, const unsigned long long* lmax,
                               const unsigned long long* slots, int nslot) {
  if (M <= 0 || N <= 0) return hipSuccess;
  constexpr int BK = sizeof(T) == 8 ? 16 : 32;
  constexpr size_t VEC = 16 / sizeof(T);
  const bool fast = lda % VEC == 0 && (uintptr_t)A % 16 == 0 && ldb % VEC == 0 && (uintptr_t)B % 16 == 0 && ldc % VEC == 0 &&
                    (uintptr_t)C % 16 == 0 && K % BK == 0 && N % VEC == 0 && M >= 64 && N >= 64;
  const dim3 grid(cdiv(N, 64), cdiv(M, 64));
  if (fast)
    hipLaunchKernelGGL((k_gemm_checked<T, true>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, first_fail,
                       gp, lmax, slots, nslot);
  else
    hipLaunchKernelGGL((k_gemm_checked<T, false>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, first_fail,
                       gp, lmax, slots, nslot);
  return hipGetLastError();
}
template hipError_t gemm<float>(int, int, int, float, const float*, size_t, const float*, size_t, float, float*, size_t, hipStream_t,
                                const int*, int, unsigned long long*);

// Workspace of one factorisation (kept with the factors: the solves read the block inverses,
// the permutation and the pivot flag from it).
template <typename T>
struct LuWorkspace {
  static constexpr int G = 1024, NB = LU_NB;
  double* cand_val; int* cand_idx; T* cand_row;   // pivot candidates of the cooperative panel
  T* Y; T* Z;                                     // scratch inverses of the current panel
  T* U12; T* U12b;                                // U12 of the current panel (two, alternating by panel)
  T* Linv; T* Uinv;                               // inverses of every diagonal block (nblk x NB x NB)
  T* Lm; T* Um;                                   // coupling blocks of the solves: inv(L_bb) L_b,b-1, inv(U_bb) U_b,b+1 (nblk x NB x NB)
  int* perm; int* flags; int* err; int* pivoted;  // permutation, solve tickets, error, any-interchange flag
  double* rhs;                                    // copy of the right-hand sides (nrhs x N, nrhs <= 64)
  double* ytmp;                                   // L^{-1} P b of the solve in progress (N)
  T* Ltmp;                                        // L21 = A21 inv(U11) of the speculative pass (2 x N x NB, by panel parity)
  T* Lp; T* Up;                                   // L21 and U12 of a panel group of up to 4: [L_q | ..] (N x 4NB, rows
                                                  // absolute) and [U_q ; ..] (4NB x ldu, columns absolute), two of each
  size_t ldu;
  T* Dbak;                                        // the current diagonal block before its factorisation (NB x NB)
  unsigned long long* lmaxv;                      // max |l| per panel of the speculative pass (nblk)
  int* first_fail;                                // first panel of that pass that failed the no-pivot test
  unsigned long long* lslots;                     // max |l| per workgroup of a panel's L21 product
  unsigned long long* lmax;
  size_t bytes;
  LuWorkspace(void* base, int N) {
    const int nblk = (N + NB - 1) / NB;
    size_t o = 0;
    auto take = [&](size_t b) { size_t at = o; o = (o + b + 255) & ~(size_t)255; return at; };
    unsigned char* p = (unsigned char*)base;
    size_t a_cv = take(2 * G * sizeof(double)), a_ci = take(2 * G * sizeof(int)), a_cr = take((2 * (size_t)G * NB + 2 * NB) * sizeof(T));
    size_t a_y = take((size_t)NB * NB * sizeof(T)), a_z = take((size_t)NB * NB * sizeof(T));
    size_t a_u = take((size_t)NB * N * sizeof(T)), a_pb = take((size_t)NB * N * sizeof(T));
    size_t a_li = take((size_t)nblk * NB * NB * sizeof(T)), a_ui = take((size_t)nblk * NB * NB * sizeof(T));
    size_t a_lc = take((size_t)nblk * NB * NB * sizeof(T)), a_uc = take((size_t)nblk * NB * NB * sizeof(T));
    size_t a_pm = take((size_t)N * sizeof(int)), a_fl = take((2 * (size_t)nblk + 2) * sizeof(int));
    size_t a_er = take(sizeof(int)), a_pv = take(sizeof(int)), a_rhs = take(64 * (size_t)N * sizeof(double)), a_lm = take(8);
    size_t a_yt = take((size_t)N * sizeof(double));
    size_t a_lt = take(2 * (size_t)N * NB * sizeof(T)), a_lv = take((size_t)nblk * 8), a_ff = take(sizeof(int));
    size_t a_db = take((size_t)NB * NB * sizeof(T));
    ldu = ((size_t)N + 31) / 32 * 32;
    size_t a_lp = take(2 * (size_t)N * 4 * NB * sizeof(T)), a_up = take(2 * (size_t)4 * NB * ldu * sizeof(T));
    size_t a_ls = take(((size_t)N / 32 + 64) * 8);
    bytes = o;
    if (!base) return;
    cand_val = (double*)(p + a_cv); cand_idx = (int*)(p + a_ci); cand_row = (T*)(p + a_cr);
    Y = (T*)(p + a_y); Z = (T*)(p + a_z); U12 = (T*)(p + a_u); U12b = (T*)(p + a_pb);
    Linv = (T*)(p + a_li); Uinv = (T*)(p + a_ui); perm = (int*)(p + a_pm); flags = (int*)(p + a_fl);
    Lm = (T*)(p + a_lc); Um = (T*)(p + a_uc);
    err = (int*)(p + a_er); pivoted = (int*)(p + a_pv); rhs = (double*)(p + a_rhs); lmax = (unsigned long long*)(p + a_lm);
    ytmp = (double*)(p + a_yt);
    Ltmp = (T*)(p + a_lt); lmaxv = (unsigned long long*)(p + a_lv); first_fail = (int*)(p + a_ff); Dbak = (T*)(p + a_db);
    Lp = (T*)(p + a_lp); Up = (T*)(p + a_up);
    lslots = (unsigned long long*)(p + a_ls);
  }
};

template <typename T>
size_t lu_workspace_bytes(int N) {
  return LuWorkspace<T>(nullptr, N).bytes;
}
template size_t lu_workspace_bytes<double>(int);
template size_t lu_workspace_bytes<float>(int);

// Second stream and events of the factorisation's lookahead, one set per device, kept for
// the life of the process (default priority).  Calls on one device are serialised.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t ev_start = nullptr, ev_r = nullptr, ev_u = nullptr, ev_end = nullptr;
  std::mutex mu;
};
static SideStream g_side[64];
static std::mutex g_side_mu;

static hipError_t side_stream(int dev, SideStream** out) {
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_side_mu);
  SideStream& x = g_side[dev];
  if (!x.s) {
    hipError_t e = hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking);
    if (e != hipSuccess) { x.s = nullptr; return e; }
    for (hipEvent_t* ev : {&x.ev_start, &x.ev_r, &x.ev_u, &x.ev_end}) {
      e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
  }
  *out = &x;
  return hipSuccess;
}

// Panel groups (one bulk update at K = G NB for G panels): groups of 4 while the trailing
// matrix has at least lu_group_rows(4) rows, of 2 while it has lu_group_rows(2), single panels
// after.  TRACEON_LU_QUAD_ROWS and TRACEON_LU_PAIR_ROWS override them (0: throughout, a value
// above N: never).
static constexpr int LU_GROUP_MAX = 4;
static_assert(LU_GROUP_MAX == 4, "the workspace's group buffers hold four panels");
static int lu_group_rows(int G) {
  const char* e = getenv(G == 4 ? "TRACEON_LU_QUAD_ROWS" : "TRACEON_LU_PAIR_ROWS");
  return e && *e ? atoi(e) : (G == 4 ? 8192 : 4096);
}

// Factorisation driver.  Most BEM panels need no row interchange, so the factorisation
// first runs speculatively without pivoting and without host round trips.  Per panel p
// (side stream A unless stated):
//   P(p):  factor the diagonal block and invert its triangles (k_diag_block, which keeps a
//          copy of the block), L21 = A21 inv(U11) into a panel buffer with max |l| folded
//          in by the GEMM, and mark p failed if some |l| > 1;
//   U(p):  U12 = inv(L11) A12 into a buffer, after the bulk stream's Tr(p-1);
//   Tc(p): the trailing update of the next panel's column only;
//   Tr(p)  (caller's stream B): the trailing update of the remaining columns, then the
//          copies of L21 and U12 into the matrix.
// The chain P(p) -> U(p) -> Tc(p) -> P(p+1) stays on one stream, and P(p+1) and Tc(p) run
// beside the large Tr(p) (one-panel lookahead); L21 and U12 alternate between two buffers
// by panel parity, so Tr(p) still reads panel p's while panel p+1 writes its own.  Every
// kernel of panel p is gated on "no panel <= p failed" (a device-side first_fail).  After a
// pass that failed at panel f, the matrix holds panels < f complete with their trailing
// updates and panel f's column untouched except its diagonal block, which is restored from
// the copy; f is then factored with partial pivoting (cooperative k_panel) and the
// speculative pass resumes at f + 1.  A panel that passes the test (every |l| <= 1) is
// exactly the partially pivoted result: its diagonal entries are the column maxima, so its
// ipiv is the identity.
// While the trailing matrix is large the pass takes panels in groups of G = 4 or 2: the side
// stream factors them one after another (P, L21, U12 of each, then its update of the group's
// later panels' columns and rows) and then updates the next group's columns and rows with the
// whole group at once (K = G NB); the bulk stream applies the group's update to the rest of the
// matrix in one product, so that trailing C is read and written once per G panels.  A failure
// at panel j > 0 of a group leaves panels 0 .. j-1's update beyond the group unapplied; the
// driver applies it before the pivoted panel.
template <typename T>
hipError_t lu_factor(T* A, size_t lda, int N, int* ipiv, void* ws, int* info_dev, hipStream_t st, int panel_groups,
                     KernelTiming* tm) {
  constexpr int NB = LU_NB;
  if (N <= 0) return hipSuccess;
  struct Events {   // destroys what an early return leaves behind
    std::vector<std::pair<hipEvent_t, hipEvent_t>> v;
    ~Events() {
      for (auto& p : v) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
    }
  } events;
  auto& ev = events.v;
  auto timed_gemm = [&](hipStream_t s, const int* gate, int gp, int M, int Nn, int K, T alpha, const T* Ap, size_t la, const T* Bp,
                        size_t lb, T beta, T* Cp, size_t lc, unsigned long long* vmax = nullptr) -> hipError_t {
    if (M <= 0 || Nn <= 0) return hipSuccess;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = tm != nullptr && !gemm_small(M, Nn);   // the large (trailing) products only
    if (timed) {
      if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return hipErrorOutOfMemory;
      ev.emplace_back(e0, e1);
      if (hipEventRecord(e0, s) != hipSuccess) return hipErrorLaunchFailure;
    }
    hipError_t r = gemm<T>(M, Nn, K, alpha, Ap, la, Bp, lb, beta, Cp, lc, s, gate, gp, vmax);
    if (timed) {
      if (hipEventRecord(e1, s) != hipSuccess) return hipErrorLaunchFailure;
      tm->launches++;
      tm->flops += 2.0 * M * (double)Nn * K;
      tm->bytes += sizeof(T) * ((double)M * K + (double)K * Nn + (beta != T(0) ? 2.0 : 1.0) * M * (double)Nn);
    }
    return r;
  };
  auto finish_timing = [&]() {
    for (auto& p : ev) {
      float ms = 0.f;
      if (hipEventSynchronize(p.second) == hipSuccess && hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) tm->ms += ms;
    }
  };
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return e;
  int G = panel_groups > 0 ? panel_groups : prop.multiProcessorCount;
  if (G > 1024) G = 1024;
  const size_t max_lds = 160 * 1024 - (2 * NB * sizeof(T) + 256 * (sizeof(double) + 2 * sizeof(int)) + 64);
  size_t Rmax = ((size_t)N + G - 1) / G;
  if (Rmax * NB * sizeof(T) > max_lds) return hipErrorInvalidValue;
  auto kpanel = k_panel<T, NB>;
  e = hipFuncSetAttribute((const void*)kpanel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(Rmax * NB * sizeof(T)));
  if (e != hipSuccess) return e;
  int occ = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kpanel, 256, Rmax * NB * sizeof(T));
  if (e != hipSuccess) return e;
  if (occ * prop.multiProcessorCount < G) return hipErrorCooperativeLaunchTooLarge;
  e = hipFuncSetAttribute((const void*)k_diag_block<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)diag_lds_bytes<T>());
  if (e != hipSuccess) return e;
  SideStream* ss = nullptr;
  e = side_stream(dev, &ss);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(ss->mu);
  const hipStream_t sa = ss->s;

  LuWorkspace<T> W(ws, N);
  const int nblk = (N + NB - 1) / NB;
  T* U12 = W.U12;
  bool any_pivot = false;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, W.err, 0);
  auto Yp = [&](int q) { return W.Linv + (size_t)q * NB * NB; };
  auto Zp = [&](int q) { return W.Uinv + (size_t)q * NB * NB; };

  // ---- speculative pass from panel s0; *kf = first failed panel, or nblk; *fix = the
  //      panel group that holds *kf at a position j > 0 (its panels 0 .. j-1's update of the
  //      matrix beyond the group is then still to apply), j = 0 if none ----
  constexpr int LDL = LU_GROUP_MAX * NB;
  const size_t ldu = W.ldu;
  const int pair_min = lu_group_rows(2), quad_min = lu_group_rows(4);
  auto Lgr = [&](int gi) { return W.Lp + (size_t)(gi & 1) * N * LDL; };   // [L_q | .. | L_q+G-1], row r at r * LDL
  auto Ugr = [&](int gi) { return W.Up + (size_t)(gi & 1) * LDL * ldu; };  // [U_q ; .. ; U_q+G-1], column c at c
  struct Group { int q, G, gi; };
  struct Fix { int q, j, gi, G; };
  auto speculate = [&](int s0, int* kf, Fix* fix) -> hipError_t {
    const int* g = W.first_fail;
    *fix = {0, 0, 0, 0};
    std::vector<Group> groups;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, W.first_fail, nblk);
    hipLaunchKernelGGL(k_zero_u64, dim3(cdiv(nblk, 256)), dim3(256), 0, st, W.lmaxv, nblk);
    hipError_t r = hipEventRecord(ss->ev_start, st);
    if (r != hipSuccess) return r;
    if ((r = hipStreamWaitEvent(sa, ss->ev_start, 0)) != hipSuccess) return r;
    auto Lbuf = [&](int q) { return W.Ltmp + (size_t)(q & 1) * N * NB; };
    auto Ubuf = [&](int q) { return (q & 1) ? W.U12b : U12; };
    // P(q) and L21(q) = A21 inv(U11) into C (ldc) with max |l| per workgroup; returns the slot count
    auto diag_l21 = [&](int q, T* C, size_t ldc, int* nslot) -> hipError_t {
      const int k0 = q * NB, nb = min(NB, N - k0), k1 = k0 + nb, m2 = N - k1;
      hipLaunchKernelGGL(k_diag_block<T>, dim3(1), dim3(256), diag_lds_bytes<T>(), sa, A, lda, k0, nb, 1, ipiv, Yp(q), Zp(q),
                         W.lmaxv + q, g, q, W.Dbak);
      *nslot = 0;
      if (m2 <= 0) return hipSuccess;
      *nslot = (int)gemm_tiles(m2, nb);
      return timed_gemm(sa, g, q, m2, nb, nb, T(1), A + (size_t)k1 * lda + k0, lda, Zp(q), NB, T(0), C, ldc, W.lslots);
    };
    int q = s0;
    // Group mode while the trailing matrix is large: the G panels q .. q+G-1 (G = 4 or 2) are
    // factored on the side stream, each followed by its update of the group's later panels'
    // columns and rows; then the next G panels' columns and rows receive the whole group's
    // update at K = G NB, and the bulk stream applies it to the rest of the matrix in one
    // product, reading and writing that C once for G panels.  Everything of the group except
    // the bulk update and the strips runs beside the previous group's bulk update.  Group
    // sizes never grow along the pass, so a group's strips always cover the next group.
    for (int gi = 0;; gi++) {
      int G = N - q * NB >= quad_min ? 4 : N - q * NB >= pair_min ? 2 : 1;
      while (G > 1 && q + G >= nblk) G /= 2;
      if (G == 1) break;
      const int b0 = q * NB, bG = b0 + G * NB, b2G = min(bG + G * NB, N);
      T* Lq = Lgr(gi);
      T* Uq = Ugr(gi);
      for (int j = 0; j < G; j++) {
        const int bj = b0 + j * NB, bj1 = bj + NB;
        int nslot = 0;
        if ((r = diag_l21(q + j, Lq + (size_t)bj1 * LDL + j * NB, LDL, &nslot)) != hipSuccess) return r;
        if ((r = gemm_checked<T>(NB, N - bj1, NB, T(1), Yp(q + j), NB, A + (size_t)bj * lda + bj1, lda, T(0),
                                 Uq + (size_t)j * NB * ldu + bj1, ldu, sa, W.first_fail, q + j, W.lmaxv + q + j, W.lslots, nslot)) !=
            hipSuccess)
          return r;
        if (j + 1 == G) break;
        // panel q+j's update of the group's later panels: their columns, then their rows
        const T* Lj = Lq + (size_t)bj1 * LDL + j * NB;
        const T* Uj = Uq + (size_t)j * NB * ldu;
        if ((r = timed_gemm(sa, g, q + j, N - bj1, bG - bj1, NB, T(-1), Lj, LDL, Uj + bj1, ldu, T(1), A + (size_t)bj1 * lda + bj1,
                            lda)) != hipSuccess)
          return r;
        if ((r = timed_gemm(sa, g, q + j, bG - bj1, N - bG, NB, T(-1), Lj, LDL, Uj + bG, ldu, T(1), A + (size_t)bj1 * lda + bG,
                            lda)) != hipSuccess)
          return r;
      }
      groups.push_back({q, G, gi});
      // the next G panels' columns and rows, after the previous group's bulk update and copies
      const int gl = q + G - 1, KG = G * NB;
      if (gi > 0 && (r = hipStreamWaitEvent(sa, ss->ev_r, 0)) != hipSuccess) return r;
      if ((r = timed_gemm(sa, g, gl, N - bG, b2G - bG, KG, T(-1), Lq + (size_t)bG * LDL, LDL, Uq + bG, ldu, T(1),
                          A + (size_t)bG * lda + bG, lda)) != hipSuccess)
        return r;
      if ((r = timed_gemm(sa, g, gl, b2G - bG, N - b2G, KG, T(-1), Lq + (size_t)bG * LDL, LDL, Uq + b2G, ldu, T(1),
                          A + (size_t)bG * lda + b2G, lda)) != hipSuccess)
        return r;
      if ((r = hipEventRecord(ss->ev_u, sa)) != hipSuccess) return r;
      // the bulk update on the caller's stream, then the group's factors into the matrix
      if ((r = hipStreamWaitEvent(st, ss->ev_u, 0)) != hipSuccess) return r;
      if ((r = timed_gemm(st, g, gl, N - b2G, N - b2G, KG, T(-1), Lq + (size_t)b2G * LDL, LDL, Uq + b2G, ldu, T(1),
                          A + (size_t)b2G * lda + b2G, lda)) != hipSuccess)
        return r;
      for (int j = 0; j < G; j++) {
        const int bj = b0 + j * NB, bj1 = bj + NB;
        hipLaunchKernelGGL(k_copy_panel<T>, dim3(cdiv((size_t)2 * (N - bj1) * NB, 256)), dim3(256), 0, st,
                           Lq + (size_t)bj1 * LDL + j * NB, (size_t)LDL, A + (size_t)bj1 * lda + bj, Uq + (size_t)j * NB * ldu + bj1, ldu,
                           A + (size_t)bj * lda + bj1, lda, N - bj1, N - bj1, NB, g, q + j);
      }
      if ((r = hipEventRecord(ss->ev_r, st)) != hipSuccess) return r;
      q += G;
    }
    for (; q < nblk; q++) {
      const int k0 = q * NB, nb = min(NB, N - k0), m2 = N - k0 - nb, n2 = m2;
      const int k1 = k0 + nb;
      int nslot = 0;
      if ((r = diag_l21(q, Lbuf(q), (size_t)nb, &nslot)) != hipSuccess) return r;
      if (n2 <= 0) {   // the last panel: its test alone (otherwise it runs in front of U12)
        hipLaunchKernelGGL(k_panel_check, dim3(1), dim3(256), 0, sa, W.lmaxv + q, W.lslots, nslot, W.first_fail, q);
        break;
      }
      const int nb1 = min(NB, n2);
      const T* A12 = A + (size_t)k0 * lda + k1;
      // U(q) after Tr(q-1) of the bulk stream (or the last pair's bulk update)
      if (q > s0 && (r = hipStreamWaitEvent(sa, ss->ev_r, 0)) != hipSuccess) return r;
      if ((r = gemm_checked<T>(nb, n2, nb, T(1), Yp(q), NB, A12, lda, T(0), Ubuf(q), (size_t)n2, sa, W.first_fail, q, W.lmaxv + q,
                               W.lslots, nslot)) != hipSuccess)
        return r;
      if ((r = hipEventRecord(ss->ev_u, sa)) != hipSuccess) return r;
      // Tc(q): the next panel's column
      if ((r = timed_gemm(sa, g, q, n2, nb1, nb, T(-1), Lbuf(q), (size_t)nb, Ubuf(q), (size_t)n2, T(1), A + (size_t)k1 * lda + k1,
                          lda)) != hipSuccess)
        return r;
      // Tr(q) on the caller's stream, then the factors of panel q into the matrix
      if ((r = hipStreamWaitEvent(st, ss->ev_u, 0)) != hipSuccess) return r;
      if ((r = timed_gemm(st, g, q, n2, n2 - nb1, nb, T(-1), Lbuf(q), (size_t)nb, Ubuf(q) + nb1, (size_t)n2, T(1),
                          A + (size_t)k1 * lda + k1 + nb1, lda)) != hipSuccess)
        return r;
      hipLaunchKernelGGL(k_copy_panel<T>, dim3(cdiv((size_t)m2 * nb + (size_t)nb * n2, 256)), dim3(256), 0, st, Lbuf(q), (size_t)nb,
                         A + (size_t)k1 * lda + k0, Ubuf(q), (size_t)n2, A + (size_t)k0 * lda + k1, lda, m2, n2, nb, g, q);
      if ((r = hipEventRecord(ss->ev_r, st)) != hipSuccess) return r;
    }
    if ((r = hipEventRecord(ss->ev_end, sa)) != hipSuccess) return r;
    if ((r = hipStreamWaitEvent(st, ss->ev_end, 0)) != hipSuccess) return r;
    int h = nblk;
    if ((r = hipMemcpyAsync(&h, W.first_fail, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess) return r;
    if ((r = hipStreamSynchronize(st)) != hipSuccess) return r;
    *kf = h;
    // a failure at panel j > 0 of a group leaves panels 0 .. j-1's update beyond the group
    for (const Group& gr : groups)
      if (h > gr.q && h < gr.q + gr.G) *fix = {gr.q, h - gr.q, gr.gi, gr.G};

    return hipGetLastError();
  };

  // ---- one panel with partial pivoting (its column already restored) ----
  auto pivoted_panel = [&](int q) -> hipError_t {
    const int k0 = q * NB, nb = min(NB, N - k0);
    T* Ap = A; size_t ldap = lda; int Np = N, k0p = k0;
    int* ipivp = ipiv; double* wsdp = W.cand_val; int* wsip = W.cand_idx; T* wsrp = W.cand_row; int* infop = info_dev;
    void* args[] = {&Ap, &ldap, &Np, &k0p, &ipivp, &wsdp, &wsip, &wsrp, &infop};
    const int m = N - k0;
    size_t R = ((size_t)m + G - 1) / G;
    hipError_t r = hipLaunchCooperativeKernel((const void*)kpanel, dim3(G), dim3(256), args, (unsigned)(R * NB * sizeof(T)), st);
    if (r != hipSuccess) return r;
    any_pivot = true;
    hipLaunchKernelGGL(k_diag_block<T>, dim3(1), dim3(256), diag_lds_bytes<T>(), st, A, lda, k0, nb, 0, ipiv, Yp(q), Zp(q), W.lmax,
                       (const int*)nullptr, 0, (T*)nullptr);
    if (N - nb > 0) hipLaunchKernelGGL(k_laswp<T>, dim3(cdiv(N - nb, 256)), dim3(256), 0, st, A, lda, N, k0, nb, ipiv);
    const int n2 = N - k0 - nb;
    if (n2 > 0) {
      T* A12 = A + (size_t)k0 * lda + k0 + nb;
      if ((r = timed_gemm(st, nullptr, 0, nb, n2, nb, T(1), Yp(q), NB, A12, lda, T(0), U12, (size_t)n2)) != hipSuccess) return r;
      hipLaunchKernelGGL(k_copy2d<T>, dim3(cdiv((size_t)nb * n2, 256)), dim3(256), 0, st, U12, (size_t)n2, A12, lda, nb, n2,
                         (const int*)nullptr, 0);
      const T* L21 = A + (size_t)(k0 + nb) * lda + k0;
      T* A22 = A + (size_t)(k0 + nb) * lda + k0 + nb;
      if ((r = timed_gemm(st, nullptr, 0, n2, n2, nb, T(-1), L21, lda, U12, (size_t)n2, T(1), A22, lda)) != hipSuccess) return r;
    }
    return hipGetLastError();
  };

  for (int q = 0; q < nblk;) {
    int kf = nblk;
    Fix fix;
    e = speculate(q, &kf, &fix);
    if (e != hipSuccess) return e;
    if (kf < q || kf >= nblk) break;
    if (fix.j > 0) {   // the update of the group's panels before kf, beyond the group
      const int bG = (fix.q + fix.G) * NB;
      if ((e = timed_gemm(st, nullptr, 0, N - bG, N - bG, fix.j * NB, T(-1), Lgr(fix.gi) + (size_t)bG * LDL, LDL, Ugr(fix.gi) + bG, ldu,
                          T(1), A + (size_t)bG * lda + bG, lda)) != hipSuccess)
        return e;
    }
    const int k0 = kf * NB, nb = min(NB, N - k0);
    hipLaunchKernelGGL(k_copy2d<T>, dim3(cdiv((size_t)nb * nb, 256)), dim3(256), 0, st, W.Dbak, (size_t)NB,
                       A + (size_t)k0 * lda + k0, lda, nb, nb, (const int*)nullptr, 0);
    e = pivoted_panel(kf);
    if (e != hipSuccess) return e;
    q = kf + 1;
  }
  hipLaunchKernelGGL(k_diag_check<T>, dim3(1), dim3(1024), 0, st, A, lda, N, info_dev);
  hipLaunchKernelGGL((k_trsv_coupling<T, true>), dim3(nblk), dim3(256), 0, st, A, lda, N, W.Linv, W.Lm);
  hipLaunchKernelGGL((k_trsv_coupling<T, false>), dim3(nblk), dim3(256), 0, st, A, lda, N, W.Uinv, W.Um);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, W.pivoted, any_pivot ? 1 : 0);
  if (any_pivot) {
    if ((size_t)N * sizeof(int) > 150 * 1024) return hipErrorInvalidValue;
    e = hipFuncSetAttribute((const void*)k_perm, hipFuncAttributeMaxDynamicSharedMemorySize, (int)((size_t)N * sizeof(int)));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_perm, dim3(1), dim3(256), (size_t)N * sizeof(int), st, ipiv, N, W.perm);
  }
  if (tm) finish_timing();
  return hipGetLastError();
}
template hipError_t lu_factor<double>(double*, size_t, int, int*, void*, int*, hipStream_t, int, KernelTiming*);
template hipError_t lu_factor<float>(float*, size_t, int, int*, void*, int*, hipStream_t, int, KernelTiming*);

template <typename T>
hipError_t lu_solve(const T* LU, size_t lda, int N, const int* ipiv, double* X, int nrhs, const void* ws, hipStream_t st) {
  if (N <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs > 64) return hipErrorInvalidValue;
  const LuWorkspace<T> W(const_cast<void*>(ws), N);
  const int nblk = (N + LU_NB - 1) / LU_NB;
  // a wait of the triangular solves that passed its spin limit records itself in W.err: cleared
  // per solve call, read back with lu_solve_status
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, W.err, 0);
  // the right-hand sides are read through the permutation into a copy
  hipMemcpyAsync(W.rhs, X, sizeof(double) * (size_t)nrhs * N, hipMemcpyDeviceToDevice, st);
  constexpr size_t lds = trsv_lds_bytes<T>();
  hipError_t e = hipFuncSetAttribute((const void*)k_trsv_blocked<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_trsv_blocked<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const unsigned pgrid = cdiv(N, 256);
  for (int r = 0; r < nrhs; r++) {
    // y = L^{-1} P b into the workspace, then x = U^{-1} y into X (each output prefilled
    // with the pending marker its readers poll)
    double* xr = X + (size_t)r * N;
    hipLaunchKernelGGL(k_trsv_prep, dim3(pgrid), dim3(256), 0, st, W.ytmp, N, W.flags);
    hipLaunchKernelGGL((k_trsv_blocked<T, true>), dim3(nblk), dim3(TRSV_THREADS), lds, st, LU, lda, N, W.perm, W.pivoted,
                       W.rhs + (size_t)r * N, W.ytmp, W.Linv, W.Lm, W.flags, W.err);
    hipLaunchKernelGGL(k_trsv_prep, dim3(pgrid), dim3(256), 0, st, xr, N, W.flags + 1);
    hipLaunchKernelGGL((k_trsv_blocked<T, false>), dim3(nblk), dim3(TRSV_THREADS), lds, st, LU, lda, N, (const int*)nullptr,
                       (const int*)nullptr, W.ytmp, xr, W.Uinv, W.Um, W.flags + 1, W.err);
  }
  return hipGetLastError();
}
template hipError_t lu_solve<double>(const double*, size_t, int, const int*, double*, int, const void*, hipStream_t);
template hipError_t lu_solve<float>(const float*, size_t, int, const int*, double*, int, const void*, hipStream_t);

hipError_t residual(const double* A, size_t lda, int N, const double* B, const double* X, double* R, int nrhs,
                    unsigned long long* rmax, hipStream_t st) {
  hipMemsetAsync(rmax, 0, sizeof(unsigned long long) * nrhs, st);
  hipLaunchKernelGGL(k_residual, dim3(cdiv(N, 4)), dim3(256), 0, st, A, lda, N, B, X, R, nrhs, rmax);
  return hipGetLastError();
}

template <typename T>
hipError_t lu_solve_status(const void* ws, int N, int* status, hipStream_t st) {
  const LuWorkspace<T> W(const_cast<void*>(ws), N);
  hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, st, (const int*)W.err, status);
  return hipGetLastError();
}
template hipError_t lu_solve_status<double>(const void*, int, int*, hipStream_t);
template hipError_t lu_solve_status<float>(const void*, int, int*, hipStream_t);

hipError_t absmax(const double* x, size_t n, unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, out, 1);
  if (n) {
    const unsigned g = (unsigned)std::min<size_t>(cdiv(n, 256), 1024);
    hipLaunchKernelGGL(k_absmax, dim3(g), dim3(256), 0, st, x, n, out);
  }
  return hipGetLastError();
}

hipError_t axpy(double* X, const double* D, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3(cdiv(n, 256)), dim3(256), 0, st, X, D, n);
  return hipGetLastError();
}

hipError_t convert_d2f(const double* S, size_t lds, float* D, size_t ldd, int rows, int cols, hipStream_t st) {
  hipLaunchKernelGGL(k_d2f, dim3(cdiv((size_t)rows * cols, 256)), dim3(256), 0, st, S, lds, D, ldd, rows, cols);
  return hipGetLastError();
}

}  // namespace tb
