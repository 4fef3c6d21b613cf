// Learned field surrogate for the tracer: a 2-128-128-128-128-5 MLP with SiLU activations.
//
// Hidden layers (128 x 128) run on MFMA in BF16 (v_mfma_f32_16x16x16bf16_1k, FP32
// accumulation); the first layer (2 -> 128), the last layer (128 -> 5), activations, loss,
// reductions and the Adam update run in FP32 on the vector units.  Master weights are FP32;
// BF16 copies of the hidden-layer weights feed the MFMA.  The surrogate has no
// counterpart in the reference: it learns the difference between the direct BEM field and
// the paraxial (axial-derivative) series off the axis, where the series is truncated.
//
// Conventions: Z = H_prev W^T + b, W stored [out][in] row-major; batches are row-major
// (B, width).  Training minimises the mean squared error over the batch and the 5 outputs.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include "mlp_kernels.h"

namespace tb {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------
// C (+)= op(A) op(B), A (M,K) and B (K,N) in BF16 through strides:
//   A(m,k) = A[m*sam + k*sak],  B(k,n) = B[k*sbk + n*sbn],  C[m*ldc + n] FP32.
// 64 x 64 tile per workgroup of 4 waves (32 x 32 each = 2 x 2 MFMA 16x16x16), K staged in
// chunks of 128 through LDS (a hidden layer's whole K in one chunk).  An operand that is contiguous along k or along its row/column
// index is staged with one 16-byte load per thread (8 elements) where the group of 8 lies
// inside the matrix and is 16-byte aligned, else element by element.  gridDim.z splits K:
// with `part` each split stores its M x N result at part + z M N (k_ksplit_reduce then adds
// them in split order: the sum does not depend on timing), without it the split results are
// added atomically into C, which the caller has zeroed.  Operand layout of
// v_mfma_f32_16x16x16bf16_1k: lane l holds A[l%16][4(l/16)+i] and B[4(l/16)+i][l%16],
// i = 0..3; the result D[4(l/16)+v][l%16].
// Epilogues (one K split only): EPI_STORE C = acc (or C += acc); EPI_BIAS_SILU
// C = acc + aux[col] (aux: the layer's bias), outbf = bf16(silu(C)); EPI_DSILU
// C = acc * silu'(aux[row][col]) (aux: the layer's pre-activations), outbf = bf16(C).
// ---------------------------------------------------------------------------------
constexpr int MG_T = 64, MG_K = 128, MG_PAD = 8;
enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_DSILU = 2 };

// eight 16-bit values f(0..7) packed into 16 bytes, element i in halfword i
template <class F>
__device__ __forceinline__ uint4 pack8(F f) {
  unsigned w[4];
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = (unsigned)f(2 * i) | ((unsigned)f(2 * i + 1) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// one 64 x MG_K operand tile into S[64][MG_K + MG_PAD] (S[r][k] = X(r0 + r, k0 + k)), X(r, k) =
// X[r * sr + k * sk], rows below R and k below kend, zero elsewhere.  The tile is staged in
// MG_K / 32 passes of 64 x 32; the vector paths issue every pass's 16-byte load before the
// first LDS store, so one memory latency covers the whole tile.
__device__ __forceinline__ void stage_tile(uint16_t (*S)[MG_K + MG_PAD], const uint16_t* __restrict__ X, long sr, long sk, int r0,
                                           int R, int k0, int kend, int tid) {
  constexpr int NP = MG_K / 32;
  const bool vec_ok = ((reinterpret_cast<uintptr_t>(X) & 15) == 0);
  if (sk == 1 && vec_ok && (sr & 7) == 0) {          // rows contiguous in k: thread = (row, 8 k) per pass
    const int r = tid >> 2, kq = (tid & 3) * 8, gr = r0 + r;
    uint4 v[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) {
      const int gk = k0 + 32 * p + kq;
      if (gr < R && gk + 8 <= kend && (gk & 7) == 0) v[p] = *reinterpret_cast<const uint4*>(X + gr * sr + gk);
      else v[p] = pack8([&](int i) { return (gr < R && gk + i < kend) ? X[gr * sr + gk + i] : (uint16_t)0; });
    }
#pragma unroll
    for (int p = 0; p < NP; p++) *reinterpret_cast<uint4*>(&S[r][32 * p + kq]) = v[p];
  } else if (sr == 1 && vec_ok && (sk & 7) == 0) {   // columns contiguous in r: thread = (k, 8 rows) per pass
    const int kl = tid >> 3, rq = (tid & 7) * 8, gr = r0 + rq;
    uint4 v[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) {
      const int gk = k0 + 32 * p + kl;
      if (gk < kend && gr + 8 <= R && (gr & 7) == 0) v[p] = *reinterpret_cast<const uint4*>(X + gk * sk + gr);
      else v[p] = pack8([&](int i) { return (gk < kend && gr + i < R) ? X[gk * sk + gr + i] : (uint16_t)0; });
    }
#pragma unroll
    for (int p = 0; p < NP; p++) {
      const unsigned w[4] = {v[p].x, v[p].y, v[p].z, v[p].w};
#pragma unroll
      for (int i = 0; i < 8; i++) S[rq + i][32 * p + kl] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
    }
  } else {
    for (int e = tid; e < MG_T * MG_K; e += 256) {
      const int r = e / MG_K, k = e % MG_K, gr = r0 + r, gk = k0 + k;
      S[r][k] = (gr < R && gk < kend) ? X[gr * sr + gk * sk] : (uint16_t)0;
    }
  }
}

__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, long sam, long sak,
                                                   const uint16_t* __restrict__ B, long sbk, long sbn, float* __restrict__ C,
                                                   long ldc, int accumulate, int kchunk, int epi, const float* __restrict__ aux,
                                                   uint16_t* __restrict__ outbf, float* __restrict__ part) {
  __shared__ __align__(16) uint16_t As[MG_T][MG_K + MG_PAD];   // [m][k]
  __shared__ __align__(16) uint16_t Bs[MG_T][MG_K + MG_PAD];   // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * MG_T, n0 = blockIdx.y * MG_T;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
  // the epilogue's operand (bias or pre-activation) loaded before the products, so its memory
  // latency overlaps the operand staging
  float pre[2][2][4];
  if (gridDim.z == 1 && epi != EPI_STORE) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + v;
          const int col = n0 + wn + j * 16 + (lane & 15);
          pre[i][j][v] = (row < M && col < N) ? (epi == EPI_DSILU ? aux[(long)row * ldc + col] : aux[col]) : 0.f;
        }
  }
  for (int k0 = kbeg; k0 < kend; k0 += MG_K) {
    stage_tile(As, A, sam, sak, m0, M, k0, kend, tid);
    stage_tile(Bs, B, sbn, sbk, n0, N, k0, kend, tid);     // B^T rows: n, contiguous along k when sbk == 1
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < MG_K; kk += 16) {
      s4v a[2], b[2];
      const int kb = kk + 4 * (lane >> 4);
#pragma unroll
      for (int i = 0; i < 2; i++) {
        a[i] = *reinterpret_cast<const s4v*>(&As[wm + i * 16 + (lane & 15)][kb]);
        b[i] = *reinterpret_cast<const s4v*>(&Bs[wn + i * 16 + (lane & 15)][kb]);
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + v;
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (row < M && col < N) {
          const long o = (long)row * ldc + col;
          const float x = acc[i][j][v];
          if (split) {
            if (part) part[((long)blockIdx.z * M + row) * N + col] = x;
            else atomicAdd(C + o, x);
          } else if (epi == EPI_BIAS_SILU) {
            const float z = x + pre[i][j][v];
            C[o] = z;
            outbf[o] = f2bf(silu(z));
          } else if (epi == EPI_DSILU) {
            const float dz = x * dsilu(pre[i][j][v]);
            C[o] = dz;
            outbf[o] = f2bf(dz);
          } else {
            C[o] = accumulate ? C[o] + x : x;
          }
        }
      }
}

// C = (accumulate ? C : 0) + part[0] + part[1] + ... + part[kz - 1], element by element in
// that order
__global__ __launch_bounds__(256) void k_ksplit_reduce(int M, int N, int kz, const float* __restrict__ part, float* __restrict__ C,
                                                       long ldc, int accumulate) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)M * N) return;
  const int row = (int)(e / N), col = (int)(e % N);
  float x = accumulate ? C[(long)row * ldc + col] : 0.f;
  // eight loads in flight, then their sums in order (a load-add chain waits one memory latency
  // per split)
  int z = 0;
  for (; z + 8 <= kz; z += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = part[(long)(z + u) * M * N + e];
#pragma unroll
    for (int u = 0; u < 8; u++) x += v[u];
  }
  for (; z < kz; z++) x += part[(long)z * M * N + e];
  C[(long)row * ldc + col] = x;
}

// first layer: Z1 = X W1^T + b1, H1 = silu(Z1) (BF16)
__global__ __launch_bounds__(256) void k_mlp_in(int B, const float* __restrict__ X, const float* __restrict__ P,
                                                float* __restrict__ Z, uint16_t* __restrict__ H) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)B * MLP_W) return;
  const int b = (int)(e / MLP_W), j = (int)(e % MLP_W);
  const float* W1 = P + MLP_OFF_W1;
  const float z = fmaf(W1[2 * j], X[2 * b], fmaf(W1[2 * j + 1], X[2 * b + 1], P[MLP_OFF_B1 + j]));
  Z[e] = z;
  H[e] = f2bf(silu(z));
}

// last layer from the FP32 pre-activations of layer 4: Y = silu(Z4) W5^T + b5.
// With targets T: dY = 2 (Y - T) / (5 B) and the squared error summed over the block's rows
// into loss[blockIdx.x].
// output layer Y = silu(Z4) W5^T + b5 (5 outputs) and, with targets T, dY = 2 (Y - T) / (5 B),
// the squared error of each workgroup's samples and, with dZ, the backward step through the
// last layer and the layer-4 activation, dZ4 = (dY W5) * silu'(Z4) in FP32 and BF16.
// A wave takes OUT_SPW samples: lane l reads z[l] and z[l + 64] of each (coalesced 256-byte
// rows) and stages silu(z) in LDS; lane 5u + o then forms output o of sample u as
// b5[o] + sum_n W5[o][n] h[n], one FMA per neuron in the order n = 0..127 (the order of the
// tracer's in-wave evaluation, so that both give the same correction); dZ4 comes from the
// z and W5 values the lanes already hold.
constexpr int OUT_SPW = 4, OUT_SPB = 4 * OUT_SPW;   // samples per wave, per 256-thread workgroup
__global__ __launch_bounds__(256) void k_mlp_out(int B, const float* __restrict__ Z4, const float* __restrict__ P,
                                                 float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY,
                                                 float* __restrict__ loss, float* __restrict__ dZ, uint16_t* __restrict__ dZbf) {
  __shared__ float hs[4][OUT_SPW][MLP_W];
  __shared__ float red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float* W5 = P + MLP_OFF_W5;
  const int b0 = blockIdx.x * OUT_SPB + wave * OUT_SPW;
  float zv[OUT_SPW][2];
#pragma unroll
  for (int u = 0; u < OUT_SPW; u++) {
    const int b = b0 + u;
    zv[u][0] = b < B ? Z4[(long)b * MLP_W + lane] : 0.f;
    zv[u][1] = b < B ? Z4[(long)b * MLP_W + 64 + lane] : 0.f;
    hs[wave][u][lane] = silu(zv[u][0]);
    hs[wave][u][64 + lane] = silu(zv[u][1]);
  }
  __syncthreads();
  // lane 5u + o: output o of sample u
  const int tu = lane / MLP_OUT, to = lane % MLP_OUT, tb = b0 + tu;
  const bool active = lane < OUT_SPW * MLP_OUT && tb < B;
  float se = 0.f, dyv = 0.f;
  if (active) {
    float y = P[MLP_OFF_B5 + to];
    const float* w = W5 + to * MLP_W;
    const float* h = hs[wave][tu];
    for (int n = 0; n < MLP_W; n += 4) {
      const float4 h4 = *reinterpret_cast<const float4*>(h + n);   // LDS, 16-byte aligned
      y = fmaf(w[n], h4.x, y);
      y = fmaf(w[n + 1], h4.y, y);
      y = fmaf(w[n + 2], h4.z, y);
      y = fmaf(w[n + 3], h4.w, y);
    }
    if (Y) Y[(long)tb * MLP_OUT + to] = y;
    if (T) {
      const float d = y - T[(long)tb * MLP_OUT + to];
      se = d * d;
      dyv = 2.0f * d / (float)(MLP_OUT * (long)B);
      dY[(long)tb * MLP_OUT + to] = dyv;
    }
  }
  if (!T) return;
  if (dZ) {
    float w0[MLP_OUT], w1[MLP_OUT];
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) {
      w0[o] = W5[o * MLP_W + lane];
      w1[o] = W5[o * MLP_W + 64 + lane];
    }
#pragma unroll
    for (int u = 0; u < OUT_SPW; u++) {
      // sample u's dY from lanes 5u..5u+4, then dZ4 for columns lane and lane + 64
      float dh0 = 0.f, dh1 = 0.f;
#pragma unroll
      for (int o = 0; o < MLP_OUT; o++) {
        const float dyo = __shfl(dyv, MLP_OUT * u + o);
        dh0 = fmaf(dyo, w0[o], dh0);
        dh1 = fmaf(dyo, w1[o], dh1);
      }
      const int b = b0 + u;
      if (b < B) {
        const float dz0 = dh0 * dsilu(zv[u][0]), dz1 = dh1 * dsilu(zv[u][1]);
        const long e = (long)b * MLP_W + lane;
        dZ[e] = dz0;
        dZ[e + 64] = dz1;
        dZbf[e] = f2bf(dz0);
        dZbf[e + 64] = f2bf(dz1);
      }
    }
  }
  // the block's squared error: lanes 0..19 of each wave in order, then the four waves
  float x = 0.f;
#pragma unroll
  for (int t = 0; t < OUT_SPW * MLP_OUT; t++) x += __shfl(se, t);
  if (lane == 0) red[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) loss[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// *loss = the n block sums of k_mlp_out in a fixed order: thread t adds blocks t, t + 256, ...,
// then a fixed tree over the 256 threads
__global__ __launch_bounds__(256) void k_loss_reduce(int n, const float* __restrict__ part, float* __restrict__ loss) {
  __shared__ float red[256];
  float x = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) x += part[i];
  red[threadIdx.x] = x;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0];
}

// small weight gradients, sums over the batch per column j of the width-128 layer:
//   mode 0 (layer 5): dW5[o][j] = sum_b dY[b][o] silu(Z4[b][j]),  db5[j] = sum_b dY[b][j] (j < 5)
//   mode 1 (layer 1): dW1[j][c] = sum_b dZ1[b][j] X[b][c],        db1[j] = sum_b dZ1[b][j]
//   mode 2 (bias):    db[j]     = sum_b dZ[b][j]
// k_col_partials: workgroup g sums rows [g R, (g + 1) R) (R = COL_ROWS), thread t column
// t % 128 over the rows of parity t / 128 (coalesced rows), into part[g][j][0..5];
// k_col_reduce adds the partials of every column in a fixed order (deterministic): a
// workgroup per 16 columns, 64 threads per column each taking every 64th workgroup's partial,
// then the 64 sums in a fixed tree (eight groups of eight).
constexpr int COL_ROWS = 64, COL_Q = MLP_OUT + 1, COL_SPLIT = 64, COL_RJ = 16, COL_RB = 8;
__global__ __launch_bounds__(256) void k_col_partials(int B, int mode, const float* __restrict__ dY, const float* __restrict__ Z4,
                                                      const float* __restrict__ dZ, const float* __restrict__ X,
                                                      float* __restrict__ part) {
  __shared__ float red[128][COL_Q];
  const int tid = threadIdx.x, j = tid & (MLP_W - 1), par = tid >> 7;
  const int b0 = blockIdx.x * COL_ROWS, b1 = min(B, b0 + COL_ROWS);
  float s[COL_Q];
#pragma unroll
  for (int q = 0; q < COL_Q; q++) s[q] = 0.f;
  // rows in batches of COL_RB: the batch's loads issued together, then added row by row in
  // order (the same sums as the plain loop, without a memory latency per row)
  int b = b0 + par;
  if (mode == 0) {
    for (; b + 2 * (COL_RB - 1) < b1; b += 2 * COL_RB) {
      float h[COL_RB], y[COL_RB][MLP_OUT], yj[COL_RB];
#pragma unroll
      for (int u = 0; u < COL_RB; u++) {
        const long bb = b + 2 * u;
        h[u] = Z4[bb * MLP_W + j];
#pragma unroll
        for (int o = 0; o < MLP_OUT; o++) y[u][o] = dY[bb * MLP_OUT + o];
        yj[u] = j < MLP_OUT ? dY[bb * MLP_OUT + j] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < COL_RB; u++) {
        const float hs = silu(h[u]);
#pragma unroll
        for (int o = 0; o < MLP_OUT; o++) s[o] = fmaf(y[u][o], hs, s[o]);
        if (j < MLP_OUT) s[MLP_OUT] += yj[u];
      }
    }
  } else {
    for (; b + 2 * (COL_RB - 1) < b1; b += 2 * COL_RB) {
      float d[COL_RB], x0[COL_RB], x1[COL_RB];
#pragma unroll
      for (int u = 0; u < COL_RB; u++) {
        const long bb = b + 2 * u;
        d[u] = dZ[bb * MLP_W + j];
        if (mode == 1) {
          x0[u] = X[2 * bb];
          x1[u] = X[2 * bb + 1];
        }
      }
#pragma unroll
      for (int u = 0; u < COL_RB; u++) {
        if (mode == 1) {
          s[0] = fmaf(d[u], x0[u], s[0]);
          s[1] = fmaf(d[u], x1[u], s[1]);
        }
        s[2] += d[u];
      }
    }
  }
  for (; b < b1; b += 2) {
    if (mode == 0) {
      const float h = silu(Z4[(long)b * MLP_W + j]);
#pragma unroll
      for (int o = 0; o < MLP_OUT; o++) s[o] = fmaf(dY[(long)b * MLP_OUT + o], h, s[o]);
      if (j < MLP_OUT) s[MLP_OUT] += dY[(long)b * MLP_OUT + j];
    } else {
      const float d = dZ[(long)b * MLP_W + j];
      if (mode == 1) {
        s[0] = fmaf(d, X[2 * b], s[0]);
        s[1] = fmaf(d, X[2 * b + 1], s[1]);
      }
      s[2] += d;
    }
  }
  if (par == 1)
#pragma unroll
    for (int q = 0; q < COL_Q; q++) red[j][q] = s[q];
  __syncthreads();
  if (par == 0)
#pragma unroll
    for (int q = 0; q < COL_Q; q++) part[((long)blockIdx.x * MLP_W + j) * COL_Q + q] = s[q] + red[j][q];
}

__global__ __launch_bounds__(COL_RJ * COL_SPLIT) void k_col_reduce(int nblk, int mode, const float* __restrict__ part,
                                                                     float* __restrict__ G, int bias_off) {
  __shared__ float red[COL_SPLIT][COL_RJ][COL_Q];
  const int jl = threadIdx.x % COL_RJ, h = threadIdx.x / COL_RJ, j = blockIdx.x * COL_RJ + jl;
  float s[COL_Q];
#pragma unroll
  for (int q = 0; q < COL_Q; q++) s[q] = 0.f;
  // chain h adds the partials of workgroups h, h + COL_SPLIT, ... in that order, four loaded
  // together
  int g = h;
  for (; g + 3 * COL_SPLIT < nblk; g += 4 * COL_SPLIT) {
    float v[4][COL_Q];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int q = 0; q < COL_Q; q++) v[u][q] = part[((long)(g + u * COL_SPLIT) * MLP_W + j) * COL_Q + q];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int q = 0; q < COL_Q; q++) s[q] += v[u][q];
  }
  for (; g < nblk; g += COL_SPLIT)
#pragma unroll
    for (int q = 0; q < COL_Q; q++) s[q] += part[((long)g * MLP_W + j) * COL_Q + q];
#pragma unroll
  for (int q = 0; q < COL_Q; q++) red[h][jl][q] = s[q];
  __syncthreads();
  // the chains' sums in a fixed tree: eight groups of eight in order, then the eight groups
  if (h < COL_SPLIT / 8) {
#pragma unroll
    for (int q = 0; q < COL_Q; q++) {
      float x = red[8 * h][jl][q];
#pragma unroll
      for (int k = 1; k < 8; k++) x += red[8 * h + k][jl][q];
      red[8 * h][jl][q] = x;
    }
  }
  __syncthreads();
  if (h != 0) return;
#pragma unroll
  for (int q = 0; q < COL_Q; q++) {
    float x = red[0][jl][q];
#pragma unroll
    for (int k = 1; k < COL_SPLIT / 8; k++) x += red[8 * k][jl][q];
    s[q] = x;
  }
  if (mode == 0) {
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) G[MLP_OFF_W5 + o * MLP_W + j] = s[o];
    if (j < MLP_OUT) G[MLP_OFF_B5 + j] = s[MLP_OUT];
  } else if (mode == 1) {
    G[MLP_OFF_W1 + 2 * j] = s[0];
    G[MLP_OFF_W1 + 2 * j + 1] = s[1];
    G[MLP_OFF_B1 + j] = s[2];
  } else {
    G[bias_off + j] = s[2];
  }
}

// Adam on the FP32 master parameters; BF16 copies of the three hidden weight matrices.
__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ P, const float* __restrict__ G, float* __restrict__ Mo,
                                              float* __restrict__ V, float lr, float b1, float b2, float eps, float c1, float c2,
                                              uint16_t* __restrict__ Wbf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = fmaf(b1, Mo[i], (1.f - b1) * g);
  const float v = fmaf(b2, V[i], (1.f - b2) * g * g);
  Mo[i] = m;
  V[i] = v;
  const float p = P[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
  P[i] = p;
  // BF16 copies: W2 at 0, W3 at W*W, W4 at 2*W*W
  if (i >= MLP_OFF_W2 && i < MLP_OFF_W2 + MLP_W * MLP_W) Wbf[i - MLP_OFF_W2] = f2bf(p);
  else if (i >= MLP_OFF_W3 && i < MLP_OFF_W3 + MLP_W * MLP_W) Wbf[MLP_W * MLP_W + i - MLP_OFF_W3] = f2bf(p);
  else if (i >= MLP_OFF_W4 && i < MLP_OFF_W4 + MLP_W * MLP_W) Wbf[2 * MLP_W * MLP_W + i - MLP_OFF_W4] = f2bf(p);
}

__global__ __launch_bounds__(256) void k_to_bf16(int n, const float* __restrict__ P, uint16_t* __restrict__ Wbf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * MLP_W * MLP_W) return;
  const int l = i / (MLP_W * MLP_W), r = i % (MLP_W * MLP_W);
  const int off = l == 0 ? MLP_OFF_W2 : (l == 1 ? MLP_OFF_W3 : MLP_OFF_W4);
  (void)n;
  Wbf[i] = f2bf(P[off + r]);
}

// rows idx[b] of (X, T) into the batch buffers
__global__ __launch_bounds__(256) void k_gather(int B, const int* __restrict__ idx, const float* __restrict__ X, const float* __restrict__ T,
                                                float* __restrict__ Xb, float* __restrict__ Tb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const long i = idx[b];
  Xb[2 * b] = X[2 * i];
  Xb[2 * b + 1] = X[2 * i + 1];
#pragma unroll
  for (int o = 0; o < MLP_OUT; o++) Tb[(long)b * MLP_OUT + o] = T[i * MLP_OUT + o];
}

// ---------------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------------
static inline int cdivi(long a, long b) { return (int)((a + b - 1) / b); }

// K splits of a product: ksplit rounded so that every split is a whole number of K chunks
static inline int ksplit_count(int K, int ksplit) {
  int kchunk = cdivi(K, ksplit < 1 ? 1 : ksplit);
  kchunk = ((kchunk + MG_K - 1) / MG_K) * MG_K;
  return cdivi(K, kchunk);
}

// epilogue epi with aux / outbf (EPI_STORE: accumulate selects C += ...).  A K split
// (EPI_STORE only) stores its partial products in `part` (ksplit_count(K, ksplit) x M x N
// floats) and adds them to C in split order; with part null they are added atomically into
// C, which must then hold zeros or the sum to add to.
static hipError_t gemm_bf16_epi(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn,
                                float* C, long ldc, int accumulate, int ksplit, int epi, const float* aux, uint16_t* outbf,
                                hipStream_t st, float* part = nullptr) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (ksplit < 1 || epi != EPI_STORE) ksplit = 1;
  int kchunk = cdivi(K, ksplit);
  kchunk = ((kchunk + MG_K - 1) / MG_K) * MG_K;
  const int kz = cdivi(K, kchunk);
  if (kz == 1) part = nullptr;
  hipLaunchKernelGGL(k_gemm_bf16, dim3(cdivi(M, MG_T), cdivi(N, MG_T), kz), dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C,
                     ldc, accumulate, kchunk, epi, aux, outbf, part);
  if (part)
    hipLaunchKernelGGL(k_ksplit_reduce, dim3(cdivi((long)M * N, 256)), dim3(256), 0, st, M, N, kz, (const float*)part, C, ldc,
                       accumulate);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_zero_f32(float* __restrict__ x, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = 0.f;
}

hipError_t gemm_bf16(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C,
                     long ldc, int accumulate, int ksplit, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (ksplit > 1 && !accumulate && cdivi(K, ((cdivi(K, ksplit) + MG_K - 1) / MG_K) * MG_K) > 1) {
    const long n = (long)M * ldc;
    hipLaunchKernelGGL(k_zero_f32, dim3((int)std::min<long>(cdivi(n, 256), 1024)), dim3(256), 0, st, C, n);
  }
  return gemm_bf16_epi(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, accumulate, ksplit, EPI_STORE, nullptr, nullptr, st);
}

static inline int col_blocks(int B) { return cdivi(B, COL_ROWS); }
// K splits of the weight-gradient products dW = dZ^T H (K = the batch)
static inline int grad_ksplit(int B) { return cdivi(B, 256); }

size_t mlp_workspace_bytes(int B) {
  const size_t f = sizeof(float), h = sizeof(uint16_t), W = MLP_W;
  return f * (size_t)B * W * 4          // Z1..Z4
         + h * (size_t)B * W * 3        // H1..H3 (BF16)
         + f * (size_t)B * W            // dZ
         + h * (size_t)B * W * 2        // dZ (BF16), two buffers: a layer's dZ_prev GEMM reads one, writes the other
         + f * (size_t)B * (2 + 2 * MLP_OUT)   // X batch, T batch, dY
         + f * (size_t)B * MLP_OUT      // Y
         + f * (size_t)col_blocks(B) * W * COL_Q   // column partial sums
         + f * (size_t)ksplit_count(B, grad_ksplit(B)) * W * W   // K-split partials of a weight gradient
         + f * (size_t)cdivi(B, OUT_SPB)    // loss per block of k_mlp_out
         + 64;
}

struct MlpWs {
  float *Z[4], *dZ, *Xb, *Tb, *dY, *Y, *part, *kpart, *lpart, *loss;
  uint16_t *H[3], *dZbf[2];
  MlpWs(void* base, int B) {
    char* p = (char*)base;
    const size_t W = MLP_W;
    for (int l = 0; l < 4; l++) { Z[l] = (float*)p; p += sizeof(float) * B * W; }
    for (int l = 0; l < 3; l++) { H[l] = (uint16_t*)p; p += sizeof(uint16_t) * B * W; }
    dZ = (float*)p; p += sizeof(float) * B * W;
    for (int l = 0; l < 2; l++) { dZbf[l] = (uint16_t*)p; p += sizeof(uint16_t) * B * W; }
    Xb = (float*)p; p += sizeof(float) * B * 2;
    Tb = (float*)p; p += sizeof(float) * B * MLP_OUT;
    dY = (float*)p; p += sizeof(float) * B * MLP_OUT;
    Y = (float*)p; p += sizeof(float) * B * MLP_OUT;
    part = (float*)p; p += sizeof(float) * (size_t)col_blocks(B) * W * COL_Q;
    kpart = (float*)p; p += sizeof(float) * (size_t)ksplit_count(B, grad_ksplit(B)) * W * W;
    lpart = (float*)p; p += sizeof(float) * (size_t)cdivi(B, OUT_SPB);
    loss = (float*)p;
  }
};

static void col_sums(int B, int mode, const float* dY, const float* Z4, const float* dZ, const float* X, float* part, float* G,
                     int bias_off, hipStream_t st) {
  hipLaunchKernelGGL(k_col_partials, dim3(col_blocks(B)), dim3(256), 0, st, B, mode, dY, Z4, dZ, X, part);
  hipLaunchKernelGGL(k_col_reduce, dim3(MLP_W / COL_RJ), dim3(COL_RJ * COL_SPLIT), 0, st, col_blocks(B), mode, part, G, bias_off);
}

static hipError_t forward(const float* P, const uint16_t* Wbf, const float* X, int B, MlpWs& w, hipStream_t st) {
  const int g = cdivi((long)B * MLP_W, 256);
  hipLaunchKernelGGL(k_mlp_in, dim3(g), dim3(256), 0, st, B, X, P, w.Z[0], w.H[0]);
  const int boff[3] = {MLP_OFF_B2, MLP_OFF_B3, MLP_OFF_B4};
  for (int l = 0; l < 3; l++) {
    // Z_{l+2} = H_{l+1} W_{l+2}^T + b: A = H (B x W, row-major), B(k, n) = W[n][k]; bias and
    // SiLU in the epilogue (layer 4's BF16 activations are not used: they land in dZbf[0])
    uint16_t* Hout = l < 2 ? w.H[l + 1] : w.dZbf[0];
    hipError_t e = gemm_bf16_epi(B, MLP_W, MLP_W, w.H[l], MLP_W, 1, Wbf + (size_t)l * MLP_W * MLP_W, 1, MLP_W, w.Z[l + 1], MLP_W, 0,
                                 1, EPI_BIAS_SILU, P + boff[l], Hout, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t mlp_forward(const float* P, const uint16_t* Wbf, const float* X, int B, float* Y, void* ws, hipStream_t st) {
  if (B <= 0) return hipSuccess;
  MlpWs w(ws, B);
  hipError_t e = forward(P, Wbf, X, B, w, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_out, dim3(cdivi(B, OUT_SPB)), dim3(256), 0, st, B, w.Z[3], P, Y, (const float*)nullptr, (float*)nullptr,
                     (float*)nullptr, (float*)nullptr, (uint16_t*)nullptr);
  return hipGetLastError();
}

hipError_t mlp_grad(const float* P, const uint16_t* Wbf, const float* X, const float* T, int B, float* G, float* loss, void* ws,
                    hipStream_t st) {
  if (B <= 0) return hipSuccess;
  MlpWs w(ws, B);
  hipError_t e = forward(P, Wbf, X, B, w, st);
  if (e != hipSuccess) return e;
  // G starts at zero (the hidden weight gradients are added into it)
  hipLaunchKernelGGL(k_zero_f32, dim3(cdivi(MLP_NPARAM, 256)), dim3(256), 0, st, G, (long)MLP_NPARAM);
  // layer 5 forward, loss, and backward to dZ4 (dZbf[0])
  hipLaunchKernelGGL(k_mlp_out, dim3(cdivi(B, OUT_SPB)), dim3(256), 0, st, B, w.Z[3], P, w.Y, T, w.dY, w.lpart, w.dZ, w.dZbf[0]);
  hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, st, cdivi(B, OUT_SPB), (const float*)w.lpart, loss);
  // layer 5
  col_sums(B, 0, w.dY, w.Z[3], nullptr, nullptr, w.part, G, 0, st);   // dW5 and db5
  // layers 4, 3, 2
  const int woff[3] = {MLP_OFF_W2, MLP_OFF_W3, MLP_OFF_W4};
  const int boff[3] = {MLP_OFF_B2, MLP_OFF_B3, MLP_OFF_B4};
  const int ksplit = grad_ksplit(B);
  int cur = 0;
  for (int l = 2; l >= 0; l--) {
    // dW = dZ^T H_prev: A(m, k) = dZ[k][m], B(k, n) = H_prev[k][n]; K = batch, split over
    // workgroups, the splits added in a fixed order (training runs repeat bit for bit)
    e = gemm_bf16_epi(MLP_W, MLP_W, B, w.dZbf[cur], 1, MLP_W, w.H[l], MLP_W, 1, G + woff[l], MLP_W, 1, ksplit, EPI_STORE, nullptr,
                      nullptr, st, w.kpart);
    if (e != hipSuccess) return e;
    col_sums(B, 2, nullptr, nullptr, w.dZ, nullptr, w.part, G, boff[l], st);
    // dZ_prev = (dZ W) * silu'(Z_prev): A = dZ (row-major), B(k, n) = W[k][n], the SiLU derivative in
    // the epilogue; it reads dZbf[cur] and writes dZbf[1 - cur]
    e = gemm_bf16_epi(B, MLP_W, MLP_W, w.dZbf[cur], MLP_W, 1, Wbf + (size_t)l * MLP_W * MLP_W, MLP_W, 1, w.dZ, MLP_W, 0, 1,
                      EPI_DSILU, w.Z[l], w.dZbf[1 - cur], st);
    if (e != hipSuccess) return e;
    cur = 1 - cur;
  }
  // layer 1
  col_sums(B, 1, nullptr, nullptr, w.dZ, X, w.part, G, 0, st);
  return hipGetLastError();
}

hipError_t mlp_adam(float* P, const float* G, float* M, float* V, uint16_t* Wbf, float lr, float b1, float b2, float eps, int t,
                    hipStream_t st) {
  const float c1 = 1.0f / (1.0f - powf(b1, (float)t)), c2 = 1.0f / (1.0f - powf(b2, (float)t));
  hipLaunchKernelGGL(k_adam, dim3(cdivi(MLP_NPARAM, 256)), dim3(256), 0, st, MLP_NPARAM, P, G, M, V, lr, b1, b2, eps, c1, c2, Wbf);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(const float* P, uint16_t* Wbf, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3(cdivi(3 * MLP_W * MLP_W, 256)), dim3(256), 0, st, MLP_NPARAM, P, Wbf);
  return hipGetLastError();
}

hipError_t mlp_gather(int B, const int* idx, const float* X, const float* T, float* Xb, float* Tb, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3(cdivi(B, 256)), dim3(256), 0, st, B, idx, X, T, Xb, Tb);
  return hipGetLastError();
}

}  // namespace tb
