// Field-surrogate MLP (mlp.hip): sizes, FP32 parameter layout and host drivers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace tb {

constexpr int MLP_W = 128;     // hidden width
constexpr int MLP_OUT = 5;     // outputs
// FP32 master parameters, one contiguous vector: W1 (128,2) b1 | W2 (128,128) b2 | W3 b3 |
// W4 b4 | W5 (5,128) b5, every W stored [out][in].
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_W * 2;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_W;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_W * MLP_W;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_W;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_W * MLP_W;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_W;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_W * MLP_W;
constexpr int MLP_OFF_W5 = MLP_OFF_B4 + MLP_W;
constexpr int MLP_OFF_B5 = MLP_OFF_W5 + MLP_OUT * MLP_W;
constexpr int MLP_NPARAM = MLP_OFF_B5 + MLP_OUT;   // 50565

// FP32 -> BF16, round to nearest even (NaN kept quiet), and back; SiLU and its derivative
__host__ __device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}
// logistic function with the hardware exponential and reciprocal (about 1 ulp each; an IEEE
// division here costs more than the rest of a hidden-layer element in the tracer)
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

// C (+)= op(A) op(B) in BF16 on MFMA with FP32 accumulation (strided operands, optional split-K)
hipError_t gemm_bf16(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C,
                     long ldc, int accumulate, int ksplit, hipStream_t st);
size_t mlp_workspace_bytes(int B);
// Y (B,5) = MLP(X (B,2)); Wbf: BF16 copies of W2..W4 (3 x 128 x 128)
hipError_t mlp_forward(const float* P, const uint16_t* Wbf, const float* X, int B, float* Y, void* ws, hipStream_t st);
// gradient of the mean squared error against T (B,5) into G (MLP_NPARAM), loss sum into *loss
hipError_t mlp_grad(const float* P, const uint16_t* Wbf, const float* X, const float* T, int B, float* G, float* loss, void* ws,
                    hipStream_t st);
// one Adam step (t >= 1) on P, refreshing the BF16 copies
hipError_t mlp_adam(float* P, const float* G, float* M, float* V, uint16_t* Wbf, float lr, float b1, float b2, float eps, int t,
                    hipStream_t st);
hipError_t mlp_to_bf16(const float* P, uint16_t* Wbf, hipStream_t st);
hipError_t mlp_gather(int B, const int* idx, const float* X, const float* T, float* Xb, float* Tb, hipStream_t st);

}  // namespace tb
