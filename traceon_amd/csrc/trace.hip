// RK45 (Fehlberg 4(5)) particle tracing on gfx950 — replaces trace_particle and its
// field adaptors (traceon/backend/tracing.c:20-174) and the Python block loop
// (traceon/backend/__init__.py:382-420).
//
// Persistent kernel, one lane per particle.  Every lane of a wave runs the same
// attempted step (6 stages, 6 field evaluations, error norm, controller), so the only
// divergence is the per-lane accept/reject select.  A lane whose particle leaves the
// bounds (or reaches its step cap / trajectory capacity) retires: it writes its results
// and takes the next particle index from a global counter (one atomic per wave, handed
// out by prefix popcount), so the wave stays full until the queue is empty.  For the
// direct BEM field the source loop is wave-uniform (scalar loads shared by all lanes).
//
// The landing state on a plane is tracked on the fly: after every accepted step the
// signed distances of the previous and new point are compared exactly as
// plane_intersection (tracing.c:184-223) does when it scans the stored trajectory
// backwards; each crossing writes its interpolated state to the particle's landing row
// (NaN from the particle's start), so the row ends with the last crossing.  An optional
// second plane is tracked the same way.
#include <cstdlib>
#include <mutex>
#include "radial_math.h"
#include "trace_kernels.h"
#include "mlp_kernels.h"

namespace tb {

__constant__ static const double RK_B[6][5] = {
    {0, 0, 0, 0, 0},
    {2. / 9., 0, 0, 0, 0},
    {1. / 12., 1. / 4., 0, 0, 0},
    {69. / 128., -243. / 128., 135. / 64., 0, 0},
    {-17. / 12., 27. / 4., -27. / 5., 16. / 15., 0},
    {65. / 432., -5. / 16., 13. / 16., 4. / 27., 5. / 144.}};
__constant__ static const double RK_CH[6] = {47. / 450., 0., 12. / 25., 32. / 225., 1. / 30., 6. / 25.};
__constant__ static const double RK_CT[6] = {-1. / 150., 0., 3. / 100., -16. / 75., -1. / 20., 6. / 25.};

// ---------------------------------------------------------------------------------
// field evaluation
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void bem_field(const SourceSet& S, double x, double y, double z, double out[3], int s0, int ds) {
  const double r0 = sqrt(x * x + y * y);
  double Er = 0.0, Ez = 0.0;
  for (int s = s0; s < S.n; s += ds) {
    double gr, gz;
    ring_grad(r0, S.r[s] - r0, S.z[s] - z, gr, gz);
    Er = fma(-S.w[s], gr, Er);
    Ez = fma(-S.w[s], gz, Ez);
  }
  if (r0 >= AXIS_EPS) { out[0] = x / r0 * Er; out[1] = y / r0 * Er; }
  else { out[0] = 0.0; out[1] = 0.0; }
  out[2] = Ez;
}

__device__ __forceinline__ void coil_field(const SourceSet& S, double x, double y, double z, double out[3], int s0, int ds) {
  const double r = sqrt(x * x + y * y);
  double Hr = 0.0, Hz = 0.0;
  for (int s = s0; s < S.n; s += ds) {
    double hr, hz;
    current_ring_field(r, z, S.r[s], S.z[s], hr, hz);
    Hr = fma(S.w[s], hr, Hr);
    Hz = fma(S.w[s], hz, Hz);
  }
  if (r >= AXIS_EPS) { out[0] = x / r * Hr; out[1] = y / r * Hr; }
  else { out[0] = 0.0; out[1] = 0.0; }
  out[2] = Hz;
}

__device__ __forceinline__ void to_local(const FieldComp& c, const double p[3], double q[3]) {
  for (int i = 0; i < 3; i++) q[i] = c.inv[4 * i] * p[0] + c.inv[4 * i + 1] * p[1] + c.inv[4 * i + 2] * p[2] + c.inv[4 * i + 3];
}
__device__ __forceinline__ void rotate(const FieldComp& c, double v[3]) {
  double a = c.basis[0] * v[0] + c.basis[1] * v[1] + c.basis[2] * v[2];
  double b = c.basis[3] * v[0] + c.basis[4] * v[1] + c.basis[5] * v[2];
  double d = c.basis[6] * v[0] + c.basis[7] * v[1] + c.basis[8] * v[2];
  v[0] = a; v[1] = b; v[2] = d;
}
__device__ __forceinline__ bool in_bounds(const double* b, int mode, const double q[3]) {
  if (mode == BOUNDS_NONE) return true;
  if (mode == BOUNDS_STRICT)   // field_radial_traceable (tracing.c:136-138)
    return b[0] < q[0] && q[0] < b[1] && b[2] < q[1] && q[1] < b[3] && b[4] < q[2] && q[2] < b[5];
  // Field._within_field_bounds (field.py:220-221)
  return b[0] <= q[0] && q[0] <= b[1] && b[2] <= q[1] && q[1] <= b[3] && b[4] <= q[2] && q[2] <= b[5];
}

// Ring-source part of a component at local point q (sources s0, s0+ds, ...).
__device__ void bem_component(const FieldComp& c, const double q[3], double e[3], double h[3], int s0, int ds) {
  if (c.elec.n) bem_field(c.elec, q[0], q[1], q[2], e, s0, ds);
  if (c.mag.n) bem_field(c.mag, q[0], q[1], q[2], h, s0, ds);
  if (c.cur.n) {
    double hc[3];
    if (c.transform) {
      // FieldRadialBEM.magnetostatic_field_at_local_point (field.py:672-695) evaluates the
      // coil field through current_field_at_point on the already local point: that maps
      // it once more, checks the bounds there and rotates by the basis.
      double q2[3];
      to_local(c, q, q2);
      if (in_bounds(c.bounds, c.bounds_mode, q2)) { coil_field(c.cur, q2[0], q2[1], q2[2], hc, s0, ds); rotate(c, hc); }
      else { hc[0] = hc[1] = hc[2] = 0.0; }
    } else {
      coil_field(c.cur, q[0], q[1], q[2], hc, s0, ds);
    }
    h[0] += hc[0]; h[1] += hc[1]; h[2] += hc[2];
  }
}

// Learned off-axis correction of an axial component (FieldSurrogate): per lane, FP32
// arithmetic with FP32 weights read as wave-uniform scalars, BF16 activations in this
// wave's LDS columns g_mlp_act[layer parity][neuron][lane] (as the training forward stores
// them; the workgroups that evaluate corrections are one wave);
// layer 5 is folded into layer 4 so that layer 4's output stays FP32.  Adds
// y_k * w(r) * s_k(z), w = (r/r_max)^p, to (E_r, E_z, H_r, H_z) for r >= r_lo.
constexpr int TMLP_B1 = 256, TMLP_W2 = 384, TMLP_W3 = TMLP_W2 + 16512, TMLP_W4 = TMLP_W3 + 16512, TMLP_W5 = TMLP_W4 + 16512;
// LDS of the learned correction (the workgroups that evaluate corrections are one wave):
// mlp_correct: g_mlp_act[layer parity][neuron][lane], BF16; mlp_wave: activations A0 and
// A1 (64 points x 136 BF16, the row stride that keeps the MFMA operand reads conflict-free)
// and, in place of A1, the FP32 output of the last hidden layer, 64 of its neurons at a time
// (64 x 65 floats).  34 KiB: four one-wave workgroups per CU.
constexpr int MW_S = 136, MW_SF = 65;
static_assert(64 * MW_SF * 4 <= 64 * MW_S * 2, "the last layer's half fits in place of A1");
__shared__ __attribute__((aligned(16))) unsigned char g_mlp_lds[64 * MW_S * 2 * 2];
#define g_mlp_act (reinterpret_cast<uint16_t*>(g_mlp_lds))
typedef const __attribute__((address_space(4))) float cfloat_t;   // read-only weights: scalar loads

__device__ __forceinline__ cfloat_t* uniform_weights(const float* p) {
  // the component is the same for every lane of the wave: make the pointer wave-uniform
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<cfloat_t*>(((uint64_t)hi << 32) | lo);
}

__device__ void mlp_correct(const FieldComp& c, const double q[3], double e[3], double h[3], int lane) {
  const double r = sqrt(q[0] * q[0] + q[1] * q[1]);
  const double r_lo = c.mlp_geom[0], r_max = c.mlp_geom[1], z0 = c.mlp_geom[2], z1 = c.mlp_geom[3];
  if (r < r_lo || q[2] < z0 || q[2] > z1) return;
  const float u = (float)(2.0 * (r - r_lo) / (r_max - r_lo) - 1.0);
  const float v = (float)(2.0 * (q[2] - z0) / (z1 - z0) - 1.0);
  cfloat_t* W = uniform_weights(c.mlp);
  for (int j = 0; j < 128; j++)
    g_mlp_act[j * 64 + lane] = f2bf(silu(fmaf(W[2 * j], u, fmaf(W[2 * j + 1], v, W[TMLP_B1 + j]))));
  float y[5];
#pragma unroll
  for (int o = 0; o < 5; o++) y[o] = W[TMLP_W5 + 640 + o];
  for (int layer = 0; layer < 3; layer++) {
    cfloat_t* WT = W + TMLP_W2 + layer * 16512;   // (in, out), then the 128 biases
    const int in0 = (layer & 1) * 128 * 64, out0 = ((layer + 1) & 1) * 128 * 64;
    for (int jc = 0; jc < 128; jc += 16) {
      float acc[16];
#pragma unroll
      for (int t = 0; t < 16; t++) acc[t] = WT[16384 + jc + t];
      for (int i = 0; i < 128; i++) {
        const float x = bf2f(g_mlp_act[in0 + i * 64 + lane]);
#pragma unroll
        for (int t = 0; t < 16; t++) acc[t] = fmaf(WT[i * 128 + jc + t], x, acc[t]);
      }
      if (layer < 2) {
#pragma unroll
        for (int t = 0; t < 16; t++) g_mlp_act[out0 + (jc + t) * 64 + lane] = f2bf(silu(acc[t]));
      } else {
#pragma unroll
        for (int t = 0; t < 16; t++) {
          const float hs = silu(acc[t]);
#pragma unroll
          for (int o = 0; o < 5; o++) y[o] = fmaf(W[TMLP_W5 + o * 128 + jc + t], hs, y[o]);
        }
      }
    }
  }
  // s_k(z): linear between bin centres; w(r)
  const int nz = c.mlp_nz;
  double tz = (q[2] - z0) / (z1 - z0) * nz - 0.5;
  tz = tz < 0.0 ? 0.0 : (tz > nz - 1.0 ? nz - 1.0 : tz);
  int i0 = (int)tz;
  if (i0 > nz - 2) i0 = nz > 1 ? nz - 2 : 0;
  const double f = tz - i0;
  const int i1 = nz > 1 ? i0 + 1 : 0;
  const double w = pow(r / r_max, c.mlp_geom[4]);
  double d[5];
#pragma unroll
  for (int k = 1; k < 5; k++) d[k] = (double)y[k] * w * (c.mlp_zs[k * nz + i0] * (1.0 - f) + c.mlp_zs[k * nz + i1] * f);
  const double cx = r > 0.0 ? q[0] / r : 0.0, cy = r > 0.0 ? q[1] / r : 0.0;
  e[0] += d[1] * cx; e[1] += d[1] * cy; e[2] += d[2];
  h[0] += d[3] * cx; h[1] += d[3] * cy; h[2] += d[4];
}

// The surrogate's embedded error check (FieldSurrogate.calibrate): its predicted error at
// (r, z) relative to the field's scale, w(r) max_k sigma_k(z), over the tolerance (rows 5..9 of
// mlp_zs hold sigma_k / tolerance); above 1 the point is evaluated by the ring sources.
__device__ __forceinline__ bool mlp_error_exceeds(const FieldComp& c, const double q[3], double r) {
  const double r_lo = c.mlp_geom[0], r_max = c.mlp_geom[1], z0 = c.mlp_geom[2], z1 = c.mlp_geom[3];
  if (r < r_lo || q[2] < z0 || q[2] > z1) return false;
  const int nz = c.mlp_nz;
  double tz = (q[2] - z0) / (z1 - z0) * nz - 0.5;
  tz = tz < 0.0 ? 0.0 : (tz > nz - 1.0 ? nz - 1.0 : tz);
  int i0 = (int)tz;
  if (i0 > nz - 2) i0 = nz > 1 ? nz - 2 : 0;
  const double f = tz - i0;
  const int i1 = nz > 1 ? i0 + 1 : 0;
  const double* Se = c.mlp_zs + 5 * nz;
  double m = 0.0;
#pragma unroll
  for (int k = 1; k < 5; k++) {
    const double v = Se[k * nz + i0] * (1.0 - f) + Se[k * nz + i1] * f;
    m = v > m ? v : m;
  }
  return pow(r / r_max, c.mlp_geom[4]) * m > 1.0;
}

typedef short mw_s4 __attribute__((ext_vector_type(4)));
typedef float mw_f4 __attribute__((ext_vector_type(4)));

// The learned correction for all 64 lanes of a wave at once (each lane's own point q),
// the hidden layers on MFMA (v_mfma_f32_16x16x16bf16_1k, the training's layout and BF16
// rounding points): the input layer per lane in FP32 with BF16 activations, three 64 x 128
// x 128 products with FP32 accumulation, bias and SiLU, BF16 activations in between, the last
// SiLU kept in FP32 and the output layer per lane.  Must be called by every lane of the wave
// (uniform control flow); lanes with need = false add nothing.  The points that need the
// correction are packed into the first rows (a lane's row is its rank among them), and only
// the 16-row tiles that hold them are multiplied: each row's result does not depend on its
// position, so a wave with 16 needing lanes or fewer runs a quarter of the products.
// Operand layout of the instruction: lane l holds A[l%16][4(l/16)+i] (points x inputs),
// B[4(l/16)+i][l%16] (inputs x outputs: W[out][in] read row-wise) and D[4(l/16)+v][l%16].
__device__ void mlp_wave(const FieldComp& c, const double q[3], bool need, int lane, double e[3], double h[3]) {
  uint16_t* A0 = reinterpret_cast<uint16_t*>(g_mlp_lds);
  uint16_t* A1 = reinterpret_cast<uint16_t*>(g_mlp_lds + 64 * MW_S * 2);
  float* Fo = reinterpret_cast<float*>(g_mlp_lds + 64 * MW_S * 2);
  const double r = sqrt(q[0] * q[0] + q[1] * q[1]);
  const double r_lo = c.mlp_geom[0], r_max = c.mlp_geom[1], z0 = c.mlp_geom[2], z1 = c.mlp_geom[3];
  const float u = need ? (float)(2.0 * (r - r_lo) / (r_max - r_lo) - 1.0) : 0.f;
  const float v = need ? (float)(2.0 * (q[2] - z0) / (z1 - z0) - 1.0) : 0.f;
  cfloat_t* W = uniform_weights(c.mlp);
  const unsigned long long needers = __ballot(need);
  const int count = __popcll(needers);
  const int row = need ? __popcll(needers & ((1ull << lane) - 1ull)) : 0;   // rank among the needing lanes
  const int mt = (count + 15) >> 4;                                          // 16-row tiles in use
  // input layer: this lane's row of A0
  if (need) {
#pragma unroll 4
    for (int j = 0; j < 128; j += 4) {
      uint64_t pk = 0;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const float z = fmaf(W[2 * (j + t)], u, fmaf(W[2 * (j + t) + 1], v, W[TMLP_B1 + j + t]));
        pk |= (uint64_t)f2bf(silu(z)) << (16 * t);
      }
      *reinterpret_cast<uint64_t*>(&A0[row * MW_S + j]) = pk;
    }
  }
  __syncthreads();
  const int lr = lane & 15, lq = lane >> 4;
  float y[5];
#pragma unroll
  for (int o = 0; o < 5; o++) y[o] = W[TMLP_W5 + 640 + o];
  // layers 0-1: all 128 neurons; layer 2 in two passes of 64 neurons, each followed by its
  // part of the output layer (per lane on its own point, neurons in order 0..127)
  for (int pass = 0; pass < 4; pass++) {
    const int layer = pass < 2 ? pass : 2;
    const uint16_t* Ain = layer == 1 ? A1 : A0;
    uint16_t* Aout = layer == 0 ? A1 : A0;
    const uint16_t* Wl = c.mlp_wbf + (size_t)layer * 16384;
    const float* bias = c.mlp + TMLP_W2 + layer * 16512 + 16384;
    const int nb0 = pass == 3 ? 4 : 0, nb1 = pass == 2 ? 4 : 8;
    for (int nb = nb0; nb < nb1; nb++) {
      mw_f4 acc[4];
#pragma unroll
      for (int mb = 0; mb < 4; mb++) acc[mb] = mw_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
      for (int kb = 0; kb < 8; kb++) {
        const mw_s4 bfr = *reinterpret_cast<const mw_s4*>(Wl + (16 * nb + lr) * 128 + 16 * kb + 4 * lq);
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
          if (mb < mt) {   // wave-uniform
            const mw_s4 afr = *reinterpret_cast<const mw_s4*>(Ain + (16 * mb + lr) * MW_S + 16 * kb + 4 * lq);
            acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afr, bfr, acc[mb], 0, 0, 0);
          }
        }
      }
      const int n = 16 * nb + lr;
      const float bn = bias[n];
#pragma unroll
      for (int mb = 0; mb < 4; mb++)
        if (mb < mt)
#pragma unroll
          for (int vv = 0; vv < 4; vv++) {
            const int pt = 16 * mb + 4 * lq + vv;
            const float a = silu(acc[mb][vv] + bn);
            if (layer < 2) Aout[pt * MW_S + n] = f2bf(a);
            else Fo[pt * MW_SF + n - 16 * nb0] = a;
          }
    }
    __syncthreads();
    if (layer == 2) {   // output layer over this pass's 64 neurons
      const int n0 = 16 * nb0;
      for (int n = 0; n < 64; n++) {
        const float hv = Fo[row * MW_SF + n];
#pragma unroll
        for (int o = 0; o < 5; o++) y[o] = fmaf(W[TMLP_W5 + o * 128 + n0 + n], hv, y[o]);
      }
      __syncthreads();   // the LDS is reused by the next pass or evaluation
    }
  }
  if (!need) return;
  // s_k(z): linear between bin centres; w(r)
  const int nz = c.mlp_nz;
  double tz = (q[2] - z0) / (z1 - z0) * nz - 0.5;
  tz = tz < 0.0 ? 0.0 : (tz > nz - 1.0 ? nz - 1.0 : tz);
  int i0 = (int)tz;
  if (i0 > nz - 2) i0 = nz > 1 ? nz - 2 : 0;
  const double f = tz - i0;
  const int i1 = nz > 1 ? i0 + 1 : 0;
  const double w = pow(r / r_max, c.mlp_geom[4]);
  double d[5];
#pragma unroll
  for (int k = 1; k < 5; k++) d[k] = (double)y[k] * w * (c.mlp_zs[k * nz + i0] * (1.0 - f) + c.mlp_zs[k * nz + i1] * f);
  const double cx = r > 0.0 ? q[0] / r : 0.0, cy = r > 0.0 ? q[1] / r : 0.0;
  e[0] += d[1] * cx; e[1] += d[1] * cy; e[2] += d[2];
  h[0] += d[3] * cx; h[1] += d[3] * cy; h[2] += d[4];
}

// Field of a program at p.  Ring-source sums run over sources s0, s0+ds, ...; terms without
// a source sum (axial, uniform) are added only when `lead` is set, so that partial results
// of several threads add up to the field and (0, 1, true) gives the whole field in one lane.
// MLP = true (one-wave workgroups) applies learned corrections of axial components inside
// their trust region and their ring sources outside it; MLP = false evaluates such a
// component as its plain axial field (programs with corrections are traced with MLP = true).
// SOURCES = false compiles no source sums (programs flagged PROGRAM_NO_SOURCES); a BEM
// component met there gives NaN.
// WAVE (with MLP): the correction is left to mlp_wave; *mlp_need reports whether this
// point gets it (one learned component, untransformed, as the tracer checks).  With
// bem_defer non-null, a point that the trust region or the error check sends to the ring
// sources gets no source sum here: *bem_defer is set and the caller sums them (bem_wave).
// LEAD = false compiles the lead terms out (lead is then false whatever is passed), so that
// a kernel that only sums sources carries no learned-correction code or LDS.
template <bool MLP = false, bool SOURCES = true, bool WAVE = false, bool LEAD = true>
__device__ void eval_field(const FieldProgram& F, const double p[3], double E[3], double H[3], int s0 = 0, int ds = 1,
                           bool lead = true, int lane = 0, bool* mlp_need = nullptr, bool* bem_defer = nullptr) {
  if (!LEAD) lead = false;
  E[0] = E[1] = E[2] = 0.0;
  H[0] = H[1] = H[2] = 0.0;
  if (WAVE) *mlp_need = false;
  if (WAVE && bem_defer) *bem_defer = false;
  if (!in_bounds(F.bounds, F.bounds_mode, p)) return;
  for (int ci = 0; ci < F.ncomp; ci++) {
    const FieldComp& c = F.comp[ci];
    double q[3] = {p[0], p[1], p[2]};
    if (c.transform) to_local(c, p, q);
    if (!in_bounds(c.bounds, c.bounds_mode, q)) continue;
    double e[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    if (c.kind == COMP_BEM) {
      if constexpr (SOURCES) bem_component(c, q, e, h, s0, ds);
      else e[0] = e[1] = e[2] = h[0] = h[1] = h[2] = __builtin_nan("");
    } else if (c.kind == COMP_AXIAL) {
      bool use_bem = false;
      if (MLP && c.mlp) {
        // the axial series is trusted up to r_max inside its interpolation range; the learned
        // correction adds to it where it was trained (mlp_correct) unless its predicted error
        // exceeds the tolerance; elsewhere: ring sources
        const double r = sqrt(q[0] * q[0] + q[1] * q[1]);
        const bool trust = r <= c.mlp_geom[5] && q[2] >= c.zs[0] && q[2] <= c.zs[c.nz - 1] && !mlp_error_exceeds(c, q, r);
        use_bem = !trust && (c.elec.n || c.mag.n || c.cur.n);
      }
      if (MLP && use_bem) {
        if (WAVE && bem_defer) *bem_defer = true;
        else if constexpr (SOURCES) bem_component(c, q, e, h, s0, ds);
        else e[0] = e[1] = e[2] = h[0] = h[1] = h[2] = __builtin_nan("");
      } else if (lead) {
        const int parts = c.axial_parts ? c.axial_parts : 3;
        if (parts & 1) axial_field(c.zs, c.nz, c.ce, q[0], q[1], q[2], e);
        if (parts & 2) axial_field(c.zs, c.nz, c.cm, q[0], q[1], q[2], h);
        if (LEAD && MLP && c.mlp) {
          if constexpr (WAVE) {
            const double r = sqrt(q[0] * q[0] + q[1] * q[1]);
            *mlp_need = r >= c.mlp_geom[0] && q[2] >= c.mlp_geom[2] && q[2] <= c.mlp_geom[3];
          } else if constexpr (LEAD) {
            mlp_correct(c, q, e, h, lane);
          }
        }
      }
    } else if (lead) {
      for (int i = 0; i < 3; i++) { e[i] = c.uE[i]; h[i] = c.uH[i]; }
    }
    if (c.transform) { rotate(c, e); rotate(c, h); }
    for (int i = 0; i < 3; i++) { E[i] = fma(c.factor, e[i], E[i]); H[i] = fma(c.factor, h[i], H[i]); }
  }
}

// ---------------------------------------------------------------------------------
// the persistent tracer
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double plane_dist(const double* p0, const double* n, double nn, const double y[6]) {
  return (n[0] * (p0[0] - y[0]) + n[1] * (p0[1] - y[1]) + n[2] * (p0[2] - y[2])) / nn;
}
__device__ __forceinline__ double signed_dist(const TraceParams& P, const double y[6]) {
  return plane_dist(P.plane_p, P.plane_n, P.plane_nn, y);
}
__device__ __forceinline__ double signed_dist2(const TraceParams& P, const double y[6]) {
  return plane_dist(P.plane2_p, P.plane2_n, P.plane2_nn, y);
}
// a crossing between y (distance kp) and yn (distance kn): the interpolated state into row
__device__ __forceinline__ void land_if_crossed(double kp, double kn, const double y[6], const double yn[6], double* row) {
  if ((kn > 0) != (kp > 0)) {
    const double diff = kp - kn;   // kappa(i) - kappa(i+1) with i the previous point
    const double f = -kn / diff, fp = kp / diff;
#pragma unroll
    for (int c = 0; c < 6; c++) row[c] = fp * yn[c] + f * y[c];
  }
}

__device__ __forceinline__ bool inside_trace_bounds(const double* b, const double y[6]) {
  // tracing.c:79-81, inclusive
  return b[0] <= y[0] && y[0] <= b[1] && b[2] <= y[1] && y[1] <= b[3] && b[4] <= y[2] && y[2] <= b[5];
}

// Ring-source sums of the points that the lanes in `mask` hold (component c, untransformed,
// so the local point is the point itself), one point at a time by the whole wave: lane l
// takes sources l, l + 64, ..., and the six partial sums meet in a butterfly over the wave.
// The owning lane adds factor * sum to its E, H.  Every lane of the wave calls it with the
// same mask (uniform control flow).  A lane's own sum would cost the wave a walk over every
// source however few of its lanes need one; here each point costs 1/64 of that.
__device__ void bem_wave(const FieldComp& c, unsigned long long mask, const double p[3], int lane, double E[3], double H[3]) {
  while (mask) {
    const int j = __ffsll((long long)mask) - 1;
    mask &= mask - 1ull;
    const double q[3] = {__shfl(p[0], j, 64), __shfl(p[1], j, 64), __shfl(p[2], j, 64)};
    double e[3] = {0.0, 0.0, 0.0}, h[3] = {0.0, 0.0, 0.0};
    bem_component(c, q, e, h, lane, 64);
    double v[6] = {e[0], e[1], e[2], h[0], h[1], h[2]};
#pragma unroll
    for (int k = 0; k < 6; k++)
      for (int o = 32; o >= 1; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    if (lane == j) {
      for (int k = 0; k < 3; k++) { E[k] = fma(c.factor, v[k], E[k]); H[k] = fma(c.factor, v[3 + k], H[k]); }
    }
  }
}

template <int N>
struct StageIndex { static constexpr int value = N; };

template <bool MLP, bool SOURCES>
__device__ __forceinline__ void trace_body(const TraceParams& P) {
  const int lane = threadIdx.x & 63;
  // one learned, untransformed component with BF16 weights: corrections on MFMA per wave
  const bool wave_mlp = MLP && P.field.ncomp == 1 && P.field.comp[0].mlp && P.field.comp[0].mlp_wbf && !P.field.comp[0].transform;
  // lane state
  int pid = -1;
  double y[6], h = 0.0, hmax = 0.0, t = 0.0, qm = 0.0;
  long long nacc = 0, natt = 0;
  double kprev = 0.0, kprev2 = 0.0;
  long long traj_n = 0;

  auto next_particle = [&](bool need) {
    // wave-aggregated queue pop for every lane with need == true
    unsigned long long mask = __ballot(need);
    int cnt = __popcll(mask);
    int base = 0;
    if (cnt) {
      const int leader = __ffsll((long long)mask) - 1;
      if (lane == leader) base = atomicAdd(P.counter, cnt);
      base = __shfl(base, leader, 64);
    }
    if (need) {
      const int rank = __popcll(mask & ((1ull << lane) - 1ull));
      const int k = base + rank;
      const int id = k < P.n ? (P.order ? P.order[k] : k) : -1;
      if ((unsigned)id < (unsigned)P.n) {   // an order entry outside 0..n-1 is skipped
        pid = id;
        for (int c = 0; c < 6; c++) y[c] = P.state0[6 * (size_t)id + c];
        t = P.t0 ? P.t0[id] : 0.0;
        qm = P.qm[P.qm_stride ? id : 0];
        const double speed = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
        hmax = (P.hmax0 && P.hmax0[id] > 0.0) ? P.hmax0[id] : STEP_MAX / speed;   // TRACING_STEP_MAX, tracing.c:5, 68-70
        h = (P.h0 && P.h0[id] > 0.0) ? P.h0[id] : hmax;
        nacc = 0; natt = 0;
        kprev = signed_dist(P, y);
        if (P.landing)
          for (int c = 0; c < 6; c++) P.landing[6 * (size_t)id + c] = __builtin_nan("");
        if (P.landing2) {
          kprev2 = signed_dist2(P, y);
          for (int c = 0; c < 6; c++) P.landing2[6 * (size_t)id + c] = __builtin_nan("");
        }
        traj_n = 1;
        if (P.traj) {
          double* row = P.traj + (size_t)id * P.traj_cap * 6;
          for (int c = 0; c < 6; c++) row[c] = y[c];
          if (P.times) P.times[(size_t)id * P.traj_cap] = t;
        }
      } else {
        pid = -1;
      }
    }
  };

  auto retire = [&](int status) {
    const size_t id = (size_t)pid;
    for (int c = 0; c < 6; c++) P.final_state[6 * id + c] = y[c];
    if (P.status) P.status[id] = status;
    if (P.n_accepted) P.n_accepted[id] = nacc;
    if (P.n_attempted) P.n_attempted[id] = natt;
    if (P.t_final) P.t_final[id] = t;
    if (P.h_final) P.h_final[id] = h;
    if (P.hmax_final) P.hmax_final[id] = hmax;
    if (P.traj_len) P.traj_len[id] = traj_n;
  };

  for (int c = 0; c < 6; c++) y[c] = 0.0;
  next_particle(true);
  while (true) {
    // a fresh particle that starts outside the bounds takes no step (tracing.c:79 loop not entered)
    while (true) {
      const bool out0 = (pid >= 0) && natt == 0 && !inside_trace_bounds(P.bounds, y);
      if (out0) retire(STATUS_LEFT_BOUNDS);
      if (__ballot(out0) == 0ull) break;
      if (out0) pid = -1;
      next_particle(P.batch ? __ballot(pid >= 0) == 0ull : out0);
    }
    if (__ballot(pid >= 0) == 0ull) break;
    const bool active = pid >= 0;
    double k[6][6];
    unsigned long long parked = 0ull;   // park = 1: the lanes whose step needs a ring-source sum
    // the six stages, each inlined with its index a constant (k stays in registers)
    auto stage = [&](auto stage_index) __attribute__((always_inline)) {
      constexpr int s = decltype(stage_index)::value;
      double ys[6];
#pragma unroll
      for (int c = 0; c < 6; c++) {
        double acc = y[c];
#pragma unroll
        for (int j = 0; j < s; j++) acc = fma(RK_B[s][j], k[j][c], acc);
        ys[c] = acc;
      }
      double E[3], H[3];
      bool need = false, bem = false;
      if (active) {
        if (MLP && wave_mlp) eval_field<MLP, SOURCES, true>(P.field, ys, E, H, 0, 1, true, lane, &need, &bem);
        else eval_field<MLP, SOURCES>(P.field, ys, E, H, 0, 1, true, lane);
      } else {   // a lane without a particle: no field work (its state is stale or out of range)
        E[0] = E[1] = E[2] = H[0] = H[1] = H[2] = 0.0;
      }
      if (MLP && wave_mlp && __any(need)) {   // the learned correction of the whole wave on MFMA
        const FieldComp& c0 = P.field.comp[0];
        double ce[3] = {0.0, 0.0, 0.0}, ch[3] = {0.0, 0.0, 0.0};
        mlp_wave(c0, ys, need, lane, ce, ch);
        for (int i = 0; i < 3; i++) { E[i] = fma(c0.factor, ce[i], E[i]); H[i] = fma(c0.factor, ch[i], H[i]); }
      }
      if constexpr (MLP && SOURCES) {   // points sent to the ring sources, summed by the whole wave
        if (wave_mlp) {
          const unsigned long long bm = __ballot(bem);
          if (P.park) parked |= bm;   // their steps are dropped below; their k are not used
          else if (bm) bem_wave(P.field.comp[0], bm, ys, lane, E, H);
        }
      } else if constexpr (MLP) {   // no source-sum code: such points always park
        if (wave_mlp) parked |= __ballot(bem);
      }
      const double vxh0 = ys[4] * H[2] - ys[5] * H[1];
      const double vxh1 = ys[5] * H[0] - ys[3] * H[2];
      const double vxh2 = ys[3] * H[1] - ys[4] * H[0];
      k[s][0] = h * ys[3]; k[s][1] = h * ys[4]; k[s][2] = h * ys[5];
      k[s][3] = h * qm * (E[0] + MU0 * vxh0);
      k[s][4] = h * qm * (E[1] + MU0 * vxh1);
      k[s][5] = h * qm * (E[2] + MU0 * vxh2);
    };
    stage(StageIndex<0>{}); stage(StageIndex<1>{}); stage(StageIndex<2>{});
    stage(StageIndex<3>{}); stage(StageIndex<4>{}); stage(StageIndex<5>{});
    double ep = 0.0, ev = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) {
      double e = 0.0;
#pragma unroll
      for (int s = 0; s < 6; s++) e = fma(RK_CT[s], k[s][c], e);
      e = fabs(e);
      if (c < 3) ep = e > ep ? e : ep;
      else ev = e > ev ? e : ev;
    }
    const double err = ep + h * ev;
    int st = -1;
    if (active && ((parked >> lane) & 1ull)) {   // retired at the start of the step, to be resumed
      st = STATUS_PARKED;
      retire(st);
    } else if (active) {
      natt++;
      if (err <= P.atol) {
        double yn[6];
#pragma unroll
        for (int c = 0; c < 6; c++)
          yn[c] = y[c] + (RK_CH[0] * k[0][c] + RK_CH[1] * k[1][c] + RK_CH[2] * k[2][c] + RK_CH[3] * k[3][c] +
                          RK_CH[4] * k[4][c] + RK_CH[5] * k[5][c]);
        t += h;
        nacc++;
        if (P.landing) {
          const double kn = signed_dist(P, yn);
          land_if_crossed(kprev, kn, y, yn, P.landing + 6 * (size_t)pid);
          kprev = kn;
        }
        if (P.landing2) {
          const double kn = signed_dist2(P, yn);
          land_if_crossed(kprev2, kn, y, yn, P.landing2 + 6 * (size_t)pid);
          kprev2 = kn;
        }
#pragma unroll
        for (int c = 0; c < 6; c++) y[c] = yn[c];
        if (P.traj) {
          double* row = P.traj + ((size_t)pid * P.traj_cap + traj_n) * 6;
          for (int c = 0; c < 6; c++) row[c] = y[c];
          if (P.times) P.times[(size_t)pid * P.traj_cap + traj_n] = t;
        }
        traj_n++;
        if (!inside_trace_bounds(P.bounds, y)) st = STATUS_LEFT_BOUNDS;
        else if (P.traj && traj_n == P.traj_cap) st = STATUS_BLOCK_FULL;
      }
      const double hn = 0.9 * h * pow(P.atol / err, 0.2);   // tracing.c:119
      h = hn < hmax ? hn : hmax;
      if (st < 0 && natt >= P.max_attempts) st = STATUS_STEP_CAP;
      if (st >= 0) retire(st);
    }
    const bool need = st >= 0;
    if (need) pid = -1;
    next_particle(P.batch ? __ballot(pid >= 0) == 0ull : need);
  }
}

template <bool MLP>
__global__ __launch_bounds__(64) void k_trace(TraceParams P) { trace_body<MLP, true>(P); }
// Axial and uniform fields only: no source-sum code.  W waves per SIMD: at W = 1 the
// compiler keeps every coefficient load of a field evaluation in flight (256 VGPRs); W = 2
// fits in 218 VGPRs without spills; W = 3 and 4 spill (about 200 and 320 bytes per lane)
// and trade scratch traffic for latency hiding.  The launch picks W (occupancy).
template <int W>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) void k_trace_axial(TraceParams P) {
  trace_body<false, false>(P);
}

// A learned field whose ring sources serve only its fallback (PROGRAM_LEARNED_ONLY), traced
// with park = 1: the points that need them park their particles, so the kernel carries no
// source-sum code, and W waves per SIMD as the axial-only kernel.
template <int W>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) void k_trace_learned(TraceParams P) {
  trace_body<true, false>(P);
}

// ---------------------------------------------------------------------------------
// Stage-synchronous tracer for a few particles in fields with ring-source sums.
//
// With fewer particles than the chip has lanes, one lane per particle leaves most CUs
// idle while each lane walks 10^5 sources.  Here every RK stage is one launch that spreads
// the source sum of each particle over `nchunk` workgroups (thread j of chunk c takes
// sources c*256 + j, + nchunk*256, ...), followed by a per-particle launch that adds the
// chunk partials in a fixed order, forms k_s and the next stage point.  A control launch
// per attempted step does the error norm, accept/reject, step-size update, landing and
// trajectory bookkeeping with exactly the arithmetic of k_trace.  The host issues
// SPLIT_BATCH attempts between reads of the active count.
// Programs with a learned correction add a lead launch per stage, one wave per 64
// particles: the terms without a source sum (axial series, learned correction on MFMA as
// in k_trace, uniform fields) go to W.lead, and the chunk launch then sums the ring sources
// only where the trust region or the error check sends a point to them.  A point that
// keeps the learned field gets k_trace's value exactly; a point sent to the sources gets
// the same sum in the chunked order.
// ---------------------------------------------------------------------------------
struct SplitWs {
  double *y, *ys, *k, *part, *lead, *h, *hmax, *t, *kprev, *kprev2, *qmv;
  long long *nacc, *natt, *tlen;
  int *active, *count;
  size_t used = 0;   // bytes of the layout (a null base only measures it)
  SplitWs(void* base, int P, int nchunk) {
    char* b = (char*)base;
    auto take = [&](size_t bytes) {
      char* p = b ? b + used : nullptr;
      used += (bytes + 255) & ~(size_t)255;
      return p;
    };
    y = (double*)take(sizeof(double) * 6 * P);
    ys = (double*)take(sizeof(double) * 6 * P);
    k = (double*)take(sizeof(double) * 36 * P);
    part = (double*)take(sizeof(double) * 6 * (size_t)P * nchunk);
    lead = (double*)take(sizeof(double) * 6 * (size_t)P);
    h = (double*)take(sizeof(double) * P);
    hmax = (double*)take(sizeof(double) * P);
    t = (double*)take(sizeof(double) * P);
    kprev = (double*)take(sizeof(double) * P);
    kprev2 = (double*)take(sizeof(double) * P);
    qmv = (double*)take(sizeof(double) * P);
    nacc = (long long*)take(sizeof(long long) * P);
    natt = (long long*)take(sizeof(long long) * P);
    tlen = (long long*)take(sizeof(long long) * P);
    active = (int*)take(sizeof(int) * P);
    count = (int*)take(sizeof(int) * 2);
  }
  static size_t bytes(int P, int nchunk) { return SplitWs(nullptr, P, nchunk).used; }
};
constexpr int SPLIT_THREADS = 256;
constexpr int SPLIT_NO_SOURCE = 0x7fffffff;   // a first source index past every source set
constexpr int SPLIT_BATCH = 8;
constexpr long long SPLIT_MAX_ATTEMPTS = 1ll << 22;

__device__ __forceinline__ void split_retire(const TraceParams& P, const SplitWs& W, int i, int status) {
  for (int c = 0; c < 6; c++) P.final_state[6 * (size_t)i + c] = W.y[6 * (size_t)i + c];
  if (P.status) P.status[i] = status;
  if (P.n_accepted) P.n_accepted[i] = W.nacc[i];
  if (P.n_attempted) P.n_attempted[i] = W.natt[i];
  if (P.t_final) P.t_final[i] = W.t[i];
  if (P.h_final) P.h_final[i] = W.h[i];
  if (P.hmax_final) P.hmax_final[i] = W.hmax[i];
  if (P.traj_len) P.traj_len[i] = W.tlen[i];
  W.active[i] = 0;
}

__global__ __launch_bounds__(256) void k_split_init(TraceParams P, SplitWs W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double y[6];
  for (int c = 0; c < 6; c++) {
    y[c] = P.state0[6 * (size_t)i + c];
    W.y[6 * (size_t)i + c] = y[c];
    W.ys[6 * (size_t)i + c] = y[c];
    if (P.landing) P.landing[6 * (size_t)i + c] = __builtin_nan("");
    if (P.landing2) P.landing2[6 * (size_t)i + c] = __builtin_nan("");
  }
  W.t[i] = P.t0 ? P.t0[i] : 0.0;
  W.qmv[i] = P.qm[P.qm_stride ? i : 0];
  const double hm = (P.hmax0 && P.hmax0[i] > 0.0) ? P.hmax0[i] : STEP_MAX / sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);   // tracing.c:68-70
  W.hmax[i] = hm;
  W.h[i] = (P.h0 && P.h0[i] > 0.0) ? P.h0[i] : hm;
  W.kprev[i] = signed_dist(P, y);
  W.kprev2[i] = P.landing2 ? signed_dist2(P, y) : 0.0;
  W.nacc[i] = 0; W.natt[i] = 0; W.tlen[i] = 1;
  if (P.traj) {
    for (int c = 0; c < 6; c++) P.traj[(size_t)i * P.traj_cap * 6 + c] = y[c];
    if (P.times) P.times[(size_t)i * P.traj_cap] = W.t[i];
  }
  W.active[i] = 1;
  if (!inside_trace_bounds(P.bounds, y)) split_retire(P, W, i, STATUS_LEFT_BOUNDS);   // tracing.c:79: no step taken
}

// Lead terms of a program with learned corrections at each particle's stage point: every
// term without a source sum (s0 past the last source), the learned corrections of the
// wave's points on MFMA where k_trace takes that path, per lane otherwise.  64 threads, one
// particle per lane; a wave without an active particle does nothing.
__global__ __launch_bounds__(64) void k_split_lead(TraceParams P, SplitWs W) {
  const int i = blockIdx.x * 64 + threadIdx.x, lane = threadIdx.x;
  const bool active = i < P.n && W.active[i];
  if (__ballot(active) == 0ull) return;
  const bool wave_mlp = P.field.ncomp == 1 && P.field.comp[0].mlp && P.field.comp[0].mlp_wbf && !P.field.comp[0].transform;
  double p[3] = {0.0, 0.0, 0.0};
  if (active)
    for (int c = 0; c < 3; c++) p[c] = W.ys[6 * (size_t)i + c];
  double E[3] = {0.0, 0.0, 0.0}, H[3] = {0.0, 0.0, 0.0};
  bool need = false;
  if (active) {
    if (wave_mlp) eval_field<true, true, true>(P.field, p, E, H, SPLIT_NO_SOURCE, 1, true, lane, &need);
    else eval_field<true, true>(P.field, p, E, H, SPLIT_NO_SOURCE, 1, true, lane);
  }
  if (wave_mlp && __any(need)) {
    const FieldComp& c0 = P.field.comp[0];
    double ce[3] = {0.0, 0.0, 0.0}, ch[3] = {0.0, 0.0, 0.0};
    mlp_wave(c0, p, need, lane, ce, ch);
    for (int c = 0; c < 3; c++) { E[c] = fma(c0.factor, ce[c], E[c]); H[c] = fma(c0.factor, ch[c], H[c]); }
  }
  if (active) {
    double* o = W.lead + 6 * (size_t)i;
    o[0] = E[0]; o[1] = E[1]; o[2] = E[2]; o[3] = H[0]; o[4] = H[1]; o[5] = H[2];
  }
}

// partial field sums: blockIdx.y = particle, blockIdx.x = source chunk.  LEARNED: the lead
// terms come from k_split_lead, and a learned component adds its ring sources only at
// points outside its trust region or error check.
template <bool LEARNED>
__global__ __launch_bounds__(SPLIT_THREADS) void k_split_field(TraceParams P, SplitWs W, int nchunk) {
  const int i = blockIdx.y;
  if (!W.active[i]) return;
  __shared__ double red[SPLIT_THREADS / 64][6];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double p[3] = {W.ys[6 * (size_t)i], W.ys[6 * (size_t)i + 1], W.ys[6 * (size_t)i + 2]};
  double E[3], H[3];
  if (LEARNED) eval_field<true, true, false, false>(P.field, p, E, H, blockIdx.x * SPLIT_THREADS + tid, nchunk * SPLIT_THREADS, false);
  else eval_field(P.field, p, E, H, blockIdx.x * SPLIT_THREADS + tid, nchunk * SPLIT_THREADS, blockIdx.x == 0 && tid == 0);
  double v[6] = {E[0], E[1], E[2], H[0], H[1], H[2]};
#pragma unroll
  for (int c = 0; c < 6; c++) {
    double x = v[c];
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[wave][c] = x;
  }
  __syncthreads();
  if (tid < 6) {
    double x = 0.0;
    for (int w = 0; w < SPLIT_THREADS / 64; w++) x += red[w][tid];
    W.part[((size_t)i * nchunk + blockIdx.x) * 6 + tid] = x;
  }
}

// k_s from the chunk partials, then the point of stage s + 1
__global__ __launch_bounds__(256) void k_split_stage(TraceParams P, SplitWs W, int nchunk, int s, int learned) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n || !W.active[i]) return;
  double f[6] = {0, 0, 0, 0, 0, 0};
  for (int c = 0; c < nchunk; c++)
    for (int q = 0; q < 6; q++) f[q] += W.part[((size_t)i * nchunk + c) * 6 + q];
  if (learned)
    for (int q = 0; q < 6; q++) f[q] += W.lead[6 * (size_t)i + q];
  const double* ys = W.ys + 6 * (size_t)i;
  const double h = W.h[i], qm = W.qmv[i];
  const double vxh0 = ys[4] * f[5] - ys[5] * f[4];
  const double vxh1 = ys[5] * f[3] - ys[3] * f[5];
  const double vxh2 = ys[3] * f[4] - ys[4] * f[3];
  double* k = W.k + 36 * (size_t)i;
  k[6 * s + 0] = h * ys[3]; k[6 * s + 1] = h * ys[4]; k[6 * s + 2] = h * ys[5];
  k[6 * s + 3] = h * qm * (f[0] + MU0 * vxh0);
  k[6 * s + 4] = h * qm * (f[1] + MU0 * vxh1);
  k[6 * s + 5] = h * qm * (f[2] + MU0 * vxh2);
  if (s < 5) {
    const double* y = W.y + 6 * (size_t)i;
    double* yn = W.ys + 6 * (size_t)i;
    for (int c = 0; c < 6; c++) {
      double acc = y[c];
      for (int j = 0; j <= s; j++) acc = fma(RK_B[s + 1][j], k[6 * j + c], acc);
      yn[c] = acc;
    }
  }
}

// error norm, accept/reject and step control of one attempted step (as in k_trace)
__global__ __launch_bounds__(256) void k_split_control(TraceParams P, SplitWs W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n || !W.active[i]) return;
  const double* k = W.k + 36 * (size_t)i;
  double* y = W.y + 6 * (size_t)i;
  double h = W.h[i];
  double ep = 0.0, ev = 0.0;
  for (int c = 0; c < 6; c++) {
    double e = 0.0;
    for (int s = 0; s < 6; s++) e = fma(RK_CT[s], k[6 * s + c], e);
    e = fabs(e);
    if (c < 3) ep = e > ep ? e : ep;
    else ev = e > ev ? e : ev;
  }
  const double err = ep + h * ev;
  int st = -1;
  const long long natt = ++W.natt[i];
  if (err <= P.atol) {
    double yn[6];
    for (int c = 0; c < 6; c++)
      yn[c] = y[c] + (RK_CH[0] * k[c] + RK_CH[1] * k[6 + c] + RK_CH[2] * k[12 + c] + RK_CH[3] * k[18 + c] + RK_CH[4] * k[24 + c] +
                      RK_CH[5] * k[30 + c]);
    W.t[i] += h;
    W.nacc[i]++;
    if (P.landing) {
      const double kn = signed_dist(P, yn);
      land_if_crossed(W.kprev[i], kn, y, yn, P.landing + 6 * (size_t)i);
      W.kprev[i] = kn;
    }
    if (P.landing2) {
      const double kn = signed_dist2(P, yn);
      land_if_crossed(W.kprev2[i], kn, y, yn, P.landing2 + 6 * (size_t)i);
      W.kprev2[i] = kn;
    }
    for (int c = 0; c < 6; c++) y[c] = yn[c];
    const long long tn = W.tlen[i];
    if (P.traj) {
      double* row = P.traj + ((size_t)i * P.traj_cap + tn) * 6;
      for (int c = 0; c < 6; c++) row[c] = y[c];
      if (P.times) P.times[(size_t)i * P.traj_cap + tn] = W.t[i];
    }
    W.tlen[i] = tn + 1;
    if (!inside_trace_bounds(P.bounds, y)) st = STATUS_LEFT_BOUNDS;
    else if (P.traj && tn + 1 == P.traj_cap) st = STATUS_BLOCK_FULL;
  }
  const double hn = 0.9 * h * pow(P.atol / err, 0.2);   // tracing.c:119
  W.h[i] = hn < W.hmax[i] ? hn : W.hmax[i];
  if (st < 0 && natt >= P.max_attempts) st = STATUS_STEP_CAP;
  if (st >= 0) {
    split_retire(P, W, i, st);
  } else {
    for (int c = 0; c < 6; c++) W.ys[6 * (size_t)i + c] = y[c];
  }
}

__global__ __launch_bounds__(256) void k_split_count(const int* __restrict__ active, int n, int* __restrict__ count) {
  __shared__ int cs;
  if (threadIdx.x == 0) cs = 0;
  __syncthreads();
  int c = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += active[i];
  if (c) atomicAdd(&cs, c);
  __syncthreads();
  if (threadIdx.x == 0 && cs) atomicAdd(count, cs);
}

int split_chunks(int n_particles, long long n_sources) {
  if (n_particles <= 0) return 1;   // nothing to trace; any valid chunk count
  // enough workgroups to cover the chip (~2048), at least 8 sources per thread
  long long by_src = n_sources / (SPLIT_THREADS * 8);
  long long by_occ = (2048 + n_particles - 1) / n_particles;
  long long c = by_src < by_occ ? by_src : by_occ;
  if (c < 1) c = 1;
  if (c > 1024) c = 1024;
  return (int)c;
}

size_t split_workspace_bytes(int n_particles, int nchunk) { return SplitWs::bytes(n_particles, nchunk); }

__global__ void k_split_zero_count(int* c) {
  if (threadIdx.x == 0) c[0] = 0;
}

// SPLIT_BATCH attempted steps (13 launches each) and the count of particles still active
static void split_batch(const TraceParams& P, const SplitWs& W, int nchunk, hipStream_t st) {
  const dim3 gp((P.n + 255) / 256), bp(256), gf(nchunk, P.n);
  const int count_blocks = (P.n + 255) / 256 < 64 ? (P.n + 255) / 256 : 64;
  const int learned = (P.field.flags & PROGRAM_HAS_MLP) ? 1 : 0;
  for (int a = 0; a < SPLIT_BATCH; a++) {
    for (int s = 0; s < 6; s++) {
      if (learned) {
        hipLaunchKernelGGL(k_split_lead, dim3((P.n + 63) / 64), dim3(64), 0, st, P, W);
        hipLaunchKernelGGL(k_split_field<true>, gf, dim3(SPLIT_THREADS), 0, st, P, W, nchunk);
      } else {
        hipLaunchKernelGGL(k_split_field<false>, gf, dim3(SPLIT_THREADS), 0, st, P, W, nchunk);
      }
      hipLaunchKernelGGL(k_split_stage, gp, bp, 0, st, P, W, nchunk, s, learned);
    }
    hipLaunchKernelGGL(k_split_control, gp, bp, 0, st, P, W);
  }
  hipLaunchKernelGGL(k_split_zero_count, dim3(1), dim3(64), 0, st, W.count);   // a kernel: graph-safe
  hipLaunchKernelGGL(k_split_count, dim3(count_blocks), dim3(256), 0, st, W.active, P.n, W.count);
}

// One batch captured once into a HIP graph (on a stream of this call, default priority) and
// replayed per batch on the caller's stream: one launch instead of 106 (154 with lead
// launches).  TRACEON_SPLIT_GRAPH=0
// launches the kernels directly instead; the results are the same.  Returns null when
// capture is off or fails (the batches are then launched directly).
static hipGraphExec_t capture_split_batch(const TraceParams& P, const SplitWs& W, int nchunk) {
  const char* env = getenv("TRACEON_SPLIT_GRAPH");
  if (env && env[0] == '0') return nullptr;
  hipStream_t cs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    split_batch(P, W, nchunk, cs);
    if (hipStreamEndCapture(cs, &graph) == hipSuccess && graph) {
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) exec = nullptr;
      hipGraphDestroy(graph);
    }
  }
  hipStreamDestroy(cs);
  (void)hipGetLastError();   // a failed capture leaves no error behind: the direct launches follow
  return exec;
}

hipError_t trace_split(const TraceParams& P, int nchunk, void* ws, hipStream_t st) {
  if (P.n <= 0) return hipSuccess;
  if (nchunk < 1 || P.n > 65535) return hipErrorInvalidValue;
  SplitWs W(ws, P.n, nchunk);
  hipLaunchKernelGGL(k_split_init, dim3((P.n + 255) / 256), dim3(256), 0, st, P, W);
  hipGraphExec_t exec = capture_split_batch(P, W, nchunk);
  hipError_t e = hipSuccess;
  for (long long attempts = 0;; attempts += SPLIT_BATCH) {
    // a trace that needs more attempted steps than this has a non-finite field or step
    if (attempts > SPLIT_MAX_ATTEMPTS) { e = hipErrorLaunchTimeOut; break; }
    if (exec) {
      e = hipGraphLaunch(exec, st);
      if (e != hipSuccess) break;
    } else {
      split_batch(P, W, nchunk, st);
    }
    int remaining = 0;
    e = hipMemcpyAsync(&remaining, W.count, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) break;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess || remaining == 0) break;
  }
  if (exec) hipGraphExecDestroy(exec);
  return e != hipSuccess ? e : hipGetLastError();
}

// single field evaluation at a batch of points through a field program
template <bool MLP>
__global__ __launch_bounds__(64) void k_eval_program(FieldProgram F, const double* __restrict__ pts, int M, double* __restrict__ E,
                                                     double* __restrict__ H) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  // as the tracer: one learned, untransformed component with BF16 weights takes the MFMA path,
  // for which every lane of the wave stays (lanes past M on a dummy point)
  const bool wave_mlp = MLP && F.ncomp == 1 && F.comp[0].mlp && F.comp[0].mlp_wbf && !F.comp[0].transform;
  if (!wave_mlp && t >= M) return;
  const int tt = t < M ? t : M - 1;
  double p[3] = {pts[3 * (size_t)tt], pts[3 * (size_t)tt + 1], pts[3 * (size_t)tt + 2]};
  double e[3], h[3];
  if (MLP && wave_mlp) {
    bool need = false;
    eval_field<MLP, true, true>(F, p, e, h, 0, 1, true, lane, &need);
    need = need && t < M;
    if (__any(need)) {
      const FieldComp& c0 = F.comp[0];
      double ce[3] = {0.0, 0.0, 0.0}, ch[3] = {0.0, 0.0, 0.0};
      mlp_wave(c0, p, need, lane, ce, ch);
      for (int i = 0; i < 3; i++) { e[i] = fma(c0.factor, ce[i], e[i]); h[i] = fma(c0.factor, ch[i], h[i]); }
    }
    if (t >= M) return;
  } else {
    eval_field<MLP>(F, p, e, h, 0, 1, true, lane);
  }
  for (int i = 0; i < 3; i++) { if (E) E[3 * (size_t)t + i] = e[i]; if (H) H[3 * (size_t)t + i] = h[i]; }
}

hipError_t launch_trace(const TraceParams& P, int waves, hipStream_t st) {
  if (P.n <= 0) return hipSuccess;
  int nw = waves > 0 ? waves : 1;
  const int need = (P.n + 63) / 64;
  if (nw > need) nw = need;
  if ((P.field.flags & PROGRAM_HAS_MLP) && (P.field.flags & PROGRAM_LEARNED_ONLY) && P.park) {
    if (P.occupancy == 1) hipLaunchKernelGGL(k_trace_learned<1>, dim3(nw), dim3(64), 0, st, P);
    else hipLaunchKernelGGL(k_trace_learned<2>, dim3(nw), dim3(64), 0, st, P);
  } else if (P.field.flags & PROGRAM_HAS_MLP) hipLaunchKernelGGL(k_trace<true>, dim3(nw), dim3(64), 0, st, P);
  else if (P.field.flags & PROGRAM_NO_SOURCES) {
    switch (P.occupancy) {
      case 1: hipLaunchKernelGGL(k_trace_axial<1>, dim3(nw), dim3(64), 0, st, P); break;
      case 3: hipLaunchKernelGGL(k_trace_axial<3>, dim3(nw), dim3(64), 0, st, P); break;
      case 4: hipLaunchKernelGGL(k_trace_axial<4>, dim3(nw), dim3(64), 0, st, P); break;
      default: hipLaunchKernelGGL(k_trace_axial<2>, dim3(nw), dim3(64), 0, st, P); break;
    }
  } else hipLaunchKernelGGL(k_trace<false>, dim3(nw), dim3(64), 0, st, P);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Queue order: counting sort by a 16-bit key (start z major, start v_z minor).
// ws layout: key range (4 x u64: ordered-int min/max of z and v_z) | histogram/offsets
// (ORDER_BINS ints) | per-particle keys (n ints).
// ---------------------------------------------------------------------------------
constexpr int ORDER_BINS = 1 << 16, ORDER_ZBINS = 16, ORDER_SCAN_THREADS = 1024;

// doubles as unsigned integers with the same order (NaN excluded by the caller)
__device__ __forceinline__ unsigned long long ordered_bits(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_ordered_bits(unsigned long long u) {
  return __longlong_as_double((long long)((u >> 63) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u));
}

__global__ __launch_bounds__(256) void k_order_init(unsigned long long* range, int* hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2) { range[2 * i] = ~0ull; range[2 * i + 1] = 0ull; }   // (min, max) of z, then of v_z
  for (int b = i; b < ORDER_BINS; b += gridDim.x * blockDim.x) hist[b] = 0;
}

__global__ __launch_bounds__(256) void k_order_range(const double* s, int n, unsigned long long* range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long zmin = ~0ull, zmax = 0ull, vmin = ~0ull, vmax = 0ull;
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    const double z = s[6 * (size_t)k + 2], v = s[6 * (size_t)k + 5];
    if (z == z) { const unsigned long long u = ordered_bits(z); zmin = u < zmin ? u : zmin; zmax = u > zmax ? u : zmax; }
    if (v == v) { const unsigned long long u = ordered_bits(v); vmin = u < vmin ? u : vmin; vmax = u > vmax ? u : vmax; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long t;
    t = __shfl_xor(zmin, o, 64); zmin = t < zmin ? t : zmin;
    t = __shfl_xor(zmax, o, 64); zmax = t > zmax ? t : zmax;
    t = __shfl_xor(vmin, o, 64); vmin = t < vmin ? t : vmin;
    t = __shfl_xor(vmax, o, 64); vmax = t > vmax ? t : vmax;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&range[0], zmin); atomicMax(&range[1], zmax);
    atomicMin(&range[2], vmin); atomicMax(&range[3], vmax);
  }
}

__device__ __forceinline__ int order_bin(double x, double lo, double hi, int nb) {
  if (!(x == x) || !(hi > lo)) return 0;
  const double f = (x - lo) / (hi - lo) * nb;
  const int b = (int)f;
  return b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
}

__global__ __launch_bounds__(256) void k_order_hist(const double* s, int n, const unsigned long long* range, int* hist, int* key) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double zlo = 0.0, zhi = 0.0, vlo = 0.0, vhi = 0.0;
  if (range[0] <= range[1]) { zlo = from_ordered_bits(range[0]); zhi = from_ordered_bits(range[1]); }
  if (range[2] <= range[3]) { vlo = from_ordered_bits(range[2]); vhi = from_ordered_bits(range[3]); }
  const int nzb = zhi > zlo ? ORDER_ZBINS : 1, nvb = ORDER_BINS / nzb;
  const int k = order_bin(s[6 * (size_t)i + 2], zlo, zhi, nzb) * nvb + order_bin(s[6 * (size_t)i + 5], vlo, vhi, nvb);
  key[i] = k;
  atomicAdd(&hist[k], 1);
}

// exclusive prefix sum of the histogram in place, one workgroup of ORDER_SCAN_THREADS
__global__ __launch_bounds__(ORDER_SCAN_THREADS) void k_order_scan(int* hist) {
  constexpr int PER = ORDER_BINS / ORDER_SCAN_THREADS;
  __shared__ int part[ORDER_SCAN_THREADS];
  const int t = threadIdx.x;
  int sum = 0;
  for (int j = 0; j < PER; j++) sum += hist[t * PER + j];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < ORDER_SCAN_THREADS; o <<= 1) {   // Hillis-Steele inclusive scan
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  for (int j = 0; j < PER; j++) {
    const int c = hist[t * PER + j];
    hist[t * PER + j] = run;
    run += c;
  }
}

__global__ __launch_bounds__(256) void k_order_scatter(const int* key, int n, int* offsets, int* order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  order[atomicAdd(&offsets[key[i]], 1)] = i;
}

size_t trace_order_workspace_bytes(int n) {
  return 4 * sizeof(unsigned long long) + sizeof(int) * (size_t)ORDER_BINS + sizeof(int) * (size_t)(n > 0 ? n : 0);
}

hipError_t trace_order(const double* state0, int n, int* order, void* ws, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  unsigned long long* range = (unsigned long long*)ws;
  int* hist = (int*)(range + 4);
  int* key = hist + ORDER_BINS;
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_order_init, dim3(ORDER_BINS / 256), dim3(256), 0, st, range, hist);
  hipLaunchKernelGGL(k_order_range, dim3(nb < 1024 ? nb : 1024), dim3(256), 0, st, state0, n, range);
  hipLaunchKernelGGL(k_order_hist, dim3(nb), dim3(256), 0, st, state0, n, range, hist, key);
  hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(ORDER_SCAN_THREADS), 0, st, hist);
  hipLaunchKernelGGL(k_order_scatter, dim3(nb), dim3(256), 0, st, key, n, hist, order);
  return hipGetLastError();
}

hipError_t launch_eval_program(const FieldProgram& F, const double* pts, int M, double* E, double* H, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  if (F.flags & PROGRAM_HAS_MLP) hipLaunchKernelGGL(k_eval_program<true>, dim3((M + 63) / 64), dim3(64), 0, st, F, pts, M, E, H);
  else hipLaunchKernelGGL(k_eval_program<false>, dim3((M + 63) / 64), dim3(64), 0, st, F, pts, M, E, H);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Beam focus (focus.py:9-43): the least-squares point closest to the final rays
// x_i + a_i (Z - z_i), y_i + b_i (Z - z_i) with a = vx / vz, b = vy / vz.  numpy.lstsq of the
// reference's 2N x 3 system has the closed form
//   Z = -(S_ac + S_bd) / (S_aa + S_bb),  X = mean(c) + mean(a) Z,  Y = mean(d) + mean(b) Z,
// with c = x - z a, d = y - z b and S the centred sums of products; the centred form keeps the
// conditioning of the rays themselves.  Two passes over the states (means, then centred
// sums), each into FOCUS_BLOCKS partials added in a fixed order: the result does not depend
// on scheduling.  ws: FOCUS_BLOCKS * 4 + 8 doubles.
// ---------------------------------------------------------------------------------
constexpr int FOCUS_BLOCKS = 256, FOCUS_THREADS = 256;
static_assert(FOCUS_BLOCKS == FOCUS_THREADS, "k_focus_reduce takes one partial per thread");

__device__ __forceinline__ void focus_terms(const double* s, double& a, double& b, double& c, double& d) {
  a = s[3] / s[5];
  b = s[4] / s[5];
  c = s[0] - s[2] * a;
  d = s[1] - s[2] * b;
}

// four sums of a workgroup in a fixed tree order into out[0..3]
__device__ __forceinline__ void block_sum4(double v[4], double* out) {
  __shared__ double red[4][FOCUS_THREADS];
  const int t = threadIdx.x;
  for (int k = 0; k < 4; k++) red[k][t] = v[k];
  __syncthreads();
  for (int o = FOCUS_THREADS / 2; o > 0; o >>= 1) {
    if (t < o)
      for (int k = 0; k < 4; k++) red[k][t] += red[k][t + o];
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 4; k++) out[k] = red[k][0];
}

// pass 0: sums of a, b, c, d; pass 1: S_aa + S_bb, S_ac + S_bd (means in m[0..3])
__global__ __launch_bounds__(FOCUS_THREADS) void k_focus_partial(const double* __restrict__ st, int n, long long stride, int pass,
                                                                 const double* __restrict__ m, double* __restrict__ part) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * FOCUS_THREADS + threadIdx.x; i < n; i += FOCUS_BLOCKS * FOCUS_THREADS) {
    double a, b, c, d;
    focus_terms(st + (size_t)i * stride, a, b, c, d);
    if (pass == 0) {
      v[0] += a; v[1] += b; v[2] += c; v[3] += d;
    } else {
      const double da = a - m[0], db = b - m[1], dc = c - m[2], dd = d - m[3];
      v[0] += da * da + db * db;
      v[1] += da * dc + db * dd;
    }
  }
  block_sum4(v, part + 4 * blockIdx.x);
}

// adds the partials in block order; pass 0 leaves the means in m, pass 1 writes (X, Y, Z)
__global__ __launch_bounds__(FOCUS_THREADS) void k_focus_reduce(int n, int pass, const double* __restrict__ part, double* m,
                                                                double* __restrict__ out) {
  double v[4] = {part[4 * threadIdx.x], part[4 * threadIdx.x + 1], part[4 * threadIdx.x + 2], part[4 * threadIdx.x + 3]};
  __shared__ double tot[4];
  block_sum4(v, tot);
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (pass == 0) {
    for (int k = 0; k < 4; k++) m[k] = tot[k] / n;
  } else {
    const double Z = -tot[1] / tot[0];
    out[0] = m[2] + m[0] * Z;
    out[1] = m[3] + m[1] * Z;
    out[2] = Z;
  }
}

size_t focus_workspace_bytes() { return sizeof(double) * (FOCUS_BLOCKS * 4 + 8); }

hipError_t focus_position(const double* states, int n, long long stride, double* out, void* ws, hipStream_t st) {
  double* part = static_cast<double*>(ws);
  double* m = part + FOCUS_BLOCKS * 4;
  for (int pass = 0; pass < 2; pass++) {
    hipLaunchKernelGGL(k_focus_partial, dim3(FOCUS_BLOCKS), dim3(FOCUS_THREADS), 0, st, states, n, stride, pass, m, part);
    hipLaunchKernelGGL(k_focus_reduce, dim3(1), dim3(FOCUS_THREADS), 0, st, n, pass, part, m, out);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Tracing without allocation: the work counter, the queue order and the order's sort
// workspace live in one static device array (each GPU holds its own copy until the process
// ends).  Host code reaches it only through the address hipGetSymbolAddress returns on that
// device; calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------------
constexpr size_t STATIC_ORDER_OFFSET = 256;
constexpr size_t STATIC_WS_OFFSET = STATIC_ORDER_OFFSET + (((size_t)TRACE_STATIC_MAX_PARTICLES * sizeof(int) + 255) & ~(size_t)255);
constexpr size_t STATIC_SCRATCH_BYTES = STATIC_WS_OFFSET + ((4 * sizeof(unsigned long long) + sizeof(int) * (size_t)ORDER_BINS +
                                                              sizeof(int) * (size_t)TRACE_STATIC_MAX_PARTICLES + 255) & ~(size_t)255);
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES] = {};

// the device address of this GPU's scratch array, looked up once per device
static hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

__global__ void k_zero_counter(int* c) {
  if (threadIdx.x == 0) *c = 0;
}

size_t trace_static_scratch_bytes() { return STATIC_SCRATCH_BYTES; }

// SplitMix64 (Steele, Lea and Flood 2014): output j of the generator seeded with s is
// mix(s + j * golden), j = 1, 2, ...; the mixer is Stafford's variant 13.
__device__ __forceinline__ unsigned long long splitmix64_at(unsigned long long s, unsigned long long j) {
  unsigned long long z = s + j * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// uniform in [0, 1): the top 53 bits
__device__ __forceinline__ double unit_draw(unsigned long long s, long i, int k) {
  return (double)(splitmix64_at(s, 8ull * (unsigned long long)i + (unsigned long long)k + 1ull) >> 11) * 0x1.0p-53;
}

__global__ __launch_bounds__(256) void k_sample_beam(BeamSpec B, const unsigned long long* __restrict__ seed, int n,
                                                     double* __restrict__ S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long s = *seed;
  const double two_pi = 6.283185307179586476925;
  // draws 0, 1: the start point, r^2 uniform over the disc
  const double r = B.r_max * sqrt(unit_draw(s, i, 0));
  const double phi = two_pi * unit_draw(s, i, 1);
  // draws 2, 3 and 4, 5: two Box-Muller pairs (1 - u keeps the log finite); the second pair's
  // cosine branch is the energy's normal deviate
  const double m1 = sqrt(-2.0 * log(1.0 - unit_draw(s, i, 2))), a1 = two_pi * unit_draw(s, i, 3);
  const double m2 = sqrt(-2.0 * log(1.0 - unit_draw(s, i, 4))), a2 = two_pi * unit_draw(s, i, 5);
  const double tx = tan(B.angle_sigma * m1 * cos(a1)), ty = tan(B.angle_sigma * m1 * sin(a1));
  const double energy = B.e_median * exp(B.e_sigma * m2 * cos(a2));
  const double speed = sqrt(2.0 * energy * B.charge / B.mass);
  const double inv = speed / sqrt(tx * tx + ty * ty + 1.0);
  double* o = S + 6 * i;
  o[0] = r * cos(phi);
  o[1] = r * sin(phi);
  o[2] = B.z0;
  o[3] = tx * inv;
  o[4] = ty * inv;
  o[5] = inv;
}

hipError_t sample_beam(const BeamSpec& B, const unsigned long long* seed, int n, double* state0, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample_beam, dim3((n + 255) / 256), dim3(256), 0, st, B, seed, n, state0);
  return hipGetLastError();
}

hipError_t trace_noalloc(const TraceParams& P, int waves, int ordered, hipStream_t st) {
  if (P.n > TRACE_STATIC_MAX_PARTICLES) return hipErrorInvalidValue;
  void* base = nullptr;
  hipError_t e = scratch_address(&base);
  if (e != hipSuccess) return e;
  if (P.n <= 0) return hipSuccess;
  TraceParams Q = P;
  Q.counter = reinterpret_cast<int*>(base);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, Q.counter);
  if (ordered) {
    int* order = reinterpret_cast<int*>(static_cast<char*>(base) + STATIC_ORDER_OFFSET);
    e = trace_order(P.state0, P.n, order, static_cast<char*>(base) + STATIC_WS_OFFSET, st);
    if (e != hipSuccess) return e;
    Q.order = order;
  }
  return launch_trace(Q, waves, st);
}

}  // namespace tb
