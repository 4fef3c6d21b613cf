// Device (and host) arithmetic of Traceon's radially symmetric BEM path for gfx950.
//
// Formulas follow the reference C backend (traceon/backend/elliptic.c, utilities_2d.c,
// radial_ring.c, radial.c, tracing.c at v0.10.0); each function names the lines it
// replaces.  Differences from the reference's layout are deliberate GPU choices:
//   * K(m) and E(m) of one pair share a single log(1/p) (the reference calls log
//     once per integral, three times per field pair);
//   * 1/sqrt(S) is formed once per pair and reused for the division by S;
//   * a pair's potential, dPhi/dr0 and dPhi/dz0 come out of one evaluation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TB_HD __host__ __device__ __forceinline__

namespace tb {

constexpr int NQ = 16;        // N_QUAD_2D, utilities_2d.c:2
constexpr int NQ3 = 12;       // N_TRIANGLE_QUAD, utilities_3d.c:4
constexpr int NDERIV = 9;     // DERIV_2D_MAX, radial_ring.c:2
constexpr double AXIS_EPS = 1e-10;            // MIN_DISTANCE_AXIS, three_d.c:2
constexpr double MU0 = 1.25663706212e-06;     // definitions.c:31
constexpr double INV_PI = 0.318309886183790671537767526745;
constexpr double PI = 3.14159265358979323846;

// Excitation type codes, definitions.c:34-40
enum : uint8_t { VOLTAGE_FIXED = 1, VOLTAGE_FUN = 2, DIELECTRIC = 3, CURRENT = 4,
                 MAGNETOSTATIC_POT = 5, MAGNETIZABLE = 6 };

// 16-point Gauss-Legendre rule with the reference's 16-digit literals (utilities_2d.c:3-5).
__constant__ static const double GL_X[NQ] = {
    -0.0950125098376374, 0.0950125098376374, -0.2816035507792589, 0.2816035507792589,
    -0.4580167776572274, 0.4580167776572274, -0.6178762444026438, 0.6178762444026438,
    -0.7554044083550030, 0.7554044083550030, -0.8656312023878318, 0.8656312023878318,
    -0.9445750230732326, 0.9445750230732326, -0.9894009349916499, 0.9894009349916499};
__constant__ static const double GL_W[NQ] = {
    0.1894506104550685, 0.1894506104550685, 0.1826034150449236, 0.1826034150449236,
    0.1691565193950025, 0.1691565193950025, 0.1495959888165767, 0.1495959888165767,
    0.1246289712555339, 0.1246289712555339, 0.0951585116824928, 0.0951585116824928,
    0.0622535239386479, 0.0622535239386479, 0.0271524594117541, 0.0271524594117541};

// 12-point triangle rule (utilities_3d.c:6-47).
__constant__ static const double TRI_W[NQ3] = {
    0.0254224531851035, 0.0254224531851035, 0.0254224531851035, 0.0583931378631895, 0.0583931378631895,
    0.0583931378631895, 0.041425537809187, 0.041425537809187, 0.041425537809187, 0.041425537809187,
    0.041425537809187, 0.041425537809187};
__constant__ static const double TRI_B1[NQ3] = {
    0.873821971016996, 0.063089014491502, 0.063089014491502, 0.501426509658179, 0.249286745170910,
    0.249286745170910, 0.636502499121399, 0.636502499121399, 0.310352451033785, 0.310352451033785,
    0.053145049844816, 0.053145049844816};
__constant__ static const double TRI_B2[NQ3] = {
    0.063089014491502, 0.873821971016996, 0.063089014491502, 0.249286745170910, 0.501426509658179,
    0.249286745170910, 0.310352451033785, 0.053145049844816, 0.636502499121399, 0.053145049844816,
    0.636502499121399, 0.310352451033785};

// ---------------------------------------------------------------------------------
// Complete elliptic integrals from p = 1 - m: Cody (1965) fits, elliptic.c:9-75.
//   K = PKA(p) + ln(1/p) PKB(p),  E = PEA(p) + ln(1/p) PEB(p)
// ---------------------------------------------------------------------------------
TB_HD double horner_ka(double p) {
  double s = 1.40704915496101e-3;
  s = fma(s, p, 1.09423810688623e-2); s = fma(s, p, 1.68695685967517e-2);
  s = fma(s, p, 1.25565693543211e-2); s = fma(s, p, 1.52618320622534e-2);
  s = fma(s, p, 3.08909633861795e-2); s = fma(s, p, 9.65736020516771e-2);
  return fma(s, p, 1.3862943611198906);   // ln 4
}
TB_HD double horner_kb(double p) {
  double s = 3.42805719229748e-4;
  s = fma(s, p, 5.81807961871996e-3); s = fma(s, p, 2.09857677336790e-2);
  s = fma(s, p, 3.57218443007327e-2); s = fma(s, p, 4.87379510945218e-2);
  s = fma(s, p, 7.03114105853296e-2); s = fma(s, p, 1.24999998585309e-1);
  return fma(s, p, 0.5);
}
TB_HD double horner_ea(double p) {
  double s = 1.55618744745296e-3;
  s = fma(s, p, 1.21819481486695e-2); s = fma(s, p, 1.92284389022977e-2);
  s = fma(s, p, 1.56847700239786e-2); s = fma(s, p, 2.21862206993846e-2);
  s = fma(s, p, 5.68115681053803e-2); s = fma(s, p, 4.43147193467733e-1);
  return fma(s, p, 1.0);
}
TB_HD double horner_eb(double p) {
  double s = 3.78886487349367e-4;
  s = fma(s, p, 6.45682247315060e-3); s = fma(s, p, 2.35091602564984e-2);
  s = fma(s, p, 4.09074821593164e-2); s = fma(s, p, 5.84950297066166e-2);
  s = fma(s, p, 9.37488062098189e-2); s = fma(s, p, 2.49999998448655e-1);
  return s * p;
}

// A wave-uniform FP64 constant held in a scalar register pair where it is used: the FMA of a
// Horner step then takes it as its addend operand directly, where a constant kept in vector
// registers makes the compiler copy it into the accumulating FMA's destination at every step.
TB_HD double sconst(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+s"(c));
#endif
  return c;
}

// ln(1/p) for p > 0 (finite), from the mathematics: p = m 2^e with m in [1/sqrt2, sqrt2),
// ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716, and the atanh series
// s + s^3/3 + ... + s^19/19 in Horner form over s^2 (the first omitted term, s^21/21, is
// below 3e-17 of the sum).
// One division, no table; about a third of the instructions of the library log on gfx950.
// Non-positive, infinite or NaN p fall back to the library log.
TB_HD double neg_log(double p) {
  if (!(p > 0.0) || !(p < 1.7976931348623157e308)) return -log(p);
  int e;
  double m = frexp(p, &e);            // m in [0.5, 1)
  if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
#if defined(__HIP_DEVICE_COMPILE__)
  // (m - 1) / (m + 1) with the denominator in [1.7, 2.5]: the hardware reciprocal (relative
  // error below 2^-24 here, measured over 2^24 points of the range) refined by one Newton step
  // (error below 2^-48) and the quotient corrected once, which squares that error again (no
  // scaling or special cases in this range, which the general division sequence spends four
  // instructions on; a second Newton step changed none of the 2^24 quotients)
  const double dn = m + 1.0, nm = m - 1.0;
  double y = __builtin_amdgcn_rcp(dn);
  y = fma(fma(-dn, y, 1.0), y, y);
  double s = nm * y;
  s = fma(fma(-dn, s, nm), y, s);
#else
  const double s = (m - 1.0) / (m + 1.0);
#endif
  const double s2 = s * s;
  double q = 1.0 / 19.0;
  q = fma(q, s2, sconst(1.0 / 17.0));
  q = fma(q, s2, sconst(1.0 / 15.0)); q = fma(q, s2, sconst(1.0 / 13.0)); q = fma(q, s2, sconst(1.0 / 11.0));
  q = fma(q, s2, sconst(1.0 / 9.0));  q = fma(q, s2, sconst(1.0 / 7.0));  q = fma(q, s2, sconst(1.0 / 5.0));
  q = fma(q, s2, sconst(1.0 / 3.0));
  const double lnm = fma(2.0 * s * s2, q, 2.0 * s);
  const double de = (double)e;
  // ln 2 split in a high part exact in its leading bits and a low correction
  return -(fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lnm)));
}

// 1/sqrt(x) for positive finite x (the callers pass sums of squares of a ring geometry): the
// hardware estimate and one third-order correction, y (1 + e/2 + 3e^2/8) with
// e = 1 - x y^2, without the library's special-value selects
TB_HD double rsqrt_(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
#else
  return 1.0 / sqrt(x);
#endif
}

// The same fits as sums of c_i p^i with the powers of p formed once and shared by the fits
// of a pair (the reference evaluates its fits with pow).  On gfx950 an FP64 coefficient
// cannot be an instruction operand, so it sits in a register; in Horner's form it is the
// FMA's accumulator and each step first copies it, while in a power sum the running sum is
// the accumulator and the coefficient a multiplicand.  Terms are added from the highest
// power (the smallest) to the constant.
struct PPow { double p1, p2, p3, p4, p5, p6, p7; };
TB_HD PPow ppow(double p) {
  PPow w;
  w.p1 = p; w.p2 = p * p; w.p3 = w.p2 * p; w.p4 = w.p2 * w.p2; w.p5 = w.p4 * p; w.p6 = w.p3 * w.p3; w.p7 = w.p6 * p;
  return w;
}
TB_HD double psum8(const PPow& w, double c7, double c6, double c5, double c4, double c3, double c2, double c1, double c0) {
  double s = c7 * w.p7;
  s = fma(c6, w.p6, s); s = fma(c5, w.p5, s); s = fma(c4, w.p4, s); s = fma(c3, w.p3, s);
  s = fma(c2, w.p2, s); s = fma(c1, w.p1, s);
  return s + c0;
}
TB_HD double pka(const PPow& w) {
  return psum8(w, 1.40704915496101e-3, 1.09423810688623e-2, 1.68695685967517e-2, 1.25565693543211e-2, 1.52618320622534e-2,
               3.08909633861795e-2, 9.65736020516771e-2, 1.3862943611198906);
}
TB_HD double pkb(const PPow& w) {
  return psum8(w, 3.42805719229748e-4, 5.81807961871996e-3, 2.09857677336790e-2, 3.57218443007327e-2, 4.87379510945218e-2,
               7.03114105853296e-2, 1.24999998585309e-1, 0.5);
}
TB_HD double pea(const PPow& w) {
  return psum8(w, 1.55618744745296e-3, 1.21819481486695e-2, 1.92284389022977e-2, 1.56847700239786e-2, 2.21862206993846e-2,
               5.68115681053803e-2, 4.43147193467733e-1, 1.0);
}
TB_HD double peb(const PPow& w) {   // p (b6 p^6 + ... + b0)
  double s = 3.78886487349367e-4 * w.p6;
  s = fma(6.45682247315060e-3, w.p5, s); s = fma(2.35091602564984e-2, w.p4, s); s = fma(4.09074821593164e-2, w.p3, s);
  s = fma(5.84950297066166e-2, w.p2, s); s = fma(9.37488062098189e-2, w.p1, s);
  return (s + 2.49999998448655e-1) * w.p1;
}

TB_HD double ellipkm1_(double p) { double L = neg_log(p); const PPow w = ppow(p); return fma(L, pkb(w), pka(w)); }
TB_HD double ellipem1_(double p) { double L = neg_log(p); const PPow w = ppow(p); return fma(L, peb(w), pea(w)); }
TB_HD void ellip_km1_em1(double p, double& K, double& E) {
  double L = neg_log(p);
  const PPow w = ppow(p);
  K = fma(L, pkb(w), pka(w));
  E = fma(L, peb(w), pea(w));
}
// elliptic.c:37-41, 71-75 (parameter forms used by the current-ring kernel)
TB_HD double ellipk_(double k) { return k > -1.0 ? ellipkm1_(1.0 - k) : ellipkm1_(1.0 / (1.0 - k)) / sqrt(k); }
TB_HD double ellipe_(double k) {
  return (k >= 0.0 && k <= 1.0) ? ellipem1_(1.0 - k) : ellipem1_(-1.0 / (k - 1.0)) * sqrt(1.0 - k);
}

// ---------------------------------------------------------------------------------
// Line4 element map (utilities_2d.c:18-87).  Node order in parameter alpha is
// vertex rows 0, 2, 3, 1 of a (4,3) element (radial.c:265-268); r = x (col 0), z = col 2.
// ---------------------------------------------------------------------------------
struct Line4 { double r[4], z[4]; };   // nodes in parameter order

TB_HD Line4 load_line4(const double* line) {   // line: (4,3) row-major
  Line4 L;
  const int order[4] = {0, 2, 3, 1};
#pragma unroll
  for (int n = 0; n < 4; n++) { L.r[n] = line[3 * order[n]]; L.z[n] = line[3 * order[n] + 2]; }
  return L;
}

TB_HD void lagrange4(double a, double B[4], double dB[4]) {
  double a2 = a * a;
  B[0] = -(a - 1.0) * (9.0 * a2 - 1.0) * 0.0625;
  B[1] = 9.0 * (a2 - 1.0) * (3.0 * a - 1.0) * 0.0625;
  B[2] = -9.0 * (a2 - 1.0) * (3.0 * a + 1.0) * 0.0625;
  B[3] = (a + 1.0) * (9.0 * a2 - 1.0) * 0.0625;
  dB[0] = (-27.0 * a2 + 18.0 * a + 1.0) * 0.0625;
  dB[1] = 9.0 * (9.0 * a2 - 2.0 * a - 3.0) * 0.0625;
  dB[2] = -9.0 * (9.0 * a2 + 2.0 * a - 3.0) * 0.0625;
  dB[3] = (27.0 * a2 + 18.0 * a - 1.0) * 0.0625;
}

// position_and_jacobian_radial, utilities_2d.c:50-68
TB_HD void line4_point(const Line4& L, double a, double& r, double& z, double& jac) {
  double B[4], dB[4];
  lagrange4(a, B, dB);
  double rr = 0, zz = 0, dr = 0, dz = 0;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    rr = fma(B[n], L.r[n], rr); zz = fma(B[n], L.z[n], zz);
    dr = fma(dB[n], L.r[n], dr); dz = fma(dB[n], L.z[n], dz);
  }
  r = rr; z = zz; jac = sqrt(dr * dr + dz * dz);
}

// delta_position_and_jacobian_radial, utilities_2d.c:70-87: offset from alpha = 0,
// basis without constant terms so that small offsets keep their relative precision.
TB_HD void line4_delta(const Line4& L, double a, double& dr_out, double& dz_out, double& jac) {
  double B[4], dB[4];
  lagrange4(a, B, dB);
  double a2 = a * a;
  double D[4] = {a * (-9.0 * a2 + 9.0 * a + 1.0) * 0.0625, 9.0 * a * (3.0 * a2 - a - 3.0) * 0.0625,
                 -9.0 * a * (3.0 * a2 + a - 3.0) * 0.0625, a * (9.0 * a2 + 9.0 * a - 1.0) * 0.0625};
  double rr = 0, zz = 0, dr = 0, dz = 0;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    rr = fma(D[n], L.r[n], rr); zz = fma(D[n], L.z[n], zz);
    dr = fma(dB[n], L.r[n], dr); dz = fma(dB[n], L.z[n], dz);
  }
  dr_out = rr; dz_out = zz; jac = sqrt(dr * dr + dz * dz);
}

// higher_order_normal_radial, utilities_2d.c:32-48: tangent (dr,dz) -> (dz,-dr)/|.|
TB_HD void line4_normal(const Line4& L, double a, double& nr, double& nz) {
  double B[4], dB[4];
  lagrange4(a, B, dB);
  double dr = 0, dz = 0;
#pragma unroll
  for (int n = 0; n < 4; n++) { dr = fma(dB[n], L.r[n], dr); dz = fma(dB[n], L.z[n], dz); }
  double len = sqrt(dr * dr + dz * dz);
  nr = dz / len; nz = -dr / len;
}

// flux_density_to_charge_factor, radial_ring.c:4-18
TB_HD double flux_factor(double K) { return 2.0 * (K - 1.0) / (1.0 + K); }

// ---------------------------------------------------------------------------------
// Ring kernels (radial_ring.c:21-53).  Target ring (r0, z0), source ring offset
// (dr, dz); source radius r = r0 + dr.  S = dz^2 + (r + r0)^2, p = (dr^2 + dz^2)/S.
// ---------------------------------------------------------------------------------

// potential_radial_ring, radial_ring.c:37-43.  One reciprocal square root gives both
// 1/sqrt(S) and 1/S.
TB_HD double ring_potential(double r0, double dr, double dz) {
  const double r = r0 + dr;
  const double s = r + r0;
  const double dz2 = dz * dz;
  const double S = fma(s, s, dz2);
  const double rs = rsqrt_(S);
  const double p = fma(dr, dr, dz2) * (rs * rs);
  return ellipkm1_(p) * r * rs * INV_PI;
}

// dr1/dz1_potential_radial_ring, radial_ring.c:21-35, 45-53:
//   d/dz0 = dz r E / (pi d2 sqrt S)
//   d/dr0 = -r [d2 K + (dr(dr - 2r) - dz^2) E] / (2 pi r0 d2 sqrt S), 0 for r0 < 1e-10
TB_HD void ring_grad(double r0, double dr, double dz, double& g_r, double& g_z) {
  const double r = r0 + dr;
  const double s = r + r0;
  const double dz2 = dz * dz;
  const double d2 = fma(dr, dr, dz2);
  const double S = fma(s, s, dz2);
  const double rs = rsqrt_(S);
  double K, E;
  ellip_km1_em1(d2 * (rs * rs), K, E);
  const double c = r * rs * INV_PI / d2;
  g_z = dz * E * c;
  g_r = (r0 < AXIS_EPS) ? 0.0 : -c * fma(d2, K, (dr * (dr - 2.0 * r) - dz2) * E) / (2.0 * r0);
}

// All three of a pair at once: potential, d/dr0, d/dz0 (one log, one rsqrt, one division).
TB_HD void ring_all(double r0, double dr, double dz, double& phi, double& g_r, double& g_z) {
  const double r = r0 + dr;
  const double s = r + r0;
  const double dz2 = dz * dz;
  const double d2 = fma(dr, dr, dz2);
  const double S = fma(s, s, dz2);
  const double rs = rsqrt_(S);
  double K, E;
  ellip_km1_em1(d2 * (rs * rs), K, E);
  const double rq = r * rs * INV_PI;   // r / (pi sqrt S)
  phi = K * rq;
  const double c = rq / d2;
  g_z = dz * E * c;
  g_r = (r0 < AXIS_EPS) ? 0.0 : -c * fma(d2, K, (dr * (dr - 2.0 * r) - dz2) * E) / (2.0 * r0);
}

// current_field_radial_ring, radial_ring.c:78-107: H of a unit current ring of radius a
// at height y, evaluated at (r, z0).  Returns (Hr, Hz).
TB_HD void current_ring_field(double r, double z0, double a, double y, double& hr, double& hz) {
  double z = z0 - y;
  if (a < AXIS_EPS) { hr = 0.0; hz = 0.0; return; }
  double A = z * z + (a + r) * (a + r);
  double B = z * z + (r - a) * (r - a);
  double k = 4.0 * r * a / A;
  double Kk = ellipk_(k), Ek = ellipe_(k);
  double sA = sqrt(A);
  hr = (r < AXIS_EPS) ? 0.0 : z / (2.0 * PI * r * sA) * ((z * z + r * r + a * a) / B * Ek - Kk);
  hz = 1.0 / (2.0 * PI * sA) * ((a * a - z * z - r * r) / B * Ek + Kk);
}

// current_potential_axial_radial_ring, radial_ring.c:73-76
TB_HD double current_ring_axial_potential(double z0, double r, double z) {
  double dz = z0 - z;
  return -dz / (2.0 * sqrt(dz * dz + r * r));
}

// axial_derivatives_radial_ring, radial_ring.c:109-122 (times r/2 applied by caller's weight)
TB_HD void ring_axial_derivs(double z0, double r, double z, double d[NDERIV]) {
  double u = z0 - z;
  double R2 = fma(u, u, r * r);
  double R = sqrt(R2);
  double iR2 = 1.0 / R2;
  d[0] = 1.0 / R;
  d[1] = -u * d[0] * iR2;
#pragma unroll
  for (int n = 1; n < NDERIV - 1; n++) d[n + 1] = -((2 * n + 1) * u * d[n] + double(n * n) * d[n - 1]) * iR2;
}

// current_axial_derivatives_radial_ring, radial_ring.c:124-136
TB_HD void current_ring_axial_derivs(double z0, double r, double z, double d[NDERIV]) {
  double u = z0 - z;
  double R2 = u * u + r * r;
  double R = sqrt(R2);
  double mu = u / R;
  d[0] = -u / (2.0 * R);
  d[1] = -r * r / (2.0 * R2 * R);
#pragma unroll
  for (int n = 2; n < NDERIV; n++) d[n] = -(2 * n - 1) * mu / R * d[n - 1] - double(n * n - 2 * n) / R2 * d[n - 2];
}

// ---------------------------------------------------------------------------------
// Paraxial series evaluation (radial.c:313-384).  coeffs (Nz-1, 9, 6) highest power first.
// ---------------------------------------------------------------------------------
// d[l] for l = L0 .. NDERIV-1 (the field needs no d[0])
template <int L0 = 0>
TB_HD bool axial_local(const double* zs, int Nz, const double* coeffs, double z, double d[NDERIV]) {
  double z0 = zs[0], zl = zs[Nz - 1];
  if (!(z0 <= z && z <= zl)) return false;
  double dz = zs[1] - zs[0];
  int idx = int((z - z0) / dz);
  idx = idx < 0 ? 0 : idx;
  idx = idx > Nz - 2 ? Nz - 2 : idx;
  double x = z - zs[idx];
  const double* C = coeffs + (size_t)idx * (NDERIV * 6);
#pragma unroll
  for (int l = L0; l < NDERIV; l++) {
    const double* c = C + 6 * l;
    d[l] = fma(fma(fma(fma(fma(c[0], x, c[1]), x, c[2]), x, c[3]), x, c[4]), x, c[5]);
  }
  return true;
}

TB_HD double axial_potential(const double* zs, int Nz, const double* coeffs, double x, double y, double z) {
  double d[NDERIV];
  if (!axial_local(zs, Nz, coeffs, z, d)) return 0.0;
  double r2 = x * x + y * y;
  return d[0] - r2 / 4.0 * d[2] + r2 * r2 / 64.0 * d[4] - r2 * r2 * r2 / 2304.0 * d[6] + r2 * r2 * r2 * r2 / 147456.0 * d[8];
}

TB_HD void axial_field(const double* zs, int Nz, const double* coeffs, double x, double y, double z, double out[3]) {
  double d[NDERIV];
  if (!axial_local<1>(zs, Nz, coeffs, z, d)) { out[0] = out[1] = out[2] = 0.0; return; }
  double r2 = x * x + y * y;
  double fr = 0.5 * (d[2] - r2 / 8.0 * d[4] + r2 * r2 / 192.0 * d[6] - r2 * r2 * r2 / 9216.0 * d[8]);
  double fz = -d[1] + r2 / 4.0 * d[3] - r2 * r2 / 64.0 * d[5] + r2 * r2 * r2 / 2304.0 * d[7];
  out[0] = x * fr; out[1] = y * fr; out[2] = fz;
}

}  // namespace tb
