// Device field programs and the tracer launch interface (trace.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tb {

constexpr double STEP_MAX = 0.01;   // TRACING_STEP_MAX, tracing.c:5

enum { COMP_BEM = 0, COMP_AXIAL = 1, COMP_UNIFORM = 2 };
enum { BOUNDS_NONE = 0, BOUNDS_STRICT = 1, BOUNDS_INCLUSIVE = 2 };
enum { STATUS_LEFT_BOUNDS = 0, STATUS_BLOCK_FULL = 1, STATUS_STEP_CAP = 2, STATUS_PARKED = 3 };

// Flattened ring sources: weight (charge * Gauss weight * jacobian, or current density
// * weight * area), ring radius, ring height.
struct SourceSet {
  const double* w;
  const double* r;
  const double* z;
  int n, pad;
};

// One field of a superposition.  A plain FieldRadialBEM / FieldRadialAxial is one
// component without transform; FieldSuperposition members carry their affine map.
struct FieldComp {
  int kind, transform, bounds_mode, pad;
  double factor;
  double inv[12];    // global -> local, 3x4 row-major
  double basis[9];   // local field vector -> global, row-major
  double bounds[6];
  SourceSet elec, mag, cur;
  const double* zs;
  const double* ce;
  const double* cm;
  int nz, axial_parts;     // bit 0: electrostatic series present, bit 1: magnetostatic; 0 = both
  double uE[3], uH[3];
  // learned correction of an axial component (null: none).  mlp: FP32 tracer weights
  // W1 (128,2) b1 | W2^T (in,out) b2 | W3^T b3 | W4^T b4 | W5 (5,128) b5; mlp_zs: s_k(z)
  // (5, mlp_nz) at the z-bin centres; mlp_geom: r_lo, r_max, z_min, z_max, weight power,
  // trust radius (the BEM sources serve r > trust radius).
  // Outside r <= r_max, z in [z_min, z_max] the component uses its ring sources (BEM).
  const float* mlp;
  const double* mlp_zs;
  int mlp_nz, pad3;
  double mlp_geom[6];
  // BF16 hidden weights W2 W3 W4 as trained, (out, in) each (3 x 128 x 128); null: the
  // correction runs per lane in FP32 (mlp_correct) instead of on MFMA (mlp_wave)
  const uint16_t* mlp_wbf;
};

// PROGRAM_HAS_MLP: some component carries a learned correction.  PROGRAM_NO_SOURCES: no
// component has ring sources (axial and uniform fields only): the tracer then runs a
// variant without the source-sum code (far fewer registers); a BEM component in such a
// program evaluates to NaN rather than being skipped.  PROGRAM_LEARNED_ONLY: the program is
// one untransformed learned component with BF16 weights, whose ring sources serve only its
// fallback; with TraceParams::park set the persistent tracer parks every particle that needs
// them and runs a variant without the source-sum code.
enum { PROGRAM_HAS_MLP = 1, PROGRAM_NO_SOURCES = 2, PROGRAM_LEARNED_ONLY = 4 };
struct FieldProgram {
  const FieldComp* comp;   // device array
  int ncomp, bounds_mode;
  int flags, pad;          // PROGRAM_HAS_MLP, PROGRAM_NO_SOURCES, PROGRAM_LEARNED_ONLY
  double bounds[6];
};

struct TraceParams {
  FieldProgram field;
  const double* state0;   // (n, 6)
  const double* t0;       // (n) or null
  const double* h0;       // (n) or null: first step, default hmax
  const double* qm;       // charge / mass, (n) or (1)
  int qm_stride, n;
  int* counter;           // zeroed work counter
  double bounds[6];
  double atol;
  long long max_attempts;
  double plane_p[3], plane_n[3], plane_nn;
  double* final_state;    // (n, 6)
  double* landing;        // (n, 6) or null
  int* status;
  long long* n_accepted;
  long long* n_attempted;
  double* t_final;
  double* h_final;
  double* traj;           // (n, traj_cap, 6) or null
  double* times;          // (n, traj_cap) or null
  long long traj_cap;
  long long* traj_len;
  // a second landing plane (landing2 non-null): e.g. the entrance plane that reflected
  // particles cross on their way back
  double plane2_p[3], plane2_n[3], plane2_nn;
  double* landing2;       // (n, 6) or null
  // Launch shape of the persistent tracer (the split tracer ignores these).  order: the
  // k-th particle taken from the queue is order[k] (null: k); a queue ordered by axial
  // start and speed lets the lanes of a wave travel together, so that their loads of the
  // axial coefficients fall in few cache lines.  batch = 1: a wave takes 64 particles at
  // once, when all its lanes are free, instead of refilling lanes one by one.
  // occupancy: waves per SIMD the axial-only tracer is compiled for (1-4; 0 picks 2).
  const int* order;
  int batch, occupancy;
  // hmax0: each particle's step cap (null or <= 0: STEP_MAX / |v| of state0, tracing.c:68-70);
  // hmax_final: the cap each particle traced with (null: not written), so that a trace
  // resumed from final_state, t_final and h_final keeps it.
  const double* hmax0;
  double* hmax_final;
  // park = 1 (persistent tracer, a field with one learned component): a particle whose step
  // needs a ring-source sum (outside the trust region, or where the error check fails) is
  // retired at the start of that step with STATUS_PARKED (its state, t, h and counts at that
  // point), instead of its wave summing the sources; the caller finishes the parked particles
  // on the stage-synchronous tracer, which spreads each sum over the GPU.
  int park;
};

// One lane per particle; `waves` 64-lane workgroups drain the particle queue.
hipError_t launch_trace(const TraceParams& P, int waves, hipStream_t st);
// Queue order for the persistent tracer: a counting sort of the particles by a 16-bit key,
// the start z (major, 16 bins when the starts differ in z) and the start v_z (minor).
// Particles with equal keys keep no particular order (their results do not depend on it).
// ws: trace_order_workspace_bytes(n).
size_t trace_order_workspace_bytes(int n);
hipError_t trace_order(const double* state0, int n, int* order, void* ws, hipStream_t st);
// launch_trace with the work counter (and, ordered != 0, the queue order and its sort
// workspace) in a static device array instead of caller memory: no allocation.  At most
// TRACE_STATIC_MAX_PARTICLES particles (else hipErrorInvalidValue).  Overlapping calls on
// one GPU share the array and race.
constexpr int TRACE_STATIC_MAX_PARTICLES = 1 << 21;
size_t trace_static_scratch_bytes();
hipError_t trace_noalloc(const TraceParams& P, int waves, int ordered, hipStream_t st);
// Seeded beam of n particles on the device (layout of tb_beam_t): start points uniform in a
// disc of radius r_max on the plane z = z0, kinetic energy e_median exp(e_sigma g), angles
// tan^-1 of the direction's x/z and y/z slopes N(0, angle_sigma), non-relativistic speed.
// Draw k of particle i is output 8 i + k + 1 of SplitMix64 seeded with *seed (device memory,
// so a captured launch takes a new seed per replay).
struct BeamSpec {
  double z0, r_max, e_median, e_sigma, angle_sigma, mass, charge;
};
hipError_t sample_beam(const BeamSpec& B, const unsigned long long* seed, int n, double* state0, hipStream_t st);
// Least-squares focus of n rays given by states (x, y, z, vx, vy, vz) `stride` doubles apart
// (focus.py:9-43): out = (X, Y, Z).  ws: focus_workspace_bytes().
size_t focus_workspace_bytes();
hipError_t focus_position(const double* states, int n, long long stride, double* out, void* ws, hipStream_t st);
// Stage-synchronous tracer (few particles, n <= 65535): every RK stage spreads each
// particle's source sums over nchunk workgroups.  ws: split_workspace_bytes(n, nchunk).
int split_chunks(int n_particles, long long n_sources);
size_t split_workspace_bytes(int n_particles, int nchunk);
hipError_t trace_split(const TraceParams& P, int nchunk, void* ws, hipStream_t st);
hipError_t launch_eval_program(const FieldProgram& F, const double* pts, int M, double* E, double* H, hipStream_t st);

}  // namespace tb
