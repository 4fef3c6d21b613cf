/*
 * traceon_hip.h — C ABI of libtraceon_hip.so, the MI355X (gfx950) implementation of
 * Traceon's radially symmetric BEM path.
 *
 * Two tiers.
 *
 * Tier R — the reference's own FFI.  Same symbol names, argument order, array layouts
 * and accumulate semantics as the C backend that traceon/backend/__init__.py binds
 * (signature table at traceon/backend/__init__.py:287-328).  Arguments are HOST pointers
 * (numpy arrays); every call copies its inputs to the GPU, runs there and copies the
 * results back before returning.  A maintainer switches Traceon to this library by
 * pointing the loader at it (traceon/backend/__init__.py:23-55); see INTEGRATION.md.
 * Where the reference has no error channel (void functions, exit(1) on a bad type) the
 * call records a message readable with tb_last_error() and leaves its outputs untouched.
 *
 * Tier B — batched, device-resident entry points the reference lacks (the GPU's way of
 * doing the same work).  Pointers are DEVICE pointers, `stream` is a hipStream_t (NULL:
 * the default stream), calls are asynchronous on that stream and return 0 on success or
 * -1 with tb_last_error() set.
 *
 * Units are SI; charges are stored as sigma/epsilon_0 (traceon/__init__.py:35).
 */
#ifndef TRACEON_HIP_H
#define TRACEON_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------ */
/* Tier R — data symbols read with ctypes in_dll (traceon/backend/__init__.py:58-66)    */
/* ------------------------------------------------------------------------------------ */
extern const size_t TRACING_BLOCK_SIZE;   /* tracing.c:7 */
extern const int N_QUAD_2D_SYM;           /* utilities_2d.c:3 */
extern const int N_TRIANGLE_QUAD_SYM;     /* utilities_3d.c:5 */
extern const int DERIV_2D_MAX_SYM;        /* radial.c:3 */
extern const int NU_MAX_SYM;              /* three_d.c:10 */
extern const int M_MAX_SYM;               /* three_d.c:11 */

/* Reference structs carried through void* arguments (radial.c:10-15, three_d.c:18-31,
   radial.c:175-180).  Pointers inside are host pointers. */
struct effective_point_charges_2d { double *charges; double *jacobians; double *positions; size_t N; };
struct effective_point_charges_3d { double *charges; double *jacobians; double *positions; size_t N; };
struct field_evaluation_args { void *elec_charges; void *mag_charges; void *currents; double *bounds; };
struct field_derivs_args { double *z_interpolation; double *electrostatic_axial_coeffs; double *magnetostatic_axial_coeffs; size_t N_z; };

typedef void (*field_fun)(double *position, double *velocity, void *args, double *elec_out, double *mag_out);

/* elliptic.c:9-75 — Cody (1965) complete elliptic integrals (backend/__init__.py:291-294) */
double ellipkm1(double p);
double ellipem1(double p);
double ellipk(double k);
double ellipe(double k);

/* utilities_2d.c:18-87 (backend/__init__.py:295-300) */
void normal_2d(const double *p1, const double *p2, double *normal);
void higher_order_normal_radial(double alpha, const double *v1, const double *v2, const double *v3, const double *v4, double *normal);
void position_and_jacobian_radial(double alpha, const double *v1, const double *v2, const double *v3, const double *v4,
                                  double *pos_out, double *jac);
void delta_position_and_jacobian_radial(double alpha, const double *v1, const double *v2, const double *v3, const double *v4,
                                        double *pos_out, double *jac);

/* radial_ring.c:4-107 (backend/__init__.py:302-304, 308, 316, 318) */
double potential_radial_ring(double r0, double z0, double delta_r, double delta_z, void *unused);
double dr1_potential_radial_ring(double r0, double z0, double delta_r, double delta_z, void *unused);
double dz1_potential_radial_ring(double r0, double z0, double delta_r, double delta_z, void *unused);
double flux_density_to_charge_factor(double K);
double current_potential_axial_radial_ring(double z0, double r, double z);
void current_field_radial_ring(double x0, double y0, double x, double y, double *result);

/* radial.c:224-246, three_d.c:34-62 (backend/__init__.py:321, 325) */
void fill_jacobian_buffer_radial(double *jacobian_buffer, double *pos_buffer, const double *line_points, size_t N_lines);
void fill_jacobian_buffer_3d(double *jacobian_buffer, double *pos_buffer, const double *triangles, size_t N_triangles);

/* radial.c:249-310 (backend/__init__.py:324): rows [start,end] inclusive, ACCUMULATES into matrix
   (row stride N_matrix).  An unknown excitation type records an error and changes nothing. */
void fill_matrix_radial(double *matrix, const double *line_points, const uint8_t *excitation_types,
                        const double *excitation_values, const double *jacobian_buffer, const double *pos_buffer,
                        size_t N_lines, size_t N_matrix, int lines_range_start, int lines_range_end);

/* The singular diagonal's integrands (radial.c:182-222; backend/__init__.py:322-323), with
   the reference's signatures, for scipy.integrate.quad through LowLevelCallable
   (backend/__init__.py:688-708): user data is the (4,3) element, or the struct below.  Each
   call evaluates one alpha on the GPU. */
struct self_field_dot_normal_radial_args { double *line_points; double K; };   /* radial.c:198-201 */
double self_potential_radial(double alpha, void *line_points);
double self_field_dot_normal_radial(double alpha, void *args);
/* The same integrals in one call each (no reference counterpart): a graded Gauss rule over
   [-1, 1] on the GPU.  self_field_dot_normal_radial_integral returns the integral itself (the
   solver subtracts 1, solver.py:245). */
double self_potential_radial_integral(const double *line_points);
double self_field_dot_normal_radial_integral(const double *line_points, double K);

/* The flat-triangle helpers of the reference's FFI (utilities_3d.c:95-134, three_d.c:310-328;
   backend/__init__.py:297-298, 327), evaluated on the GPU.  triangle_areas is used by the
   mesher for coil triangles (mesher.py:249). */
void normal_3d(const double *triangle, double *normal);
void position_and_jacobian_3d(double alpha, double beta, const double *triangle, double *pos_out, double *jac);
void triangle_areas(const double *triangles, double *out, size_t N);

/* Entries of the reference's FFI table that lie off the radially symmetric path
   (backend/__init__.py:289, 312-315): exported so the reference's loader binds every name of
   its table; each records an error readable with tb_last_error() and returns NaN or leaves its
   outputs untouched.  kronrod_adaptive (kronrod.c:43) integrates a host callback, which cannot
   run on the GPU; the other four belong to 3D BEM. */
double kronrod_adaptive(double (*f)(double, void *), double a, double b, void *args, double abs_tol, double rel_tol);
void axial_coefficients_3d(const void *charges, const void *jacobian_buffer, const void *pos_buffer, void *trig_cos_buffer,
                           void *trig_sin_buffer, size_t N_charges, const double *z, void *output_coeffs, size_t N_z);
void fill_jacobian_buffer_current_three_d(const double *lines, double *jacobians, double *positions, double *directions, size_t N_lines);
double potential_3d_derivs(const double *point, const double *z, const double *coeffs, size_t N_z);
void field_3d_derivs(const double *point, double *field, const double *z, const double *coeffs, size_t N_z);

/* radial.c:19-173, 449-521 (backend/__init__.py:305-307, 309-311, 317, 319-320) */
double potential_radial(const double *point, const double *charges, const double *jacobian_buffer, const double *pos_buffer, size_t N);
void field_radial(const double *point, double *result, const double *charges, const double *jacobian_buffer, const double *pos_buffer,
                  size_t N);
double current_potential_axial(double z0, const double *currents, const double *jacobian_buffer, const double *pos_buffer, size_t N);
void current_field_radial(const double *point, double *result, const double *currents, const double *jacobian_buffer,
                          const double *pos_buffer, size_t N);
void axial_derivatives_radial(double *derivs, const double *charges, const double *jacobian_buffer, const double *pos_buffer,
                              size_t N_lines, const double *z, size_t N_z);   /* accumulates */
void current_axial_derivatives_radial(double *derivs, const double *currents, const double *jacobian_buffer,
                                      const double *pos_buffer, size_t N, const double *z, size_t N_z);   /* accumulates */
double potential_radial_derivs(const double *point, const double *z_inter, const double *coeff, size_t N_z);
void field_radial_derivs(const double *point, double *field, const double *z_inter, const double *coeff, size_t N_z);
double charge_radial(const double *vertices, double charge);

/* tracing.c:60-223 (backend/__init__.py:301, 326).  trace_particle runs on the GPU when `field`
   is this library's field_radial_traceable or field_radial_derivs_traceable (the two C field
   functions the reference passes, field.py:814-816, 1064-1066); any other function pointer
   (a Python callback) records an error and returns 0. */
size_t trace_particle(double *times_array, double *pos_array, double charge_over_mass, field_fun field,
                      const double *bounds, double atol, void *args);
void field_radial_traceable(double *position, double *velocity, void *args, double *elec_out, double *mag_out);
void field_radial_derivs_traceable(double *position, double *velocity, void *args, double *elec_out, double *mag_out);
bool plane_intersection(const double *p0, const double *normal, const double *positions, size_t N_p, double *result);

/* ------------------------------------------------------------------------------------ */
/* Tier B — device-resident batched entry points (no reference counterpart unless noted) */
/* ------------------------------------------------------------------------------------ */
const char *tb_last_error(void);
void tb_clear_error(void);
int tb_version(void);
/* sizes of tb_field_comp_t, tb_field_program_t, tb_trace_params_t and one row record, for binding checks (host only) */
void tb_struct_sizes(size_t *out4);
int tb_synchronize(void *stream);

/* Element preparation for the fill (fill_jacobian_buffer_radial + the per-row data of
   fill_matrix_radial, radial.c:224-310).  src: 3 x (16, N) SoA (r, z, weight*jac);
   rows: N x tb_rowinfo_bytes(); bad_type (device int, zeroed by caller) is set to 1 if a
   type code is not 1, 2, 3, 5 or 6.  jac/pos may be NULL. */
size_t tb_rowinfo_bytes(void);
int tb_prepare_elements(const double *lines, const uint8_t *types, const double *values, int N, double *jac, double *pos,
                        double *src, void *rows, int *bad_type, void *stream);
/* Dense fill of rows [row0,row1]: stores (accumulate = 0) or adds; M32 (nullable) receives an
   FP32 copy.  Replaces the threaded fill_matrix_radial calls of solver.py:223-236. */
int tb_fill_matrix(double *M, size_t ld, float *M32, size_t ld32, const void *rows, const double *src, int N, int row0, int row1,
                   int accumulate, void *stream);
/* Singular diagonal of rows [row0,row1] (solver.py:239-247): M[i][i] = self potential, or
   self flux - 1 for dielectric/magnetizable rows.  M, M32, diag_out nullable. */
int tb_fill_diagonal(double *M, size_t ld, float *M32, size_t ld32, const double *lines, const void *rows, int N, int row0,
                     int row1, double *diag_out, void *stream);

/* Dense LU with partial pivoting, row-major (numpy.linalg.solve, solver.py:173).
   ipiv: N ints, 0-based pivot rows; info: device int zeroed by caller, set to 1 + first zero pivot. */
size_t tb_lu_workspace_bytes(int N, int fp32);
int tb_lu_factor(double *A, size_t lda, int N, int *ipiv, void *workspace, int *info, void *stream);
int tb_lu_factor_f32(float *A, size_t lda, int N, int *ipiv, void *workspace, int *info, void *stream);
/* Instrumented factorisation (no counterpart in the reference: measurement only).  As
   tb_lu_factor / tb_lu_factor_f32 (fp32 selects the type of A), and accumulates into
   *gemm_timing the number, HIP-event duration (ms, on the stream each is launched on:
   `stream` or the factorisation's second stream, so launches that overlap count in full),
   algorithmic FLOP (2 M N K each) and algorithmic HBM bytes (A and B read once, C read
   and written) of the large GEMM launches of this factorisation: those on 128 x 128 tiles
   (the bulk updates of the panel groups and their strips; FP32: split-BF16 MFMA), not
   the per-panel products. */
typedef struct { long long launches; double ms; double flops; double bytes; } tb_kernel_timing_t;
int tb_lu_factor_timed(void *A, size_t lda, int N, int *ipiv, void *workspace, int *info, int fp32,
                       tb_kernel_timing_t *gemm_timing, void *stream);
/* X (nrhs x N, nrhs <= 64) <- A^{-1} X with the factors and the workspace of the same
   factorisation, which holds the diagonal-block inverses and the row permutation
   (FP64 vectors; FP32 factors promoted on load) */
int tb_lu_solve(const double *LU, size_t lda, int N, const int *ipiv, double *X, int nrhs, const void *workspace, void *stream);
int tb_lu_solve_f32(const float *LU, size_t lda, int N, const int *ipiv, double *X, int nrhs, const void *workspace, void *stream);
/* R = B - A X in FP64, rmax[r] = bit pattern of max |R[r]| (nonnegative double) */
int tb_residual(const double *A, size_t lda, int N, const double *B, const double *X, double *R, int nrhs,
                unsigned long long *rmax, void *stream);
int tb_axpy(double *X, const double *D, size_t n, void *stream);
/* status (device int) <- 1 if a dependency wait of the last tb_lu_solve / tb_lu_solve_f32 on
   this workspace passed its spin limit (its solution is then NaN), else 0; fp32 as the
   factorisation.  The refinement loop reads it with the norms below (one host sync per step). */
int tb_lu_solve_status(const void *workspace, int N, int fp32, int *status, void *stream);
/* *out (device) = bit pattern of max |x[i]| over n doubles (NaN: the largest pattern) */
int tb_absmax(const double *x, size_t n, unsigned long long *out, void *stream);
int tb_convert_f64_f32(const double *S, size_t lds, float *D, size_t ldd, int rows, int cols, void *stream);
/* C = beta C + alpha A B (row-major), the MFMA GEMM used by the factorisation */
int tb_gemm_f64(int M, int N, int K, double alpha, const double *A, size_t lda, const double *B, size_t ldb, double beta, double *C,
                size_t ldc, void *stream);
int tb_gemm_f32(int M, int N, int K, float alpha, const float *A, size_t lda, const float *B, size_t ldb, float beta, float *C,
                size_t ldc, void *stream);

/* Ring sources for the field sums: w = q * jac, r, z flattened over (element, quad point);
   nq = 16 (lines, pstride 2, zcol 1) or 12 (coil triangles, pstride 3, zcol 2). */
int tb_sources(const double *q, const double *jac, const double *pos, int N, int nq, int pstride, int zcol, double *w, double *r,
               double *z, void *stream);
/* potential_radial / field_radial at M points (M x 3); pot or field nullable */
int tb_field_points(const double *pts, int M, const double *w, const double *r, const double *z, int Ns, double *pot, double *field,
                    int accumulate, void *stream);
/* current_field_radial at M points */
int tb_current_points(const double *pts, int M, const double *w, const double *r, const double *z, int Ns, double *field,
                      int accumulate, void *stream);
int tb_current_axial_potential(const double *zq, int M, const double *w, const double *r, const double *z, int Ns, double *out,
                               void *stream);
/* axial_derivatives_radial / current_axial_derivatives_radial (accumulate into out (Nz, 9)) */
int tb_axial_derivatives(double *out, const double *zq, int Nz, const double *w, const double *r, const double *z, int Ns,
                         int current, void *stream);
/* FieldRadialAxial coefficients (field.py:906-937): C (Nz-1, 9, 6); scratch >= 6*Nz doubles */
int tb_axial_coefficients(const double *derivs, const double *zq, int Nz, double *C, double *scratch, void *stream);
int tb_axial_points(const double *pts, int M, const double *zq, int Nz, const double *C, double *pot, double *field, void *stream);
int tb_jacobian_3d(const double *tri, int N, double *jac, double *pos, void *stream);
int tb_charge_radial(const double *lines, const double *q, int N, double *out, void *stream);

/* Field programs and batched tracing (tracing.c:60-174 for many particles at once). */
enum { TB_COMP_BEM = 0, TB_COMP_AXIAL = 1, TB_COMP_UNIFORM = 2 };
enum { TB_BOUNDS_NONE = 0, TB_BOUNDS_STRICT = 1, TB_BOUNDS_INCLUSIVE = 2 };
enum { TB_STATUS_LEFT_BOUNDS = 0, TB_STATUS_BLOCK_FULL = 1, TB_STATUS_STEP_CAP = 2, TB_STATUS_PARKED = 3 };

typedef struct { const double *w; const double *r; const double *z; int n, pad; } tb_sources_t;
typedef struct {
  int kind, transform, bounds_mode, pad;
  double factor;
  double inv[12];
  double basis[9];
  double bounds[6];
  tb_sources_t elec, mag, cur;
  const double *zs; const double *ce; const double *cm;
  int nz, axial_parts;   /* bit 0: electrostatic coefficients present, bit 1: magnetostatic;
                            0 = both.  An absent series is not evaluated (its field is zero). */
  double uE[3], uH[3];
  /* learned correction of an axial component (null: none): FP32 weights W1 b1 W2^T b2 W3^T b3
     W4^T b4 W5 b5, z-scale table (5, mlp_nz), geometry {r_lo, r_max, z_min, z_max, power,
     trust radius}; outside the trust region (r above the trust radius, or z outside the axial
     range) the component's ring sources give the field */
  const float *mlp; const double *mlp_zs; int mlp_nz, pad3; double mlp_geom[6];
  /* BF16 hidden weights W2 W3 W4 as trained, (out, in) each (3 x 128 x 128): with them a
     component traced alone evaluates its correction for a whole wave on MFMA; null: per lane */
  const uint16_t *mlp_wbf;
} tb_field_comp_t;
/* flags: 1 = some component carries a learned correction; 2 = no component has ring sources
   (axial and uniform fields only: the tracer runs its variant without source sums, and a BEM
   component in such a program evaluates to NaN); 4 = the program is one untransformed learned
   component with BF16 weights (its ring sources serve only its fallback): with park = 1,
   tb_trace runs a variant without source sums and parks every particle that needs them */
typedef struct { const tb_field_comp_t *comp; int ncomp, bounds_mode; int flags, pad; double bounds[6]; } tb_field_program_t;
typedef struct {
  tb_field_program_t field;
  const double *state0; const double *t0; const double *h0; const double *qm;
  int qm_stride, n;
  int *counter;
  double bounds[6];
  double atol;
  long long max_attempts;
  double plane_p[3], plane_n[3], plane_nn;
  double *final_state; double *landing; int *status;
  long long *n_accepted; long long *n_attempted;
  double *t_final; double *h_final;
  double *traj; double *times; long long traj_cap; long long *traj_len;
  /* second landing plane (landing2 non-null; NaN rows where never crossed): with the entrance
     plane it records where reflected particles leave, beside the exit plane's landing */
  double plane2_p[3], plane2_n[3], plane2_nn;
  double *landing2;
  /* launch shape of tb_trace (tb_trace_split ignores it): order (n ints, null: 0..n-1) is the
     order in which the queue hands out particles (tb_trace_order makes one in which the
     lanes of a wave travel together); batch = 1: a wave takes 64 particles at once when all
     its lanes are free; occupancy: waves per SIMD of the axial-only tracer (1-4, 0: 2).
     None of them changes any particle's result. */
  const int *order;
  int batch, occupancy;
  /* hmax0 (n, or null): each particle's step cap (null or <= 0: TRACING_STEP_MAX / |v| of
     state0, tracing.c:68-70); hmax_final (n, or null): the cap each particle traced with, so
     that a trace resumed from final_state, t_final, h_final and hmax_final continues with the
     same steps. */
  const double *hmax0; double *hmax_final;
  /* park = 1 (tb_trace, a program with one learned component): a particle whose step needs a
     ring-source sum (outside the trust region, or the error check fails) is retired at the
     start of that step with TB_STATUS_PARKED, its state, t, h and counts at that point,
     instead of its wave summing the sources; the caller finishes such particles with
     tb_trace_split, which spreads every sum over the GPU.  tb_trace_split ignores it. */
  int park;
} tb_trace_params_t;

/* Persistent RK45 tracer over params->n particles (all pointers device pointers).
   One lane per particle; waves: number of 64-lane workgroups that drain the particle
   queue (0: 4 x occupancy per CU for axial-only fields, 8 per CU otherwise).  Ring-source sums of direct BEM fields
   run wave-uniformly, every lane of a wave sharing each source. */
int tb_trace(const tb_trace_params_t *params, int waves, void *stream);
/* Queue order for tb_trace: a counting sort of the n particles of state0 (n x 6, device) by
   start z (major) and start v_z (minor) into order (n ints, device); workspace of
   tb_trace_order_workspace_bytes(n) bytes. */
size_t tb_trace_order_workspace_bytes(int n);
int tb_trace_order(const double *state0, int n, int *order, void *workspace, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): tb_trace
   with its work counter, and for ordered != 0 the queue order of tb_trace_order and its
   sort workspace, in a static device array of tb_trace_noalloc_scratch_bytes() bytes (about
   16.3 MiB, held by each GPU the library runs on until the process ends) instead of caller
   memory; params->counter is ignored, and params->order too when ordered != 0.  At most
   tb_trace_noalloc_max_particles() = 2^21 particles (else -1 with an error).  Allocates
   nothing; the caller still passes every input and output.  Calls that overlap on one GPU
   share the array and race. */
int tb_trace_noalloc(const tb_trace_params_t *params, int waves, int ordered, void *stream);
int tb_trace_noalloc_max_particles(void);
size_t tb_trace_noalloc_scratch_bytes(void);
/* Beam focus, replaces focus_position (focus.py:9-43) on device-resident final states: the
   least-squares point (out[0..2] = x, y, z) of the n >= 2 rays (x, y, z) + t (vx/vz, vy/vz, 1)
   whose states (x, y, z, vx, vy, vz) lie `stride` doubles apart (6 for a (n, 6) array), in the
   closed form of the reference's numpy.linalg.lstsq system (centred sums, fixed summation
   order).  Rays that are all parallel have no focus: the result is then not finite.
   Workspace of tb_focus_workspace_bytes() bytes. */
/* No counterpart in the reference (which takes start states from the caller; the beams of its
   examples are drawn with numpy on the host).  A seeded beam drawn on the device, so that a
   sample -> trace sequence can be captured once as a HIP graph and replayed per seed:
   state0 (n, 6) = (x, y, z0, vx, vy, vz) with (x, y) uniform in the disc of radius r_max,
   kinetic energy e_median * exp(e_sigma * g) eV (g ~ N(0, 1)), the direction (tan ax, tan ay, 1)
   normalised with ax, ay ~ N(0, angle_sigma), speed sqrt(2 E charge / mass) (charge: the
   magnitude in C).  Draw k of particle i is output 8 i + k + 1 of SplitMix64 seeded with *seed,
   read from device memory when the launch runs.  Allocates nothing. */
typedef struct {
  double z0, r_max, e_median, e_sigma, angle_sigma, mass, charge;
} tb_beam_t;
int tb_sample_beam(const tb_beam_t *beam, const unsigned long long *seed, int n, double *state0, void *stream);
size_t tb_focus_workspace_bytes(void);
int tb_focus_position(const double *states, int n, long long stride, double *out, void *workspace, void *stream);
/* Stage-synchronous tracer for a few particles (n <= 65535) in fields with ring-source
   sums: each RK stage spreads every particle's source sum over nchunk workgroups and adds
   the partials in a fixed order; the RK arithmetic per particle is that of tb_trace.
   nchunk from tb_trace_split_chunks(n, total ring sources); workspace of
   tb_trace_split_workspace_bytes(n, nchunk) bytes.  Synchronises `stream`.  Returns -1
   (tb_last_error) after 2^22 attempted steps without every particle done. */
int tb_trace_split_chunks(int n_particles, long long n_sources);
size_t tb_trace_split_workspace_bytes(int n_particles, int nchunk);
int tb_trace_split(const tb_trace_params_t *params, int nchunk, void *workspace, void *stream);
/* Field program at M points (E, H nullable). */
int tb_eval_program(const tb_field_program_t *program, const double *pts, int M, double *E, double *H, void *stream);

/* ---- Learned field surrogate (no counterpart in the reference) -----------------------
   A 2-128-128-128-128-5 SiLU MLP; hidden layers in BF16 on MFMA with FP32 accumulation,
   FP32 master parameters (tb_mlp_nparams() floats: W1 (128,2) b1 | W2 (128,128) b2 | W3 b3 |
   W4 b4 | W5 (5,128) b5, W stored [out][in]); wbf = BF16 copies of W2..W4 (3*128*128 u16).
   Workspaces of tb_mlp_workspace_bytes(B) bytes.  All pointers device pointers. */
int tb_mlp_nparams(void);
size_t tb_mlp_workspace_bytes(int B);
/* C (+)= op(A) op(B): A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], BF16 in, FP32 out;
   ksplit > 1 splits K over workgroups (C zeroed first unless accumulate). */
int tb_gemm_bf16(int M, int N, int K, const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, float *C,
                 long ldc, int accumulate, int ksplit, void *stream);
int tb_mlp_forward(const float *params, const uint16_t *wbf, const float *X, int B, float *Y, void *workspace, void *stream);
/* gradient of mean((Y - T)^2) over the batch and the 5 outputs into grad; *loss = sum of squares */
int tb_mlp_grad(const float *params, const uint16_t *wbf, const float *X, const float *T, int B, float *grad, float *loss,
                void *workspace, void *stream);
int tb_mlp_adam(float *params, const float *grad, float *m, float *v, uint16_t *wbf, float lr, float beta1, float beta2, float eps,
                int step, void *stream);
int tb_mlp_to_bf16(const float *params, uint16_t *wbf, void *stream);
int tb_mlp_gather(int B, const int *idx, const float *X, const float *T, float *Xb, float *Tb, void *stream);

#ifdef __cplusplus
}
#endif
#endif
