// C ABI of libtraceon_hip.so (declared in include/traceon_hip.h).
//
// Tier R entry points keep the reference's names and host-pointer signatures
// (traceon/backend/__init__.py:287-328): each stages its inputs into device memory,
// runs the GPU kernels on the default stream, waits and copies the outputs back.
// Tier B entry points take device pointers and a stream and only enqueue work.
// No entry point computes on the host: everything numeric runs in the HIP kernels.
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <cstring>
#include <cstdio>
#include <memory>
#include "../../include/traceon_hip.h"
#include "bem_kernels.h"
#include "lu_kernels.h"
#include "trace_kernels.h"
#include "mlp_kernels.h"

static_assert(sizeof(tb_sources_t) == sizeof(tb::SourceSet), "source set layout");
static_assert(sizeof(tb_field_comp_t) == sizeof(tb::FieldComp), "field comp layout");
static_assert(sizeof(tb_field_program_t) == sizeof(tb::FieldProgram), "field program layout");
static_assert(sizeof(tb_trace_params_t) == sizeof(tb::TraceParams), "trace params layout");

extern "C" {
const size_t TRACING_BLOCK_SIZE = 100000;
const int N_QUAD_2D_SYM = 16;
const int N_TRIANGLE_QUAD_SYM = 12;
const int DERIV_2D_MAX_SYM = 9;
const int NU_MAX_SYM = 4;
const int M_MAX_SYM = 8;
}

namespace {

thread_local std::string g_err;

void set_err(const std::string& s) { g_err = s; }

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  set_err(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

hipStream_t S(void* stream) { return reinterpret_cast<hipStream_t>(stream); }

// Device buffer owned by a Tier R call.
template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  bool good = true;
  explicit Dev(size_t count) : n(count) {
    if (n && hipMalloc(&p, n * sizeof(T)) != hipSuccess) { p = nullptr; good = false; set_err("hipMalloc failed"); }
  }
  Dev(const T* host, size_t count) : Dev(count) {
    if (good && n && host && hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
      good = false; set_err("hipMemcpy H2D failed");
    }
  }
  ~Dev() { if (p) (void)hipFree(p); }
  bool get(T* host, size_t count) const {
    if (!count) return true;
    return ok(hipMemcpy(host, p, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
};

bool finish(const char* what) {
  if (!ok(hipGetLastError(), what)) return false;
  return ok(hipDeviceSynchronize(), what);
}

// Device staging of the one-value entry points (scalar kernels, element geometry, the
// diagonal's integrands, which QUADPACK calls about a thousand times per element): one
// 64-double buffer per host thread and device, allocated at its first use and kept until the
// process ends, so such a call costs a copy in, a launch and a copy out, and no allocation.
constexpr size_t STAGE = 64;
struct Staging {
  int dev = -1;
  double* p = nullptr;
};
thread_local Staging g_stage;

double* staging() {
  int dev = 0;
  if (!ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
  if (g_stage.p && g_stage.dev == dev) return g_stage.p;
  double* p = nullptr;
  if (!ok(hipMalloc(&p, STAGE * sizeof(double)), "hipMalloc (staging)")) return nullptr;
  g_stage.dev = dev;
  g_stage.p = p;   // a buffer of another device stays allocated: the thread may return to it
  return p;
}

// nin inputs in, `launch(din, dout)` on the default stream, nout outputs back (nin + nout <= STAGE)
template <typename F>
bool staged_call(const char* what, const double* in, size_t nin, double* out, size_t nout, F launch) {
  double* d = staging();
  if (!d) return false;
  if (!ok(hipMemcpy(d, in, nin * sizeof(double), hipMemcpyHostToDevice), what)) return false;
  launch(d, d + nin);
  if (!finish(what)) return false;
  return ok(hipMemcpy(out, d + nin, nout * sizeof(double), hipMemcpyDeviceToHost), what);
}

double scalar_op(int op, double a0, double a1 = 0, double a2 = 0, double a3 = 0, double* second = nullptr) {
  const double in[4] = {a0, a1, a2, a3};
  double out[4] = {0, 0, 0, 0};
  if (!staged_call("scalar", in, 4, out, 4, [&](double* di, double* dout) { tb::launch_scalar(op, di, 1, dout, 0); }))
    return __builtin_nan("");
  if (second) *second = out[1];
  return out[0];
}

// (alpha, v1..v4) -> (r, z, jac) or normal, through k_line_eval; v in parameter order
void line_op(int op, double alpha, const double* v1, const double* v2, const double* v3, const double* v4, double out3[3]) {
  double in[13];
  in[0] = alpha;
  // rows 0, 2, 3, 1 of the (4,3) element are the nodes at alpha = -1, -1/3, 1/3, 1
  const double* rowsrc[4] = {v1, v4, v2, v3};
  for (int r = 0; r < 4; r++) { in[1 + 3 * r] = rowsrc[r][0]; in[2 + 3 * r] = 0.0; in[3 + 3 * r] = rowsrc[r][2]; }
  staged_call("line_eval", in, 13, out3, 3, [&](double* di, double* dout) { tb::launch_line_eval(op, di, 1, dout, 0); });
}

// (alpha, beta, triangle (3,3)) -> 4 values of k_tri_eval
bool tri_op(int op, double alpha, double beta, const double* tri, double out4[4]) {
  double in[11];
  in[0] = alpha;
  in[1] = beta;
  std::memcpy(in + 2, tri, 9 * sizeof(double));
  return staged_call("triangle", in, 11, out4, 4, [&](double* di, double* dout) { tb::launch_tri_eval(op, di, 1, dout, 0); });
}

// one value of a diagonal integrand (kind 0 potential, 1 flux with K)
double self_integrand(int kind, double alpha, double K, const double* line) {
  double in[14], out = __builtin_nan("");
  in[0] = alpha;
  in[1] = K;
  std::memcpy(in + 2, line, 12 * sizeof(double));
  staged_call("self integrand", in, 14, &out, 1, [&](double* di, double* dout) { tb::launch_self_integrand(kind, di, 1, dout, 0); });
  return out;
}

// entry points of the reference's FFI that lie off the radial path: they record why and
// return NaN / leave their outputs untouched
void off_path(const char* name, const char* why) { set_err(std::string(name) + ": " + why); }

bool valid_type(uint8_t t) { return t == 1 || t == 2 || t == 3 || t == 5 || t == 6; }

// Sources (w, r, z) for a set of point charges (host arrays); caller owns the Dev buffers
struct DevSources {
  Dev<double> w, r, z;
  int n;
  DevSources(const double* q, const double* jac, const double* pos, size_t N, int nq, int pstride, int zcol)
      : w(N * nq), r(N * nq), z(N * nq), n((int)(N * nq)) {
    if (!N) return;
    Dev<double> dq(q, N), dj(jac, N * nq), dp(pos, N * nq * pstride);
    if (!dq.good || !dj.good || !dp.good || !w.good || !r.good || !z.good) { n = -1; return; }
    tb::launch_sources(dq.p, dj.p, dp.p, (int)N, nq, pstride, zcol, w.p, r.p, z.p, 0);
    if (!finish("sources")) n = -1;
  }
};

}  // namespace

extern "C" {

const char* tb_last_error(void) { return g_err.c_str(); }
void tb_clear_error(void) { g_err.clear(); }
int tb_version(void) { return 1; }
void tb_struct_sizes(size_t* out4) {
  out4[0] = sizeof(tb_field_comp_t);
  out4[1] = sizeof(tb_field_program_t);
  out4[2] = sizeof(tb_trace_params_t);
  out4[3] = sizeof(tb::RowInfo);
}
int tb_synchronize(void* stream) { return ok(hipStreamSynchronize(S(stream)), "synchronize") ? 0 : -1; }

// ====================================================================================
// Tier R
// ====================================================================================
double ellipkm1(double p) { return scalar_op(0, p); }
double ellipem1(double p) { return scalar_op(1, p); }
double ellipk(double k) { return scalar_op(2, k); }
double ellipe(double k) { return scalar_op(3, k); }

void normal_2d(const double* p1, const double* p2, double* normal) {
  double second = 0;
  double first = scalar_op(10, p1[0], p1[1], p2[0], p2[1], &second);
  normal[0] = first; normal[1] = second;
}

void higher_order_normal_radial(double alpha, const double* v1, const double* v2, const double* v3, const double* v4, double* normal) {
  double o[3] = {0, 0, 0};
  line_op(2, alpha, v1, v2, v3, v4, o);
  normal[0] = o[0]; normal[1] = o[1];
}

void position_and_jacobian_radial(double alpha, const double* v1, const double* v2, const double* v3, const double* v4,
                                  double* pos_out, double* jac) {
  double o[3] = {0, 0, 0};
  line_op(0, alpha, v1, v2, v3, v4, o);
  pos_out[0] = o[0]; pos_out[1] = o[1]; *jac = o[2];
}

void delta_position_and_jacobian_radial(double alpha, const double* v1, const double* v2, const double* v3, const double* v4,
                                        double* pos_out, double* jac) {
  double o[3] = {0, 0, 0};
  line_op(1, alpha, v1, v2, v3, v4, o);
  pos_out[0] = o[0]; pos_out[1] = o[1]; *jac = o[2];
}

double potential_radial_ring(double r0, double z0, double dr, double dz, void*) { return scalar_op(4, r0, z0, dr, dz); }
double dr1_potential_radial_ring(double r0, double z0, double dr, double dz, void*) { return scalar_op(5, r0, z0, dr, dz); }
double dz1_potential_radial_ring(double r0, double z0, double dr, double dz, void*) { return scalar_op(6, r0, z0, dr, dz); }
double flux_density_to_charge_factor(double K) { return scalar_op(7, K); }
double current_potential_axial_radial_ring(double z0, double r, double z) { return scalar_op(8, z0, r, z); }
void current_field_radial_ring(double x0, double y0, double x, double y, double* result) {
  double second = 0;
  result[0] = scalar_op(9, x0, y0, x, y, &second);
  result[1] = second;
}

void fill_jacobian_buffer_radial(double* jac, double* pos, const double* lines, size_t N) {
  if (!N) return;
  Dev<double> dl(lines, N * 12), dj(N * 16), dp(N * 32);
  if (!dl.good || !dj.good || !dp.good) return;
  tb::launch_elements(dl.p, nullptr, nullptr, (int)N, dj.p, dp.p, nullptr, nullptr, nullptr, 0);
  if (finish("fill_jacobian_buffer_radial")) { dj.get(jac, N * 16); dp.get(pos, N * 32); }
}

void fill_jacobian_buffer_3d(double* jac, double* pos, const double* tri, size_t N) {
  if (!N) return;
  Dev<double> dt(tri, N * 9), dj(N * 12), dp(N * 36);
  if (!dt.good || !dj.good || !dp.good) return;
  tb::launch_jac3d(dt.p, (int)N, dj.p, dp.p, 0);
  if (finish("fill_jacobian_buffer_3d")) { dj.get(jac, N * 12); dp.get(pos, N * 36); }
}

void fill_matrix_radial(double* matrix, const double* lines, const uint8_t* types, const double* values, const double* jac,
                        const double* pos, size_t N_lines, size_t N_matrix, int start, int end) {
  if (start < 0 || end < start || (size_t)end >= N_lines || N_matrix < N_lines) {
    set_err("fill_matrix_radial: invalid row range or matrix size");
    return;
  }
  for (int i = start; i <= end; i++)
    if (!valid_type(types[i])) { set_err("ExcitationType unknown"); return; }
  const size_t nrows = (size_t)(end - start + 1);
  Dev<double> dl(lines, N_lines * 12), dv(values, N_lines), dj(jac, N_lines * 16), dp(pos, N_lines * 32);
  Dev<uint8_t> dt(types, N_lines);
  Dev<double> src(N_lines * 48);
  Dev<char> rows(N_lines * sizeof(tb::RowInfo));
  Dev<double> dM(matrix + (size_t)start * N_matrix, nrows * N_matrix);
  if (!dl.good || !dv.good || !dj.good || !dp.good || !dt.good || !src.good || !rows.good || !dM.good) return;
  tb::launch_elements(dl.p, dt.p, dv.p, (int)N_lines, nullptr, nullptr, nullptr, (tb::RowInfo*)rows.p, nullptr, 0);
  tb::launch_src_from_buffers(dj.p, dp.p, (int)N_lines, src.p, 0);
  // rows of the device block are offset by `start`
  tb::launch_fill(dM.p - (size_t)start * N_matrix, N_matrix, nullptr, 0, (tb::RowInfo*)rows.p, src.p, (int)N_lines, start, end, 1, 0);
  if (finish("fill_matrix_radial")) dM.get(matrix + (size_t)start * N_matrix, nrows * N_matrix);
}

static double self_integral(const double* line, int type, double K) {
  uint8_t t = (uint8_t)type;
  Dev<double> dl(line, 12), dv(&K, 1), out(1);
  Dev<uint8_t> dt(&t, 1);
  Dev<char> rows(sizeof(tb::RowInfo));
  if (!dl.good || !dv.good || !out.good || !dt.good || !rows.good) return __builtin_nan("");
  tb::launch_elements(dl.p, dt.p, dv.p, 1, nullptr, nullptr, nullptr, (tb::RowInfo*)rows.p, nullptr, 0);
  tb::launch_self(nullptr, 0, nullptr, 0, dl.p, (tb::RowInfo*)rows.p, 1, 0, 0, out.p, 0);
  double v = __builtin_nan("");
  if (finish("self integral")) out.get(&v, 1);
  return v;
}
double self_potential_radial_integral(const double* line) { return self_integral(line, 1, 0.0); }
double self_field_dot_normal_radial_integral(const double* line, double K) { return self_integral(line, 3, K); }

// The integrands themselves, with the reference's signatures (radial.c:182-222), for
// scipy.integrate.quad through LowLevelCallable (backend/__init__.py:688-708): each call
// evaluates one alpha on the GPU.
double self_potential_radial(double alpha, void* line_points) {
  if (!line_points) { set_err("self_potential_radial: null element"); return __builtin_nan(""); }
  return self_integrand(0, alpha, 0.0, (const double*)line_points);
}
double self_field_dot_normal_radial(double alpha, void* args) {
  const auto* a = (const self_field_dot_normal_radial_args*)args;
  if (!a || !a->line_points) { set_err("self_field_dot_normal_radial: null arguments"); return __builtin_nan(""); }
  return self_integrand(1, alpha, a->K, a->line_points);
}

// ---- the reference FFI's 3D and generic entries (backend/__init__.py:289, 297-298, 312-315, 327) ----
double kronrod_adaptive(double (*f)(double, void*), double a, double b, void* args, double abs_tol, double rel_tol) {
  (void)f; (void)a; (void)b; (void)args; (void)abs_tol; (void)rel_tol;
  off_path("kronrod_adaptive", "integrates a host callback, which cannot run on the GPU; not on the radial path");
  return __builtin_nan("");
}
void normal_3d(const double* tri, double* normal) {
  double o[4];
  if (tri_op(0, 0.0, 0.0, tri, o)) for (int k = 0; k < 3; k++) normal[k] = o[k];
}
void position_and_jacobian_3d(double alpha, double beta, const double* tri, double* pos_out, double* jac) {
  double o[4];
  if (tri_op(1, alpha, beta, tri, o)) {
    for (int k = 0; k < 3; k++) pos_out[k] = o[k];
    *jac = o[3];
  }
}
void triangle_areas(const double* triangles, double* out, size_t N) {
  if (!N) return;
  std::vector<double> in(11 * N, 0.0);
  for (size_t i = 0; i < N; i++) std::memcpy(&in[11 * i + 2], triangles + 9 * i, 9 * sizeof(double));
  Dev<double> din(in.data(), in.size()), dout(4 * N);
  if (!din.good || !dout.good) return;
  tb::launch_tri_eval(2, din.p, (int)N, dout.p, 0);
  std::vector<double> o(4 * N);
  if (!finish("triangle_areas") || !dout.get(o.data(), o.size())) return;
  for (size_t i = 0; i < N; i++) out[i] = o[4 * i];
}
void axial_coefficients_3d(const void*, const void*, const void*, void*, void*, size_t, const double*, void*, size_t) {
  off_path("axial_coefficients_3d", "3D BEM is outside the radially symmetric path of this library");
}
void fill_jacobian_buffer_current_three_d(const double*, double*, double*, double*, size_t) {
  off_path("fill_jacobian_buffer_current_three_d", "3D current lines are outside the radially symmetric path of this library");
}
double potential_3d_derivs(const double*, const double*, const double*, size_t) {
  off_path("potential_3d_derivs", "3D BEM is outside the radially symmetric path of this library");
  return __builtin_nan("");
}
void field_3d_derivs(const double*, double*, const double*, const double*, size_t) {
  off_path("field_3d_derivs", "3D BEM is outside the radially symmetric path of this library");
}

double potential_radial(const double* point, const double* q, const double* jac, const double* pos, size_t N) {
  DevSources s(q, jac, pos, N, 16, 2, 1);
  if (s.n < 0) return __builtin_nan("");
  Dev<double> dpt(point, 3), out(1);
  tb::launch_field_points(dpt.p, 1, s.w.p, s.r.p, s.z.p, s.n, out.p, nullptr, 0, 0);
  double v = __builtin_nan("");
  if (finish("potential_radial")) out.get(&v, 1);
  return v;
}

void field_radial(const double* point, double* result, const double* q, const double* jac, const double* pos, size_t N) {
  DevSources s(q, jac, pos, N, 16, 2, 1);
  if (s.n < 0) return;
  Dev<double> dpt(point, 3), out(3);
  tb::launch_field_points(dpt.p, 1, s.w.p, s.r.p, s.z.p, s.n, nullptr, out.p, 0, 0);
  if (finish("field_radial")) out.get(result, 3);
}

double current_potential_axial(double z0, const double* I, const double* jac, const double* pos, size_t N) {
  DevSources s(I, jac, pos, N, 12, 3, 2);
  if (s.n < 0) return __builtin_nan("");
  Dev<double> dz(&z0, 1), out(1);
  tb::launch_current_axial_potential(dz.p, 1, s.w.p, s.r.p, s.z.p, s.n, out.p, 0);
  double v = __builtin_nan("");
  if (finish("current_potential_axial")) out.get(&v, 1);
  return v;
}

void current_field_radial(const double* point, double* result, const double* I, const double* jac, const double* pos, size_t N) {
  DevSources s(I, jac, pos, N, 12, 3, 2);
  if (s.n < 0) return;
  Dev<double> dpt(point, 3), out(3);
  tb::launch_current_points(dpt.p, 1, s.w.p, s.r.p, s.z.p, s.n, out.p, 0, 0);
  if (finish("current_field_radial")) out.get(result, 3);
}

static void axial_derivs_common(double* derivs, const double* q, const double* jac, const double* pos, size_t N, const double* z,
                                size_t Nz, int current) {
  if (!Nz) return;
  DevSources s(q, jac, pos, N, current ? 12 : 16, current ? 3 : 2, current ? 2 : 1);
  if (s.n < 0) return;
  Dev<double> dz(z, Nz), dd(derivs, Nz * 9);
  tb::launch_axial_derivs(dd.p, dz.p, (int)Nz, s.w.p, s.r.p, s.z.p, s.n, current, 0);
  if (finish("axial_derivatives")) dd.get(derivs, Nz * 9);
}
void axial_derivatives_radial(double* derivs, const double* q, const double* jac, const double* pos, size_t N, const double* z,
                              size_t Nz) {
  axial_derivs_common(derivs, q, jac, pos, N, z, Nz, 0);
}
void current_axial_derivatives_radial(double* derivs, const double* I, const double* jac, const double* pos, size_t N,
                                      const double* z, size_t Nz) {
  axial_derivs_common(derivs, I, jac, pos, N, z, Nz, 1);
}

double potential_radial_derivs(const double* point, const double* zi, const double* coeff, size_t Nz) {
  Dev<double> dpt(point, 3), dz(zi, Nz), dc(coeff, (Nz - 1) * 54), out(1);
  tb::launch_axial_points(dpt.p, 1, dz.p, (int)Nz, dc.p, out.p, nullptr, 0);
  double v = __builtin_nan("");
  if (finish("potential_radial_derivs")) out.get(&v, 1);
  return v;
}

void field_radial_derivs(const double* point, double* field, const double* zi, const double* coeff, size_t Nz) {
  Dev<double> dpt(point, 3), dz(zi, Nz), dc(coeff, (Nz - 1) * 54), out(3);
  tb::launch_axial_points(dpt.p, 1, dz.p, (int)Nz, dc.p, nullptr, out.p, 0);
  if (finish("field_radial_derivs")) out.get(field, 3);
}

double charge_radial(const double* vertices, double charge) {
  Dev<double> dl(vertices, 12), dq(&charge, 1), out(1);
  tb::launch_charge(dl.p, dq.p, 1, out.p, 0);
  double v = __builtin_nan("");
  if (finish("charge_radial")) out.get(&v, 1);
  return v;
}

// ---- field programs from the reference's argument structs ---------------------------
namespace {
struct HostProgram {
  std::vector<std::unique_ptr<Dev<double>>> keep;
  tb::FieldComp comp{};
  Dev<tb::FieldComp>* dcomp = nullptr;
  bool good = true;
  ~HostProgram() { delete dcomp; }

  void add_sources(tb::SourceSet& set, const effective_point_charges_2d* e, int nq, int pstride, int zcol) {
    if (!e || !e->N) { set = tb::SourceSet{nullptr, nullptr, nullptr, 0, 0}; return; }
    auto s = new DevSources(e->charges, e->jacobians, e->positions, e->N, nq, pstride, zcol);
    if (s->n < 0) good = false;
    set = tb::SourceSet{s->w.p, s->r.p, s->z.p, s->n, 0};
    owned.emplace_back(s);
  }
  std::vector<std::unique_ptr<DevSources>> owned;

  bool upload() {
    dcomp = new Dev<tb::FieldComp>(&comp, 1);
    return good && dcomp->good;
  }
};

bool program_from_args(field_fun f, void* args, HostProgram& hp, int& kind) {
  std::memset(&hp.comp, 0, sizeof(hp.comp));
  hp.comp.factor = 1.0;
  if (f == &field_radial_traceable) {
    auto* a = (field_evaluation_args*)args;
    hp.comp.kind = tb::COMP_BEM;
    hp.add_sources(hp.comp.elec, (effective_point_charges_2d*)a->elec_charges, 16, 2, 1);
    hp.add_sources(hp.comp.mag, (effective_point_charges_2d*)a->mag_charges, 16, 2, 1);
    hp.add_sources(hp.comp.cur, (effective_point_charges_2d*)a->currents, 12, 3, 2);
    if (a->bounds) {
      hp.comp.bounds_mode = tb::BOUNDS_STRICT;
      for (int i = 0; i < 6; i++) hp.comp.bounds[i] = a->bounds[i];
    }
    kind = 0;
  } else if (f == &field_radial_derivs_traceable) {
    auto* a = (field_derivs_args*)args;
    hp.comp.kind = tb::COMP_AXIAL;
    size_t nz = a->N_z;
    auto z = new Dev<double>(a->z_interpolation, nz);
    auto ce = new Dev<double>(a->electrostatic_axial_coeffs, (nz - 1) * 54);
    auto cm = new Dev<double>(a->magnetostatic_axial_coeffs, (nz - 1) * 54);
    hp.keep.emplace_back(z); hp.keep.emplace_back(ce); hp.keep.emplace_back(cm);
    if (!z->good || !ce->good || !cm->good) hp.good = false;
    hp.comp.zs = z->p; hp.comp.ce = ce->p; hp.comp.cm = cm->p; hp.comp.nz = (int)nz;
    kind = 1;
  } else {
    set_err("trace_particle: field function is not one of this library's traceable fields "
            "(field_radial_traceable, field_radial_derivs_traceable)");
    return false;
  }
  return hp.upload();
}
}  // namespace

void field_radial_traceable(double* position, double* velocity, void* args, double* elec_out, double* mag_out) {
  (void)velocity;
  HostProgram hp;
  int kind;
  if (!program_from_args(&field_radial_traceable, args, hp, kind)) return;
  tb::FieldProgram F{hp.dcomp->p, 1, tb::BOUNDS_NONE, 0, 0, {0, 0, 0, 0, 0, 0}};
  Dev<double> dp(position, 3), E(3), H(3);
  tb::launch_eval_program(F, dp.p, 1, E.p, H.p, 0);
  if (finish("field_radial_traceable")) { E.get(elec_out, 3); H.get(mag_out, 3); }
}

void field_radial_derivs_traceable(double* position, double* velocity, void* args, double* elec_out, double* mag_out) {
  (void)velocity;
  HostProgram hp;
  int kind;
  if (!program_from_args(&field_radial_derivs_traceable, args, hp, kind)) return;
  tb::FieldProgram F{hp.dcomp->p, 1, tb::BOUNDS_NONE, 0, 0, {0, 0, 0, 0, 0, 0}};
  Dev<double> dp(position, 3), E(3), H(3);
  tb::launch_eval_program(F, dp.p, 1, E.p, H.p, 0);
  if (finish("field_radial_derivs_traceable")) { E.get(elec_out, 3); H.get(mag_out, 3); }
}

size_t trace_particle(double* times, double* states, double qm, field_fun field, const double* bounds, double atol, void* args) {
  HostProgram hp;
  int kind;
  if (!program_from_args(field, args, hp, kind)) return 0;
  const long long cap = (long long)TRACING_BLOCK_SIZE;
  Dev<double> s0(states, 6), t0(times, 1), dq(&qm, 1);
  Dev<double> traj((size_t)cap * 6), tt((size_t)cap), fin(6), tf(1), hf(1);
  Dev<int> counter(1), status(1);
  Dev<long long> nacc(1), natt(1), tlen(1);
  if (!s0.good || !traj.good || !tt.good || !fin.good || !counter.good) return 0;
  hipMemset(counter.p, 0, sizeof(int));
  tb::TraceParams P{};
  const long long nsrc = (long long)hp.comp.elec.n + hp.comp.mag.n + hp.comp.cur.n;
  P.field = tb::FieldProgram{hp.dcomp->p, 1, tb::BOUNDS_NONE, nsrc > 0 ? 0 : (int)tb::PROGRAM_NO_SOURCES, 0, {0, 0, 0, 0, 0, 0}};
  P.state0 = s0.p; P.t0 = t0.p; P.h0 = nullptr; P.qm = dq.p; P.qm_stride = 0; P.n = 1; P.counter = counter.p;
  for (int i = 0; i < 6; i++) P.bounds[i] = bounds[i];
  P.atol = atol;
  P.max_attempts = (long long)1 << 62;
  P.plane_n[2] = 1.0; P.plane_nn = 1.0;
  P.final_state = fin.p; P.landing = nullptr; P.status = status.p; P.n_accepted = nacc.p; P.n_attempted = natt.p;
  P.t_final = tf.p; P.h_final = hf.p; P.traj = traj.p; P.times = tt.p; P.traj_cap = cap; P.traj_len = tlen.p;
  if (nsrc > 0) {
    // one particle: spread each stage's source sums over the chip
    const int nchunk = tb::split_chunks(1, nsrc);
    Dev<unsigned char> ws(tb::split_workspace_bytes(1, nchunk));
    if (!ws.good) return 0;
    const hipError_t e = tb::trace_split(P, nchunk, ws.p, 0);
    if (e == hipErrorLaunchTimeOut) { set_err("trace_particle: more than 2^22 attempted steps (non-finite field?)"); return 0; }
    if (!ok(e, "trace_particle")) return 0;
  } else {
    tb::launch_trace(P, 1, 0);
  }
  long long n = 0;
  if (!finish("trace_particle") || !tlen.get(&n, 1)) return 0;
  // row 0 is the caller's initial state; rows 1..n-1 the accepted steps
  if (n > 1) {
    if (!ok(hipMemcpy(states + 6, traj.p + 6, (size_t)(n - 1) * 6 * sizeof(double), hipMemcpyDeviceToHost), "D2H")) return 0;
    if (!ok(hipMemcpy(times + 1, tt.p + 1, (size_t)(n - 1) * sizeof(double), hipMemcpyDeviceToHost), "D2H")) return 0;
  }
  return (size_t)n;
}

bool plane_intersection(const double* p0, const double* normal, const double* positions, size_t N_p, double* result) {
  if (N_p < 2) return false;
  Dev<double> dp0(p0, 3), dn(normal, 3), ds(positions, N_p * 6), out(6);
  int init = -1;
  Dev<int> found(&init, 1);
  tb::launch_plane(dp0.p, dn.p, ds.p, (int)N_p, found.p, out.p, 0);
  int f = -1;
  if (!finish("plane_intersection") || !found.get(&f, 1)) return false;
  if (f < 0) return false;
  return out.get(result, 6);
}

// ====================================================================================
// Tier B
// ====================================================================================
#define TB_CALL(expr, name) return ok((expr), name) ? 0 : -1

size_t tb_rowinfo_bytes(void) { return sizeof(tb::RowInfo); }

int tb_prepare_elements(const double* lines, const uint8_t* types, const double* values, int N, double* jac, double* pos,
                        double* src, void* rows, int* bad_type, void* stream) {
  TB_CALL(tb::launch_elements(lines, types, values, N, jac, pos, src, (tb::RowInfo*)rows, bad_type, S(stream)), "prepare_elements");
}
int tb_fill_matrix(double* M, size_t ld, float* M32, size_t ld32, const void* rows, const double* src, int N, int row0, int row1,
                   int accumulate, void* stream) {
  TB_CALL(tb::launch_fill(M, ld, M32, ld32, (const tb::RowInfo*)rows, src, N, row0, row1, accumulate, S(stream)), "fill_matrix");
}
int tb_fill_diagonal(double* M, size_t ld, float* M32, size_t ld32, const double* lines, const void* rows, int N, int row0, int row1,
                     double* diag_out, void* stream) {
  TB_CALL(tb::launch_self(M, ld, M32, ld32, lines, (const tb::RowInfo*)rows, N, row0, row1, diag_out, S(stream)), "fill_diagonal");
}
size_t tb_lu_workspace_bytes(int N, int fp32) { return fp32 ? tb::lu_workspace_bytes<float>(N) : tb::lu_workspace_bytes<double>(N); }
int tb_lu_factor(double* A, size_t lda, int N, int* ipiv, void* ws, int* info, void* stream) {
  TB_CALL(tb::lu_factor<double>(A, lda, N, ipiv, ws, info, S(stream), 0), "lu_factor");
}
int tb_lu_factor_f32(float* A, size_t lda, int N, int* ipiv, void* ws, int* info, void* stream) {
  TB_CALL(tb::lu_factor<float>(A, lda, N, ipiv, ws, info, S(stream), 0), "lu_factor_f32");
}
static_assert(sizeof(tb_kernel_timing_t) == sizeof(tb::KernelTiming), "timing layout");
int tb_lu_factor_timed(void* A, size_t lda, int N, int* ipiv, void* ws, int* info, int fp32, tb_kernel_timing_t* gemm_timing,
                       void* stream) {
  auto* tm = reinterpret_cast<tb::KernelTiming*>(gemm_timing);
  if (fp32) TB_CALL(tb::lu_factor<float>((float*)A, lda, N, ipiv, ws, info, S(stream), 0, tm), "lu_factor_f32");
  TB_CALL(tb::lu_factor<double>((double*)A, lda, N, ipiv, ws, info, S(stream), 0, tm), "lu_factor");
}
int tb_lu_solve(const double* LU, size_t lda, int N, const int* ipiv, double* X, int nrhs, const void* ws, void* stream) {
  TB_CALL(tb::lu_solve<double>(LU, lda, N, ipiv, X, nrhs, ws, S(stream)), "lu_solve");
}
int tb_lu_solve_f32(const float* LU, size_t lda, int N, const int* ipiv, double* X, int nrhs, const void* ws, void* stream) {
  TB_CALL(tb::lu_solve<float>(LU, lda, N, ipiv, X, nrhs, ws, S(stream)), "lu_solve_f32");
}
int tb_residual(const double* A, size_t lda, int N, const double* B, const double* X, double* R, int nrhs, unsigned long long* rmax,
                void* stream) {
  TB_CALL(tb::residual(A, lda, N, B, X, R, nrhs, rmax, S(stream)), "residual");
}
int tb_axpy(double* X, const double* D, size_t n, void* stream) { TB_CALL(tb::axpy(X, D, n, S(stream)), "axpy"); }
int tb_lu_solve_status(const void* ws, int N, int fp32, int* status, void* stream) {
  if (fp32) TB_CALL(tb::lu_solve_status<float>(ws, N, status, S(stream)), "lu_solve_status");
  TB_CALL(tb::lu_solve_status<double>(ws, N, status, S(stream)), "lu_solve_status");
}
int tb_absmax(const double* x, size_t n, unsigned long long* out, void* stream) { TB_CALL(tb::absmax(x, n, out, S(stream)), "absmax"); }
int tb_convert_f64_f32(const double* Sx, size_t lds, float* D, size_t ldd, int rows, int cols, void* stream) {
  TB_CALL(tb::convert_d2f(Sx, lds, D, ldd, rows, cols, S(stream)), "convert");
}
int tb_gemm_f64(int M, int N, int K, double alpha, const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C,
                size_t ldc, void* stream) {
  TB_CALL(tb::gemm<double>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, S(stream)), "gemm_f64");
}
int tb_gemm_f32(int M, int N, int K, float alpha, const float* A, size_t lda, const float* B, size_t ldb, float beta, float* C,
                size_t ldc, void* stream) {
  TB_CALL(tb::gemm<float>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, S(stream)), "gemm_f32");
}
int tb_sources(const double* q, const double* jac, const double* pos, int N, int nq, int pstride, int zcol, double* w, double* r,
               double* z, void* stream) {
  TB_CALL(tb::launch_sources(q, jac, pos, N, nq, pstride, zcol, w, r, z, S(stream)), "sources");
}
int tb_field_points(const double* pts, int M, const double* w, const double* r, const double* z, int Ns, double* pot, double* field,
                    int accumulate, void* stream) {
  TB_CALL(tb::launch_field_points(pts, M, w, r, z, Ns, pot, field, accumulate, S(stream)), "field_points");
}
int tb_current_points(const double* pts, int M, const double* w, const double* r, const double* z, int Ns, double* field,
                      int accumulate, void* stream) {
  TB_CALL(tb::launch_current_points(pts, M, w, r, z, Ns, field, accumulate, S(stream)), "current_points");
}
int tb_current_axial_potential(const double* zq, int M, const double* w, const double* r, const double* z, int Ns, double* out,
                               void* stream) {
  TB_CALL(tb::launch_current_axial_potential(zq, M, w, r, z, Ns, out, S(stream)), "current_axial_potential");
}
int tb_axial_derivatives(double* out, const double* zq, int Nz, const double* w, const double* r, const double* z, int Ns,
                         int current, void* stream) {
  TB_CALL(tb::launch_axial_derivs(out, zq, Nz, w, r, z, Ns, current, S(stream)), "axial_derivatives");
}
int tb_axial_coefficients(const double* derivs, const double* zq, int Nz, double* C, double* scratch, void* stream) {
  if (Nz < 4) { set_err("axial coefficients need at least 4 samples"); return -1; }
  TB_CALL(tb::launch_axial_coeffs(derivs, zq, Nz, C, scratch, S(stream)), "axial_coefficients");
}
int tb_axial_points(const double* pts, int M, const double* zq, int Nz, const double* C, double* pot, double* field, void* stream) {
  TB_CALL(tb::launch_axial_points(pts, M, zq, Nz, C, pot, field, S(stream)), "axial_points");
}
int tb_jacobian_3d(const double* tri, int N, double* jac, double* pos, void* stream) {
  TB_CALL(tb::launch_jac3d(tri, N, jac, pos, S(stream)), "jacobian_3d");
}
int tb_charge_radial(const double* lines, const double* q, int N, double* out, void* stream) {
  TB_CALL(tb::launch_charge(lines, q, N, out, S(stream)), "charge_radial");
}
// workgroups of a persistent tracer launch: `waves`, or by default as many as the CUs hold
// (4 per CU per wave of occupancy for the axial-only kernel, 8 per CU otherwise); -1 on error
static int trace_waves(const tb_trace_params_t* params, int waves) {
  if (waves > 0) return waves;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    set_err("device attributes");
    return -1;
  }
  const int occ = params->occupancy >= 1 && params->occupancy <= 4 ? params->occupancy : 2;
  return cus * ((params->field.flags & tb::PROGRAM_NO_SOURCES) ? 4 * occ : 8);
}
int tb_trace(const tb_trace_params_t* params, int waves, void* stream) {
  if (!params) { set_err("tb_trace: null params"); return -1; }
  if (params->traj && params->traj_cap < 2) { set_err("tb_trace: traj_cap must be >= 2"); return -1; }
  const int nw = trace_waves(params, waves);
  if (nw < 0) return -1;
  TB_CALL(tb::launch_trace(*reinterpret_cast<const tb::TraceParams*>(params), nw, S(stream)), "trace");
}
int tb_trace_noalloc(const tb_trace_params_t* params, int waves, int ordered, void* stream) {
  if (!params) { set_err("tb_trace_noalloc: null params"); return -1; }
  if (params->traj && params->traj_cap < 2) { set_err("tb_trace_noalloc: traj_cap must be >= 2"); return -1; }
  if (params->n > tb::TRACE_STATIC_MAX_PARTICLES) {
    set_err("tb_trace_noalloc: " + std::to_string(params->n) + " particles exceed the static maximum of " +
            std::to_string(tb::TRACE_STATIC_MAX_PARTICLES));
    return -1;
  }
  const int nw = trace_waves(params, waves);
  if (nw < 0) return -1;
  TB_CALL(tb::trace_noalloc(*reinterpret_cast<const tb::TraceParams*>(params), nw, ordered, S(stream)), "trace_noalloc");
}
int tb_sample_beam(const tb_beam_t* beam, const unsigned long long* seed, int n, double* state0, void* stream) {
  if (!beam || !seed || (n > 0 && !state0) || n < 0) { set_err("tb_sample_beam: null beam, seed or state0, or n < 0"); return -1; }
  if (!(beam->mass > 0.0) || !(beam->charge > 0.0) || !(beam->e_median > 0.0) || !(beam->r_max >= 0.0) ||
      !(beam->e_sigma >= 0.0) || !(beam->angle_sigma >= 0.0)) {
    set_err("tb_sample_beam: mass, |charge| and e_median must be > 0; r_max, e_sigma and angle_sigma >= 0");
    return -1;
  }
  static_assert(sizeof(tb_beam_t) == sizeof(tb::BeamSpec), "tb_beam_t layout");
  TB_CALL(tb::sample_beam(*reinterpret_cast<const tb::BeamSpec*>(beam), seed, n, state0, S(stream)), "sample_beam");
}
size_t tb_focus_workspace_bytes(void) { return tb::focus_workspace_bytes(); }
int tb_focus_position(const double* states, int n, long long stride, double* out, void* workspace, void* stream) {
  if (n < 2 || stride < 6 || !states || !out || !workspace) { set_err("tb_focus_position: needs n >= 2 rays, stride >= 6"); return -1; }
  TB_CALL(tb::focus_position(states, n, stride, out, workspace, S(stream)), "focus_position");
}
int tb_trace_noalloc_max_particles(void) { return tb::TRACE_STATIC_MAX_PARTICLES; }
size_t tb_trace_noalloc_scratch_bytes(void) { return tb::trace_static_scratch_bytes(); }
size_t tb_trace_order_workspace_bytes(int n) { return tb::trace_order_workspace_bytes(n); }
int tb_trace_order(const double* state0, int n, int* order, void* workspace, void* stream) {
  if (n > 0 && (!state0 || !order || !workspace)) { set_err("tb_trace_order: null pointer"); return -1; }
  TB_CALL(tb::trace_order(state0, n, order, workspace, S(stream)), "trace_order");
}
int tb_trace_split_chunks(int n_particles, long long n_sources) { return tb::split_chunks(n_particles, n_sources); }
size_t tb_trace_split_workspace_bytes(int n_particles, int nchunk) { return tb::split_workspace_bytes(n_particles, nchunk); }
int tb_trace_split(const tb_trace_params_t* params, int nchunk, void* workspace, void* stream) {
  if (!params || !workspace) { set_err("tb_trace_split: null params or workspace"); return -1; }
  if (params->traj && params->traj_cap < 2) { set_err("tb_trace_split: traj_cap must be >= 2"); return -1; }
  const hipError_t e = tb::trace_split(*reinterpret_cast<const tb::TraceParams*>(params), nchunk, workspace, S(stream));
  if (e == hipErrorLaunchTimeOut) { set_err("trace_split: more than 2^22 attempted steps (non-finite field?)"); return -1; }
  return ok(e, "trace_split") ? 0 : -1;
}
int tb_mlp_nparams(void) { return tb::MLP_NPARAM; }
size_t tb_mlp_workspace_bytes(int B) { return tb::mlp_workspace_bytes(B); }
int tb_gemm_bf16(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C, long ldc,
                 int accumulate, int ksplit, void* stream) {
  TB_CALL(tb::gemm_bf16(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, accumulate, ksplit, S(stream)), "gemm_bf16");
}
int tb_mlp_forward(const float* params, const uint16_t* wbf, const float* X, int B, float* Y, void* workspace, void* stream) {
  TB_CALL(tb::mlp_forward(params, wbf, X, B, Y, workspace, S(stream)), "mlp_forward");
}
int tb_mlp_grad(const float* params, const uint16_t* wbf, const float* X, const float* T, int B, float* grad, float* loss,
                void* workspace, void* stream) {
  TB_CALL(tb::mlp_grad(params, wbf, X, T, B, grad, loss, workspace, S(stream)), "mlp_grad");
}
int tb_mlp_adam(float* params, const float* grad, float* m, float* v, uint16_t* wbf, float lr, float beta1, float beta2, float eps,
                int step, void* stream) {
  if (step < 1) { set_err("tb_mlp_adam: step must be >= 1"); return -1; }
  TB_CALL(tb::mlp_adam(params, grad, m, v, wbf, lr, beta1, beta2, eps, step, S(stream)), "mlp_adam");
}
int tb_mlp_to_bf16(const float* params, uint16_t* wbf, void* stream) { TB_CALL(tb::mlp_to_bf16(params, wbf, S(stream)), "mlp_to_bf16"); }
int tb_mlp_gather(int B, const int* idx, const float* X, const float* T, float* Xb, float* Tb, void* stream) {
  TB_CALL(tb::mlp_gather(B, idx, X, T, Xb, Tb, S(stream)), "mlp_gather");
}
int tb_eval_program(const tb_field_program_t* program, const double* pts, int M, double* E, double* H, void* stream) {
  TB_CALL(tb::launch_eval_program(*reinterpret_cast<const tb::FieldProgram*>(program), pts, M, E, H, S(stream)), "eval_program");
}

}  // extern "C"
