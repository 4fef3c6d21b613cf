/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, single-threaded restatement of Traceon's radially symmetric BEM path
 * (reference v0.10.0, traceon/backend/ C files).  It is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * It is never part of the product path.
 *
 * Exported symbols use the reference's own names and argument orders
 * (traceon/backend/__init__.py:287-328) so the same ctypes harness can drive
 * this oracle, the reference build in oracle/_ref and the GPU library alike.
 *
 * Every function names the reference lines it restates.  Formulas are written
 * from the mathematics (Lagrange basis, ring-potential closed forms, Fehlberg
 * tableau); constants that must be identical (Cody 1965 coefficients, 16-point
 * Gauss-Legendre and 12-point triangle tables, MU_0) carry the same decimal
 * literals as the reference so results agree to rounding.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <stdbool.h>
#include <stdio.h>

#define ORC_PI 3.14159265358979323846

/* definitions.c:31 — scipy.constants.mu_0 at the reference's precision */
static const double MU0 = 1.25663706212e-06;
/* three_d.c:2 */
static const double AXIS_EPS = 1e-10;

/* Exported constants read with ctypes in_dll (backend/__init__.py:58-66). */
const int N_QUAD_2D_SYM = 16;          /* utilities_2d.c:2 */
const int N_TRIANGLE_QUAD_SYM = 12;    /* utilities_3d.c:4 */
const int DERIV_2D_MAX_SYM = 9;        /* radial_ring.c:2 */
const int NU_MAX_SYM = 4;              /* three_d.c:3 */
const int M_MAX_SYM = 8;               /* three_d.c:4 */
const size_t TRACING_BLOCK_SIZE = 100000; /* tracing.c:7 */

enum { T_VOLTAGE_FIXED = 1, T_VOLTAGE_FUN = 2, T_DIELECTRIC = 3, T_CURRENT = 4,
       T_MAGNETOSTATIC_POT = 5, T_MAGNETIZABLE = 6 };   /* definitions.c:34-40 */

/* ------------------------------------------------------------------------- */
/* Complete elliptic integrals, Cody (1965) Chebyshev fits: elliptic.c:9-75.   */
/* K(1-p) = sum_i (a_i + ln(1/p) b_i) p^i, same for E with its own table.      */
/* ------------------------------------------------------------------------- */
static const double CODY_KA[8] = {1.3862943611198906 /* ln 4 */, 9.65736020516771e-2, 3.08909633861795e-2,
    1.52618320622534e-2, 1.25565693543211e-2, 1.68695685967517e-2, 1.09423810688623e-2, 1.40704915496101e-3};
static const double CODY_KB[8] = {0.5, 1.24999998585309e-1, 7.03114105853296e-2, 4.87379510945218e-2,
    3.57218443007327e-2, 2.09857677336790e-2, 5.81807961871996e-3, 3.42805719229748e-4};
static const double CODY_EA[8] = {1.0, 4.43147193467733e-1, 5.68115681053803e-2, 2.21862206993846e-2,
    1.56847700239786e-2, 1.92284389022977e-2, 1.21819481486695e-2, 1.55618744745296e-3};
static const double CODY_EB[8] = {0.0, 2.49999998448655e-1, 9.37488062098189e-2, 5.84950297066166e-2,
    4.09074821593164e-2, 2.35091602564984e-2, 6.45682247315060e-3, 3.78886487349367e-4};

static double horner8(const double *c, double x) {
    double s = c[7];
    for (int i = 6; i >= 0; i--) s = s * x + c[i];
    return s;
}

/* elliptic.c:9-35 */
double ellipkm1(double p) {
    double lg = -log(p);
    return horner8(CODY_KA, p) + lg * horner8(CODY_KB, p);
}
/* elliptic.c:43-69 */
double ellipem1(double p) {
    double lg = -log(p);
    return horner8(CODY_EA, p) + lg * horner8(CODY_EB, p);
}
/* elliptic.c:37-41: parameter form with the k<=-1 transform */
double ellipk(double k) {
    if (k > -1.0) return ellipkm1(1.0 - k);
    return ellipkm1(1.0 / (1.0 - k)) / sqrt(k);
}
/* elliptic.c:71-75 */
double ellipe(double k) {
    if (k >= 0.0 && k <= 1.0) return ellipem1(1.0 - k);
    return ellipem1(-1.0 / (k - 1.0)) * sqrt(1.0 - k);
}

/* ------------------------------------------------------------------------- */
/* Quadrature tables (utilities_2d.c:3-5, utilities_3d.c:6-47).                */
/* ------------------------------------------------------------------------- */
static const double GL16_X[16] = {-0.0950125098376374, 0.0950125098376374, -0.2816035507792589, 0.2816035507792589,
    -0.4580167776572274, 0.4580167776572274, -0.6178762444026438, 0.6178762444026438,
    -0.7554044083550030, 0.7554044083550030, -0.8656312023878318, 0.8656312023878318,
    -0.9445750230732326, 0.9445750230732326, -0.9894009349916499, 0.9894009349916499};
static const double GL16_W[16] = {0.1894506104550685, 0.1894506104550685, 0.1826034150449236, 0.1826034150449236,
    0.1691565193950025, 0.1691565193950025, 0.1495959888165767, 0.1495959888165767,
    0.1246289712555339, 0.1246289712555339, 0.0951585116824928, 0.0951585116824928,
    0.0622535239386479, 0.0622535239386479, 0.0271524594117541, 0.0271524594117541};

static const double TRI_W[12] = {0.0254224531851035, 0.0254224531851035, 0.0254224531851035,
    0.0583931378631895, 0.0583931378631895, 0.0583931378631895,
    0.041425537809187, 0.041425537809187, 0.041425537809187, 0.041425537809187, 0.041425537809187, 0.041425537809187};
static const double TRI_B1[12] = {0.873821971016996, 0.063089014491502, 0.063089014491502,
    0.501426509658179, 0.249286745170910, 0.249286745170910,
    0.636502499121399, 0.636502499121399, 0.310352451033785, 0.310352451033785, 0.053145049844816, 0.053145049844816};
static const double TRI_B2[12] = {0.063089014491502, 0.873821971016996, 0.063089014491502,
    0.249286745170910, 0.501426509658179, 0.249286745170910,
    0.310352451033785, 0.053145049844816, 0.636502499121399, 0.053145049844816, 0.636502499121399, 0.310352451033785};

/* ------------------------------------------------------------------------- */
/* Cubic (line4) element map, utilities_2d.c:18-87.                           */
/* Nodes at alpha = -1, -1/3, 1/3, 1 are the element vertices [0],[2],[3],[1]  */
/* (gmsh line4 order, radial.c:265-268).  Coordinates used: x (index 0) as r,  */
/* z (index 2).  Written as a Lagrange basis.                                  */
/* ------------------------------------------------------------------------- */
static void lagrange4(double a, double L[4], double dL[4]) {
    double a2 = a * a;
    /* L1 = -(a-1)(3a-1)(3a+1)/16, L2 = 9(a^2-1)(3a-1)/16, L3 = -9(a^2-1)(3a+1)/16, L4 = (a+1)(3a-1)(3a+1)/16 */
    L[0] = -(a - 1.0) * (9.0 * a2 - 1.0) / 16.0;
    L[1] = 9.0 * (a2 - 1.0) * (3.0 * a - 1.0) / 16.0;
    L[2] = -9.0 * (a2 - 1.0) * (3.0 * a + 1.0) / 16.0;
    L[3] = (a + 1.0) * (9.0 * a2 - 1.0) / 16.0;
    dL[0] = (-27.0 * a2 + 18.0 * a + 1.0) / 16.0;
    dL[1] = 9.0 * (9.0 * a2 - 2.0 * a - 3.0) / 16.0;
    dL[2] = -9.0 * (9.0 * a2 + 2.0 * a - 3.0) / 16.0;
    dL[3] = (27.0 * a2 + 18.0 * a - 1.0) / 16.0;
}

/* utilities_2d.c:50-68; v1..v4 are the nodes in parameter order, coordinates (x=[0], z=[2]) */
void position_and_jacobian_radial(double a, const double *v1, const double *v2, const double *v3, const double *v4,
                                  double pos[2], double *jac) {
    double L[4], dL[4];
    lagrange4(a, L, dL);
    const double *v[4] = {v1, v2, v3, v4};
    double r = 0, z = 0, dr = 0, dz = 0;
    for (int n = 0; n < 4; n++) {
        r += L[n] * v[n][0]; z += L[n] * v[n][2];
        dr += dL[n] * v[n][0]; dz += dL[n] * v[n][2];
    }
    pos[0] = r; pos[1] = z;
    *jac = sqrt(dr * dr + dz * dz);
}

/* utilities_2d.c:70-87: offset from the alpha=0 point.  The basis is written
   without its constant term, L_n(a) - L_n(0), so small offsets keep full relative precision. */
void delta_position_and_jacobian_radial(double a, const double *v1, const double *v2, const double *v3, const double *v4,
                                        double pos[2], double *jac) {
    double L[4], dL[4];
    lagrange4(a, L, dL);
    double a2 = a * a;
    double D[4] = {a * (-9.0 * a2 + 9.0 * a + 1.0) / 16.0, 9.0 * a * (3.0 * a2 - a - 3.0) / 16.0,
                   -9.0 * a * (3.0 * a2 + a - 3.0) / 16.0, a * (9.0 * a2 + 9.0 * a - 1.0) / 16.0};
    const double *v[4] = {v1, v2, v3, v4};
    double r = 0, z = 0, dr = 0, dz = 0;
    for (int n = 0; n < 4; n++) {
        r += D[n] * v[n][0]; z += D[n] * v[n][2];
        dr += dL[n] * v[n][0]; dz += dL[n] * v[n][2];
    }
    pos[0] = r; pos[1] = z;
    *jac = sqrt(dr * dr + dz * dz);
}

/* utilities_2d.c:18-30: normal = tangent rotated clockwise, normalised */
void normal_2d(const double *p1, const double *p2, double n[2]) {
    double tx = p2[0] - p1[0], ty = p2[1] - p1[1];
    double len = sqrt(tx * tx + ty * ty);
    n[0] = ty / len;
    n[1] = -tx / len;
}

/* utilities_2d.c:32-48 */
void higher_order_normal_radial(double a, const double *v1, const double *v2, const double *v3, const double *v4, double n[2]) {
    double L[4], dL[4];
    lagrange4(a, L, dL);
    const double *v[4] = {v1, v2, v3, v4};
    double dr = 0, dz = 0;
    for (int k = 0; k < 4; k++) { dr += dL[k] * v[k][0]; dz += dL[k] * v[k][2]; }
    double len = sqrt(dr * dr + dz * dz);
    n[0] = dz / len;
    n[1] = -dr / len;
}

/* element vertex array (4,3) -> nodes in parameter order (radial.c:265-268) */
static void nodes_of(const double *line, const double **v) {
    v[0] = line + 0; v[1] = line + 6; v[2] = line + 9; v[3] = line + 3;
}

/* ------------------------------------------------------------------------- */
/* Ring kernels, radial_ring.c.  Target ring (r0,z0); source ring at           */
/* (r0+dr, z0+dz).  S = (r+r0)^2+dz^2 with r the source radius,                */
/* p = (dr^2+dz^2)/S = 1-m.                                                    */
/* ------------------------------------------------------------------------- */

/* radial_ring.c:4-18 */
double flux_density_to_charge_factor(double K) { return 2.0 * (K - 1.0) / (1.0 + K); }

/* radial_ring.c:37-43 */
double potential_radial_ring(double r0, double z0, double dr, double dz, void *unused) {
    (void)z0; (void)unused;
    double r = r0 + dr;
    double S = dz * dz + (r + r0) * (r + r0);
    double p = (dz * dz + dr * dr) / S;
    return ellipkm1(p) * r / (ORC_PI * sqrt(S));
}

/* radial_ring.c:21-35: d/dr0 of the ring potential; 0 on the axis */
double dr1_potential_radial_ring(double r0, double z0, double dr, double dz, void *unused) {
    (void)z0; (void)unused;
    if (r0 < AXIS_EPS) return 0.0;
    double r = r0 + dr;
    double d2 = dz * dz + dr * dr;
    double S = dz * dz + (r + r0) * (r + r0);
    double p = d2 / S;
    /* -(r / (2 pi r0 d2 sqrt S)) * [ d2 K + (dr (dr - 2r) - dz^2) E ] */
    double bracket = d2 * ellipkm1(p) + (dr * (dr - 2.0 * r) - dz * dz) * ellipem1(p);
    return -r * bracket / (2.0 * ORC_PI * r0 * d2 * sqrt(S));
}

/* radial_ring.c:45-53 */
double dz1_potential_radial_ring(double r0, double z0, double dr, double dz, void *unused) {
    (void)z0; (void)unused;
    double r = r0 + dr;
    double d2 = dz * dz + dr * dr;
    double S = dz * dz + (r + r0) * (r + r0);
    return dz * r * ellipem1(d2 / S) / (ORC_PI * d2 * sqrt(S));
}

/* radial_ring.c:55-71 */
static double field_dot_normal_radial(double r0, double z0, double dr, double dz, const double n[2], double K) {
    double Er = -dr1_potential_radial_ring(r0, z0, dr, dz, NULL);
    double Ez = -dz1_potential_radial_ring(r0, z0, dr, dz, NULL);
    return flux_density_to_charge_factor(K) * (n[0] * Er + n[1] * Ez);
}

/* radial_ring.c:73-76: on-axis magnetic scalar potential of a unit current ring */
double current_potential_axial_radial_ring(double z0, double r, double z) {
    double dz = z0 - z;
    return -dz / (2.0 * sqrt(dz * dz + r * r));
}

/* radial_ring.c:78-107: H of a unit current ring of radius x at height y, evaluated at (x0, y0) */
void current_field_radial_ring(double x0, double y0, double x, double y, double out[2]) {
    double a = x, r = x0, z = y0 - y;
    if (a < AXIS_EPS) { out[0] = 0.0; out[1] = 0.0; return; }
    double A = z * z + (a + r) * (a + r);
    double B = z * z + (r - a) * (r - a);
    double k = 4.0 * r * a / A;   /* parameter m */
    double Kk = ellipk(k), Ek = ellipe(k);
    double sA = sqrt(A);
    out[0] = (r < AXIS_EPS) ? 0.0 : z / (2.0 * ORC_PI * r * sA) * ((z * z + r * r + a * a) / B * Ek - Kk);
    out[1] = 1.0 / (2.0 * ORC_PI * sA) * ((a * a - z * z - r * r) / B * Ek + Kk);
}

/* radial_ring.c:109-122: on-axis derivatives of the ring potential (times r/2),
   d_0 = 1/R, d_1 = -(z0-z)/R^3, d_{n+1} = -((2n+1)(z0-z) d_n + n^2 d_{n-1})/R^2 */
void axial_derivatives_radial_ring(double z0, double r, double z, double d[9]) {
    double u = z0 - z;
    double R2 = u * u + r * r;
    double R = sqrt(R2);
    d[0] = 1.0 / R;
    d[1] = -u / (R2 * R);
    for (int n = 1; n < 8; n++) d[n + 1] = -((2 * n + 1) * u * d[n] + (double)(n * n) * d[n - 1]) / R2;
    for (int n = 0; n < 9; n++) d[n] *= 0.5 * r;
}

/* radial_ring.c:124-136: same for a current ring, via mu = (z0-z)/R */
void current_axial_derivatives_radial_ring(double z0, double r, double z, double d[9]) {
    double u = z0 - z;
    double R2 = u * u + r * r;
    double R = sqrt(R2);
    double mu = u / R;
    d[0] = -u / (2.0 * R);
    d[1] = -r * r / (2.0 * R2 * R);
    for (int n = 2; n < 9; n++) d[n] = -(2 * n - 1) * mu / R * d[n - 1] - (double)(n * n - 2 * n) / R2 * d[n - 2];
}

/* ------------------------------------------------------------------------- */
/* Buffers, radial.c:224-246 and three_d.c:34-62                               */
/* ------------------------------------------------------------------------- */
void fill_jacobian_buffer_radial(double *jac, double *pos, const double *lines, size_t N) {
    for (size_t i = 0; i < N; i++) {
        const double *v[4];
        nodes_of(lines + 12 * i, v);
        for (int k = 0; k < 16; k++) {
            double p[2], j;
            position_and_jacobian_radial(GL16_X[k], v[0], v[1], v[2], v[3], p, &j);
            jac[16 * i + k] = GL16_W[k] * j;
            pos[32 * i + 2 * k] = p[0];
            pos[32 * i + 2 * k + 1] = p[1];
        }
    }
}

void fill_jacobian_buffer_3d(double *jac, double *pos, const double *tri, size_t N) {
    for (size_t i = 0; i < N; i++) {
        const double *t = tri + 9 * i;
        double e1[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
        double e2[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
        double c0 = e1[1] * e2[2] - e1[2] * e2[1];
        double c1 = e1[2] * e2[0] - e1[0] * e2[2];
        double c2 = e1[0] * e2[1] - e1[1] * e2[0];
        double area = 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        for (int k = 0; k < 12; k++) {
            jac[12 * i + k] = 2.0 * TRI_W[k] * area;
            for (int c = 0; c < 3; c++) pos[36 * i + 3 * k + c] = t[c] + TRI_B1[k] * e1[c] + TRI_B2[k] * e2[c];
        }
    }
}

/* radial.c:124-148: total charge of one element with uniform density */
double charge_radial(const double *line, double q) {
    const double *v[4];
    nodes_of(line, v);
    double s = 0.0;
    for (int k = 0; k < 16; k++) {
        double p[2], j;
        position_and_jacobian_radial(GL16_X[k], v[0], v[1], v[2], v[3], p, &j);
        s += 2.0 * ORC_PI * p[0] * GL16_W[k] * j * q;
    }
    return s;
}

/* ------------------------------------------------------------------------- */
/* Matrix fill, radial.c:249-310.  Rows [row_start,row_end] inclusive,         */
/* accumulating into M (row stride N_matrix).  Unknown type -> returns -1      */
/* (the reference prints and exits; an oracle must not kill the test run).     */
/* ------------------------------------------------------------------------- */
int fill_matrix_radial(double *M, const double *lines, const uint8_t *types, const double *values,
                       const double *jac, const double *pos, size_t N_lines, size_t N_matrix,
                       int row_start, int row_end) {
    for (int i = row_start; i <= row_end; i++) {
        const double *v[4];
        nodes_of(lines + 12 * (size_t)i, v);
        double tgt[2], tj;
        position_and_jacobian_radial(0.0, v[0], v[1], v[2], v[3], tgt, &tj);
        int t = types[i];
        double *row = M + (size_t)i * N_matrix;
        if (t == T_VOLTAGE_FIXED || t == T_VOLTAGE_FUN || t == T_MAGNETOSTATIC_POT) {
            for (size_t j = 0; j < N_lines; j++) {
                double acc = 0.0;
                for (int k = 0; k < 16; k++) {
                    const double *p = pos + 32 * j + 2 * k;
                    acc += jac[16 * j + k] * potential_radial_ring(tgt[0], tgt[1], p[0] - tgt[0], p[1] - tgt[1], NULL);
                }
                row[j] += acc;
            }
        } else if (t == T_DIELECTRIC || t == T_MAGNETIZABLE) {
            double n[2];
            higher_order_normal_radial(0.0, v[0], v[1], v[2], v[3], n);
            for (size_t j = 0; j < N_lines; j++) {
                double acc = 0.0;
                for (int k = 0; k < 16; k++) {
                    const double *p = pos + 32 * j + 2 * k;
                    acc += jac[16 * j + k] * field_dot_normal_radial(tgt[0], tgt[1], p[0] - tgt[0], p[1] - tgt[1], n, values[i]);
                }
                row[j] += acc;
            }
        } else {
            return -1;
        }
    }
    return 0;
}

/* Self-integrands, radial.c:182-222.  Called by scipy.integrate.quad through
   LowLevelCallable with signature double f(double, void*) exactly as the
   reference (backend/__init__.py:688-708). */
double self_potential_radial(double a, void *user) {
    const double *line = (const double *)user;
    const double *v[4];
    nodes_of(line, v);
    double d[2], j, tgt[2], j0;
    delta_position_and_jacobian_radial(a, v[0], v[1], v[2], v[3], d, &j);
    position_and_jacobian_radial(0.0, v[0], v[1], v[2], v[3], tgt, &j0);
    return j * potential_radial_ring(tgt[0], tgt[1], d[0], d[1], NULL);
}

struct self_fdn_args { const double *line; double K; };   /* radial.c:198-201 */

double self_field_dot_normal_radial(double a, void *user) {
    const struct self_fdn_args *args = (const struct self_fdn_args *)user;
    const double *v[4];
    nodes_of(args->line, v);
    double d[2], j, tgt[2], j0, n[2];
    delta_position_and_jacobian_radial(a, v[0], v[1], v[2], v[3], d, &j);
    position_and_jacobian_radial(0.0, v[0], v[1], v[2], v[3], tgt, &j0);
    higher_order_normal_radial(0.0, v[0], v[1], v[2], v[3], n);
    return j * field_dot_normal_radial(tgt[0], tgt[1], d[0], d[1], n, args->K);
}

/* ------------------------------------------------------------------------- */
/* Field sums, radial.c:19-173                                                 */
/* ------------------------------------------------------------------------- */
double potential_radial(const double *pt, const double *q, const double *jac, const double *pos, size_t N) {
    double r0 = sqrt(pt[0] * pt[0] + pt[1] * pt[1]), z0 = pt[2];
    double s = 0.0;
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 16; k++) {
            const double *p = pos + 32 * i + 2 * k;
            s += q[i] * jac[16 * i + k] * potential_radial_ring(r0, z0, p[0] - r0, p[1] - z0, NULL);
        }
    return s;
}

void field_radial(const double *pt, double *out, const double *q, const double *jac, const double *pos, size_t N) {
    double r = sqrt(pt[0] * pt[0] + pt[1] * pt[1]), z0 = pt[2];
    double Er = 0.0, Ez = 0.0;
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 16; k++) {
            const double *p = pos + 32 * i + 2 * k;
            double w = q[i] * jac[16 * i + k];
            Er -= w * dr1_potential_radial_ring(r, z0, p[0] - r, p[1] - z0, NULL);
            Ez -= w * dz1_potential_radial_ring(r, z0, p[0] - r, p[1] - z0, NULL);
        }
    if (r >= AXIS_EPS) { out[0] = pt[0] / r * Er; out[1] = pt[1] / r * Er; }
    else { out[0] = 0.0; out[1] = 0.0; }
    out[2] = Ez;
}

/* radial.c:39-53 (current rings are stored as (r, 0, z) with 12 quad points) */
double current_potential_axial(double z0, const double *I, const double *jac, const double *pos, size_t N) {
    double s = 0.0;
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 12; k++) {
            const double *p = pos + 36 * i + 3 * k;
            s += I[i] * jac[12 * i + k] * current_potential_axial_radial_ring(z0, p[0], p[2]);
        }
    return s;
}

/* radial.c:56-83 */
void current_field_radial(const double *pt, double *out, const double *I, const double *jac, const double *pos, size_t N) {
    double r = sqrt(pt[0] * pt[0] + pt[1] * pt[1]);
    double Br = 0.0, Bz = 0.0;
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 12; k++) {
            const double *p = pos + 36 * i + 3 * k;
            double f[2];
            current_field_radial_ring(r, pt[2], p[0], p[2], f);
            double w = I[i] * jac[12 * i + k];
            Br += w * f[0];
            Bz += w * f[1];
        }
    if (r >= AXIS_EPS) { out[0] = pt[0] / r * Br; out[1] = pt[1] / r * Br; }
    else { out[0] = 0.0; out[1] = 0.0; }
    out[2] = Bz;
}

/* radial.c:19-36, accumulating into d (Nz,9) */
void axial_derivatives_radial(double *d, const double *q, const double *jac, const double *pos, size_t N,
                              const double *z, size_t Nz) {
    for (size_t i = 0; i < Nz; i++)
        for (size_t j = 0; j < N; j++)
            for (int k = 0; k < 16; k++) {
                double D[9];
                axial_derivatives_radial_ring(z[i], pos[32 * j + 2 * k], pos[32 * j + 2 * k + 1], D);
                double w = jac[16 * j + k] * q[j];
                for (int l = 0; l < 9; l++) d[9 * i + l] += w * D[l];
            }
}

/* radial.c:86-103, accumulating */
void current_axial_derivatives_radial(double *d, const double *I, const double *jac, const double *pos, size_t N,
                                      const double *z, size_t Nz) {
    for (size_t i = 0; i < Nz; i++)
        for (size_t j = 0; j < N; j++)
            for (int k = 0; k < 12; k++) {
                double D[9];
                current_axial_derivatives_radial_ring(z[i], pos[36 * j + 3 * k], pos[36 * j + 3 * k + 2], D);
                double w = jac[12 * j + k] * I[j];
                for (int l = 0; l < 9; l++) d[9 * i + l] += w * D[l];
            }
}

/* ------------------------------------------------------------------------- */
/* Axial (paraxial series) evaluation, radial.c:313-384.                       */
/* coeffs (Nz-1, 9, 6), highest power first.                                   */
/* ------------------------------------------------------------------------- */
static int axial_interval(const double *zs, size_t Nz, double z, double *dz_out) {
    double dz = zs[1] - zs[0];
    int idx = (int)((z - zs[0]) / dz);
    if (idx < 0) idx = 0;
    if (idx > (int)Nz - 2) idx = (int)Nz - 2;
    *dz_out = z - zs[idx];
    return idx;
}

static void axial_local(const double *C, double x, double d[9]) {
    for (int l = 0; l < 9; l++) {
        const double *c = C + 6 * l;
        d[l] = ((((c[0] * x + c[1]) * x + c[2]) * x + c[3]) * x + c[4]) * x + c[5];
    }
}

double potential_radial_derivs(const double *pt, const double *zs, const double *coeffs, size_t Nz) {
    double r = sqrt(pt[0] * pt[0] + pt[1] * pt[1]), z = pt[2];
    if (!(zs[0] <= z && z <= zs[Nz - 1])) return 0.0;
    double x;
    int idx = axial_interval(zs, Nz, z, &x);
    double d[9];
    axial_local(coeffs + 54 * (size_t)idx, x, d);
    double r2 = r * r;
    return d[0] - r2 / 4.0 * d[2] + r2 * r2 / 64.0 * d[4] - r2 * r2 * r2 / 2304.0 * d[6]
           + r2 * r2 * r2 * r2 / 147456.0 * d[8];
}

void field_radial_derivs(const double *pt, double *out, const double *zs, const double *coeffs, size_t Nz) {
    double r = sqrt(pt[0] * pt[0] + pt[1] * pt[1]), z = pt[2];
    if (!(zs[0] <= z && z <= zs[Nz - 1])) { out[0] = out[1] = out[2] = 0.0; return; }
    double x;
    int idx = axial_interval(zs, Nz, z, &x);
    double d[9];
    axial_local(coeffs + 54 * (size_t)idx, x, d);
    double r2 = r * r;
    double fr = 0.5 * (d[2] - r2 / 8.0 * d[4] + r2 * r2 / 192.0 * d[6] - r2 * r2 * r2 / 9216.0 * d[8]);
    double fz = -d[1] + r2 / 4.0 * d[3] - r2 * r2 / 64.0 * d[5] + r2 * r2 * r2 / 2304.0 * d[7];
    out[0] = pt[0] * fr;
    out[1] = pt[1] * fr;
    out[2] = fz;
}

/* ------------------------------------------------------------------------- */
/* Tracing, tracing.c.  Field adaptors keep the reference's ctypes structs.    */
/* ------------------------------------------------------------------------- */
struct charges_2d { double *q; double *jac; double *pos; size_t N; };            /* radial.c:10-15 */
struct charges_3d { double *q; double *jac; double *pos; size_t N; };            /* three_d.c:18-23 */
struct field_evaluation_args { void *elec; void *mag; void *cur; double *bounds; }; /* radial.c:175-180 */
struct field_derivs_args { double *z; double *elec; double *mag; size_t Nz; };    /* three_d.c:26-31 */

typedef void (*field_fn)(double *pos, double *vel, void *args, double *E, double *H);

/* tracing.c:125-165: strict bounds; E from electrostatic charges, H from magnetostatic charges + coils */
void field_radial_traceable(double *p, double *v, void *args_p, double *E, double *H) {
    (void)v;
    struct field_evaluation_args *a = (struct field_evaluation_args *)args_p;
    const double *b = a->bounds;
    bool inside = (b == NULL) || (b[0] < p[0] && p[0] < b[1] && b[2] < p[1] && p[1] < b[3] && b[4] < p[2] && p[2] < b[5]);
    if (!inside) { E[0] = E[1] = E[2] = H[0] = H[1] = H[2] = 0.0; return; }
    struct charges_2d *e = (struct charges_2d *)a->elec, *m = (struct charges_2d *)a->mag;
    struct charges_3d *c = (struct charges_3d *)a->cur;
    field_radial(p, E, e->q, e->jac, e->pos, e->N);
    field_radial(p, H, m->q, m->jac, m->pos, m->N);
    double hc[3] = {0, 0, 0};
    current_field_radial(p, hc, c->q, c->jac, c->pos, c->N);
    H[0] += hc[0]; H[1] += hc[1]; H[2] += hc[2];
}

/* tracing.c:168-174 */
void field_radial_derivs_traceable(double *p, double *v, void *args_p, double *E, double *H) {
    (void)v;
    struct field_derivs_args *a = (struct field_derivs_args *)args_p;
    field_radial_derivs(p, E, a->z, a->elec, a->Nz);
    field_radial_derivs(p, H, a->z, a->mag, a->Nz);
}

/* A uniform (E, H) field for tests of the integrator; args = double[6]. */
void field_uniform_traceable(double *p, double *v, void *args_p, double *E, double *H) {
    (void)p; (void)v;
    const double *u = (const double *)args_p;
    for (int i = 0; i < 3; i++) { E[i] = u[i]; H[i] = u[3 + i]; }
}

/* Fehlberg 4(5) tableau (tracing.c:9-16; Fehlberg 1969, NASA TR R-315). */
static const double RK_B[6][5] = {
    {0, 0, 0, 0, 0},
    {2. / 9., 0, 0, 0, 0},
    {1. / 12., 1. / 4., 0, 0, 0},
    {69. / 128., -243. / 128., 135. / 64., 0, 0},
    {-17. / 12., 27. / 4., -27. / 5., 16. / 15., 0},
    {65. / 432., -5. / 16., 13. / 16., 4. / 27., 5. / 144.}};
static const double RK_CH[6] = {47. / 450., 0., 12. / 25., 32. / 225., 1. / 30., 6. / 25.};
static const double RK_CT[6] = {-1. / 150., 0., 3. / 100., -16. / 75., -1. / 20., 6. / 25.};

/* tracing.c:60-123.  states row 0 holds the initial state; returns the number of rows written. */
size_t trace_particle(double *times, double *states, double qm, field_fn ff, const double *bounds, double atol, void *args) {
    double y[6];
    for (int i = 0; i < 6; i++) y[i] = states[i];
    double speed = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    double hmax = 0.01 / speed;   /* TRACING_STEP_MAX, tracing.c:5 */
    double h = hmax;
    size_t N = 1;
    while (bounds[0] <= y[0] && y[0] <= bounds[1] && bounds[2] <= y[1] && y[1] <= bounds[3] &&
           bounds[4] <= y[2] && y[2] <= bounds[5]) {
        double k[6][6];
        for (int s = 0; s < 6; s++) {
            double ys[6];
            for (int c = 0; c < 6; c++) {
                double acc = y[c];
                for (int j = 0; j < s; j++) acc += RK_B[s][j] * k[j][c];
                ys[c] = acc;
            }
            double E[3] = {0, 0, 0}, H[3] = {0, 0, 0};
            ff(ys, ys + 3, args, E, H);
            double vxh0 = ys[4] * H[2] - ys[5] * H[1];
            double vxh1 = ys[5] * H[0] - ys[3] * H[2];
            double vxh2 = ys[3] * H[1] - ys[4] * H[0];
            k[s][0] = h * ys[3]; k[s][1] = h * ys[4]; k[s][2] = h * ys[5];
            k[s][3] = h * qm * (E[0] + MU0 * vxh0);
            k[s][4] = h * qm * (E[1] + MU0 * vxh1);
            k[s][5] = h * qm * (E[2] + MU0 * vxh2);
        }
        double ep = 0.0, ev = 0.0;
        for (int c = 0; c < 6; c++) {
            double e = 0.0;
            for (int s = 0; s < 6; s++) e += RK_CT[s] * k[s][c];
            e = fabs(e);
            if (c < 3) { if (e > ep) ep = e; }
            else { if (e > ev) ev = e; }
        }
        double err = ep + h * ev;
        if (err <= atol) {
            for (int c = 0; c < 6; c++) {
                y[c] += RK_CH[0] * k[0][c] + RK_CH[1] * k[1][c] + RK_CH[2] * k[2][c] + RK_CH[3] * k[3][c]
                        + RK_CH[4] * k[4][c] + RK_CH[5] * k[5][c];
                states[6 * N + c] = y[c];
            }
            times[N] = times[N - 1] + h;
            N += 1;
            if (N == TRACING_BLOCK_SIZE) return N;
        }
        double hn = 0.9 * h * pow(atol / err, 0.2);
        h = hn < hmax ? hn : hmax;
    }
    return N;
}

/* tracing.c:184-223: scan backwards for a sign change of the signed distance,
   interpolate all six components linearly.  Zero counts as negative. */
bool plane_intersection(const double *p0, const double *n, const double *states, size_t Np, double out[6]) {
    if (Np < 2) return false;
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double *s = states + 6 * (Np - 1);
    double prev = (n[0] * (p0[0] - s[0]) + n[1] * (p0[1] - s[1]) + n[2] * (p0[2] - s[2])) / nn;
    for (long i = (long)Np - 2; i >= 0; i--) {
        s = states + 6 * i;
        double kap = (n[0] * (p0[0] - s[0]) + n[1] * (p0[1] - s[1]) + n[2] * (p0[2] - s[2])) / nn;
        int sk = kap > 0 ? 1 : -1, sp = prev > 0 ? 1 : -1;
        if (sk != sp) {
            double diff = kap - prev;
            double f = -prev / diff, fp = kap / diff;
            for (int c = 0; c < 6; c++) out[c] = fp * states[6 * (i + 1) + c] + f * states[6 * i + c];
            return true;
        }
        prev = kap;
    }
    return false;
}
