/*
 * akantu_oracle.c -- CPU FP64 restatement of Akantu's explicit
 * central-difference / lumped-mass step (SolidMechanicsModel::solveStep with
 * _explicit_lumped_mass), used as the parity oracle.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product library.
 *
 * Build: gcc -O2 -ffp-contract=off (see oracle/Makefile).  Floating-point
 * contraction is off so that every multiply and add rounds separately, as in
 * the reference compiled for plain x86-64.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Where the reference delegates a reduction
 * to Eigen 3.4 (not vendored in the reference, so unbuildable here), this file
 * uses Eigen's published algorithm for the operation (cofactor 3x3 inverse and
 * determinant for fixed-size matrices; partial-pivot LU for the dynamic-size
 * Jacobian) and sums small products in index order.  Agreement with the
 * reference is therefore pinned to the tolerances of the reference's own tests
 * (5e-13 on gradients, 1e-13 on stresses), not to the last bit.
 */
#include "akantu_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* -------------------------------------------------------------------------- */
/* Threads of the element and DOF loops (the CPU baseline's multi-core run).
 * Every element's arithmetic and the assembly order are those of the
 * one-thread run, so the results are bit-identical for any count. */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

int orc_nnpe(int type) { return type == ORC_TET4 ? 4 : 8; }
int orc_nqp(int type) { return type == ORC_TET4 ? 1 : 8; }

/* Gauss tables: gauss_integration.cc:42-43 (segment, 2 points at -+1/sqrt(3),
 * weights 1), :109-110 (tetrahedron, 1 point at 1/4, weight 1/6).  The
 * hexahedron points are the tensor product built by
 * gauss_integration_tmpl.hh:136-181, which orders them with xi fastest. */
void orc_gauss_points(int type, double *xi, double *w) {
  if (type == ORC_TET4) {
    xi[0] = 1. / 4.;
    xi[1] = 1. / 4.;
    xi[2] = 1. / 4.;
    w[0] = 1. / 6.;
    return;
  }
  double pos[2] = {-1. / sqrt(3.), 1. / sqrt(3.)};
  double wts[2] = {1., 1.};
  for (int q = 0; q < 8; ++q) {
    xi[3 * q + 0] = pos[q & 1];
    xi[3 * q + 1] = pos[(q >> 1) & 1];
    xi[3 * q + 2] = pos[(q >> 2) & 1];
    /* quads_weights.fill(1) then *= weight per direction */
    double wq = 1.;
    wq *= wts[q & 1];
    wq *= wts[(q >> 1) & 1];
    wq *= wts[(q >> 2) & 1];
    w[q] = wq;
  }
}

/* N: element_class_tetrahedron_4_inline_impl.hh:61-77,
 *    element_class_hexahedron_8_inline_impl.hh:111-124 */
void orc_shapes(int type, const double *c, double *N) {
  if (type == ORC_TET4) {
    double c0 = 1 - c[0] - c[1] - c[2];
    N[0] = c0;
    N[1] = c[1];
    N[2] = c[2];
    N[3] = c[0];
    return;
  }
  N[0] = .125 * (1 - c[0]) * (1 - c[1]) * (1 - c[2]);
  N[1] = .125 * (1 + c[0]) * (1 - c[1]) * (1 - c[2]);
  N[2] = .125 * (1 + c[0]) * (1 + c[1]) * (1 - c[2]);
  N[3] = .125 * (1 - c[0]) * (1 + c[1]) * (1 - c[2]);
  N[4] = .125 * (1 - c[0]) * (1 - c[1]) * (1 + c[2]);
  N[5] = .125 * (1 + c[0]) * (1 - c[1]) * (1 + c[2]);
  N[6] = .125 * (1 + c[0]) * (1 + c[1]) * (1 + c[2]);
  N[7] = .125 * (1 - c[0]) * (1 + c[1]) * (1 + c[2]);
}

/* dN/dxi stored row-major dnds[i*nnpe + n] = dN_n / dxi_i:
 * element_class_tetrahedron_4_inline_impl.hh:81-121,
 * element_class_hexahedron_8_inline_impl.hh:128-188 */
void orc_dnds(int type, const double *c, double *d) {
  if (type == ORC_TET4) {
    static const double t[12] = {-1., 0., 0., 1., /* d/dxi   */
                                 -1., 1., 0., 0., /* d/deta  */
                                 -1., 0., 1., 0.}; /* d/dzeta */
    memcpy(d, t, sizeof(t));
    return;
  }
  d[0 * 8 + 0] = -.125 * (1 - c[1]) * (1 - c[2]);
  d[0 * 8 + 1] = .125 * (1 - c[1]) * (1 - c[2]);
  d[0 * 8 + 2] = .125 * (1 + c[1]) * (1 - c[2]);
  d[0 * 8 + 3] = -.125 * (1 + c[1]) * (1 - c[2]);
  d[0 * 8 + 4] = -.125 * (1 - c[1]) * (1 + c[2]);
  d[0 * 8 + 5] = .125 * (1 - c[1]) * (1 + c[2]);
  d[0 * 8 + 6] = .125 * (1 + c[1]) * (1 + c[2]);
  d[0 * 8 + 7] = -.125 * (1 + c[1]) * (1 + c[2]);

  d[1 * 8 + 0] = -.125 * (1 - c[0]) * (1 - c[2]);
  d[1 * 8 + 1] = -.125 * (1 + c[0]) * (1 - c[2]);
  d[1 * 8 + 2] = .125 * (1 + c[0]) * (1 - c[2]);
  d[1 * 8 + 3] = .125 * (1 - c[0]) * (1 - c[2]);
  d[1 * 8 + 4] = -.125 * (1 - c[0]) * (1 + c[2]);
  d[1 * 8 + 5] = -.125 * (1 + c[0]) * (1 + c[2]);
  d[1 * 8 + 6] = .125 * (1 + c[0]) * (1 + c[2]);
  d[1 * 8 + 7] = .125 * (1 - c[0]) * (1 + c[2]);

  d[2 * 8 + 0] = -.125 * (1 - c[0]) * (1 - c[1]);
  d[2 * 8 + 1] = -.125 * (1 + c[0]) * (1 - c[1]);
  d[2 * 8 + 2] = -.125 * (1 + c[0]) * (1 + c[1]);
  d[2 * 8 + 3] = -.125 * (1 - c[0]) * (1 + c[1]);
  d[2 * 8 + 4] = .125 * (1 - c[0]) * (1 - c[1]);
  d[2 * 8 + 5] = .125 * (1 + c[0]) * (1 - c[1]);
  d[2 * 8 + 6] = .125 * (1 + c[0]) * (1 + c[1]);
  d[2 * 8 + 7] = .125 * (1 - c[0]) * (1 + c[1]);
}

/* -------------------------------------------------------------------------- */
/* 3x3 helpers, m[i][j] */
typedef double mat3[3][3];

/* Fixed-size 3x3 determinant (cofactor expansion along the first row, Eigen's
 * determinant_impl<.,3>); used by Material::gradUToF/neo-Hookean
 * (material_neohookean_inline_impl.hh:45) and tetrahedron_volume
 * (aka_math_tmpl.hh:71-93). */
static double det3_fixed(const mat3 m) {
  double a = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
  double b = m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
  double c = m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  return a - b + c;
}

/* Fixed-size 3x3 inverse (adjugate scaled by 1/det, Eigen's
 * compute_inverse<.,.,3>); used for C^{-1} in the neo-Hookean law
 * (material_neohookean_inline_impl.hh:47-48). */
static double cof3(const mat3 m, int i, int j) {
  int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
}
static void inv3_fixed(const mat3 m, mat3 r) {
  double c0 = cof3(m, 0, 0), c1 = cof3(m, 1, 0), c2 = cof3(m, 2, 0);
  double det = c0 * m[0][0] + c1 * m[1][0] + c2 * m[2][0];
  double invdet = 1. / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r[i][j] = cof3(m, j, i) * invdet;
}

/* Dynamic-size 3x3 determinant and inverse through a partial-pivot LU
 * (Eigen's PartialPivLU: unblocked LU with row pivoting on the largest
 * magnitude, det = sign * prod(diag U), inverse = U^{-1} L^{-1} P solved
 * column by column).  The reference's Jacobian matrices are dynamic-size
 * (element_class_tmpl.hh:327 Matrix<Real> J; :345-346 J.determinant();
 * :381-385 J_.inverse()). */
static void lu3(const mat3 A, mat3 lu, int perm[3], int *nb_transp) {
  memcpy(lu, A, sizeof(mat3));
  *nb_transp = 0;
  for (int k = 0; k < 3; ++k) {
    int p = k;
    double big = fabs(lu[k][k]);
    for (int i = k + 1; i < 3; ++i)
      if (fabs(lu[i][k]) > big) {
        big = fabs(lu[i][k]);
        p = i;
      }
    perm[k] = p;
    if (big != 0.) {
      if (p != k) {
        for (int j = 0; j < 3; ++j) {
          double t = lu[k][j];
          lu[k][j] = lu[p][j];
          lu[p][j] = t;
        }
        ++*nb_transp;
      }
      for (int i = k + 1; i < 3; ++i)
        lu[i][k] = lu[i][k] / lu[k][k];
    }
    for (int i = k + 1; i < 3; ++i)
      for (int j = k + 1; j < 3; ++j)
        lu[i][j] = lu[i][j] - lu[i][k] * lu[k][j];
  }
}

static double det3_dyn(const mat3 A) {
  mat3 lu;
  int perm[3], nt;
  lu3(A, lu, perm, &nt);
  double s = (nt % 2) ? -1. : 1.;
  return s * (lu[0][0] * lu[1][1] * lu[2][2]);
}

static void inv3_dyn(const mat3 A, mat3 R) {
  mat3 lu;
  int perm[3], nt;
  lu3(A, lu, perm, &nt);
  /* dst = P * I : apply the row transpositions in order */
  mat3 x = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
  for (int k = 0; k < 3; ++k)
    if (perm[k] != k)
      for (int j = 0; j < 3; ++j) {
        double t = x[k][j];
        x[k][j] = x[perm[k]][j];
        x[perm[k]][j] = t;
      }
  for (int j = 0; j < 3; ++j) {
    /* unit lower: forward substitution, column-oriented updates */
    for (int i = 0; i < 3; ++i)
      for (int r = i + 1; r < 3; ++r)
        x[r][j] = x[r][j] - x[i][j] * lu[r][i];
    /* upper: backward substitution, scale by the reciprocal pivot */
    for (int i = 2; i >= 0; --i) {
      double a = 1. / lu[i][i];
      x[i][j] = x[i][j] * a;
      for (int r = 0; r < i; ++r)
        x[r][j] = x[r][j] - x[i][j] * lu[r][i];
    }
  }
  memcpy(R, x, sizeof(mat3));
}

/* -------------------------------------------------------------------------- */
/* Shape derivatives and w*detJ of one element at all quadrature points:
 * shape_lagrange_inline_impl.hh:61-83 (dnds, J = dnds X^T, dNdX = J^{-1} dnds),
 * element_class_tmpl.hh:297-301 (computeJMat), :319-335 + :345-346 (detJ),
 * integrator_gauss_inline_impl.hh:176-195 (negative-jacobian check),
 * :379-398 (J *= w).  Returns -1 - q for the first qp with detJ < 0. */
static int element_shape_derivatives(int type, const double *Xe /*[nnpe][3]*/,
                                     double *B /*[nqp][3*nnpe]*/,
                                     double *wJ /*[nqp]*/) {
  int nnpe = orc_nnpe(type), nqp = orc_nqp(type);
  double xi[24], w[8];
  orc_gauss_points(type, xi, w);
  int bad = 0;
  for (int q = 0; q < nqp; ++q) {
    double d[24];
    orc_dnds(type, xi + 3 * q, d);
    mat3 J;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.;
        for (int n = 0; n < nnpe; ++n)
          s += d[i * nnpe + n] * Xe[3 * n + j];
        J[i][j] = s;
      }
    double detJ = det3_dyn(J);
    if (detJ < 0 && !bad)
      bad = -1 - q;
    if (wJ)
      wJ[q] = detJ * w[q];
    if (B) {
      mat3 Ji;
      inv3_dyn(J, Ji);
      for (int n = 0; n < nnpe; ++n)
        for (int i = 0; i < 3; ++i) {
          double s = 0.;
          for (int k = 0; k < 3; ++k)
            s += Ji[i][k] * d[k * nnpe + n];
          B[q * 3 * nnpe + i + 3 * n] = s;
        }
    }
  }
  return bad;
}

static void gather_element(const orc_model *m, long e, const double *field,
                           double *fe) {
  int nnpe = orc_nnpe(m->type);
  for (int n = 0; n < nnpe; ++n) {
    long node = m->conn[e * nnpe + n];
    fe[3 * n + 0] = field[3 * node + 0];
    fe[3 * n + 1] = field[3 * node + 1];
    fe[3 * n + 2] = field[3 * node + 2];
  }
}

/* a13: fe_engine_template_tmpl.hh:98-120 precompute on the mesh nodes
 * (total Lagrangian reference configuration). */
long orc_precompute(orc_model *m) {
  int nnpe = orc_nnpe(m->type), nqp = orc_nqp(m->type);
  long first_bad = 0;
  for (long e = 0; e < m->nel; ++e) {
    double Xe[24];
    gather_element(m, e, m->X, Xe);
    int bad = element_shape_derivatives(
        m->type, Xe, m->B ? m->B + e * nqp * 3 * nnpe : NULL,
        m->wJ + e * nqp);
    if (bad && !first_bad)
      first_bad = 1 + e * nqp + (-bad - 1);
  }
  return first_bad;
}

/* -------------------------------------------------------------------------- */
/* Lame constants, material_elastic.cc:53-62 / material_neohookean.cc:62-68 */
void orc_lame(double E, double nu, double *lambda, double *mu) {
  *lambda = nu * E / ((1 + nu) * (1 - 2 * nu));
  *mu = E / (2 * (1 + nu));
}

static void element_params(const orc_model *m, long e, double *lambda,
                           double *mu, double *sy, double *h) {
  double E = m->E_el ? m->E_el[e] : m->E;
  double nu = m->nu_el ? m->nu_el[e] : m->nu;
  orc_lame(E, nu, lambda, mu);
  *sy = m->sigma_y_el ? m->sigma_y_el[e] : m->sigma_y;
  *h = m->h_el ? m->h_el[e] : m->h;
}

/* a14: solid_mechanics_model_mass.cc:49-129 -> fe_engine_template_tmpl_field.hh
 * :112-138 (assembleLumpedRowSum: rho N per qp, integrate, assemble) ->
 * dof_manager.cc:147-168 (element order scatter into a zeroed array). */
void orc_lumped_mass(orc_model *m) {
  int nnpe = orc_nnpe(m->type), nqp = orc_nqp(m->type);
  double xi[24], w[8], N[8 * 8];
  orc_gauss_points(m->type, xi, w);
  for (int q = 0; q < nqp; ++q)
    orc_shapes(m->type, xi + 3 * q, N + q * nnpe);
  memset(m->mass, 0, sizeof(double) * 3 * m->nn);
  for (long e = 0; e < m->nel; ++e) {
    const double *wJ = m->wJ + e * nqp;
    for (int n = 0; n < nnpe; ++n) {
      /* int_f = f * J with f(c, q) = rho * N_n(q) for every component c */
      double s = 0.;
      for (int q = 0; q < nqp; ++q)
        s += (m->rho * N[q * nnpe + n]) * wJ[q];
      long node = m->conn[e * nnpe + n];
      for (int c = 0; c < 3; ++c)
        m->mass[3 * node + c] += 1. * s;
    }
  }
}

/* -------------------------------------------------------------------------- */
/* a15: element inradius -- Tet4: 2 * Math::tetrahedron_inradius
 * (element_class_tetrahedron_4_inline_impl.hh:127-131, aka_math_tmpl.hh:71-124);
 * Hex8: min of the 12 edges (element_class_hexahedron_8_inline_impl.hh:194-212).
 */
static double dist3(const double *a, const double *b) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

static double tet_volume(const double *c1, const double *c2, const double *c3,
                         const double *c4) {
  mat3 xx;
  /* xx.col(k) = coord */
#define SETCOL(k, c)                                                           \
  do {                                                                         \
    xx[0][k] = (c)[0];                                                         \
    xx[1][k] = (c)[1];                                                         \
    xx[2][k] = (c)[2];                                                         \
  } while (0)
  SETCOL(0, c2);
  SETCOL(1, c3);
  SETCOL(2, c4);
  double vol = det3_fixed(xx);
  SETCOL(0, c1);
  vol -= det3_fixed(xx);
  SETCOL(1, c2);
  vol += det3_fixed(xx);
  SETCOL(2, c3);
  vol -= det3_fixed(xx);
#undef SETCOL
  vol /= 6;
  return vol;
}

double orc_element_inradius(int type, const double *X) {
  if (type == ORC_TET4) {
    const double *c1 = X, *c2 = X + 3, *c3 = X + 6, *c4 = X + 9;
    double l12 = dist3(c1, c2), l13 = dist3(c1, c3), l14 = dist3(c1, c4);
    double l23 = dist3(c2, c3), l24 = dist3(c2, c4), l34 = dist3(c3, c4);
    double s1 = (l12 + l23 + l13) * 0.5;
    s1 = sqrt(s1 * (s1 - l12) * (s1 - l23) * (s1 - l13));
    double s2 = (l12 + l24 + l14) * 0.5;
    s2 = sqrt(s2 * (s2 - l12) * (s2 - l24) * (s2 - l14));
    double s3 = (l23 + l34 + l24) * 0.5;
    s3 = sqrt(s3 * (s3 - l23) * (s3 - l34) * (s3 - l24));
    double s4 = (l13 + l34 + l14) * 0.5;
    s4 = sqrt(s4 * (s4 - l13) * (s4 - l34) * (s4 - l14));
    double volume = tet_volume(c1, c2, c3, c4);
    return 2. * (3 * volume / (s1 + s2 + s3 + s4));
  }
  static const int edges[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0},
                                   {0, 4}, {1, 5}, {2, 6}, {3, 7},
                                   {4, 5}, {5, 6}, {6, 7}, {7, 4}};
  double p = 0.;
  for (int k = 0; k < 12; ++k) {
    double l = dist3(X + 3 * edges[k][0], X + 3 * edges[k][1]);
    if (k == 0 || l < p)
      p = l;
  }
  return p;
}

/* solid_mechanics_model.cc:423-473 with Material::getCelerity = push wave
 * speed (material_elastic.cc:125-128, material_neohookean.cc:145-148). */
double orc_stable_dt(const orc_model *m) {
  int nnpe = orc_nnpe(m->type);
  double min_dt = DBL_MAX;
  for (long e = 0; e < m->nel; ++e) {
    double Xe[24];
    for (int n = 0; n < nnpe; ++n) {
      long node = m->conn[e * nnpe + n];
      for (int c = 0; c < 3; ++c) /* current_position = X + u */
        Xe[3 * n + c] = m->X[3 * node + c] + (m->u ? m->u[3 * node + c] : 0.);
    }
    double lambda, mu, sy, h;
    element_params(m, e, &lambda, &mu, &sy, &h);
    double el_h = orc_element_inradius(m->type, Xe);
    double el_c = sqrt((lambda + 2 * mu) / m->rho);
    double el_dt = el_h / el_c;
    if (el_dt < min_dt)
      min_dt = el_dt;
  }
  return min_dt;
}

/* -------------------------------------------------------------------------- */
/* Constitutive laws, per quadrature point, 3x3 tensors column-major t[i+3j].
 */
#define T(a, i, j) (a)[(i) + 3 * (j)]

/* material_elastic_inline_impl.hh:33-49: sigma = 2 mu eps + (lambda tr + s_th) I
 * with eps = 0.5 (g^T + g) (material_inline_impl.hh:109-113). */
static void elastic_on_quad(double lambda, double mu, const double *g,
                            double sigma_th, double *sigma) {
  double trace = T(g, 0, 0) + T(g, 1, 1) + T(g, 2, 2);
  double two_mu = 2. * mu;
  double s = lambda * trace + sigma_th;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      double eps = .5 * (T(g, j, i) + T(g, i, j));
      T(sigma, i, j) = two_mu * eps + s * (i == j ? 1. : 0.);
    }
}

/* material_inline_impl.hh:125-130 gradUToE: E = (g^T g + g^T + g) / 2 */
static void grad_u_to_green(const double *g, double *E) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      double p = 0.;
      for (int k = 0; k < 3; ++k)
        p += T(g, k, i) * T(g, k, j);
      T(E, i, j) = (p + T(g, j, i) + T(g, i, j)) / 2.;
    }
}

/* material_neohookean_inline_impl.hh:39-50: F = I + g, C = F^T F,
 * J = det F sqrt(C33) (C33 = 1 in 3-D), S = mu I + (lambda ln J - mu) C^{-1} */
static void neohookean_on_quad(double lambda, double mu, const double *g,
                               double *S) {
  mat3 F, C, Ci;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      F[i][j] = (i == j ? 1. : 0.) + T(g, i, j);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.;
      for (int k = 0; k < 3; ++k)
        s += F[k][i] * F[k][j];
      C[i][j] = s;
    }
  double J = det3_fixed(F) * sqrt(1.);
  inv3_fixed(C, Ci);
  double t = lambda * log(J) - mu;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      T(S, i, j) = (i == j ? 1. : 0.) * mu + t * Ci[i][j];
}

/* material_linear_isotropic_hardening_inline_impl.hh:31-86 and
 * material_plastic.hh:65-87 (small strain, 3-D). */
static void plastic_lih_on_quad(double lambda, double mu, double sigma_y,
                                double h, const double *g, const double *pg,
                                const double *psigma, double piso,
                                const double *peps, double dsth, double *sigma,
                                double *iso, double *eps) {
  double dg[9], str[9], dev[9];
  for (int k = 0; k < 9; ++k)
    dg[k] = g[k] - pg[k];
  /* trial stress: elastic on grad(delta u) with delta sigma_th = previous
   * sigma_th - sigma_th (:33-36; 0 without a temperature field) */
  elastic_on_quad(lambda, mu, dg, dsth, str);
  for (int k = 0; k < 9; ++k)
    str[k] += psigma[k];
  double tr = T(str, 0, 0) + T(str, 1, 1) + T(str, 2, 2);
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      T(dev, i, j) = T(str, i, j) - (i == j ? 1. : 0.) * tr / 3.0;
  double s = 0.;
  for (int k = 0; k < 9; ++k)
    s += dev[k] * dev[k];
  double eff = sqrt(3. / 2. * s);
  /* the current iso_hardening equals the previous one after
   * savePreviousState (material.cc:737) */
  int yielding = ((eff - piso - sigma_y) > 0);
  double dp = yielding ? (eff - sigma_y - piso) / (3 * mu + h) : 0;
  *iso = piso + h * dp;
  double deps[9] = {0};
  double tol = 1e2 * DBL_EPSILON; /* aka_extern.cc:70 Math::tolerance */
  if (fabs(eff) > sqrt(s) * tol)
    for (int k = 0; k < 9; ++k)
      deps[k] = dev[k] * 3. / 2. * dp / eff;
  /* computeStressAndInelasticStrainOnQuad */
  double dge[9], sel[9];
  for (int k = 0; k < 9; ++k)
    dge[k] = (g[k] - pg[k]) - deps[k];
  elastic_on_quad(lambda, mu, dge, 0., sel);
  for (int k = 0; k < 9; ++k) {
    sigma[k] = psigma[k] + sel[k];
    eps[k] = peps[k] + deps[k];
  }
}

void orc_law_stress_th(int law, double lambda, double mu, double sigma_y,
                       double h, const double *g, const double *pg,
                       const double *psigma, double piso, const double *peps,
                       double sth, double dsth, double *sigma, double *iso,
                       double *eps) {
  switch (law) {
  case ORC_ELASTIC:
    elastic_on_quad(lambda, mu, g, sth, sigma);
    break;
  case ORC_ELASTIC_FD: {
    /* material_elastic.cc:92-96: the law applied to the Green strain */
    double E[9];
    grad_u_to_green(g, E);
    elastic_on_quad(lambda, mu, E, sth, sigma);
    break;
  }
  case ORC_NEOHOOKEAN:
    /* material_neohookean_inline_impl.hh:39-50 does not read sigma_th */
    neohookean_on_quad(lambda, mu, g, sigma);
    break;
  case ORC_PLASTIC_LIH:
    plastic_lih_on_quad(lambda, mu, sigma_y, h, g, pg, psigma, piso, peps,
                        dsth, sigma, iso, eps);
    break;
  }
}

/* sigma_th = -E/(1-2nu) alpha dT = -0 with the default alpha = 0
 * (material_thermal.hh:91-95), and the plastic trial stress's
 * previous - current sigma_th is then 0 */
void orc_law_stress(int law, double lambda, double mu, double sigma_y,
                    double h, const double *g, const double *pg,
                    const double *psigma, double piso, const double *peps,
                    double *sigma, double *iso, double *eps) {
  orc_law_stress_th(law, lambda, mu, sigma_y, h, g, pg, psigma, piso, peps, -0.,
                    0., sigma, iso, eps);
}

/* MaterialThermal::computeStressOnQuad (material_thermal.hh:91-95) with the
 * element's E and nu */
static double thermal_stress(const orc_model *m, long e, long iq) {
  double E = m->E_el ? m->E_el[e] : m->E, nu = m->nu_el ? m->nu_el[e] : m->nu;
  return -E / (1. - 2. * nu) * m->alpha * m->delta_T[iq];
}

static int law_is_fd(int law) {
  return law == ORC_ELASTIC_FD || law == ORC_NEOHOOKEAN;
}

/* -------------------------------------------------------------------------- */
/* a6 + a7: material.cc:128-154 (computeAllStresses) ->
 * shape_lagrange_inline_impl.hh:344-368 + fe_engine_inline_impl.hh:99-139
 * (gather u_e) + shape_functions_inline_impl.hh:303-310 (grad u = u_e B^T),
 * then computeStress (material.hh:78). */
void orc_compute_all_stresses(orc_model *m) {
  int nnpe = orc_nnpe(m->type), nqp = orc_nqp(m->type);
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (long e = 0; e < m->nel; ++e) {
    double Bl[8 * 24], wJl[8];
    double ue[24];
    gather_element(m, e, m->u, ue);
    const double *B;
    if (m->B) {
      B = m->B + e * nqp * 3 * nnpe;
    } else {
      double Xe[24];
      gather_element(m, e, m->X, Xe);
      element_shape_derivatives(m->type, Xe, Bl, wJl);
      B = Bl;
    }
    double lambda, mu, sy, h;
    element_params(m, e, &lambda, &mu, &sy, &h);
    for (int q = 0; q < nqp; ++q) {
      const double *Bq = B + q * 3 * nnpe;
      long iq = e * nqp + q;
      double *g = m->grad_u + 9 * iq;
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
          double s = 0.;
          for (int n = 0; n < nnpe; ++n)
            s += ue[i + 3 * n] * Bq[j + 3 * n];
          T(g, i, j) = s;
        }
      /* gradu_vect -= eigengradu (material.cc:147) */
      if (m->eigen_grad_u)
        for (int k = 0; k < 9; ++k)
          g[k] -= m->eigen_grad_u[9 * iq + k];
      /* MaterialThermal::computeStress, called first by the laws'
       * computeStress (material_elastic.cc:85-86) */
      double sth = -0., dsth = 0.;
      if (m->delta_T) {
        sth = thermal_stress(m, e, iq);
        m->sigma_th[iq] = sth;
        dsth = m->prev_sigma_th[iq] - sth;
      }
      if (m->law == ORC_PLASTIC_LIH) {
        orc_law_stress_th(m->law, lambda, mu, sy, h, g, m->prev_grad_u + 9 * iq,
                          m->prev_stress + 9 * iq, m->prev_iso[iq],
                          m->prev_eps_p + 9 * iq, sth, dsth, m->stress + 9 * iq,
                          m->iso + iq, m->eps_p + 9 * iq);
      } else {
        orc_law_stress_th(m->law, lambda, mu, sy, h, g, NULL, NULL, 0., NULL, sth,
                          dsth, m->stress + 9 * iq, NULL, NULL);
      }
    }
  }
}

/* Finite-deformation elemental force at one qp: material.cc:525-584 with
 * VoigtHelper<3>::transferBMatrixToSymVoigtBMatrix (aka_voigthelper_tmpl.hh
 * :99-140), transferBMatrixToBL2 (:180-207), Voigt order xx yy zz yz xz xy
 * (aka_voigthelper.hh:57-70): r = (B_L + B_NL)^T voigt(S). */
static void fd_qp_force(int nnpe, const double *Bq, const double *g,
                        const double *S, double *r) {
  double BL[6][24], BN[6][24];
  int ncol = 3 * nnpe;
  memset(BL, 0, sizeof(BL));
  memset(BN, 0, sizeof(BN));
#define BQ(i, n) Bq[(i) + 3 * (n)]
  for (int i = 0; i < 3; ++i)
    for (int n = 0; n < nnpe; ++n)
      BL[i][i + n * 3] = BQ(i, n);
  for (int n = 0; n < nnpe; ++n) {
    double dndx = BQ(0, n), dndy = BQ(1, n), dndz = BQ(2, n);
    BL[3][1 + n * 3] = dndz;
    BL[3][2 + n * 3] = dndy;
    BL[4][0 + n * 3] = dndz;
    BL[4][2 + n * 3] = dndx;
    BL[5][0 + n * 3] = dndy;
    BL[5][1 + n * 3] = dndx;
  }
  for (int I = 0; I < 3; ++I)
    for (int a = 0; a < nnpe; ++a)
      for (int i = 0; i < 3; ++i)
        BN[I][a * 3 + i] = T(g, i, I) * BQ(I, a);
  for (int Iv = 3; Iv < 6; ++Iv)
    for (int a = 0; a < nnpe; ++a)
      for (int k = 0; k < 3; ++k) {
        int aux = Iv - 3;
        for (int mm = 0; mm < 3; ++mm)
          if (mm != aux) {
            int index1 = mm, index2 = 3 - mm - aux;
            BN[Iv][a * 3 + k] += T(g, k, index1) * BQ(index2, a);
          }
      }
#undef BQ
  double Sv[6] = {T(S, 0, 0), T(S, 1, 1), T(S, 2, 2),
                  T(S, 1, 2), T(S, 0, 2), T(S, 0, 1)};
  for (int c = 0; c < ncol; ++c) {
    double s = 0.;
    for (int I = 0; I < 6; ++I)
      s += (BL[I][c] + BN[I][c]) * Sv[I];
    r[c] = s;
  }
}

/* a8 + a9: material.cc:467-510 (small strain: computeBtD
 * shape_lagrange_inline_impl.hh:373-410, integrate
 * integrator_gauss_inline_impl.hh:401-441) or :525-584 (finite deformation),
 * then DOFManager::assembleElementalArrayLocalArray (dof_manager.cc:65-120)
 * with scale -1 in ascending element order.  f_int is zeroed first
 * (solid_mechanics_model.cc:338). */
static void element_force(const orc_model *m, long e, double *re);

void orc_assemble_internal_forces(orc_model *m) {
  int nnpe = orc_nnpe(m->type);
  int ncol = 3 * nnpe;
  memset(m->f_int, 0, sizeof(double) * 3 * m->nn);
  if (g_threads > 1) {
    /* elemental forces in parallel, then the scatter in element order */
    double *all = (double *)malloc(sizeof(double) * (size_t)ncol * (size_t)(m->nel > 0 ? m->nel : 1));
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long e = 0; e < m->nel; ++e)
      element_force(m, e, all + e * ncol);
    for (long e = 0; e < m->nel; ++e)
      for (int n = 0; n < nnpe; ++n) {
        long node = m->conn[e * nnpe + n];
        for (int i = 0; i < 3; ++i)
          m->f_int[3 * node + i] += -1. * all[e * ncol + i + 3 * n];
      }
    free(all);
    return;
  }
  for (long e = 0; e < m->nel; ++e) {
    double re[24];
    element_force(m, e, re);
    for (int n = 0; n < nnpe; ++n) {
      long node = m->conn[e * nnpe + n];
      for (int i = 0; i < 3; ++i)
        m->f_int[3 * node + i] += -1. * re[i + 3 * n];
    }
  }
}

/* r_e = sum_q B_q^T sigma_q wJ_q (or the finite-deformation form) */
static void element_force(const orc_model *m, long e, double *re) {
  int nnpe = orc_nnpe(m->type), nqp = orc_nqp(m->type);
  int ncol = 3 * nnpe;
  double Bl[8 * 24], wJl[8];
  int fd = law_is_fd(m->law);
  {
    const double *B;
    const double *wJ = m->wJ + e * nqp;
    if (m->B) {
      B = m->B + e * nqp * ncol;
    } else {
      double Xe[24];
      gather_element(m, e, m->X, Xe);
      element_shape_derivatives(m->type, Xe, Bl, wJl);
      B = Bl;
    }
    double fq[8][24];
    for (int q = 0; q < nqp; ++q) {
      long iq = e * nqp + q;
      const double *Bq = B + q * ncol;
      const double *S = m->stress + 9 * iq;
      if (fd) {
        fd_qp_force(nnpe, Bq, m->grad_u + 9 * iq, S, fq[q]);
      } else {
        /* BtD(i, n) = sum_j D(i, j) B(j, n) */
        for (int n = 0; n < nnpe; ++n)
          for (int i = 0; i < 3; ++i) {
            double s = 0.;
            for (int j = 0; j < 3; ++j)
              s += T(S, i, j) * Bq[j + 3 * n];
            fq[q][i + 3 * n] = s;
          }
      }
    }
    for (int c = 0; c < ncol; ++c) {
      double s = 0.;
      for (int q = 0; q < nqp; ++q)
        s += fq[q][c] * wJ[q];
      re[c] = s;
    }
  }
}

void orc_internal_forces(orc_model *m) {
  orc_compute_all_stresses(m);
  orc_assemble_internal_forces(m);
}

/* -------------------------------------------------------------------------- */
/* The lumped central-difference step:
 *  predictor  newmark-beta.cc:47-77 (alpha = 0, beta = 1/2, k = h = 0:
 *             u += dt v + 0.5 dt (dt a); v = v + dt a; a = a)
 *  residual   time_step_solver.cc:176-183 (zero), then the callback's
 *             f_ext + f_int (solid_mechanics_model.cc:235-249), then
 *             integration_scheme_2nd_order.cc:70-93 +
 *             dof_manager_default.cc:260-278 (r += M * (-a))
 *  solve      non_linear_solver_lumped.cc:76-89 (x = 1 * (r / M), unblocked)
 *  corrector  newmark-beta.cc:164-197 with _acceleration coefficients
 *             (c = 1, d = beta dt, e = alpha beta dt dt = 0, :112-161). */
static void predictor(long ndof, double dt, double *u, double *v, double *a,
                      const unsigned char *blocked) {
  const double k = 0., h = 0., alpha = 0., beta = .5;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (long d = 0; d < ndof; ++d) {
    if (!blocked[d]) {
      double dt_a_n = dt * a[d];
      u[d] += (1 - k * alpha) * dt * v[d] + (.5 - h * alpha * beta) * dt * dt_a_n;
      v[d] = (1 - k) * v[d] + (1 - h * beta) * dt_a_n;
      a[d] = (1 - h) * a[d];
    }
  }
}

static void solve_and_correct(long ndof, double dt, double *u, double *v,
                              double *a, const unsigned char *blocked,
                              const double *mass, double *r) {
  const double alpha = 0., beta = .5;
  /* r += M * cache, cache = 0 + (-1) * a */
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (long d = 0; d < ndof; ++d) {
    double cache = 0.;
    cache += -1. * a[d];
    r[d] += mass[d] * cache;
  }
  double c = 1., dd = beta * dt, e = alpha * beta * dt * dt;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (long d = 0; d < ndof; ++d) {
    double x = 0.;
    if (!blocked[d])
      x = 1. * (r[d] / mass[d]);
    if (!blocked[d]) {
      u[d] += e * x;
      v[d] += dd * x;
      a[d] += c * x;
    }
  }
}

void orc_lumped_step_cb(long ndof, double dt, double *u, double *v, double *a,
                        const unsigned char *blocked, const double *mass,
                        orc_residual_cb cb, void *ctx) {
  predictor(ndof, dt, u, v, a, blocked);
  double *r = (double *)calloc((size_t)ndof, sizeof(double));
  cb(ctx, r);
  solve_and_correct(ndof, dt, u, v, a, blocked, mass, r);
  free(r);
}

/* a1..a12 for the solid mechanics model: non_linear_solver_lumped.cc:48-73 */
void orc_solve_step(orc_model *m) {
  long ndof = 3 * m->nn;
  /* beforeSolveStep: Material::savePreviousState (material.cc:737,
   * constitutive_law_tmpl.hh:80-86); only the plastic law reads it */
  if (m->law == ORC_PLASTIC_LIH) {
    long nq = m->nel * orc_nqp(m->type);
    memcpy(m->prev_grad_u, m->grad_u, sizeof(double) * 9 * nq);
    memcpy(m->prev_stress, m->stress, sizeof(double) * 9 * nq);
    memcpy(m->prev_eps_p, m->eps_p, sizeof(double) * 9 * nq);
    memcpy(m->prev_iso, m->iso, sizeof(double) * nq);
  }
  if (m->delta_T) /* sigma_th has a history (material_thermal.cc:34) */
    memcpy(m->prev_sigma_th, m->sigma_th,
           sizeof(double) * m->nel * orc_nqp(m->type));
  /* savePreviousDOFs: u_prev = u (time_step_solver_default.cc:175-177) */
  if (m->u_prev)
    memcpy(m->u_prev, m->u, sizeof(double) * ndof);
  predictor(ndof, m->dt, m->u, m->v, m->a, m->blocked);
  orc_internal_forces(m);
  double *r = (double *)calloc((size_t)ndof, sizeof(double));
  for (long d = 0; d < ndof; ++d) {
    r[d] += 1. * m->f_ext[d];
    r[d] += 1. * m->f_int[d];
  }
  solve_and_correct(ndof, m->dt, m->u, m->v, m->a, m->blocked, m->mass, r);
  free(r);
  /* afterSolveStep -> Material::updateEnergies (material.cc:740-750,
   * material_plastic.cc:86-99): wp = (sigma + sigma_prev) : (eps_p -
   * eps_p_prev) / 2, pe += wp (doubleDot = cwiseProduct().sum()) */
  if (m->law == ORC_PLASTIC_LIH && m->plastic_energy) {
    long nq = m->nel * orc_nqp(m->type);
    for (long iq = 0; iq < nq; ++iq) {
      const double *S = m->stress + 9 * iq, *pS = m->prev_stress + 9 * iq;
      const double *ep = m->eps_p + 9 * iq, *pep = m->prev_eps_p + 9 * iq;
      double dd = 0.;
      for (int k = 0; k < 9; ++k) {
        double t = (S[k] + pS[k]) * (ep[k] - pep[k]);
        dd = (k == 0) ? t : dd + t;
      }
      m->plastic_energy[iq] += dd / 2.;
    }
  }
  ++m->step;
}

/* -------------------------------------------------------------------------- */
/* Neumann boundary condition, constant traction on facets.
 * Facet shapes: element_class_triangle_3_inline_impl.hh:52-90 (N = 1 - xi -
 * eta, xi, eta; one Gauss point (1/3, 1/3), weight 1/2, gauss_integration.cc
 * :56-57) and element_class_quadrangle_4_inline_impl.hh:64-110 (bilinear, 2 x 2
 * Gauss points +-1/sqrt(3), xi fastest, weights 1).  The surface jacobian is
 * |J_xi x J_eta| (element_class_tmpl.hh:352-371), times the weight
 * (integrator_gauss_inline_impl.hh:379-398).  Per facet: computeNtb gives
 * N_a t_i at i + 3a, integrate sums over q (q order), the result is added to
 * f_ext node by node (dof_manager.cc:65-120). */
void orc_apply_traction(orc_model *m, int facet_type, long nb_facets, const int *conn,
                        const double *t) {
  const int nn = facet_type == 2 ? 3 : 4, nq = facet_type == 2 ? 1 : 4;
  double xi[4][2], w[4];
  if (facet_type == 2) {
    xi[0][0] = 1. / 3.;
    xi[0][1] = 1. / 3.;
    w[0] = 1. / 2.;
  } else {
    const double g = 1. / sqrt(3.), pos[2] = {-g, g};
    for (int q = 0; q < 4; ++q) {
      xi[q][0] = pos[q & 1];
      xi[q][1] = pos[q >> 1];
      w[q] = 1. * 1.;
    }
  }
  for (long f = 0; f < nb_facets; ++f) {
    double Xe[4][3];
    for (int a = 0; a < nn; ++a)
      for (int c = 0; c < 3; ++c)
        Xe[a][c] = m->X[3 * (long)conn[f * nn + a] + c];
    double intf[12];
    for (int q = 0; q < nq; ++q) {
      double N[4], d[2][4];
      const double x = xi[q][0], y = xi[q][1];
      if (facet_type == 2) {
        N[0] = 1 - x - y;
        N[1] = x;
        N[2] = y;
        d[0][0] = -1.; d[0][1] = 1.; d[0][2] = 0.;
        d[1][0] = -1.; d[1][1] = 0.; d[1][2] = 1.;
      } else {
        N[0] = 1. / 4. * (1. - x) * (1. - y);
        N[1] = 1. / 4. * (1. + x) * (1. - y);
        N[2] = 1. / 4. * (1. + x) * (1. + y);
        N[3] = 1. / 4. * (1. - x) * (1. + y);
        d[0][0] = -1. / 4. * (1. - y);
        d[0][1] = 1. / 4. * (1. - y);
        d[0][2] = 1. / 4. * (1. + y);
        d[0][3] = -1. / 4. * (1. + y);
        d[1][0] = -1. / 4. * (1. - x);
        d[1][1] = -1. / 4. * (1. + x);
        d[1][2] = 1. / 4. * (1. + x);
        d[1][3] = 1. / 4. * (1. - x);
      }
      double J[2][3];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) {
          double s = d[r][0] * Xe[0][c];
          for (int a = 1; a < nn; ++a)
            s += d[r][a] * Xe[a][c];
          J[r][c] = s;
        }
      const double cx = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      const double cy = J[0][2] * J[1][0] - J[0][0] * J[1][2];
      const double cz = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double wJ = sqrt(cx * cx + cy * cy + cz * cz) * w[q];
      for (int a = 0; a < nn; ++a)
        for (int i = 0; i < 3; ++i) {
          const double v = N[a] * t[i] * wJ;
          intf[i + 3 * a] = (q == 0) ? v : intf[i + 3 * a] + v;
        }
    }
    for (int a = 0; a < nn; ++a)
      for (int i = 0; i < 3; ++i)
        m->f_ext[3 * (long)conn[f * nn + a] + i] += 1. * intf[i + 3 * a];
  }
}

/* -------------------------------------------------------------------------- */
/* Energies */

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

double orc_math_reduce(double *array, long nb_values) {
  if (nb_values == 0)
    return 0.;
  long nb_values_to_sum = nb_values >> 1;
  qsort(array, (size_t)nb_values, sizeof(double), cmp_double);
  while (nb_values_to_sum != 0) {
    long remaining = nb_values - 2 * nb_values_to_sum;
    if (remaining != 0)
      array[nb_values - 2] += array[nb_values - 1];
    for (long i = 0; i < nb_values_to_sum; ++i)
      array[i] = array[2 * i] + array[2 * i + 1];
    nb_values = nb_values_to_sum;
    nb_values_to_sum >>= 1;
  }
  return array[0];
}

/* solid_mechanics_model.cc:476-522 (lumped mass branch) */
double orc_kinetic_energy(const orc_model *m) {
  double ekin = 0.;
  for (long n = 0; n < m->nn; ++n) {
    double mv2 = 0.;
    for (int i = 0; i < 3; ++i)
      if (m->mass[3 * n + i] > DBL_EPSILON)
        mv2 += m->v[3 * n + i] * m->v[3 * n + i] * m->mass[3 * n + i];
    ekin += mv2;
  }
  return ekin / 2.;
}

/* solid_mechanics_model.cc:545-583 (dynamic: velocities, times dt) */
double orc_external_work(const orc_model *m) {
  double work = 0.;
  for (long n = 0; n < m->nn; ++n)
    for (int i = 0; i < 3; ++i) {
      long d = 3 * n + i;
      if (m->blocked[d])
        work -= m->f_int[d] * m->v[d];
      else
        work += m->f_ext[d] * m->v[d];
    }
  return work * m->dt;
}

static double double_dot(const double *a, const double *b) {
  double s = a[0] * b[0];
  for (int k = 1; k < 9; ++k)
    s += a[k] * b[k];
  return s;
}

/* Potential energy density at one quadrature point:
 *   elastic small strain  .5 sigma : grad_u (material_elastic_inline_impl
 *                         .hh:113-116)
 *   elastic finite        .5 S : E(grad_u) (material_elastic.cc:138-160)
 *   neo-Hookean           lambda (ln J)^2 + mu (-ln J + (tr C - 3) / 2) / 2
 *                         (material_neohookean_inline_impl.hh:119-126)
 *   plastic               .5 sigma : (grad_u - eps_p) (material_plastic.cc
 *                         :71-83) */
static double epot_on_quad(const orc_model *m, long iq, double lambda, double mu) {
  const double *g = m->grad_u + 9 * iq, *S = m->stress + 9 * iq;
  switch (m->law) {
  case ORC_ELASTIC:
    return .5 * double_dot(S, g);
  case ORC_ELASTIC_FD: {
    double E[9];
    grad_u_to_green(g, E);
    return .5 * double_dot(S, E);
  }
  case ORC_NEOHOOKEAN: {
    mat3 F, C;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        F[i][j] = (i == j ? 1. : 0.) + T(g, i, j);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.;
        for (int k = 0; k < 3; ++k)
          s += F[k][i] * F[k][j];
        C[i][j] = s;
      }
    double J = det3_fixed(F);
    double trC = C[0][0] + C[1][1] + C[2][2];
    return lambda * pow(log(J), 2.) + mu * (-log(J) + (trC - 3) / 2.) / 2.;
  }
  default: {
    double el[9];
    for (int k = 0; k < 9; ++k)
      el[k] = g[k] - m->eps_p[9 * iq + k];
    return .5 * double_dot(S, el);
  }
  }
}

/* FEEngine::integrate (integrator_gauss_inline_impl.hh:401-487) of a
 * per-quadrature-point density: per element sum_q f_q wJ_q, then
 * Math::reduce over the elements. */
static double integrate_density(const orc_model *m, int plastic) {
  int nnpe = orc_nnpe(m->type), nqp = orc_nqp(m->type);
  double *per_el = (double *)malloc(sizeof(double) * (size_t)(m->nel > 0 ? m->nel : 1));
  double Bl[8 * 24], wJl[8];
  for (long e = 0; e < m->nel; ++e) {
    const double *wJ;
    if (m->wJ) {
      wJ = m->wJ + e * nqp;
    } else {
      double Xe[24];
      gather_element(m, e, m->X, Xe);
      element_shape_derivatives(m->type, Xe, Bl, wJl);
      wJ = wJl;
    }
    double lambda, mu, sy, h;
    element_params(m, e, &lambda, &mu, &sy, &h);
    double s = 0.;
    for (int q = 0; q < nqp; ++q) {
      long iq = e * nqp + q;
      double f = plastic ? m->plastic_energy[iq] : epot_on_quad(m, iq, lambda, mu);
      s = (q == 0) ? f * wJ[q] : s + f * wJ[q];
    }
    per_el[e] = s;
  }
  double r = orc_math_reduce(per_el, m->nel);
  free(per_el);
  (void)nnpe;
  return r;
}

double orc_potential_energy(const orc_model *m) { return integrate_density(m, 0); }

double orc_plastic_energy(const orc_model *m) {
  if (m->law != ORC_PLASTIC_LIH || !m->plastic_energy)
    return 0.;
  return integrate_density(m, 1);
}

void orc_solve_steps(orc_model *m, long nsteps) {
  for (long s = 0; s < nsteps; ++s)
    orc_solve_step(m);
}
