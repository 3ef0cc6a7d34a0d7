// akx_math.hpp -- per-quadrature-point arithmetic of the explicit internal
// force, written for gfx950 device code (and usable on the host).
//
// The FP64 instantiation reproduces the reference's operation order so that
// the plain path matches the CPU oracle bit for bit, apart from the last bit
// of log() in the neo-Hookean law (device libm vs glibc).  All code here is
// compiled with -ffp-contract=off; the FP32 precision form enables
// contraction in its own kernels.
//
// Tensors are 3x3 arrays t[i][j] (row index i, column index j).
//
// Terms that are exactly zero in the reference's expressions (adding the
// off-diagonal of lambda tr I, the 0 that starts a sum, 0 * mu) are left
// out: that changes at most the sign of a zero result, never a value.
#pragma once

#ifndef __HIPCC_RTC__ // run-time compilation provides these
#include <hip/hip_runtime.h>
#endif

#define AKX_HD __host__ __device__ __forceinline__

// The namespace is a macro so that a translation unit compiled with a
// different floating-point contraction setting (the fast FP64 / FP32 element
// kernels) gets its own copies of these inline functions.
#ifndef AKX_MATH_NS
#define AKX_MATH_NS akx
#endif

namespace AKX_MATH_NS {

template <class R> struct M3 {
  R m[3][3];
};

// ---------------------------------------------------------------------------
// Jacobian inverse and determinant through a partial-pivot LU.  The
// reference's Jacobian is a dynamic-size Eigen matrix
// (element_class_tmpl.hh:327, :345-346, :381-385), so Eigen evaluates
// det and inverse with PartialPivLU: row pivoting on the largest magnitude
// (first one on ties), det = sign * prod(diag U), inverse by a permuted unit
// lower solve followed by an upper solve that scales by the reciprocal
// pivot.  Indices are compile-time after unrolling; row swaps are selects.
// ---------------------------------------------------------------------------
template <class R> AKX_HD R akx_abs(R x) { return x < R(0) ? -x : x; }

template <class R> AKX_HD void swap_rows_if(R (&a)[3][3], int r0, int r1, bool c) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    R x = a[r0][j], y = a[r1][j];
    a[r0][j] = c ? y : x;
    a[r1][j] = c ? x : y;
  }
}

template <class R>
AKX_HD void jacobian_lu(const R (&A)[3][3], R (&lu)[3][3], int (&perm)[3],
                        int &nb_transp) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      lu[i][j] = A[i][j];
  nb_transp = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int p = k;
    R big = akx_abs(lu[k][k]);
#pragma unroll
    for (int i = k + 1; i < 3; ++i) {
      R v = akx_abs(lu[i][k]);
      bool gt = v > big;
      big = gt ? v : big;
      p = gt ? i : p;
    }
    perm[k] = p;
    if (big != R(0)) {
      if (p != k) { // most elements never pivot: skip the selects
#pragma unroll
        for (int i = k + 1; i < 3; ++i)
          swap_rows_if(lu, k, i, p == i);
        nb_transp += 1;
      }
#pragma unroll
      for (int i = k + 1; i < 3; ++i)
        lu[i][k] = lu[i][k] / lu[k][k];
    }
#pragma unroll
    for (int i = k + 1; i < 3; ++i)
#pragma unroll
      for (int j = k + 1; j < 3; ++j)
        lu[i][j] = lu[i][j] - lu[i][k] * lu[k][j];
  }
}

// det J and J^{-1} of the element Jacobian
template <class R>
AKX_HD R jacobian_det_inv(const R (&J)[3][3], R (&Ji)[3][3]) {
  R lu[3][3];
  int perm[3], nt;
  jacobian_lu(J, lu, perm, nt);
  R det = ((nt & 1) ? R(-1) : R(1)) * (lu[0][0] * lu[1][1] * lu[2][2]);
  R x[3][3] = {{R(1), R(0), R(0)}, {R(0), R(1), R(0)}, {R(0), R(0), R(1)}};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (perm[k] != k) {
#pragma unroll
      for (int i = k + 1; i < 3; ++i)
        swap_rows_if(x, k, i, perm[k] == i);
    }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int r = i + 1; r < 3; ++r)
        x[r][j] = x[r][j] - x[i][j] * lu[r][i];
#pragma unroll
    for (int i = 2; i >= 0; --i) {
      R a = R(1) / lu[i][i];
      x[i][j] = x[i][j] * a;
#pragma unroll
      for (int r = 0; r < i; ++r)
        x[r][j] = x[r][j] - x[i][j] * lu[r][i];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Ji[i][j] = x[i][j];
  return det;
}

template <class R> AKX_HD R jacobian_det(const R (&J)[3][3]) {
  R lu[3][3];
  int perm[3], nt;
  jacobian_lu(J, lu, perm, nt);
  return ((nt & 1) ? R(-1) : R(1)) * (lu[0][0] * lu[1][1] * lu[2][2]);
}

// Fixed-size 3x3 determinant and inverse (Eigen's closed forms for
// Matrix<Real,3,3>): used on F and C by the neo-Hookean law
// (material_neohookean_inline_impl.hh:45-48).
template <class R> AKX_HD R det3(const R (&m)[3][3]) {
  R a = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
  R b = m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
  R c = m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  return a - b + c;
}

template <class R> AKX_HD R cof3(const R (&m)[3][3], int i, int j) {
  int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
}

template <class R> AKX_HD void inv3(const R (&m)[3][3], R (&r)[3][3]) {
  R c0 = cof3(m, 0, 0), c1 = cof3(m, 1, 0), c2 = cof3(m, 2, 0);
  R det = c0 * m[0][0] + c1 * m[1][0] + c2 * m[2][0];
  R invdet = R(1) / det;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r[i][j] = cof3(m, j, i) * invdet;
}

// ---------------------------------------------------------------------------
// Lame constants: material_elastic.cc:53-62
template <class R> AKX_HD void lame(R E, R nu, R &lambda, R &mu) {
  lambda = nu * E / ((R(1) + nu) * (R(1) - R(2) * nu));
  mu = E / (R(2) * (R(1) + nu));
}

// Linear elastic law: material_elastic_inline_impl.hh:33-49 with
// eps = .5 (g^T + g) (material_inline_impl.hh:109-113)
template <class R>
AKX_HD void elastic_stress(R lambda, R mu, const R (&g)[3][3], R sigma_th,
                           R (&s)[3][3]) {
  R trace = g[0][0] + g[1][1] + g[2][2];
  R two_mu = R(2) * mu;
  R sd = lambda * trace + sigma_th;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R eps = R(.5) * (g[j][i] + g[i][j]);
      // (lambda tr + sigma_th) * I: the off-diagonal term adds a zero
      s[i][j] = (i == j) ? two_mu * eps + sd : two_mu * eps;
    }
}

// Green strain: material_inline_impl.hh:125-130
template <class R> AKX_HD void green_strain(const R (&g)[3][3], R (&E)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R p = g[0][i] * g[0][j];
#pragma unroll
      for (int k = 1; k < 3; ++k)
        p += g[k][i] * g[k][j];
      E[i][j] = (p + g[j][i] + g[i][j]) / R(2);
    }
}

template <class R> AKX_HD R akx_log(R x);
template <> AKX_HD double akx_log<double>(double x) { return log(x); }
template <> AKX_HD float akx_log<float>(float x) { return logf(x); }
template <class R> AKX_HD R akx_sqrt(R x);
template <> AKX_HD double akx_sqrt<double>(double x) { return sqrt(x); }
template <> AKX_HD float akx_sqrt<float>(float x) { return sqrtf(x); }

// Compressible neo-Hookean PK2: material_neohookean_inline_impl.hh:39-50
template <class R>
AKX_HD void neohookean_pk2(R lambda, R mu, const R (&g)[3][3], R (&S)[3][3]) {
  R F[3][3], C[3][3], Ci[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      F[i][j] = (i == j) ? R(1) + g[i][j] : g[i][j]; // I + grad u
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R s = F[0][i] * F[0][j];
#pragma unroll
      for (int k = 1; k < 3; ++k)
        s += F[k][i] * F[k][j];
      C[i][j] = s;
    }
  R J = det3(F); // * sqrt(C33), C33 = 1 in 3-D
  inv3(C, Ci);
  R t = lambda * akx_log(J) - mu;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) // I * mu + t C^{-1}
      S[i][j] = (i == j) ? mu + t * Ci[i][j] : t * Ci[i][j];
}

// The same neo-Hookean PK2 written without cancellation, for the FP32 form:
// mu (I - C^{-1}) = 2 mu C^{-1} E with the Green strain E formed from grad u
// directly, and ln J = log1p(J - 1) with
// J - 1 = tr g + ((tr g)^2 - tr(g g)) / 2 + det g.
// At strains of 1e-3 the reference's form keeps only ~4 significant digits
// in FP32; this one keeps full FP32 accuracy.
// log1p in FP32: for |x| < 1/16 the Taylor series to x^7 (truncation below
// (1/16)^7 / 8 = 5e-10 relative, under the FP32 rounding of the result), in
// Horner form with FMAs; larger |x| (strains beyond a few percent) take the
// device library's log1pf, which costs about 120 instructions.
template <class R> AKX_HD R akx_log1p(R x);
template <> AKX_HD double akx_log1p<double>(double x) { return log1p(x); }
template <> AKX_HD float akx_log1p<float>(float x) {
  if (fabsf(x) < 0.0625f) {
    float p = fmaf(1.f / 7.f, x, -1.f / 6.f);
    p = fmaf(p, x, 0.2f);
    p = fmaf(p, x, -0.25f);
    p = fmaf(p, x, 1.f / 3.f);
    p = fmaf(p, x, -0.5f);
    return fmaf(p * x, x, x);
  }
  return log1pf(x);
}

// 1 / x: the hardware reciprocal (1 ulp) for FP32 on the device, where the
// correctly rounded division sequence costs ten instructions; exact
// division elsewhere.
template <class R> AKX_HD R akx_rcp(R x) { return R(1) / x; }
template <> AKX_HD float akx_rcp<float>(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.f / x;
#endif
}

// E, C^{-1} and S are symmetric (C^{-1} is a function of E, so C^{-1} E is
// too): each is formed once per pair (i <= j) and mirrored.
template <class R>
AKX_HD void neohookean_pk2_stable(R lambda, R mu, const R (&g)[3][3],
                                  R (&S)[3][3]) {
  constexpr int I6[6] = {0, 1, 2, 0, 0, 1}, J6[6] = {0, 1, 2, 1, 2, 2};
  R E[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    const int i = I6[v], j = J6[v];
    R p = g[0][i] * g[0][j];
#pragma unroll
    for (int k = 1; k < 3; ++k)
      p += g[k][i] * g[k][j];
    E[v] = (p + g[j][i] + g[i][j]) * R(0.5);
  }
  R tr = g[0][0] + g[1][1] + g[2][2];
  R trg2 = R(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      trg2 += g[i][k] * g[k][i];
  R jm1 = tr + R(0.5) * (tr * tr - trg2) + det3(g);
  R lnJ = akx_log1p(jm1);
  // C = I + 2 E (Voigt order 00 11 22 01 02 12) and its inverse by cofactors
  const R c00 = R(1) + R(2) * E[0], c11 = R(1) + R(2) * E[1], c22 = R(1) + R(2) * E[2];
  const R c01 = R(2) * E[3], c02 = R(2) * E[4], c12 = R(2) * E[5];
  R k00 = c11 * c22 - c12 * c12, k11 = c00 * c22 - c02 * c02, k22 = c00 * c11 - c01 * c01;
  R k01 = c02 * c12 - c01 * c22, k02 = c01 * c12 - c02 * c11, k12 = c01 * c02 - c00 * c12;
  const R id = akx_rcp(c00 * k00 + c01 * k01 + c02 * k02);
  const R Ci[3][3] = {{k00 * id, k01 * id, k02 * id},
                      {k01 * id, k11 * id, k12 * id},
                      {k02 * id, k12 * id, k22 * id}};
  // T = 2 mu E + lambda ln J I
  const R m2 = R(2) * mu, ll = lambda * lnJ;
  const R T[3][3] = {{m2 * E[0] + ll, m2 * E[3], m2 * E[4]},
                     {m2 * E[3], m2 * E[1] + ll, m2 * E[5]},
                     {m2 * E[4], m2 * E[5], m2 * E[2] + ll}};
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    const int i = I6[v], j = J6[v];
    R s = Ci[i][0] * T[0][j];
#pragma unroll
    for (int k = 1; k < 3; ++k)
      s += Ci[i][k] * T[k][j];
    S[i][j] = s;
    S[j][i] = s;
  }
}

// Linear isotropic hardening, small strain, radial return:
// material_linear_isotropic_hardening_inline_impl.hh:31-86 and
// material_plastic.hh:65-87.  Reads the previous state, writes the new one.
// dsth: previous sigma_th - sigma_th, added to the trial stress's diagonal
// (:33-47; 0 without a temperature field).
template <class R>
AKX_HD void plastic_lih(R lambda, R mu, R sigma_y, R h, const R (&g)[3][3],
                        const R (&pg)[3][3], const R (&ps)[3][3], R piso,
                        const R (&pe)[3][3], R (&sigma)[3][3], R &iso,
                        R (&eps)[3][3], R dsth = R(0)) {
  R dg[3][3], str[3][3], dev[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      dg[i][j] = g[i][j] - pg[i][j];
  elastic_stress(lambda, mu, dg, dsth, str);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      str[i][j] += ps[i][j];
  R tr = str[0][0] + str[1][1] + str[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      dev[i][j] = (i == j) ? str[i][j] - tr / R(3) : str[i][j];
  // sum of squares in column-major order (cwiseProduct().sum())
  R s = dev[0][0] * dev[0][0];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (i + j > 0)
        s += dev[i][j] * dev[i][j];
  R eff = akx_sqrt(R(3) / R(2) * s);
  bool yielding = ((eff - piso - sigma_y) > R(0));
  R dp = yielding ? (eff - sigma_y - piso) / (R(3) * mu + h) : R(0);
  iso = piso + h * dp;
  R tol = sizeof(R) == 8 ? R(1e2 * 2.220446049250313e-16) : R(1e2 * 1.1920929e-07f);
  bool flow = akx_abs(eff) > akx_sqrt(s) * tol;
  R dge[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R de = flow ? dev[i][j] * R(3) / R(2) * dp / eff : R(0);
      eps[i][j] = pe[i][j] + de;
      dge[i][j] = (g[i][j] - pg[i][j]) - de;
    }
  R sel[3][3];
  elastic_stress(lambda, mu, dge, R(0), sel);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      sigma[i][j] = ps[i][j] + sel[i][j];
}

// The same radial return written on the total strain, for the FP32 form.
// The reference's update is incremental (sigma = sigma_prev + C : (d eps -
// d eps_p)); from a zero initial state it telescopes to
// sigma = C : (eps - eps_p), so the state reduces to the plastic strain eps_p
// (symmetric, Voigt order xx yy zz yz xz xy) and the hardening: the trial
// stress is C : (eps - eps_p_prev), and since d eps_p is deviatoric,
// sigma = trial - 2 mu d eps_p.  7 numbers per Gauss point instead of 28,
// and no accumulation of the stress in FP32.
template <class R>
AKX_HD void plastic_lih_total(R lambda, R mu, R sigma_y, R h, const R (&g)[3][3],
                              const R (&pe)[6], R piso, R (&sigma)[3][3],
                              R (&pe_new)[6], R &iso) {
  const int vi[6] = {0, 1, 2, 1, 0, 0}, vj[6] = {0, 1, 2, 2, 2, 1};
  R e[3][3];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int i = vi[k], j = vj[k];
    const R x = R(.5) * (g[i][j] + g[j][i]) - pe[k];
    e[i][j] = x;
    e[j][i] = x;
  }
  R tr = e[0][0] + e[1][1] + e[2][2];
  R two_mu = R(2) * mu;
  R str[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      str[i][j] = (i == j) ? two_mu * e[i][j] + lambda * tr : two_mu * e[i][j];
  R trs = str[0][0] + str[1][1] + str[2][2];
  R dev[3][3], s = R(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      dev[i][j] = (i == j) ? str[i][j] - trs / R(3) : str[i][j];
      s += dev[i][j] * dev[i][j];
    }
  R eff = akx_sqrt(R(3) / R(2) * s);
  bool yielding = ((eff - piso - sigma_y) > R(0));
  R dp = yielding ? (eff - sigma_y - piso) / (R(3) * mu + h) : R(0);
  iso = piso + h * dp;
  R tol = sizeof(R) == 8 ? R(1e2 * 2.220446049250313e-16) : R(1e2 * 1.1920929e-07f);
  bool flow = akx_abs(eff) > akx_sqrt(s) * tol;
  R c = flow ? R(3) / R(2) * dp / eff : R(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      sigma[i][j] = str[i][j] - two_mu * (c * dev[i][j]);
#pragma unroll
  for (int k = 0; k < 6; ++k)
    pe_new[k] = pe[k] + c * dev[vi[k]][vj[k]];
}

// ---------------------------------------------------------------------------
// Nodal update of one free DOF component (blocked components are not
// touched).  Reference order, the plain FP64 path: newmark-beta.cc:47-77
// (predictor) and :164-197 (corrector) with alpha = 0, beta = 1/2, the
// residual of solid_mechanics_model.cc:235-249 and the lumped solve of
// non_linear_solver_lumped.cc:76-89:
//   gather     r = (0 + f_ext) + f_int + M (0 + -a); x = r / M;
//              a += x; v += (dt / 2) x
//   predictor  u += dt v + (dt / 2)(dt a); v += dt a
// FP32 form: the same central-difference scheme with compensated FP64
// state (FP32 compensation terms).  Between the steps of one solve call the
// velocity holds the half-step velocity w = v + dt/2 a, so the acceleration
// is neither read nor stored until the call's last step:
//   predictor  w = v + dt/2 a; u += dt w
//   gather     a = (f_ext + f_int) / M (the reference's a + x without the
//              cancelling terms); v = w + dt/2 a; then, when more steps
//              follow in the call, w = v + dt/2 a and u += dt w
// The predictor at the start of a call and the gather that ends the call
// before it perform the same two compensated additions as a step inside a
// call, so splitting a run of steps over calls keeps every bit.
// ---------------------------------------------------------------------------
AKX_HD void kahan_add(double &s, float &c, double inc) {
  double y = inc - double(c);
  double t = s + y;
  c = float((t - s) - y);
  s = t;
}

AKX_HD void reference_predict(double a, double dt, double &v, double &u) {
  const double dt_a = dt * a;
  u = u + (dt * v + (.5 * dt) * dt_a);
  v = v + dt_a;
}

template <bool PREDICT>
AKX_HD void reference_gather(double f_ext, double f_int, double M, double dt, double &a,
                             double &v, double &u) {
  double r = 0.;
  r += 1. * f_ext;
  r += 1. * f_int;
  double cache = 0.;
  cache += -1. * a;
  r += M * cache;
  const double x = 1. * (r / M);
  a += 1. * x;
  v += (.5 * dt) * x;
  if constexpr (PREDICT)
    reference_predict(a, dt, v, u);
}

AKX_HD void half_step_predict(double a, double dt, double &v, float &cv, double &u,
                              float &cu) {
  kahan_add(v, cv, (.5 * dt) * a);
  kahan_add(u, cu, dt * v);
}

// w, cw: the half-step velocity on entry; the velocity (PREDICT = false)
// or the next half-step velocity on return.  a: written.
template <bool PREDICT>
AKX_HD void half_step_gather(double f, double M, double dt, double &w, float &cw, double &u,
                             float &cu, double &a) {
  a = f / M;
  kahan_add(w, cw, (.5 * dt) * a);
  if constexpr (PREDICT) {
    kahan_add(w, cw, (.5 * dt) * a);
    kahan_add(u, cu, dt * w);
  }
}

} // namespace AKX_MATH_NS
