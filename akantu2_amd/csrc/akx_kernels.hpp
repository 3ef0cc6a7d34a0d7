// akx_kernels.hpp -- gfx950 kernels of the explicit internal-force step.
//
// Device layout (structure of arrays, all coalesced on the fast index):
//   nodal fields     X, u, v, a, f_int, f_ext : [3][nn] doubles
//   lumped mass      M                         : [nn] doubles (the three
//                                                components are equal)
//   blocked DOFs     bits 0..2 of a byte       : [nn]
//   connectivity                               : [nel][nnpe] int32 (16 B loads)
//   elemental force  r_e of element position p, local node a, is stored at
//                    entry pos[p*nnpe + a] of a node-major sliced-ELL table:
//                    nodes are cut in slices of 256; entry (node, d) of
//                    slice s is ell_off[s] + d*256 + (node % 256), where d is
//                    the rank of the element in the node's list in ascending
//                    assembly order (the reference's sequential scatter
//                    order, dof_manager.cc:92-114).  The gather then reads
//                    slot d of 256 consecutive nodes as one coalesced block.
#pragma once

#include "akx_args.hpp"
#include "akx_math.hpp"
#ifndef __HIPCC_RTC__ // run-time compilation provides these
#include <stdint.h>
#endif

#ifndef __HIPCC_RTC__ // run-time compilation provides these
#include <cmath>
#endif

namespace akx {

constexpr int ET_TET4 = 0;
constexpr int ET_HEX8 = 1;
constexpr int LAW_ELASTIC = 0;
constexpr int LAW_ELASTIC_FD = 1;
constexpr int LAW_NEOHOOKEAN = 2;
constexpr int LAW_PLASTIC = 3;
// A law computed by the caller (Material::computeStress plugin), in two
// passes: LAW_USER_GRAD forms grad u at every Gauss point and stores it (no
// forces); after the caller filled the stress array, LAW_USER (sigma, small
// strain) or LAW_USER_FD (PK2, finite deformation) reads it and forms the
// forces.
constexpr int LAW_USER_GRAD = 4;
constexpr int LAW_USER = 5;
constexpr int LAW_USER_FD = 6;
// a user law compiled with the element pass at run time (akx_model_set_device_
// stress): the program defines akx_compute_stress and akx_law_params and
// AKX_DEVICE_LAW before including this header; sigma (LAW_DEVICE) or PK2
// (LAW_DEVICE_FD) of one Gauss point from grad u, column-major
constexpr int LAW_DEVICE = 7;
constexpr int LAW_DEVICE_FD = 8;
template <int LAW> constexpr bool finite_law() {
  return LAW == LAW_ELASTIC_FD || LAW == LAW_NEOHOOKEAN || LAW == LAW_USER_FD ||
         LAW == LAW_DEVICE_FD;
}

template <int ET> struct Elem;
template <> struct Elem<ET_TET4> {
  static constexpr int nnpe = 4, nqp = 1;
};
template <> struct Elem<ET_HEX8> {
  static constexpr int nnpe = 8, nqp = 8;
};

// Hex8 natural derivatives, shapes and weights at the 8 Gauss points,
// computed on the host with the reference's formulas
// (element_class_hexahedron_8_inline_impl.hh:111-188,
// gauss_integration_tmpl.hh:136-181) and uploaded once.
struct Hex8Tables {
  double dnds[8][3][8];
  double N[8][8];
  double w[8];
  float dnds_f[8][3][8];
};
extern __constant__ Hex8Tables c_hex8;

// Hex8 natural derivatives, shapes and weights at the tensor-product Gauss
// points +-1/sqrt(3), xi fastest (element_class_hexahedron_8_inline_impl.hh
// :111-188, gauss_integration.cc:42-43); host code, shared by the
// single-model and batched engines so both upload the same bits.
#ifndef __HIPCC_RTC__ // host code
inline void fill_hex8_tables(Hex8Tables &t) {
  double pos[2] = {-1. / std::sqrt(3.), 1. / std::sqrt(3.)};
  for (int q = 0; q < 8; ++q) {
    double c[3] = {pos[q & 1], pos[(q >> 1) & 1], pos[(q >> 2) & 1]};
    double(*d)[8] = t.dnds[q];
    d[0][0] = -.125 * (1 - c[1]) * (1 - c[2]);
    d[0][1] = .125 * (1 - c[1]) * (1 - c[2]);
    d[0][2] = .125 * (1 + c[1]) * (1 - c[2]);
    d[0][3] = -.125 * (1 + c[1]) * (1 - c[2]);
    d[0][4] = -.125 * (1 - c[1]) * (1 + c[2]);
    d[0][5] = .125 * (1 - c[1]) * (1 + c[2]);
    d[0][6] = .125 * (1 + c[1]) * (1 + c[2]);
    d[0][7] = -.125 * (1 + c[1]) * (1 + c[2]);
    d[1][0] = -.125 * (1 - c[0]) * (1 - c[2]);
    d[1][1] = -.125 * (1 + c[0]) * (1 - c[2]);
    d[1][2] = .125 * (1 + c[0]) * (1 - c[2]);
    d[1][3] = .125 * (1 - c[0]) * (1 - c[2]);
    d[1][4] = -.125 * (1 - c[0]) * (1 + c[2]);
    d[1][5] = -.125 * (1 + c[0]) * (1 + c[2]);
    d[1][6] = .125 * (1 + c[0]) * (1 + c[2]);
    d[1][7] = .125 * (1 - c[0]) * (1 + c[2]);
    d[2][0] = -.125 * (1 - c[0]) * (1 - c[1]);
    d[2][1] = -.125 * (1 + c[0]) * (1 - c[1]);
    d[2][2] = -.125 * (1 + c[0]) * (1 + c[1]);
    d[2][3] = -.125 * (1 - c[0]) * (1 + c[1]);
    d[2][4] = .125 * (1 - c[0]) * (1 - c[1]);
    d[2][5] = .125 * (1 + c[0]) * (1 - c[1]);
    d[2][6] = .125 * (1 + c[0]) * (1 + c[1]);
    d[2][7] = .125 * (1 - c[0]) * (1 + c[1]);
    t.N[q][0] = .125 * (1 - c[0]) * (1 - c[1]) * (1 - c[2]);
    t.N[q][1] = .125 * (1 + c[0]) * (1 - c[1]) * (1 - c[2]);
    t.N[q][2] = .125 * (1 + c[0]) * (1 + c[1]) * (1 - c[2]);
    t.N[q][3] = .125 * (1 - c[0]) * (1 + c[1]) * (1 - c[2]);
    t.N[q][4] = .125 * (1 - c[0]) * (1 - c[1]) * (1 + c[2]);
    t.N[q][5] = .125 * (1 + c[0]) * (1 - c[1]) * (1 + c[2]);
    t.N[q][6] = .125 * (1 + c[0]) * (1 + c[1]) * (1 + c[2]);
    t.N[q][7] = .125 * (1 - c[0]) * (1 + c[1]) * (1 + c[2]);
    double wq = 1.;
    wq *= 1.;
    wq *= 1.;
    wq *= 1.;
    t.w[q] = wq;
    for (int i = 0; i < 3; ++i)
      for (int n = 0; n < 8; ++n)
        t.dnds_f[q][i][n] = (float)d[i][n];
  }
}
#endif

// blockIdx -> logical block so that each of the 8 XCDs walks a contiguous
// range of elements/nodes (blocks are dealt round-robin to the XCDs; the
// nodes shared by neighbouring elements then stay in one XCD's L2).
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks_padded) {
  int64_t b = blockIdx.x;
  int64_t per = nblocks_padded >> 3;
  return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------------------
// Jacobian at one quadrature point and the columns of dN/dX.
// Reference: shape_lagrange_inline_impl.hh:61-83 (J = dnds X^T,
// dN/dX = J^{-1} dnds), element_class_tmpl.hh:297-346, integrator_gauss
// multiplyJacobiansByWeights (integrator_gauss_inline_impl.hh:379-398).
// X is given per node, X[n][j].  jacobian_at returns w*detJ and J^{-1};
// bcol gives column n of dN/dX, b[i] = dN_n/dX_i, recomputed where it is
// needed instead of being kept (the bits are the same every time).
// ---------------------------------------------------------------------------
template <int ET, class R> struct Shape;

template <> struct Shape<ET_TET4, double> {
  __device__ static __forceinline__ double
  jacobian_at(const double (&X)[4][3], int, double (&Ji)[3][3], double &detJ) {
    double J[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      J[0][j] = X[3][j] - X[0][j]; // d/dxi   : nodes (-1, 0, 0, 1)
      J[1][j] = X[1][j] - X[0][j]; // d/deta  : nodes (-1, 1, 0, 0)
      J[2][j] = X[2][j] - X[0][j]; // d/dzeta : nodes (-1, 0, 1, 0)
    }
    detJ = jacobian_det_inv(J, Ji);
    return detJ * (1. / 6.);
  }
  __device__ static __forceinline__ void bcol(const double (&Ji)[3][3], int,
                                              int n, double (&b)[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
      b[i] = n == 0 ? -Ji[i][0] + -Ji[i][1] + -Ji[i][2]
                    : (n == 1 ? Ji[i][1] : (n == 2 ? Ji[i][2] : Ji[i][0]));
  }
};

template <> struct Shape<ET_HEX8, double> {
  __device__ static __forceinline__ double
  jacobian_at(const double (&X)[8][3], int q, double (&Ji)[3][3], double &detJ) {
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = c_hex8.dnds[q][i][0] * X[0][j];
#pragma unroll
        for (int n = 1; n < 8; ++n)
          s += c_hex8.dnds[q][i][n] * X[n][j];
        J[i][j] = s;
      }
    detJ = jacobian_det_inv(J, Ji);
    return detJ * c_hex8.w[q];
  }
  __device__ static __forceinline__ void bcol(const double (&Ji)[3][3], int q,
                                              int n, double (&b)[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = Ji[i][0] * c_hex8.dnds[q][0][n];
      s += Ji[i][1] * c_hex8.dnds[q][1][n];
      s += Ji[i][2] * c_hex8.dnds[q][2][n];
      b[i] = s;
    }
  }
};

// ---------------------------------------------------------------------------
// Inradius: Tet4 2 * Math::tetrahedron_inradius (aka_math_tmpl.hh:71-124),
// Hex8 min of the 12 edges (element_class_hexahedron_8_inline_impl.hh:194-212)
__device__ __forceinline__ double dist3(const double (&a)[3], const double (&b)[3]) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ double tet_volume(const double (&c1)[3],
                                             const double (&c2)[3],
                                             const double (&c3)[3],
                                             const double (&c4)[3]) {
  double xx[3][3];
  auto setcol = [&](int k, const double (&c)[3]) {
    xx[0][k] = c[0];
    xx[1][k] = c[1];
    xx[2][k] = c[2];
  };
  setcol(0, c2);
  setcol(1, c3);
  setcol(2, c4);
  double vol = det3(xx);
  setcol(0, c1);
  vol -= det3(xx);
  setcol(1, c2);
  vol += det3(xx);
  setcol(2, c3);
  vol -= det3(xx);
  vol /= 6;
  return vol;
}

template <int ET>
__device__ __forceinline__ double inradius(const double (&X)[Elem<ET>::nnpe][3]) {
  if constexpr (ET == ET_TET4) {
    double l12 = dist3(X[0], X[1]), l13 = dist3(X[0], X[2]), l14 = dist3(X[0], X[3]);
    double l23 = dist3(X[1], X[2]), l24 = dist3(X[1], X[3]), l34 = dist3(X[2], X[3]);
    double s1 = (l12 + l23 + l13) * 0.5;
    s1 = sqrt(s1 * (s1 - l12) * (s1 - l23) * (s1 - l13));
    double s2 = (l12 + l24 + l14) * 0.5;
    s2 = sqrt(s2 * (s2 - l12) * (s2 - l24) * (s2 - l14));
    double s3 = (l23 + l34 + l24) * 0.5;
    s3 = sqrt(s3 * (s3 - l23) * (s3 - l34) * (s3 - l24));
    double s4 = (l13 + l34 + l14) * 0.5;
    s4 = sqrt(s4 * (s4 - l13) * (s4 - l34) * (s4 - l14));
    double volume = tet_volume(X[0], X[1], X[2], X[3]);
    return 2. * (3 * volume / (s1 + s2 + s3 + s4));
  } else {
    const int ed[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                           {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    double p = dist3(X[0], X[1]);
#pragma unroll
    for (int k = 1; k < 12; ++k) {
      double l = dist3(X[ed[k][0]], X[ed[k][1]]);
      p = (l < p) ? l : p;
    }
    return p;
  }
}

// ---------------------------------------------------------------------------
// Force of one quadrature point, r_q[n][k] (before the w*detJ factor).
// Small strain: B^T sigma (shape_lagrange_inline_impl.hh:373-410).
// Finite deformation: (B_L + B_NL(grad u))^T voigt(S) (material.cc:525-584,
// aka_voigthelper_tmpl.hh:99-140,180-207), written out term by term in the
// Voigt order xx yy zz yz xz xy (aka_voigthelper.hh:57-70).
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void btd_small(const R (&S)[3][3], const R (&b)[3],
                                          R (&r)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R s = S[i][0] * b[0];
    s += S[i][1] * b[1];
    s += S[i][2] * b[2];
    r[i] = s;
  }
}

// one component k of (B_L + B_NL)^T voigt(S) for node column b; entries of
// B_L that are zero are not added (they only affect the sign of zeros)
template <int k, class R>
__device__ __forceinline__ R btd_finite_k(const R (&g)[3][3], const R (&S)[3][3],
                                          const R (&b)[3]) {
  // rows 0..2: B_L(I, 3a+k) = [k == I] B(I, a); B_NL = g(k, I) B(I, a)
  R t0 = g[k][0] * b[0];
  R t1 = g[k][1] * b[1];
  R t2 = g[k][2] * b[2];
  if constexpr (k == 0) t0 = b[0] + t0;
  if constexpr (k == 1) t1 = b[1] + t1;
  if constexpr (k == 2) t2 = b[2] + t2;
  // rows 3 (yz), 4 (xz), 5 (xy): B_NL sums two products, B_L one entry
  R n3 = g[k][1] * b[2] + g[k][2] * b[1];
  R n4 = g[k][0] * b[2] + g[k][2] * b[0];
  R n5 = g[k][0] * b[1] + g[k][1] * b[0];
  if constexpr (k == 1) n3 = b[2] + n3;
  if constexpr (k == 2) n3 = b[1] + n3;
  if constexpr (k == 0) n4 = b[2] + n4;
  if constexpr (k == 2) n4 = b[0] + n4;
  if constexpr (k == 0) n5 = b[1] + n5;
  if constexpr (k == 1) n5 = b[0] + n5;
  R s = t0 * S[0][0];
  s += t1 * S[1][1];
  s += t2 * S[2][2];
  s += n3 * S[1][2];
  s += n4 * S[0][2];
  s += n5 * S[0][1];
  return s;
}

template <class R>
__device__ __forceinline__ void btd_finite(const R (&g)[3][3], const R (&S)[3][3],
                                           const R (&b)[3], R (&r)[3]) {
  r[0] = btd_finite_k<0>(g, S, b);
  r[1] = btd_finite_k<1>(g, S, b);
  r[2] = btd_finite_k<2>(g, S, b);
}

// afterSolveStep -> Material::updateEnergies of the plastic law
// (material.cc:740-750, material_plastic.cc:86-99):
// pe + (sigma + sigma_prev) : (eps_p - eps_p_prev) / 2, the double dot summed
// in column-major order (cwiseProduct().sum()).
template <class CT>
__device__ __forceinline__ CT plastic_energy_update(CT pe_prev, const CT (&S)[3][3],
                                                    const CT (&ps)[3][3],
                                                    const CT (&eps)[3][3],
                                                    const CT (&peps)[3][3]) {
  CT dd = (S[0][0] + ps[0][0]) * (eps[0][0] - peps[0][0]);
#pragma unroll
  for (int jj = 0; jj < 3; ++jj)
#pragma unroll
    for (int ii = 0; ii < 3; ++ii)
      if (ii + jj > 0)
        dd += (S[ii][jj] + ps[ii][jj]) * (eps[ii][jj] - peps[ii][jj]);
  return pe_prev + dd / CT(2);
}

// Constitutive law of one Gauss point in the reference's operation order
// (material_elastic_inline_impl.hh:31-73, material_neohookean_inline_impl.hh
// :39-50, material_linear_isotropic_hardening_inline_impl.hh:31-86), with the
// history and output traffic of Material::computeAllStresses.
// Material::computeAllStresses's gradu -= eigengradu (material.cc:147) and
// MaterialThermal::computeStress's sigma_th (material_thermal.hh:91-95,
// computed per Gauss point at set-up from alpha and delta_T), for the
// element passes instantiated with them (EXT).  The plastic law also reads
// the previous sigma_th and records the current one (its history,
// material_thermal.cc:34).
template <int LAW, class CT, class HT>
__device__ __forceinline__ void ext_terms(const QpViewT<HT> &V, int64_t iq, CT (&g)[3][3],
                                          CT &sth, CT &dsth) {
  if (V.eig) {
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
        g[ii][jj] = g[ii][jj] - CT(V.eig[(int64_t)(ii + 3 * jj) * V.nq_all + iq]);
  }
  if (V.sth_now) {
    sth = CT(V.sth_now[iq]);
    if constexpr (LAW == LAW_PLASTIC) {
      dsth = CT(V.sth_p[iq]) - sth;
      V.sth_c[iq] = double(sth);
    }
  }
}

// sth: MaterialThermal's sigma_th (-0 with the default alpha = 0,
// material_thermal.hh:91-95), added by the elastic laws; dsth = previous -
// current sigma_th, added to the plastic trial stress.
template <int LAW, class CT, class HT>
__device__ __forceinline__ void strict_law(const QpViewT<HT> &V, int64_t iq, CT lambda,
                                           CT mu, CT sy, CT hh, const CT (&g)[3][3],
                                           CT (&S)[3][3], CT sth = CT(-0.0),
                                           CT dsth = CT(0)) {
  if constexpr (LAW == LAW_ELASTIC) {
    elastic_stress(lambda, mu, g, sth, S);
  } else if constexpr (LAW == LAW_ELASTIC_FD) {
    CT E[3][3];
    green_strain(g, E);
    elastic_stress(lambda, mu, E, sth, S);
  } else if constexpr (LAW == LAW_NEOHOOKEAN) {
    neohookean_pk2(lambda, mu, g, S);
  } else if constexpr (LAW == LAW_USER_GRAD) {
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii) {
        V.grad_u[(int64_t)(ii + 3 * jj) * V.nq_all + iq] = double(g[ii][jj]);
        S[ii][jj] = CT(0);
      }
    return;
  } else if constexpr (LAW == LAW_USER || LAW == LAW_USER_FD) {
    // the stress the caller's computeStress wrote for this Gauss point
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
        S[ii][jj] = CT(V.stress[(int64_t)(ii + 3 * jj) * V.nq_all + iq]);
    return;
#ifdef AKX_DEVICE_LAW
  } else if constexpr (LAW == LAW_DEVICE || LAW == LAW_DEVICE_FD) {
    // the caller's law, inline: grad u and the stress column-major, as the
    // host callback sees them (the stress is zero on entry)
    double gc[9], sc[9];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii) {
        gc[ii + 3 * jj] = double(g[ii][jj]);
        sc[ii + 3 * jj] = 0.;
      }
    akx_compute_stress(gc, sc, akx_law_params);
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
        S[ii][jj] = CT(sc[ii + 3 * jj]);
#endif
  } else {
    CT pg[3][3], ps[3][3], pe[3][3], eps[3][3], iso;
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii) {
        const int64_t o = (int64_t)(ii + 3 * jj) * V.nq_all + iq;
        pg[ii][jj] = CT(V.p_grad[o]);
        ps[ii][jj] = CT(V.p_sig[o]);
        pe[ii][jj] = CT(V.p_eps[o]);
      }
    plastic_lih(lambda, mu, sy, hh, g, pg, ps, CT(V.p_iso[iq]), pe, S, iso, eps, dsth);
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii) {
        const int64_t o = (int64_t)(ii + 3 * jj) * V.nq_all + iq;
        V.c_grad[o] = HT(g[ii][jj]);
        V.c_sig[o] = HT(S[ii][jj]);
        V.c_eps[o] = HT(eps[ii][jj]);
      }
    V.c_iso[iq] = HT(iso);
    if (V.c_pe)
      V.c_pe[iq] = HT(plastic_energy_update(CT(V.p_pe[iq]), S, ps, eps, pe));
  }
  if (V.grad_u) {
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii) {
        const int64_t o = (int64_t)(ii + 3 * jj) * V.nq_all + iq;
        V.grad_u[o] = double(g[ii][jj]);
        V.stress[o] = double(S[ii][jj]);
      }
  }
}

// ---------------------------------------------------------------------------
// Element pass, one thread per element: every Gauss point of element e
// (nodes conn_e, table entries pos_e, Gauss points iq0 ..), per-node sums
// over the Gauss points in ascending order.  Shared by the single-model
// launch (one material) and the batched persistent launch.
// ---------------------------------------------------------------------------
template <int ET, int LAW, class CT, class FT, class HT, class NA, int W = 3, bool EXT = false,
          bool SLOTS = false>
__device__ __forceinline__ void element_forces_one(
    const int32_t *conn_e, const NA &nodes_of,
    CT lambda, CT mu, CT sy, CT hh, const QpViewT<HT> &V, int64_t iq0,
    const int32_t *pos_e, void *fe_out, double *slots = nullptr) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  int32_t nodes[NN];
  {
    const int4 *c4 = reinterpret_cast<const int4 *>(conn_e);
#pragma unroll
    for (int k = 0; k < NN / 4; ++k) {
      int4 v = c4[k];
      nodes[4 * k] = v.x;
      nodes[4 * k + 1] = v.y;
      nodes[4 * k + 2] = v.z;
      nodes[4 * k + 3] = v.w;
    }
  }

  CT X[NN][3], U[NN][3];
#pragma unroll
  for (int n = 0; n < NN; ++n) {
    double x[3], d[3];
    nodes_of.load(nodes[n], x, d);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      X[n][c] = CT(x[c]);
      U[n][c] = CT(d[c]);
    }
  }

  CT re[NN][3];

#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    CT Ji[3][3], detJ;
    const CT wJ = Shape<ET, CT>::jacobian_at(X, q, Ji, detJ);
    // grad u (i, j) = sum_n u(i, n) B(j, n), summed over n in order
    CT g[3][3];
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      CT b[3];
      Shape<ET, CT>::bcol(Ji, q, n, b);
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
          g[ii][jj] = (n == 0) ? U[n][ii] * b[jj] : g[ii][jj] + U[n][ii] * b[jj];
    }
    CT S[3][3];
    if constexpr (EXT) {
      CT sth = CT(-0.0), dsth = CT(0);
      ext_terms<LAW, CT, HT>(V, iq0 + q, g, sth, dsth);
      strict_law<LAW, CT, HT>(V, iq0 + q, lambda, mu, sy, hh, g, S, sth, dsth);
    } else {
      strict_law<LAW, CT, HT>(V, iq0 + q, lambda, mu, sy, hh, g, S);
    }
    if constexpr (LAW == LAW_USER_GRAD)
      continue;
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      CT r[3], b[3];
      Shape<ET, CT>::bcol(Ji, q, n, b);
      if constexpr (finite_law<LAW>())
        btd_finite(g, S, b, r);
      else
        btd_small(S, b, r);
#pragma unroll
      for (int k = 0; k < 3; ++k)
        re[n][k] = (q == 0) ? r[k] * wJ : re[n][k] + r[k] * wJ;
    }
  }

  if constexpr (LAW == LAW_USER_GRAD)
    return; // the gradient pass forms no forces
  FT *fe = reinterpret_cast<FT *>(fe_out);
  const int4 *p4 = reinterpret_cast<const int4 *>(pos_e);
#pragma unroll
  for (int kk = 0; kk < NN / 4; ++kk) {
    const int4 pv = p4[kk];
    const int32_t pn[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if constexpr (SLOTS) { // a corner of its node's prefix: to its LDS slot
        if (pn[j] < 0) {
          double *o = slots + 3 * (-1 - pn[j]);
          o[0] = double(re[4 * kk + j][0]);
          o[1] = double(re[4 * kk + j][1]);
          o[2] = double(re[4 * kk + j][2]);
          continue;
        }
      }
      put_entry<W>(fe, (int64_t)pn[j], FT(re[4 * kk + j][0]), FT(re[4 * kk + j][1]),
                   FT(re[4 * kk + j][2]));
    }
  }
}

template <int ET, int LAW, class CT, class FT, bool EXT = false>
__global__ void __launch_bounds__(256)
k_element_forces(ElemArgs A) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  const int64_t i = xcd_block(A.nblocks_padded) * blockDim.x + threadIdx.x;
  if (i >= A.n)
    return;
  const int64_t e = A.filter ? (int64_t)A.filter[i] : A.first + i;
  CT lambda, mu, sy = CT(A.sigma_y), hh = CT(A.h);
  if (A.lam_el) {
    lambda = CT(A.lam_el[i]);
    mu = CT(A.mu_el[i]);
  } else {
    lambda = CT(A.lambda);
    mu = CT(A.mu);
  }
  if (A.sy_el)
    sy = CT(A.sy_el[i]);
  if (A.h_el)
    hh = CT(A.h_el[i]);
  if constexpr (ET == ET_TET4)
    element_forces_one<ET, LAW, CT, FT, double, NodesXU, 3, EXT>(
        A.conn + e * NN, NodesXU{A.xu}, lambda, mu, sy, hh, qp_view(A, NQ), i * NQ,
        A.pos + (A.pos_offset + i) * NN, A.fe);
  else
    element_forces_one<ET, LAW, CT, FT, double, NodesSoA, 3, EXT>(
        A.conn + e * NN, NodesSoA{A.X, A.u, A.nn}, lambda, mu, sy, hh, qp_view(A, NQ),
        i * NQ, A.pos + (A.pos_offset + i) * NN, A.fe);
}


// Tet4 in the reference's operation order with each 256-element block's node
// records staged in LDS (the block lists of initFull): the same arithmetic
// as k_element_forces on the same values, read from LDS instead of once per
// element corner from global memory.
template <int LAW, class CT, class FT, bool EXT = false>
__global__ void __launch_bounds__(256) k_tet4_strict_lds(ElemArgs A) {
  extern __shared__ double2 s_rec[]; // [bl_maxu][3]
  const int64_t blk = xcd_block(A.nblocks_padded);
  if (blk * 256 >= A.n) // padding block: no elements, no node list
    return;
  const int64_t i = blk * 256 + threadIdx.x;
  const bool active = i < A.n;
  // The element's own streams are loaded first, their round trip overlapping
  // the staging; lanes past the end read the last element (no branch, so no
  // load is awaited before the staging's loads are issued).
  const int64_t ic = active ? i : A.n - 1;
  const ushort4 l = *reinterpret_cast<const ushort4 *>(A.lconn + ic * 4);
  const int4 pv = *reinterpret_cast<const int4 *>(A.pos + (A.pos_offset + ic) * 4);
  CT lambda = CT(A.lambda), mu = CT(A.mu), sy = CT(A.sigma_y), hh = CT(A.h);
  if (A.lam_el) {
    lambda = CT(A.lam_el[ic]);
    mu = CT(A.mu_el[ic]);
  }
  if (A.sy_el)
    sy = CT(A.sy_el[ic]);
  if (A.h_el)
    hh = CT(A.h_el[ic]);
  stage_block_records(A, blk, s_rec);
  __syncthreads();
  if (!active)
    return;
  const int4 lc = make_int4(l.x, l.y, l.z, l.w);
  element_forces_one<ET_TET4, LAW, CT, FT, double, NodesLDS, 3, EXT>(
      reinterpret_cast<const int32_t *>(&lc), NodesLDS{reinterpret_cast<const double *>(s_rec)},
      lambda, mu, sy, hh, qp_view(A, 1), i, reinterpret_cast<const int32_t *>(&pv), A.fe);
}

// Tet4 in the reference's operation order with partial-sum assembly.  The
// gather sums node n's entries in order, f = (0 + -1 e_1) + -1 e_2 + ...
// The corners of n in one 256-element block form a run of that order; a run
// can be summed here, in LDS, from 0 and in the same order, into one entry P.
// Rounding is symmetric, so for the node's first run 0 + -1 P is the
// gather's own partial sum bit for bit (P = 0 included: both are +0):
// summing each node's first run (prefix partials) keeps f_int's bits with
// fewer entries.  Summing every run (block partials, opt-in) leaves one
// entry per (block, node) and rounds otherwise.  The LDS first holds the
// block's node records; once every lane has read its corners' records into
// registers it holds the summed runs' slots.
constexpr int PPA_TB = 256;

template <int LAW, bool EXT = false>
__global__ void __launch_bounds__(PPA_TB) k_tet4_ppa(ElemArgs A) {
  extern __shared__ double2 s_rec[]; // [bl_maxu][3] records, later [t_slots][3] slots
  const int64_t blk = xcd_block(A.nblocks_padded);
  if (blk >= A.t_nblk) // padding block: no elements, no node list
    return;
  const int64_t i = blk * PPA_TB + threadIdx.x;
  const bool active = i < A.n;
  // the element's own streams first (lanes past the end read the last one)
  const int64_t ic = active ? i : A.n - 1;
  const ushort4 l = *reinterpret_cast<const ushort4 *>(A.lconn + ic * 4);
  const int4 pv = *reinterpret_cast<const int4 *>(A.pcode + ic * 4);
  double lambda = A.lambda, mu = A.mu, sy = A.sigma_y, hh = A.h;
  if (A.lam_el) {
    lambda = A.lam_el[ic];
    mu = A.mu_el[ic];
  }
  if (A.sy_el)
    sy = A.sy_el[ic];
  if (A.h_el)
    hh = A.h_el[ic];
  // this lane's first run of the summing phase, read ahead so that its
  // loads overlap the staging instead of following the barriers
  const int32_t k0 = A.poff[blk], k1 = A.poff[blk + 1];
  const int32_t k_first = k0 + (int32_t)threadIdx.x;
  int s0_first = 0, s1_first = 0;
  int32_t pos_first = 0;
  if (k_first < k1) {
    s0_first = k_first == k0 ? 0 : A.pend[k_first - 1];
    s1_first = A.pend[k_first];
    pos_first = A.ppos[k_first];
  }
  stage_block_records(A, blk, s_rec);
  __syncthreads();
  NodesRegs4 rec;
  {
    const NodesLDS lds{reinterpret_cast<const double *>(s_rec)};
    const uint16_t lcs[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double x[3], d[3];
      lds.load(lcs[a], x, d);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rec.r[a][c] = x[c];
        rec.r[a][3 + c] = d[c];
      }
    }
  }
  __syncthreads(); // the records are dead: the LDS now takes the slots
  double *s_slot = reinterpret_cast<double *>(s_rec);
  if (active) {
    const int4 corners = make_int4(0, 1, 2, 3);
    element_forces_one<ET_TET4, LAW, double, double, double, NodesRegs4, 3, EXT, true>(
        reinterpret_cast<const int32_t *>(&corners), rec, lambda, mu, sy, hh, qp_view(A, 1),
        i, reinterpret_cast<const int32_t *>(&pv), A.fe, s_slot);
  }
  __syncthreads();
  double *fe = reinterpret_cast<double *>(A.fe);
  for (int32_t k = k_first; k < k1; k += blockDim.x) {
    const bool first = k == k_first;
    const int s0 = first ? s0_first : A.pend[k - 1], s1 = first ? s1_first : A.pend[k];
    const int32_t pos = first ? pos_first : A.ppos[k];
    double f0 = 0., f1 = 0., f2 = 0.;
    for (int sl = s0; sl < s1; ++sl) {
      f0 = f0 + s_slot[3 * sl];
      f1 = f1 + s_slot[3 * sl + 1];
      f2 = f2 + s_slot[3 * sl + 2];
    }
    put_entry<3>(fe, (int64_t)pos, f0, f1, f2);
  }
}

// ---------------------------------------------------------------------------
// Hex8 element pass in the reference's operation order, one lane per Gauss
// point (8 lanes per element, EPB elements per block).  Nodal values are
// staged in LDS.  Each lane forms J and its LU inverse from the staged
// coordinates (PRE = false, batches) or reads J^-1 and w*detJ from the
// table init fills (PRE = true, single model; only u is staged), then the
// dN/dX columns, grad u, the law and the 24 force terms of its Gauss point
// times w*detJ; the per-node sums over the Gauss points are then taken in
// ascending Gauss-point order (the oracle's and Eigen's GEMV order), two
// nodes at a time, one component per lane.  The force terms reuse the LDS
// of the nodal values, which are dead once grad u is formed: 4.6 KB of LDS
// per 8 elements, so registers, not LDS, bound the waves per SIMD.
//
// hex8_strict_group handles one group of EPB elements; every lane of the
// block calls it (it holds barriers), inactive lanes with active = false.
// sD holds dN_n/dxi_m at the Gauss points, [q][m][n]; wq is the Gauss weight;
// with PRE, jw points at the lane's Gauss point in the [5][jw_stride] table.
// ---------------------------------------------------------------------------
constexpr int HEX_STRICT_EPB = 8;

template <int LAW, int EPB, class HT, class NA, int W = 3, bool PRE = false, bool EXT = false>
__device__ __forceinline__ void hex8_strict_group(
    double (&sN)[EPB][8][6], const double (&sD)[8][3][8],
    double wq, int el, int q, bool active, const int32_t *conn_e, const NA &nodes_of,
    double lambda, double mu, double sy, double hh,
    const QpViewT<HT> &V, int64_t iq, const int32_t *pos_e, double *fe,
    const double2 *jw = nullptr, int64_t jw_stride = 0) {
  // The dN/dX columns of the gradient are kept for the force terms (24
  // registers instead of J^-1's 9 and 120 operations); the plastic law keeps
  // recomputing them, its history already holds registers.
  constexpr bool KEEP_B = LAW != LAW_PLASTIC;
  double Ji[3][3], g[3][3], S[3][3], wJ = 0.;
  double Bk[KEEP_B ? 8 : 1][3];
  if (active) {
    if constexpr (PRE) {
      // J^-1 and w det J from the table (jw: this Gauss point's column 0),
      // streamed once per step: non-temporal loads leave the caches to the
      // nodal data
      double d[3];
      nodes_of.load_u(conn_e[q], d);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        sN[el][q][3 + c] = d[c];
      typedef double dv2 __attribute__((ext_vector_type(2)));
      const dv2 *jv = reinterpret_cast<const dv2 *>(jw);
      const dv2 j0 = __builtin_nontemporal_load(jv),
                j1 = __builtin_nontemporal_load(jv + jw_stride),
                j2 = __builtin_nontemporal_load(jv + 2 * jw_stride),
                j3 = __builtin_nontemporal_load(jv + 3 * jw_stride),
                j4 = __builtin_nontemporal_load(jv + 4 * jw_stride);
      Ji[0][0] = j0.x;
      Ji[0][1] = j0.y;
      Ji[0][2] = j1.x;
      Ji[1][0] = j1.y;
      Ji[1][1] = j2.x;
      Ji[1][2] = j2.y;
      Ji[2][0] = j3.x;
      Ji[2][1] = j3.y;
      Ji[2][2] = j4.x;
      wJ = j4.y;
    } else {
      double x[3], d[3];
      nodes_of.load(conn_e[q], x, d);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sN[el][q][c] = x[c];
        sN[el][q][3 + c] = d[c];
      }
    }
  }
  __syncthreads();
  if (active) {
    if constexpr (!PRE) {
      // J(i, j) = sum_n dnds(i, n) X(n, j)
      double J[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = sD[q][a][0] * sN[el][0][j];
#pragma unroll
          for (int n = 1; n < 8; ++n)
            s += sD[q][a][n] * sN[el][n][j];
          J[a][j] = s;
        }
      double detJ = jacobian_det_inv(J, Ji);
      wJ = detJ * wq;
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      double b[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double s = Ji[a][0] * sD[q][0][n];
        s += Ji[a][1] * sD[q][1][n];
        s += Ji[a][2] * sD[q][2][n];
        b[a] = s;
        if constexpr (KEEP_B)
          Bk[n][a] = s;
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double un = sN[el][n][3 + a];
#pragma unroll
        for (int j = 0; j < 3; ++j)
          g[a][j] = (n == 0) ? un * b[j] : g[a][j] + un * b[j];
      }
    }
    if constexpr (EXT) {
      double sth = -0.0, dsth = 0.;
      ext_terms<LAW, double, HT>(V, iq, g, sth, dsth);
      strict_law<LAW, double, HT>(V, iq, lambda, mu, sy, hh, g, S, sth, dsth);
    } else {
      strict_law<LAW, double, HT>(V, iq, lambda, mu, sy, hh, g, S);
    }
  }
  if constexpr (LAW == LAW_USER_GRAD)
    return; // the gradient pass forms no forces (uniform: no lane waits)
  __syncthreads(); // every lane is done with sN: it now holds force terms
  double(&sR)[EPB][8][6] = sN;
  // four quarters of two nodes: the force terms of this Gauss point go to
  // LDS, then the lanes q < 2 sum node 2 h + q over the Gauss points in order
#pragma unroll
  for (int hh2 = 0; hh2 < 4; ++hh2) {
    if (active) {
#pragma unroll
      for (int nn = 0; nn < 2; ++nn) {
        const int n = 2 * hh2 + nn;
        double b[3], r[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if constexpr (KEEP_B) {
            b[a] = Bk[n][a];
          } else {
            double s = Ji[a][0] * sD[q][0][n];
            s += Ji[a][1] * sD[q][1][n];
            s += Ji[a][2] * sD[q][2][n];
            b[a] = s;
          }
        }
        if constexpr (finite_law<LAW>())
          btd_finite(g, S, b, r);
        else
          btd_small(S, b, r);
#pragma unroll
        for (int k = 0; k < 3; ++k)
          sR[el][q][3 * nn + k] = r[k] * wJ;
      }
    }
    __syncthreads();
    if constexpr (W == 3) {
      // lane q < 6 sums component q % 3 of node 2 hh2 + q / 3
      if (active && q < 6) {
        const int c = q >= 3 ? 1 : 0;
        double s = sR[el][0][q];
#pragma unroll
        for (int qq = 1; qq < 8; ++qq)
          s += sR[el][qq][q];
        fe[(int64_t)pos_e[2 * hh2 + c] * 3 + (q - 3 * c)] = s;
      }
    } else if (active && q < 2) {
      const int n = 2 * hh2 + q;
      double r[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double s = sR[el][0][3 * q + k];
#pragma unroll
        for (int qq = 1; qq < 8; ++qq)
          s += sR[el][qq][3 * q + k];
        r[k] = s;
      }
      put_entry<W>(fe, (int64_t)pos_e[n], r[0], r[1], r[2]);
    }
    __syncthreads();
  }
}

// The single-model launch reads J^-1 and w det J from the table that init
// fills (k_hex8_jacobian_table): they depend on the reference coordinates
// only, as the reference's FEEngine precompute does (fe_engine_template_tmpl
// .hh:98-120), so the pass loads u but not X.
template <int LAW, bool EXT = false>
__global__ void __launch_bounds__(64) k_hex8_strict(ElemArgs A) {
  __shared__ double sN[HEX_STRICT_EPB][8][6];
  __shared__ double sD[8][3][8];
  const int t = threadIdx.x, el = t >> 3, q = t & 7;
  for (int k = t; k < 192; k += 64)
    (&sD[0][0][0])[k] = (&c_hex8.dnds[0][0][0])[k];
  const int64_t i = xcd_block(A.nblocks_padded) * HEX_STRICT_EPB + el;
  const bool active = i < A.n;
  int64_t e = 0;
  double lambda = A.lambda, mu = A.mu, sy = A.sigma_y, hh = A.h;
  if (active) {
    e = A.filter ? (int64_t)A.filter[i] : A.first + i;
    if (A.lam_el) {
      lambda = A.lam_el[i];
      mu = A.mu_el[i];
    }
    if (A.sy_el)
      sy = A.sy_el[i];
    if (A.h_el)
      hh = A.h_el[i];
  }
  hex8_strict_group<LAW, HEX_STRICT_EPB, double, NodesSoA, 3, true, EXT>(
      sN, sD, c_hex8.w[q], el, q, active, A.conn + e * 8, NodesSoA{A.X, A.u, A.nn},
      lambda, mu, sy, hh, qp_view(A, 8), i * 8 + q, A.pos + (A.pos_offset + (active ? i : 0)) * 8,
      reinterpret_cast<double *>(A.fe), A.jw + e * 8 + q, A.nel * 8);
}

// ---------------------------------------------------------------------------
// Nodal gather + residual + lumped solve + corrector (+ the next step's
// predictor).  One thread per node.
//   f_int  = sum over the node's entries d = 0..deg-1 (ascending assembly
//            order) of -1 * r_e
//   r      = (0 + f_ext) + f_int + M * (0 + -1 * a)
//   x      = 1 * (r / M)                  (unblocked DOFs only)
//   a     += 1 * x ; v += (beta dt) * x    (u += 0 * x is the identity)
//   [predictor] u += dt v + (.5 dt)(dt a) ; v = v + dt a
// (reference_gather / reference_predict in akx_math.hpp; the FP32 form runs
// half_step_gather / half_step_predict there: the half-step velocity).
// References: dof_manager.cc:65-120 (scatter order), solid_mechanics_model.cc
// :235-249, integration_scheme_2nd_order.cc:70-93, dof_manager_default.cc
// :260-278, non_linear_solver_lumped.cc:76-89, newmark-beta.cc:47-77,164-197.
// ---------------------------------------------------------------------------
struct GatherArgs {
  int64_t nn;               // nodes of the mesh (stride of the [3][nn] fields)
  int64_t node0, node_end;  // nodes of this launch: [node0, node_end)
  const int64_t *ell_off; // [nn/256 + 1] first entry of each 256-node slice
  const uint8_t *deg;      // [nn] elements around each node
  const void *fe;          // [entry][3]
  double *fint;          // [3][nn], or null (not stored: steps follow)
  const double *fext;    // [3][nn] or null (zero external force)
  const double *mass;    // [nn]
  const uint8_t *blocked;
  double *u, *v, *a;     // [3][nn]
  double *xu;            // Tet4: [nn][6] {X, u} node records (u kept equal), or null
  float *cu, *cv;        // compensation terms [3][nn] (FP32 form) or null
  double dt;
  int64_t nblocks_padded;
};

template <class FT, bool UPDATE, bool PREDICT, bool COMP>
__global__ void __launch_bounds__(256) k_gather_update(GatherArgs A) {
  const int64_t n = A.node0 + xcd_block(A.nblocks_padded) * blockDim.x + threadIdx.x;
  if (n >= A.node_end)
    return;
  const FT *fe = reinterpret_cast<const FT *>(A.fe);
  double f0 = 0., f1 = 0., f2 = 0.;
  const int deg = A.deg[n];
  const FT *p = fe + (A.ell_off[n >> 8] + (n & 255)) * 3;
  // the table is read once per step: non-temporal loads keep it from
  // displacing the nodal state in the caches
  for (int d = 0; d < deg; ++d, p += 256 * 3) {
    f0 = f0 + -1. * double(__builtin_nontemporal_load(p));
    f1 = f1 + -1. * double(__builtin_nontemporal_load(p + 1));
    f2 = f2 + -1. * double(__builtin_nontemporal_load(p + 2));
  }
  if (A.fint) { // null while more steps follow in the same batch
    A.fint[n] = f0;
    A.fint[A.nn + n] = f1;
    A.fint[2 * A.nn + n] = f2;
  }
  if constexpr (!UPDATE)
    return;
  const double M = A.mass[n];
  const uint8_t blk = A.blocked[n];
  double f[3] = {f0, f1, f2};
  // u after the predictor (blocked components keep theirs), for the record
  double un[3] = {0., 0., 0.};
  if constexpr (PREDICT) {
    if (A.xu)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if ((blk >> c) & 1)
          un[c] = A.u[(int64_t)c * A.nn + n];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if ((blk >> c) & 1)
      continue;
    const int64_t o = (int64_t)c * A.nn + n;
    const double fx = A.fext ? A.fext[o] : 0.;
    if constexpr (COMP) { // FP32 form: half-step velocity (akx_math.hpp)
      double w = A.v[o], a, u = 0.;
      float cw = A.cv[o], cu = 0.f;
      if constexpr (PREDICT) {
        u = A.u[o];
        cu = A.cu[o];
      }
      half_step_gather<PREDICT>(fx + f[c], M, A.dt, w, cw, u, cu, a);
      if constexpr (PREDICT) {
        A.u[o] = u;
        un[c] = u;
        A.cu[o] = cu;
      } else {
        A.a[o] = a;
      }
      A.v[o] = w;
      A.cv[o] = cw;
    } else { // reference order (akx_math.hpp)
      double a = A.a[o], v = A.v[o], u = 0.;
      if constexpr (PREDICT)
        u = A.u[o];
      reference_gather<PREDICT>(fx, f[c], M, A.dt, a, v, u);
      if constexpr (PREDICT) {
        A.u[o] = u;
        un[c] = u;
      }
      A.a[o] = a;
      A.v[o] = v;
    }
  }
  if constexpr (PREDICT) {
    if (A.xu) { // u0 at 48 n + 24 bytes, then an aligned pair
      A.xu[6 * n + 3] = un[0];
      *reinterpret_cast<double2 *>(A.xu + 6 * n + 4) = make_double2(un[1], un[2]);
    }
  }
}

// Predictor alone (first step of a solve_step call): newmark-beta.cc:47-77
template <bool COMP>
__global__ void __launch_bounds__(256) k_predictor(GatherArgs A) {
  const int64_t n = A.node0 + xcd_block(A.nblocks_padded) * blockDim.x + threadIdx.x;
  if (n >= A.node_end)
    return;
  const uint8_t blk = A.blocked[n];
  double un[3] = {0., 0., 0.}; // u after the predictor, for the record
  if (A.xu)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if ((blk >> c) & 1)
        un[c] = A.u[(int64_t)c * A.nn + n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if ((blk >> c) & 1)
      continue;
    const int64_t o = (int64_t)c * A.nn + n;
    const double a = A.a[o];
    double v = A.v[o], u = A.u[o];
    if constexpr (COMP) { // the velocity array takes the half-step velocity
      float cu = A.cu[o], cv = A.cv[o];
      half_step_predict(a, A.dt, v, cv, u, cu);
      A.cu[o] = cu;
      A.cv[o] = cv;
    } else {
      reference_predict(a, A.dt, v, u);
    }
    A.u[o] = u;
    un[c] = u;
    A.v[o] = v;
  }
  if (A.xu) {
    A.xu[6 * n + 3] = un[0];
    *reinterpret_cast<double2 *>(A.xu + 6 * n + 4) = make_double2(un[1], un[2]);
  }
}

} // namespace akx
