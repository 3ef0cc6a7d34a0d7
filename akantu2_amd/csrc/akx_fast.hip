// akx_fast.hip -- regrouped element pass of the FP32 precision form and of
// the opt-in FP64 "fast" form.  Compiled with -ffp-contract=fast (FMA), in
// its own translation unit and math namespace.
//
// Same mechanics as the reference (material.cc:98-154 + 457-584): gather u,
// grad u at every Gauss point, constitutive law, B^T sigma (small strain) or
// (B_L + B_NL)^T S = (F S) dN/dX (total Lagrangian), times w detJ, summed per
// element.  The arithmetic is regrouped for gfx950:
//   J_q = dN/dxi_q X,  H_q = U dN/dxi_q^T              (natural gradients)
//   grad u_q = H_q J_q^{-1 T}
//   Q_q = w_q detJ_q P_q J_q^{-1}      (P = sigma, or F S for finite strain)
//   r_e(n) = sum_q Q_q dN_n/dxi_q
// which needs no dN/dX array and exchanges 9 numbers per Gauss point instead
// of 24.  In FP64 the results agree with the reference operation order to
// rounding per evaluation; the FP32 form uses a cancellation-free neo-Hookean
// stress.
//
// Hex8: 8 lanes per element (one per Gauss point; in the last phase one per
// node), 32 elements per 256-thread block, nodal data and Q in LDS.
// Tet4: one thread per element (single Gauss point).
// The same element code serves the batched FP32 form (persistent kernels
// over the tile list of many instances, per-instance parameters).
#define AKX_MATH_NS akxf
#include "akx_args.hpp"
#include "akx_math.hpp"
#include "akx_fast.hpp"

#include <hip/hip_runtime.h>

namespace akxf {

using akx::BatchElemArgs;
using akx::BInst;
using akx::NodesSoA;
using akx::NodesXU;
using akx::BBlk;
using akx::BTile;
using akx::ElemArgs;
using akx::QpCompact;
using akx::QpViewT;
using akx::batch_compact_view;
using akx::FE_W;
using akx::qp_view;

constexpr int LAW_ELASTIC = 0, LAW_ELASTIC_FD = 1, LAW_NEOHOOKEAN = 2,
              LAW_PLASTIC = 3;

// dN_n/dxi_m at the 8 Hex8 Gauss points, [q][m][n]
__constant__ double c_dnds[8][3][8];

__device__ __forceinline__ int64_t xcd_block(int64_t nblocks_padded) {
  int64_t b = blockIdx.x;
  int64_t per = nblocks_padded >> 3;
  return (b & 7) * per + (b >> 3);
}

// inverse and determinant by cofactors
template <class R>
__device__ __forceinline__ R inv_det(const R (&m)[3][3], R (&r)[3][3]) {
  R c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  R c10 = m[2][1] * m[0][2] - m[2][2] * m[0][1];
  R c20 = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  R det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
  R id = akx_rcp(det);
  r[0][0] = c00 * id;
  r[0][1] = c10 * id;
  r[0][2] = c20 * id;
  r[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
  r[1][1] = (m[2][2] * m[0][0] - m[2][0] * m[0][2]) * id;
  r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
  r[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
  r[2][1] = (m[2][0] * m[0][1] - m[2][1] * m[0][0]) * id;
  r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
  return det;
}

// Linear isotropic hardening with the full incremental history of the
// reference (grad u, stress, plastic strain, hardening).
template <class CT, class HT>
__device__ __forceinline__ void plastic_at(const QpViewT<HT> &V, int64_t iq, CT lambda,
                                           CT mu, CT sy, CT hh, const CT (&g)[3][3],
                                           CT (&S)[3][3]) {
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
  plastic_lih(lambda, mu, sy, hh, g, pg, ps, CT(V.p_iso[iq]), pe, S, iso, eps);
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
  if (V.c_pe) { // afterSolveStep -> updateEnergies (material_plastic.cc:86-99)
    CT dd = CT(0);
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
        dd += (S[ii][jj] + ps[ii][jj]) * (eps[ii][jj] - pe[ii][jj]);
    V.c_pe[iq] = HT(CT(V.p_pe[iq]) + dd / CT(2));
  }
}

// The total-strain form with the compact state (plastic strain in Voigt
// order and hardening, FP32): the batched FP32 form.
template <class CT>
__device__ __forceinline__ void plastic_at(const QpCompact &V, int64_t iq, CT lambda,
                                           CT mu, CT sy, CT hh, const CT (&g)[3][3],
                                           CT (&S)[3][3]) {
  CT pe[6], pn[6], iso;
#pragma unroll
  for (int k = 0; k < 6; ++k)
    pe[k] = CT(V.p_eps[(int64_t)k * V.nq_all + iq]);
  plastic_lih_total(lambda, mu, sy, hh, g, pe, CT(V.p_iso[iq]), S, pn, iso);
#pragma unroll
  for (int k = 0; k < 6; ++k)
    V.c_eps[(int64_t)k * V.nq_all + iq] = float(pn[k]);
  V.c_iso[iq] = float(iso);
}

// Constitutive law at one Gauss point: g -> stress S (sigma or PK2) and the
// tensor P contracted with dN/dX (sigma, or F S for finite deformation).
template <int LAW, class CT, class VT>
__device__ __forceinline__ void law_at(const VT &V, int64_t iq, CT lambda,
                                       CT mu, CT sy, CT hh, const CT (&g)[3][3],
                                       CT (&S)[3][3], CT (&P)[3][3]) {
  if constexpr (LAW == LAW_ELASTIC) {
    elastic_stress(lambda, mu, g, CT(0), S);
  } else if constexpr (LAW == LAW_ELASTIC_FD) {
    CT E[3][3];
    green_strain(g, E);
    elastic_stress(lambda, mu, E, CT(0), S);
  } else if constexpr (LAW == LAW_NEOHOOKEAN) {
    if constexpr (sizeof(CT) == 4)
      neohookean_pk2_stable(lambda, mu, g, S);
    else
      neohookean_pk2(lambda, mu, g, S);
  } else {
    plastic_at(V, iq, lambda, mu, sy, hh, g, S);
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
  if constexpr (LAW == LAW_ELASTIC_FD || LAW == LAW_NEOHOOKEAN) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        CT s = CT(0);
#pragma unroll
        for (int m = 0; m < 3; ++m)
          s += ((k == m ? CT(1) : CT(0)) + g[k][m]) * S[m][J];
        P[k][J] = s;
      }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int J = 0; J < 3; ++J)
        P[k][J] = S[k][J];
  }
}

template <class CT>
__device__ __forceinline__ void element_params(const ElemArgs &A, int64_t i,
                                               CT &lambda, CT &mu, CT &sy,
                                               CT &hh) {
  if (sizeof(CT) == 4 && A.lm_f) { // the same FP32 values, one 8-byte load
    const float2 v = A.lm_f[i];
    lambda = CT(v.x);
    mu = CT(v.y);
  } else {
    lambda = CT(A.lam_el ? A.lam_el[i] : A.lambda);
    mu = CT(A.mu_el ? A.mu_el[i] : A.mu);
  }
  sy = CT(A.sy_el ? A.sy_el[i] : A.sigma_y);
  hh = CT(A.h_el ? A.h_el[i] : A.h);
}

// ---------------------------------------------------------------------------
// Hex8: 8 lanes per element.  hex8_group handles EPB elements; every lane of
// the block calls it (it holds barriers), lanes without an element with
// active = false.  sD must hold dN_n/dxi_m ([q][m][n]) before the call.
// ---------------------------------------------------------------------------
constexpr int HEX_EPB = 32; // elements per 256-thread block

template <int LAW, class CT, class FT, class VT, int EPB, class NA, int W = 3>
__device__ __forceinline__ void hex8_group(CT (&sN)[EPB][8][6], CT (&sQ)[EPB][8][9],
                                           const CT (&sD)[8][3][8], int el, int q,
                                           bool active, const int32_t *conn_e,
                                           const NA &nodes_of, CT lambda, CT mu, CT sy,
                                           CT hh, const VT &V, int64_t iq,
                                           const int32_t *pos_e, FT *fe) {
  // lane q stages node q: absolute in FP64; in FP32 relative to node 0 of
  // the element (exact FP64 difference, rounded once; lanes 8 el .. 8 el + 7
  // are one element, so node 0 comes from the first lane of the 8)
  double x[3] = {0., 0., 0.}, d[3] = {0., 0., 0.};
  if (active)
    nodes_of.load(conn_e[q], x, d);
  if constexpr (sizeof(CT) == 4) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c] -= __shfl(x[c], 0, 8);
      d[c] -= __shfl(d[c], 0, 8);
    }
  }
  if (active) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sN[el][q][c] = CT(x[c]);
      sN[el][q][3 + c] = CT(d[c]);
    }
  }
  __syncthreads();
  if (active) {
    CT J[3][3], H[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        J[a][b] = CT(0);
        H[a][b] = CT(0);
      }
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      CT xn[3], un[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xn[c] = sN[el][n][c];
        un[c] = sN[el][n][3 + c];
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const CT d = sD[q][m][n];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          J[m][c] += d * xn[c];
          H[c][m] += un[c] * d;
        }
      }
    }
    CT Ji[3][3];
    const CT wJ = inv_det(J, Ji); // Hex8 Gauss weights are 1
    CT g[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        CT s = CT(0);
#pragma unroll
        for (int m = 0; m < 3; ++m)
          s += H[a][m] * Ji[b][m];
        g[a][b] = s;
      }
    CT S[3][3], P[3][3];
    law_at<LAW, CT>(V, iq, lambda, mu, sy, hh, g, S, P);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        CT s = CT(0);
#pragma unroll
        for (int J2 = 0; J2 < 3; ++J2)
          s += P[k][J2] * Ji[J2][m];
        sQ[el][q][3 * k + m] = wJ * s;
      }
  }
  __syncthreads();
  if (active) {
    const int n = q; // this lane now owns node n of the element
    CT r[3] = {CT(0), CT(0), CT(0)};
#pragma unroll
    for (int qq = 0; qq < 8; ++qq)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const CT d = sD[qq][m][n];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          r[k] += sQ[el][qq][3 * k + m] * d;
      }
    akx::put_entry<W>(fe, (int64_t)pos_e[n], FT(r[0]), FT(r[1]), FT(r[2]));
  }
}

// ---------------------------------------------------------------------------
// Hex8, FP32 form, sum-factorised.  The trilinear shape functions at the
// 2 x 2 x 2 Gauss points factor per natural axis: along an axis, a node pair
// (v-, v+) maps to the two Gauss points by
//   I: alpha v_own + beta v_partner          (interpolation)
//   D: (v+ - v-) / 2                           (derivative)
// with alpha, beta = (1 +- 1/sqrt(3)) / 2.  With the 8 lanes of an element
// indexed by corner bits (bit a set = + side of axis a), node and Gauss point
// of a lane share the bits, so every axis step is a butterfly with the lane
// 1, 2 or 4 away:
//   J, H:   row m of [J|H] = D on axis m, I on the other two      (nodes -> q)
//   force:  F = sum_m (D^T on axis m, I on the other two) Q_m     (q -> nodes)
// The lanes apply the raw steps  I' = v + r p  (r = beta / alpha) and
// D' = v - p (own minus partner).  Then J' = c S J and H' = c H S with
// c = 2 / alpha^2 and S = diag(s_m), s_m = the lane's sign on axis m, so
// grad u = H' J'^-1 = H J^-1 unchanged, and the raw transposed chain applied
// to Q' = (sigma / c^3) det(J') P J'^-1, sigma = s_0 s_1 s_2, gives the
// nodal forces sum_q Q_q dN/dxi_q.  No LDS and no barriers: 14 exchanges
// per component instead of 8 x 3 products over the nodes.
// Bits -> node of the reference's numbering (counter-clockwise faces):
// n(b) = b ^ ((b >> 1) & 1); the Gauss point index is the bits themselves.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_x4(float v) { // value of the lane 4 away (swizzle)
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x101F));
}

constexpr float HEX_G = 0.57735026918962576f;          // 1 / sqrt(3)
constexpr float HEX_R = 0.26794919243112270f;          // beta / alpha = (1 - G) / (1 + G)
constexpr float HEX_C3 = 33.243067814700545f;          // c^3, c = 2 / alpha^2

// One butterfly step as one instruction, the partner's value read through
// DPP: acc + r v_partner (v_fmac_f32_dpp) and v - v_partner
// (v_subrev_f32_dpp).  The compiler does not fuse a DPP move into an FMA
// here, so the steps are written out; s_nop 4 gives the wait states a DPP
// read needs after a VALU write of its source (2) or of EXEC (5), which the
// compiler cannot account for inside asm.  Partners are lanes of the same
// element, so they are active whenever the lane is.
#define AKXF_DPP_STEPS(SUFFIX, PERM)                                                      \
  __device__ __forceinline__ float fmac_##SUFFIX(float acc, float v, float r) {           \
    asm("s_nop 4\n\tv_fmac_f32_dpp %0, %1, %2 quad_perm:" PERM                            \
        " row_mask:0xf bank_mask:0xf bound_ctrl:1"                                         \
        : "+v"(acc)                                                                        \
        : "v"(v), "v"(r));                                                                 \
    return acc;                                                                            \
  }                                                                                        \
  __device__ __forceinline__ float diff_##SUFFIX(float v) {                                \
    float d;                                                                               \
    asm("s_nop 4\n\tv_subrev_f32_dpp %0, %1, %1 quad_perm:" PERM                          \
        " row_mask:0xf bank_mask:0xf bound_ctrl:1"                                         \
        : "=v"(d)                                                                          \
        : "v"(v));                                                                         \
    return d;                                                                              \
  }
AKXF_DPP_STEPS(x1, "[1,0,3,2]")
AKXF_DPP_STEPS(x2, "[2,3,0,1]")
#undef AKXF_DPP_STEPS

// r in a vector register (the source operand of v_fmac_f32_dpp)
__device__ __forceinline__ float hex_r_vgpr() {
  float r;
  asm volatile("v_mov_b32 %0, 0x3e8930a3" : "=v"(r)); // HEX_R
  return r;
}

// [J' | H'] rows xi, eta, zeta at this lane's Gauss point, from the node
// values of one component.
__device__ __forceinline__ void hex_sf_grad(float R, float v, float &rx, float &ry,
                                            float &rz) {
  const float p = lane_x4(v);
  const float a = v - p, b = v + p * R;                 // axis 2: D, I
  const float a1 = fmac_x2(a, a, R);                    // axis 1: I, D, I
  const float b1 = diff_x2(b), c1 = fmac_x2(b, b, R);
  rz = fmac_x1(a1, a1, R);                              // axis 0: I, I, D
  ry = fmac_x1(b1, b1, R);
  rx = diff_x1(c1);
}

// One force component of this lane's node from Q'_(xi, eta, zeta).
__device__ __forceinline__ float hex_sf_force(float R, float qx, float qy, float qz) {
  const float A = diff_x1(qx);                          // axis 0: D^T, I, I
  const float B = fmac_x1(qy, qy, R);
  const float C = fmac_x1(qz, qz, R);
  const float E = fmac_x2(A, A, R) + diff_x2(B);        // axis 1: I, D^T; I
  const float Cz = fmac_x2(C, C, R);
  return (E + lane_x4(E) * R) + (Cz - lane_x4(Cz));     // axis 2: I, D^T
}

// All 64 lanes of the wave call it (cross-lane steps); lanes 8 e .. 8 e + 7
// are element e, lane bits b = l % 8.  Lanes of elements beyond the range
// pass active = false and compute on zeros.  SLOT: the node's force goes to
// LDS slot ls_e[node] of `slots` (block-partial assembly) instead of table
// entry pos_e[node] of fe.
template <int LAW, class FT, class VT, class NA, int W = 3, bool SLOT = false>
__device__ __forceinline__ void hex8_group_sf(int b, bool active, const int32_t *conn_e,
                                              const NA &nodes_of, float lambda, float mu,
                                              float sy, float hh, const VT &V, int64_t iq,
                                              const int32_t *pos_e, FT *fe,
                                              float *slots = nullptr,
                                              const uint16_t *ls_e = nullptr) {
  const int node = b ^ ((b >> 1) & 1);
  double x[3] = {0., 0., 0.}, d[3] = {0., 0., 0.};
  int32_t pq = 0;
  if (active) {
    nodes_of.load(conn_e[node], x, d);
    if constexpr (SLOT)
      pq = ls_e[node];
    else
      pq = pos_e[node];
  }
  const float R = hex_r_vgpr();
  // relative to node 0 (lane bits 0): exact FP64 difference, rounded once
  float J[3][3], H[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float xr = float(x[c] - __shfl(x[c], 0, 8));
    const float ur = float(d[c] - __shfl(d[c], 0, 8));
    hex_sf_grad(R, xr, J[0][c], J[1][c], J[2][c]);
    hex_sf_grad(R, ur, H[c][0], H[c][1], H[c][2]);
  }
  float Ji[3][3];
  const float dJ = inv_det(J, Ji);
  float g[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      float s = 0.f;
#pragma unroll
      for (int m = 0; m < 3; ++m)
        s += H[a][m] * Ji[bb][m];
      g[a][bb] = s;
    }
  float S[3][3], P[3][3] = {};
  if (active)
    law_at<LAW, float>(V, iq, lambda, mu, sy, hh, g, S, P);
  // sigma = s_0 s_1 s_2 = +1 when an even number of bits are clear
  const float sig = ((3 - __builtin_popcount(b)) & 1) ? -1.f : 1.f;
  const float w = dJ * (sig / HEX_C3);
  float F[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float q[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      float s = 0.f;
#pragma unroll
      for (int J2 = 0; J2 < 3; ++J2)
        s += P[k][J2] * Ji[J2][m];
      q[m] = w * s;
    }
    F[k] = hex_sf_force(R, q[0], q[1], q[2]);
  }
  if (active) {
    if constexpr (SLOT) {
      float *o = slots + 3 * pq;
      o[0] = F[0];
      o[1] = F[1];
      o[2] = F[2];
    } else {
      akx::put_entry<W>(fe, (int64_t)pq, FT(F[0]), FT(F[1]), FT(F[2]));
    }
  }
}

template <int LAW, class FT>
__global__ void __launch_bounds__(256) k_hex8_sf(ElemArgs A) {
  const int t = threadIdx.x, el = t >> 3, b = t & 7;
  const int64_t i = xcd_block(A.nblocks_padded) * HEX_EPB + el;
  const bool active = i < A.n;
  int64_t e = 0;
  float lambda = 0.f, mu = 0.f, sy = 0.f, hh = 0.f;
  if (active) {
    e = A.filter ? (int64_t)A.filter[i] : A.first + i;
    element_params(A, i, lambda, mu, sy, hh);
  }
  hex8_group_sf<LAW, FT, QpViewT<double>>(
      b, active, A.conn + e * 8, NodesSoA{A.X, A.u, A.nn}, lambda, mu, sy, hh, qp_view(A, 8),
      i * 8 + b, A.pos + (A.pos_offset + (active ? i : 0)) * 8, reinterpret_cast<FT *>(A.fe));
}

// FP32 Hex8 with block-partial assembly: each workgroup (256 lanes, 32
// elements at a time through the sum-factorised body) does one block of
// BPA_HEX8 consecutive elements; the lanes put their nodes' forces in LDS
// slots (the corners of one distinct node of the block take consecutive
// slots in ascending element order), then one thread per distinct node adds
// its slots in that order (FP64, rounded once) and writes one table entry
// per (block, node).  The gather adds a node's partials in block order.
template <int LAW>
__global__ void __launch_bounds__(256) k_hex8_bpa(ElemArgs A) {
  __shared__ float s_f[akx::BPA_HEX8 * 8 * 3];
  const int t = threadIdx.x, el = t >> 3, b = t & 7;
  const int64_t blk = xcd_block(A.nblocks_padded);
  const int64_t i0 = blk * akx::BPA_HEX8;
  if (i0 >= A.n) // padding block: no elements, no node list
    return;
  for (int base = 0; base < akx::BPA_HEX8; base += HEX_EPB) {
    const int64_t i = i0 + base + el;
    const bool active = base + el < akx::BPA_HEX8 && i < A.n;
    const int64_t ic = active ? i : i0;
    float lambda = 0.f, mu = 0.f, sy = 0.f, hh = 0.f;
    if (active)
      element_params(A, i, lambda, mu, sy, hh);
    hex8_group_sf<LAW, float, QpViewT<double>, NodesSoA, 3, true>(
        b, active, A.conn + (A.first + ic) * 8, NodesSoA{A.X, A.u, A.nn}, lambda, mu, sy,
        hh, qp_view(A, 8), ic * 8 + b, nullptr, nullptr, s_f, A.lslot + ic * 8);
  }
  __syncthreads();
  const uint16_t *soff = A.bl_soff + blk * (A.bl_maxu + 1);
  const int32_t *dst = A.bl_pos + blk * A.bl_maxu;
  float *fe = reinterpret_cast<float *>(A.fe);
  for (int k = t; k < A.bl_maxu; k += blockDim.x) {
    const int32_t d = dst[k];
    if (d < 0)
      continue;
    const int s0 = soff[k], s1 = soff[k + 1];
    double fx = 0., fy = 0., fz = 0.;
    for (int sl = s0; sl < s1; ++sl) {
      fx += (double)s_f[3 * sl];
      fy += (double)s_f[3 * sl + 1];
      fz += (double)s_f[3 * sl + 2];
    }
    akx::put_entry<3>(fe, (int64_t)d, float(fx), float(fy), float(fz));
  }
}

template <int LAW, class CT, class FT>
__global__ void __launch_bounds__(256) k_hex8_fast(ElemArgs A) {
  __shared__ CT sN[HEX_EPB][8][6];     // X (0..2) and u (3..5) of the 8 nodes
  __shared__ CT sQ[HEX_EPB][8][9];     // Q of the 8 Gauss points
  __shared__ CT sD[8][3][8];           // dN_n/dxi_m at Gauss point q
  const int t = threadIdx.x, el = t >> 3, q = t & 7;
  for (int k = t; k < 192; k += 256)
    (&sD[0][0][0])[k] = CT((&c_dnds[0][0][0])[k]);
  const int64_t i = xcd_block(A.nblocks_padded) * HEX_EPB + el;
  const bool active = i < A.n;
  int64_t e = 0;
  CT lambda = CT(0), mu = CT(0), sy = CT(0), hh = CT(0);
  if (active) {
    e = A.filter ? (int64_t)A.filter[i] : A.first + i;
    element_params(A, i, lambda, mu, sy, hh);
  }
  hex8_group<LAW, CT, FT, QpViewT<double>, HEX_EPB>(
      sN, sQ, sD, el, q, active, A.conn + e * 8, NodesSoA{A.X, A.u, A.nn}, lambda, mu, sy,
      hh,
      qp_view(A, 8), i * 8 + q, A.pos + (A.pos_offset + (active ? i : 0)) * 8,
      reinterpret_cast<FT *>(A.fe));
}

// ---------------------------------------------------------------------------
// Tet4: one thread per element
// ---------------------------------------------------------------------------
// EARLY_POS: load the table positions beside the connectivity and keep them
// there (the scheduler otherwise places the load at its use, after the element
// math, a third memory round trip per element).  Faster for the one-element-
// per-thread FP32 pass; slower inside the persistent batch loop, and in FP64
// it costs the pass a wave per SIMD.
// Q = w detJ P J^-1 of one Tet4 element (its corner forces are the columns
// of Q and minus their sum).  BARRIER: a scheduling barrier after the node
// loads, so that a load issued before the call stays ahead of the math.
template <int LAW, class CT, class VT, class NA, bool BARRIER>
__device__ __forceinline__ void tet4_Q(const int4 nd, const NA &nodes_of, CT lambda, CT mu,
                                       CT sy, CT hh, const VT &V, int64_t iq,
                                       CT (&Q)[3][3]) {
  const int32_t nodes[4] = {nd.x, nd.y, nd.z, nd.w};
  double X[4][3], U[4][3];
#pragma unroll
  for (int n = 0; n < 4; ++n)
    nodes_of.load(nodes[n], X[n], U[n]);
  if constexpr (BARRIER)
    __builtin_amdgcn_sched_barrier(0);
  // natural directions: xi -> node 3, eta -> node 1, zeta -> node 2
  const int nm[3] = {3, 1, 2};
  CT J[3][3], H[3][3];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      J[m][c] = CT(X[nm[m]][c] - X[0][c]);
      H[c][m] = CT(U[nm[m]][c] - U[0][c]);
    }
  CT Ji[3][3];
  const CT detJ = inv_det(J, Ji);
  const CT wJ = detJ * CT(1. / 6.);
  CT g[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      CT s = CT(0);
#pragma unroll
      for (int m = 0; m < 3; ++m)
        s += H[a][m] * Ji[b][m];
      g[a][b] = s;
    }
  CT S[3][3], P[3][3];
  law_at<LAW, CT>(V, iq, lambda, mu, sy, hh, g, S, P);
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      CT s = CT(0);
#pragma unroll
      for (int J2 = 0; J2 < 3; ++J2)
        s += P[k][J2] * Ji[J2][m];
      Q[k][m] = wJ * s;
    }
}

template <int LAW, class CT, class FT, class VT, class NA, int W = 3, bool EARLY_POS = false>
__device__ __forceinline__ void tet4_one(const int4 nd, const NA &nodes_of,
                                         CT lambda, CT mu, CT sy, CT hh, const VT &V,
                                         int64_t iq, const int32_t *pos_e, FT *fe) {
  int4 pv;
  if constexpr (EARLY_POS)
    pv = *reinterpret_cast<const int4 *>(pos_e);
  CT Q[3][3];
  tet4_Q<LAW, CT, VT, NA, EARLY_POS>(nd, nodes_of, lambda, mu, sy, hh, V, iq, Q);
  if constexpr (!EARLY_POS)
    pv = *reinterpret_cast<const int4 *>(pos_e);
  akx::put_entry<W>(fe, (int64_t)pv.x, FT(-(Q[0][0] + Q[0][1] + Q[0][2])),
                    FT(-(Q[1][0] + Q[1][1] + Q[1][2])), FT(-(Q[2][0] + Q[2][1] + Q[2][2])));
  akx::put_entry<W>(fe, (int64_t)pv.y, FT(Q[0][1]), FT(Q[1][1]), FT(Q[2][1]));
  akx::put_entry<W>(fe, (int64_t)pv.z, FT(Q[0][2]), FT(Q[1][2]), FT(Q[2][2]));
  akx::put_entry<W>(fe, (int64_t)pv.w, FT(Q[0][0]), FT(Q[1][0]), FT(Q[2][0]));
}

template <int LAW, class CT, class FT>
__global__ void __launch_bounds__(256) k_tet4_fast(ElemArgs A) {
  const int64_t i = xcd_block(A.nblocks_padded) * blockDim.x + threadIdx.x;
  if (i >= A.n)
    return;
  const int64_t e = A.filter ? (int64_t)A.filter[i] : A.first + i;
  CT lambda, mu, sy, hh;
  element_params(A, i, lambda, mu, sy, hh);
  tet4_one<LAW, CT, FT, QpViewT<double>, akx::NodesXU, 3, sizeof(CT) == 4>(
      *reinterpret_cast<const int4 *>(A.conn + e * 4), akx::NodesXU{A.xu}, lambda,
                                         mu, sy, hh,
                                qp_view(A, 1), i, A.pos + (A.pos_offset + i) * 4,
                                reinterpret_cast<FT *>(A.fe));
}

// FP32 Tet4 with the block's node records staged in LDS: the 256 elements of
// a block share about 0.7 distinct nodes each, so each record is read from
// global memory once per block instead of once per element corner.
template <int LAW>
__global__ void __launch_bounds__(256) k_tet4_lds(ElemArgs A) {
  extern __shared__ double2 s_xu[]; // [bl_maxu][3]
  const int64_t blk = xcd_block(A.nblocks_padded);
  if (blk * 256 >= A.n) // padding block: no elements, no node list
    return;
  const int64_t i = blk * 256 + threadIdx.x;
  const bool active = i < A.n;
  // the element's own streams are loaded before the staging, so that their
  // round trip overlaps it (loading the last element's for the lanes past
  // the end instead, without the branch, measured 2.5% slower here)
  ushort4 lc = make_ushort4(0, 0, 0, 0);
  int4 pv = make_int4(0, 0, 0, 0);
  float lambda = 0.f, mu = 0.f, sy = 0.f, hh = 0.f;
  if (active) {
    lc = *reinterpret_cast<const ushort4 *>(A.lconn + i * 4);
    pv = *reinterpret_cast<const int4 *>(A.pos + (A.pos_offset + i) * 4);
    element_params(A, i, lambda, mu, sy, hh);
  }
  akx::stage_block_records(A, blk, s_xu);
  __syncthreads();
  if (!active)
    return;
  tet4_one<LAW, float, float, QpViewT<double>, akx::NodesLDS, 3, true>(
      make_int4(lc.x, lc.y, lc.z, lc.w), akx::NodesLDS{reinterpret_cast<const double *>(s_xu)},
      lambda, mu, sy, hh, qp_view(A, 1), i, reinterpret_cast<const int32_t *>(&pv),
      reinterpret_cast<float *>(A.fe));
}

// FP32 Tet4 with block-partial assembly.  As k_tet4_lds, then the block sums
// the forces its elements put on each of its distinct nodes and writes one
// partial per (block, distinct node) instead of one table entry per element
// corner: the corner forces go to LDS slots (the corners of one distinct node
// take consecutive slots in ascending element order), and one thread per
// distinct node adds its slots in that order (in FP64, rounded once) and
// writes the partial.  The gather then adds a node's partials in ascending
// block order.  The table traffic falls from four entries per element to
// one per (block, node) pair: about 0.7 per element on the box meshes.
template <int LAW>
__global__ void __launch_bounds__(256) k_tet4_bpa(ElemArgs A) {
  extern __shared__ double2 s_xu[]; // [bl_maxu][3] records, then [1024][3] floats
  const int64_t blk = xcd_block(A.nblocks_padded);
  if (blk * 256 >= A.n) // padding block: no elements, no node list
    return;
  const int64_t i = blk * 256 + threadIdx.x;
  const bool active = i < A.n;
  ushort4 lc = make_ushort4(0, 0, 0, 0), ls = make_ushort4(0, 0, 0, 0);
  float lambda = 0.f, mu = 0.f, sy = 0.f, hh = 0.f;
  if (active) {
    lc = *reinterpret_cast<const ushort4 *>(A.lconn + i * 4);
    ls = *reinterpret_cast<const ushort4 *>(A.lslot + i * 4);
    element_params(A, i, lambda, mu, sy, hh);
  }
  // this lane's first distinct node of the summing phase, read ahead so
  // that its loads overlap the staging instead of following the barrier
  const uint16_t *soff = A.bl_soff + blk * (A.bl_maxu + 1);
  const int32_t *dst = A.bl_pos + blk * A.bl_maxu;
  int32_t d_first = -1;
  int s0_first = 0, s1_first = 0;
  if ((int)threadIdx.x < A.bl_maxu) {
    d_first = dst[threadIdx.x];
    s0_first = soff[threadIdx.x];
    s1_first = soff[threadIdx.x + 1];
  }
  akx::stage_block_records(A, blk, s_xu);
  __syncthreads();
  float *s_f = reinterpret_cast<float *>(s_xu + 3 * A.bl_maxu);
  if (active) {
    float Q[3][3];
    tet4_Q<LAW, float, QpViewT<double>, akx::NodesLDS, false>(
        make_int4(lc.x, lc.y, lc.z, lc.w),
        akx::NodesLDS{reinterpret_cast<const double *>(s_xu)}, lambda, mu, sy, hh,
        qp_view(A, 1), i, Q);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s_f[3 * ls.x + c] = -(Q[c][0] + Q[c][1] + Q[c][2]);
      s_f[3 * ls.y + c] = Q[c][1];
      s_f[3 * ls.z + c] = Q[c][2];
      s_f[3 * ls.w + c] = Q[c][0];
    }
  }
  __syncthreads();
  float *fe = reinterpret_cast<float *>(A.fe);
  for (int k = threadIdx.x; k < A.bl_maxu; k += blockDim.x) {
    const bool first = k == (int)threadIdx.x;
    const int32_t d = first ? d_first : dst[k];
    if (d < 0)
      continue;
    const int s0 = first ? s0_first : soff[k], s1 = first ? s1_first : soff[k + 1];
    double fx = 0., fy = 0., fz = 0.;
    for (int sl = s0; sl < s1; ++sl) {
      fx += (double)s_f[3 * sl];
      fy += (double)s_f[3 * sl + 1];
      fz += (double)s_f[3 * sl + 2];
    }
    akx::put_entry<3>(fe, (int64_t)d, float(fx), float(fy), float(fz));
  }
}

// ---------------------------------------------------------------------------
// Batched FP32 form, one tile of one (element type, law) group per
// workgroup; per-instance parameters, compact FP32 plastic state (ping-pong
// by step parity), corners read from the X and u arrays.  The per-corner
// kernels (k_batch_tet4_f32, k_batch_hex8_f32) serve groups without block
// tables; FP32 batches build block tables for every group, so their tiles
// run the block-partial kernels below.
// ---------------------------------------------------------------------------
template <int LAW>
__global__ void __launch_bounds__(256) k_batch_tet4_f32(BatchElemArgs A) {
  {
    const int64_t tt = blockIdx.x; // one tile per workgroup
    if (tt >= A.ntiles)
      return;
    const BTile T = A.tiles[tt];
    const BInst P = A.inst[T.inst];
    const QpCompact V = batch_compact_view(A, (P.parity_base ^ A.parity) & 1);
    for (int j = threadIdx.x; j < T.count; j += blockDim.x) {
      const int64_t e = T.first + j;
      tet4_one<LAW, float, float, QpCompact, NodesSoA, FE_W>(
          *reinterpret_cast<const int4 *>(A.conn + e * 4), NodesSoA{A.X, A.u, A.ncap},
                                         float(P.lambda), float(P.mu), float(P.sigma_y),
                                         float(P.h), V, A.hoff + e, A.pos + e * 4,
                                         reinterpret_cast<float *>(A.fe));
    }
  }
}

// Batched FP32 Tet4 with block-partial assembly: each tile is one block of
// at most 256 consecutive elements of one instance, and each workgroup does
// one tile (a persistent loop over tiles measured 25% slower: a workgroup's
// staging, element and summing phases then ran one after another, their
// load latencies exposed).  The block's distinct
// node records are staged in LDS, the elements' corner forces go to LDS
// slots (the corners of one distinct node take consecutive slots in
// ascending element order), and one thread per distinct node adds its slots
// in that order (FP64, rounded once) and writes one table entry per (block,
// node).  The gather adds a node's partials in ascending block order.
template <int LAW>
__global__ void __launch_bounds__(256) k_batch_tet4_bpa(BatchElemArgs A) {
  extern __shared__ double2 s_xu[]; // [bmaxu][3] records, then [1024][3] floats
  float *s_f = reinterpret_cast<float *>(s_xu + 3 * A.bmaxu);
  float *fe = reinterpret_cast<float *>(A.fe);
  {
    const int64_t tt = blockIdx.x;
    if (tt >= A.ntiles)
      return;
    const BTile T = A.tiles[tt];
    const BBlk B = A.blks[tt];
    const BInst P = A.inst[T.inst];
    const QpCompact V = batch_compact_view(A, (P.parity_base ^ A.parity) & 1);
    // this lane's first distinct node of the summing phase, read ahead so
    // that its loads overlap the staging instead of following the barriers
    const int t = threadIdx.x;
    int s0_first = 0, s1_first = 0, pos_first = 0;
    if (t < B.count) {
      s0_first = t > 0 ? (int)A.bsend[B.off + t - 1] : 0;
      s1_first = A.bsend[B.off + t];
      pos_first = A.bpos[B.off + t];
    }
    // the records of the block's distinct nodes, built from X and u (FP32
    // instances keep u in the [3][ncap] array, BInst::urec = 0)
    for (int k = threadIdx.x; k < B.count; k += blockDim.x) {
      const int64_t nd = A.bnode[B.off + k];
      s_xu[3 * k] = make_double2(A.X[nd], A.X[A.ncap + nd]);
      s_xu[3 * k + 1] = make_double2(A.X[2 * A.ncap + nd], A.u[nd]);
      s_xu[3 * k + 2] = make_double2(A.u[A.ncap + nd], A.u[2 * A.ncap + nd]);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < T.count; j += blockDim.x) {
      const int64_t e = T.first + j;
      const ushort4 lc = *reinterpret_cast<const ushort4 *>(A.lc + e * 4);
      const ushort4 ls = *reinterpret_cast<const ushort4 *>(A.ls + e * 4);
      float Q[3][3];
      tet4_Q<LAW, float, QpCompact, akx::NodesLDS, false>(
          make_int4(lc.x, lc.y, lc.z, lc.w),
          akx::NodesLDS{reinterpret_cast<const double *>(s_xu)}, float(P.lambda),
          float(P.mu), float(P.sigma_y), float(P.h), V, A.hoff + e, Q);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        s_f[3 * ls.x + c] = -(Q[c][0] + Q[c][1] + Q[c][2]);
        s_f[3 * ls.y + c] = Q[c][1];
        s_f[3 * ls.z + c] = Q[c][2];
        s_f[3 * ls.w + c] = Q[c][0];
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < B.count; k += blockDim.x) {
      const bool first = k == t;
      const int s0 = first ? s0_first : (int)A.bsend[B.off + k - 1];
      const int s1 = first ? s1_first : (int)A.bsend[B.off + k];
      const int pos = first ? pos_first : A.bpos[B.off + k];
      double fx = 0., fy = 0., fz = 0.;
      for (int sl = s0; sl < s1; ++sl) {
        fx += (double)s_f[3 * sl];
        fy += (double)s_f[3 * sl + 1];
        fz += (double)s_f[3 * sl + 2];
      }
      akx::put_entry<FE_W>(fe, (int64_t)pos, float(fx), float(fy), float(fz));
    }
  }
}

template <int LAW, int EPB>
__global__ void __launch_bounds__(EPB * 8) k_batch_hex8_f32(BatchElemArgs A) {
  const int t = threadIdx.x, el = t >> 3, q = t & 7;
  {
    const int64_t tt = blockIdx.x; // one tile per workgroup
    if (tt >= A.ntiles)
      return;
    const BTile T = A.tiles[tt];
    const BInst P = A.inst[T.inst];
    const QpCompact V = batch_compact_view(A, (P.parity_base ^ A.parity) & 1);
    for (int base = 0; base < T.count; base += EPB) {
      const bool active = base + el < T.count;
      const int64_t e = T.first + (active ? base + el : 0);
      hex8_group_sf<LAW, float, QpCompact, NodesSoA, FE_W>(
          q, active, A.conn + e * 8, NodesSoA{A.X, A.u, A.ncap}, float(P.lambda),
          float(P.mu), float(P.sigma_y), float(P.h), V, A.hoff + e * 8 + q,
          A.pos + e * 8, reinterpret_cast<float *>(A.fe));
    }
  }
}

// Batched FP32 Hex8 with block-partial assembly: each tile is one block of at
// most BPA_HEX8 consecutive elements of one instance and each workgroup (256
// lanes, 32 elements at a time) does one tile.  The lanes put their nodes'
// forces in LDS slots (the corners of one distinct node of the block take
// consecutive slots in ascending element order), then one thread per
// distinct node adds its slots in that order (FP64, rounded once) and
// writes one table entry per (block, node).
template <int LAW>
__global__ void __launch_bounds__(256) k_batch_hex8_bpa(BatchElemArgs A) {
  __shared__ float s_f[akx::BPA_HEX8 * 8 * 3]; // one slot per element corner of the block
  const int t = threadIdx.x, el = t >> 3, q = t & 7;
  const int64_t tt = blockIdx.x;
  if (tt >= A.ntiles)
    return;
  const BTile T = A.tiles[tt];
  const BBlk B = A.blks[tt];
  const BInst P = A.inst[T.inst];
  const QpCompact V = batch_compact_view(A, (P.parity_base ^ A.parity) & 1);
  for (int base = 0; base < T.count; base += 32) {
    const bool active = base + el < T.count;
    const int64_t e = T.first + (active ? base + el : 0);
    hex8_group_sf<LAW, float, QpCompact, NodesSoA, FE_W, true>(
        q, active, A.conn + e * 8, NodesSoA{A.X, A.u, A.ncap}, float(P.lambda), float(P.mu),
        float(P.sigma_y), float(P.h), V, A.hoff + e * 8 + q, nullptr, nullptr, s_f,
        A.ls + e * 8);
  }
  __syncthreads();
  float *fe = reinterpret_cast<float *>(A.fe);
  for (int k = t; k < B.count; k += blockDim.x) {
    const int s0 = k > 0 ? (int)A.bsend[B.off + k - 1] : 0, s1 = A.bsend[B.off + k];
    double fx = 0., fy = 0., fz = 0.;
    for (int sl = s0; sl < s1; ++sl) {
      fx += (double)s_f[3 * sl];
      fy += (double)s_f[3 * sl + 1];
      fz += (double)s_f[3 * sl + 2];
    }
    akx::put_entry<3>(fe, (int64_t)A.bpos[B.off + k], float(fx), float(fy), float(fz));
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
hipError_t upload_hex8_dnds(const double *dnds) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_dnds), dnds, sizeof(double) * 192);
}

int64_t elements_per_block(int etype, bool bpa) {
  return etype == 1 ? (bpa ? akx::BPA_HEX8 : HEX_EPB) : 256;
}

template <int L, class CT, class FT>
static void launch_one(int etype, const ElemArgs &A, hipStream_t s) {
  dim3 grid((unsigned)A.nblocks_padded), block(256);
  if (etype != 1) {
    if constexpr (sizeof(CT) == 4)
      if (A.lconn) {
        if (A.lslot) // block-partial assembly: records, then 1024 corner slots
          hipLaunchKernelGGL((k_tet4_bpa<L>), grid, block,
                             (size_t)A.bl_maxu * 48 + 1024 * 12, s, A);
        else
          hipLaunchKernelGGL((k_tet4_lds<L>), grid, block, (size_t)A.bl_maxu * 48, s, A);
        return;
      }
    hipLaunchKernelGGL((k_tet4_fast<L, CT, FT>), grid, block, 0, s, A);
  }
  else if constexpr (sizeof(CT) == 4) { // FP32: sum-factorised Hex8
    if (A.lslot) // block-partial assembly
      hipLaunchKernelGGL((k_hex8_bpa<L>), grid, block, 0, s, A);
    else
      hipLaunchKernelGGL((k_hex8_sf<L, FT>), grid, block, 0, s, A);
  }
  else
    hipLaunchKernelGGL((k_hex8_fast<L, CT, FT>), grid, block, 0, s, A);
}

template <class CT, class FT>
static hipError_t launch_t(int etype, int law, ElemArgs A, hipStream_t s) {
  switch (law) {
  case LAW_ELASTIC: launch_one<LAW_ELASTIC, CT, FT>(etype, A, s); break;
  case LAW_ELASTIC_FD: launch_one<LAW_ELASTIC_FD, CT, FT>(etype, A, s); break;
  case LAW_NEOHOOKEAN: launch_one<LAW_NEOHOOKEAN, CT, FT>(etype, A, s); break;
  case LAW_PLASTIC: launch_one<LAW_PLASTIC, CT, FT>(etype, A, s); break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_elements(int etype, int law, bool fp32, const ElemArgs &A,
                           hipStream_t stream) {
  return fp32 ? launch_t<float, float>(etype, law, A, stream)
              : launch_t<double, double>(etype, law, A, stream);
}

template <int LAW>
static void launch_batch_law(int etype, int block, const BatchElemArgs &A,
                             hipStream_t s) {
  dim3 grid((unsigned)A.ntiles);
  if (etype == 0) {
    if (A.blks) // block-partial assembly: one tile per workgroup; records,
                // then 1024 corner slots
      hipLaunchKernelGGL(k_batch_tet4_bpa<LAW>, dim3((unsigned)A.ntiles), dim3(block),
                         (size_t)A.bmaxu * 48 + 1024 * 12, s, A);
    else
      hipLaunchKernelGGL(k_batch_tet4_f32<LAW>, grid, dim3(block), 0, s, A);
  } else if (A.blks) { // block-partial assembly: 256 lanes, LDS corner slots
    hipLaunchKernelGGL(k_batch_hex8_bpa<LAW>, grid, dim3(256), 0, s, A);
  } else if (block == 64) {
    hipLaunchKernelGGL((k_batch_hex8_f32<LAW, 8>), grid, dim3(64), 0, s, A);
  } else if (block == 128) {
    hipLaunchKernelGGL((k_batch_hex8_f32<LAW, 16>), grid, dim3(128), 0, s, A);
  } else {
    hipLaunchKernelGGL((k_batch_hex8_f32<LAW, 32>), grid, dim3(256), 0, s, A);
  }
}

hipError_t launch_batch_elements_f32(int etype, int law, int block,
                                     const BatchElemArgs &A, hipStream_t stream) {
  if (block != 64 && block != 128 && block != 256)
    return hipErrorInvalidValue;
  switch (law) {
  case LAW_ELASTIC: launch_batch_law<LAW_ELASTIC>(etype, block, A, stream); break;
  case LAW_ELASTIC_FD: launch_batch_law<LAW_ELASTIC_FD>(etype, block, A, stream); break;
  case LAW_NEOHOOKEAN: launch_batch_law<LAW_NEOHOOKEAN>(etype, block, A, stream); break;
  case LAW_PLASTIC: launch_batch_law<LAW_PLASTIC>(etype, block, A, stream); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace akxf
