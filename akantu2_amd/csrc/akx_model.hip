// akx_model.hip -- host runtime and C ABI of the MI355X explicit internal
// force path (declared in include/akantu_mi355x.h).
//
// One akx_model = one SolidMechanicsModel on one device and one stream
// (default priority).  The per-step work is two launches:
//   element pass (one kernel per material)  ->  nodal gather + update
// with the predictor of step k+1 fused into the gather of step k when
// several steps are queued together.
#include "../../include/akantu_mi355x.h"
#include "akx_fast.hpp"
#include "akx_host.hpp"
#include "akx_kernels.hpp"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace akx {

__constant__ Hex8Tables c_hex8;

thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// Layout kernels (array-of-structures at the API, structure-of-arrays on the
// device)
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_aos_to_soa(const T *aos, T *soa, int64_t n, int comps) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * comps)
    return;
  int64_t r = i / comps, c = i % comps;
  soa[c * n + r] = aos[i];
}

template <class T>
__global__ void k_soa_to_aos(const T *soa, T *aos, int64_t n, int comps) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * comps)
    return;
  int64_t r = i / comps, c = i % comps;
  aos[i] = soa[c * n + r];
}

__global__ void k_bits_from_bool3(const uint8_t *b3, uint8_t *bits, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  bits[i] = (b3[3 * i] ? 1 : 0) | (b3[3 * i + 1] ? 2 : 0) | (b3[3 * i + 2] ? 4 : 0);
}

__global__ void k_bool3_from_bits(const uint8_t *bits, uint8_t *b3, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  uint8_t b = bits[i];
  b3[3 * i] = b & 1;
  b3[3 * i + 1] = (b >> 1) & 1;
  b3[3 * i + 2] = (b >> 2) & 1;
}

__global__ void k_mass_to3(const double *m, double *m3, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  m3[3 * i] = m[i];
  m3[3 * i + 1] = m[i];
  m3[3 * i + 2] = m[i];
}

// streaming copy used to calibrate the HBM byte counters and to report the
// measured copy bandwidth next to the 8 TB/s peak
// (two 32-byte items per thread, both loads issued before the stores: the
// fastest plain copy shape measured on MI355X)
__global__ void __launch_bounds__(256) k_copy_calibration(const double4 *src,
                                                          double4 *dst,
                                                          int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 512 + threadIdx.x;
  double4 v0, v1;
  if (i < n)
    v0 = src[i];
  if (i + 256 < n)
    v1 = src[i + 256];
  if (i < n)
    dst[i] = v0;
  if (i + 256 < n)
    dst[i + 256] = v1;
}

__global__ void k_fill_f32(float *p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}
__global__ void k_fill_f64(double *p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// FP32 copies of the per-element Lame constants (the FP32 form's reads)
__global__ void k_lame_f32(const double *lam, const double *mu, float2 *lm, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    lm[i] = make_float2(float(lam[i]), float(mu[i]));
}

// Lame constants per element from E, nu (material_elastic.cc:53-62)
__global__ void k_lame(const double *E, const double *nu, double E0, double nu0,
                       double *lam, double *mu, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double l, m;
  lame(E ? E[i] : E0, nu ? nu[i] : nu0, l, m);
  lam[i] = l;
  mu[i] = m;
}

// ---------------------------------------------------------------------------
// Init kernels: jacobian check, lumped mass, stable time step
// ---------------------------------------------------------------------------
struct InitArgs {
  int64_t n, nel, first, pos_offset, nn;
  const int32_t *filter, *conn;
  const double *X, *u;
  double rho;
  double lambda, mu;
  const double *lam_el, *mu_el;
  const int32_t *pos; // [slot] entry of the node-major table
  double *me;         // [entry] mass contributions
  unsigned long long *min_bits;
  long long *first_bad;
};

template <int ET>
__device__ __forceinline__ void load_element_X(const InitArgs &A, int64_t e,
                                               bool add_u,
                                               double (&X)[Elem<ET>::nnpe][3]) {
#pragma unroll
  for (int n = 0; n < Elem<ET>::nnpe; ++n) {
    int32_t node = A.conn[e * Elem<ET>::nnpe + n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double x = A.X[(int64_t)c * A.nn + node];
      X[n][c] = add_u ? x + A.u[(int64_t)c * A.nn + node] : x;
    }
  }
}

// IntegratorGauss::checkJacobians over every element of the mesh
template <int ET> __global__ void k_check_jacobians(InitArgs A) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.nel)
    return;
  double X[NN][3];
  load_element_X<ET>(A, e, false, X);
  for (int q = 0; q < NQ; ++q) {
    double Ji[3][3], detJ;
    Shape<ET, double>::jacobian_at(X, q, Ji, detJ);
    if (detJ < 0.) {
      atomicMin(A.first_bad, (long long)(e * NQ + q));
      return;
    }
  }
}

// J^-1 and w det J of every Hex8 Gauss point of the mesh, one thread per
// Gauss point, in the arithmetic of the element pass (Shape::jacobian_at:
// J = dN/dxi X summed over the nodes in order, the LU inverse, det J * w).
__global__ void __launch_bounds__(256) k_hex8_jacobian_table(const int32_t *conn,
                                                             const double *X, int64_t nn,
                                                             int64_t nel, double2 *jw) {
  const int64_t iq = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nq = nel * 8;
  if (iq >= nq)
    return;
  const int64_t e = iq >> 3;
  const int q = (int)(iq & 7);
  double Xe[8][3];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int32_t nd = conn[e * 8 + n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      Xe[n][c] = X[(int64_t)c * nn + nd];
  }
  double Ji[3][3], detJ;
  const double wJ = Shape<ET_HEX8, double>::jacobian_at(Xe, q, Ji, detJ);
  jw[iq] = make_double2(Ji[0][0], Ji[0][1]);
  jw[nq + iq] = make_double2(Ji[0][2], Ji[1][0]);
  jw[2 * nq + iq] = make_double2(Ji[1][1], Ji[1][2]);
  jw[3 * nq + iq] = make_double2(Ji[2][0], Ji[2][1]);
  jw[4 * nq + iq] = make_double2(Ji[2][2], wJ);
}

// {X, u} node records of the Tet4 element pass from the [3][nn] arrays:
// u only (refresh after an upload of u) or X and u (init)
__global__ void __launch_bounds__(256) k_fill_xu(const double *X, const double *u, int64_t nn,
                                                 double *xu, bool with_x) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn)
    return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (with_x)
      xu[6 * n + c] = X[(int64_t)c * nn + n];
    xu[6 * n + 3 + c] = u[(int64_t)c * nn + n];
  }
}

// Lumped mass contributions m_e(a) = sum_q (rho N_a(q)) wJ_q
// (fe_engine_template_tmpl_field.hh:112-138)
template <int ET> __global__ void k_mass_elements(InitArgs A) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n)
    return;
  int64_t e = A.filter ? (int64_t)A.filter[i] : A.first + i;
  double X[NN][3];
  load_element_X<ET>(A, e, false, X);
  double m[NN];
#pragma unroll
  for (int n = 0; n < NN; ++n)
    m[n] = 0.;
  for (int q = 0; q < NQ; ++q) {
    double Ji[3][3], detJ;
    double wJ = Shape<ET, double>::jacobian_at(X, q, Ji, detJ);
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      double N = (ET == ET_TET4) ? .25 : c_hex8.N[q][n];
      m[n] += (A.rho * N) * wJ;
    }
  }
#pragma unroll
  for (int n = 0; n < NN; ++n)
    A.me[A.pos[(A.pos_offset + i) * NN + n]] = m[n];
}

// min_e h_e / c_e over one material (solid_mechanics_model.cc:423-473)
template <int ET> __global__ void __launch_bounds__(256) k_stable_dt(InitArgs A) {
  constexpr int NN = Elem<ET>::nnpe;
  __shared__ double red[256];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dt = 1.7976931348623157e308;
  if (i < A.n) {
    int64_t e = A.filter ? (int64_t)A.filter[i] : A.first + i;
    double X[NN][3];
    load_element_X<ET>(A, e, true, X);
    double lambda = A.lam_el ? A.lam_el[i] : A.lambda;
    double mu = A.mu_el ? A.mu_el[i] : A.mu;
    double h = inradius<ET>(X);
    double c = sqrt((lambda + 2 * mu) / A.rho);
    dt = h / c;
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // positive doubles order like their bit patterns
    atomicMin(A.min_bits, (unsigned long long)__double_as_longlong(red[0]));
  }
}

// Lumped mass gather: M = 0 + sum 1 * m_e over the node's entries in order
__global__ void k_gather_mass(int64_t nn, const int64_t *ell_off,
                              const uint8_t *deg, const double *me,
                              double *mass) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn)
    return;
  double m = 0.;
  const double *p = me + ell_off[n >> 8] + (n & 255);
  for (int d = 0; d < deg[n]; ++d)
    m += 1. * p[d * 256];
  mass[n] = m;
}

// ---------------------------------------------------------------------------
// Energies (solid_mechanics_model.cc:476-606, material.cc:604-649,
// material_plastic.cc:47-99).  Per-node or per-element values, then one
// deterministic tree sum (fixed shape, so repeated calls give the same bits).
// ---------------------------------------------------------------------------
template <class R> __device__ __forceinline__ R ddot3(const R (&a)[3][3], const R (&b)[3][3]) {
  R s = a[0][0] * b[0][0]; // cwiseProduct().sum(), column-major order
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (i + j > 0)
        s += a[i][j] * b[i][j];
  return s;
}

// Potential energy density of one Gauss point:
//   elastic          .5 sigma : grad u  (material_elastic_inline_impl.hh:113-116)
//   elastic finite   .5 S : E(grad u)   (material_elastic.cc:138-160)
//   neo-Hookean      lambda (ln J)^2 + mu (-ln J + (tr C - 3) / 2) / 2, as the
//                    reference writes it (material_neohookean_inline_impl.hh
//                    :119-126)
//   plastic          .5 sigma : (grad u - eps_p)  (material_plastic.cc:71-83)
template <int LAW>
__device__ __forceinline__ double epot_density(double lambda, double mu,
                                               const double (&g)[3][3],
                                               const double (&S)[3][3],
                                               const double (&ep)[3][3]) {
  if constexpr (LAW == LAW_ELASTIC) {
    return .5 * ddot3(S, g);
  } else if constexpr (LAW == LAW_ELASTIC_FD) {
    double E[3][3];
    green_strain(g, E);
    return .5 * ddot3(S, E);
  } else if constexpr (LAW == LAW_NEOHOOKEAN) {
    double F[3][3], C[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        F[i][j] = (i == j) ? 1. + g[i][j] : g[i][j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double c = F[0][i] * F[0][j];
        c += F[1][i] * F[1][j];
        c += F[2][i] * F[2][j];
        C[i][j] = c;
      }
    const double lj = log(det3(F));
    const double trC = C[0][0] + C[1][1] + C[2][2];
    return lambda * (lj * lj) + mu * (-lj + (trC - 3) / 2.) / 2.;
  } else {
    double el[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        el[i][j] = g[i][j] - ep[i][j];
    return .5 * ddot3(S, el);
  }
}

struct EnergyArgs {
  int64_t n, first, nn, nq_all;
  const int32_t *filter, *conn;
  const double *X;
  double lambda, mu;
  const double *lam_el, *mu_el;
  const double *g, *s, *eps; // [9][nq_all] current grad u, stress, plastic strain
  const double *pe;          // [nq_all] accumulated plastic energy, or null
  double *out;               // [n] integral over each element
};

// FEEngine::integrate of the density: sum_q f_q wJ_q of the i-th element of
// the material (integrator_gauss_inline_impl.hh:401-441), wJ as at init
template <int ET, int LAW>
__device__ __forceinline__ double element_energy(const EnergyArgs &A, int64_t i) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  const int64_t e = A.filter ? (int64_t)A.filter[i] : A.first + i;
  double X[NN][3];
#pragma unroll
  for (int n = 0; n < NN; ++n) {
    const int32_t node = A.conn[e * NN + n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      X[n][c] = A.X[(int64_t)c * A.nn + node];
  }
  const double lambda = A.lam_el ? A.lam_el[i] : A.lambda;
  const double mu = A.mu_el ? A.mu_el[i] : A.mu;
  double sum = 0.;
  for (int q = 0; q < NQ; ++q) {
    double Ji[3][3], detJ;
    const double wJ = Shape<ET, double>::jacobian_at(X, q, Ji, detJ);
    const int64_t iq = i * NQ + q;
    double f;
    if (A.pe) {
      f = A.pe[iq];
    } else {
      double g[3][3], S[3][3], ep[3][3];
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
          const int64_t o = (int64_t)(ii + 3 * jj) * A.nq_all + iq;
          g[ii][jj] = A.g[o];
          S[ii][jj] = A.s[o];
          ep[ii][jj] = A.eps ? A.eps[o] : 0.;
        }
      f = epot_density<LAW>(lambda, mu, g, S, ep);
    }
    sum = (q == 0) ? f * wJ : sum + f * wJ;
  }
  return sum;
}

template <int ET, int LAW> __global__ void k_energy_elements(EnergyArgs A) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n)
    return;
  A.out[i] = element_energy<ET, LAW>(A, i);
}

// kinetic energy per node: sum_i v_i v_i m (m > epsilon) (solid_mechanics_model
// .cc:476-522); external work per node: blocked ? -f_int v : f_ext v
// (solid_mechanics_model.cc:545-583)
struct NodeEnergyArgs {
  int64_t nn;
  int kind; // 0 kinetic, 1 external work
  const double *v, *mass, *fint, *fext;
  const uint8_t *blocked;
};

__device__ __forceinline__ double node_energy(const NodeEnergyArgs &A, int64_t n) {
  double w = 0.;
  for (int c = 0; c < 3; ++c) {
    const int64_t o = (int64_t)c * A.nn + n;
    if (A.kind == 0) {
      const double m = A.mass[n];
      if (m > 2.220446049250313e-16)
        w += A.v[o] * A.v[o] * m;
    } else if ((A.blocked[n] >> c) & 1) {
      w -= A.fint[o] * A.v[o];
    } else {
      w += (A.fext ? A.fext[o] : 0.) * A.v[o];
    }
  }
  return w;
}

__global__ void k_node_energy(NodeEnergyArgs A, double *out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.nn)
    return;
  out[n] = node_energy(A, n);
}

// One level of the tree sum: block b sums the values of items b*2048 ..
// b*2048+2047, thread t first adding items t, t+256, .., t+1792 in that order
// from 0, then the 256 thread sums pairwise in shared memory.  value(i) reads
// a stored value or computes it in place; both give the same bits.
template <class F>
__device__ __forceinline__ void block_tree_sum(int64_t n, double *out, F &&value) {
  __shared__ double red[256];
  const int64_t base = (int64_t)blockIdx.x * 2048;
  double s = 0.;
  for (int k = 0; k < 8; ++k) {
    const int64_t i = base + threadIdx.x + 256 * k;
    if (i < n)
      s += value(i);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_tree_sum(const double *in, int64_t n, double *out) {
  block_tree_sum(n, out, [&](int64_t i) { return in[i]; });
}

// First tree level fused with the per-node / per-element values (the
// no-allocation energy): the same sums as k_node_energy / k_energy_elements
// followed by k_tree_sum, without the [n] array of values.
__global__ void __launch_bounds__(256) k_node_energy_sum(NodeEnergyArgs A, double *out) {
  block_tree_sum(A.nn, out, [&](int64_t n) { return node_energy(A, n); });
}

template <int ET, int LAW>
__global__ void __launch_bounds__(256) k_energy_elements_sum(EnergyArgs A, double *out) {
  block_tree_sum(A.n, out, [&](int64_t i) { return element_energy<ET, LAW>(A, i); });
}

} // namespace akx

using namespace akx;

// ---------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------
// A user law compiled at run time into the material's element pass
// (akx_model_set_device_stress): the pass without and with the eigen-gradient
// terms
struct DeviceLaw {
  hipModule_t module = nullptr;
  hipFunction_t kernel[2] = {nullptr, nullptr};
  DeviceLaw() = default;
  DeviceLaw(const DeviceLaw &) = delete;
  DeviceLaw &operator=(const DeviceLaw &) = delete;
  ~DeviceLaw() {
    if (module)
      (void)hipModuleUnload(module);
  }
};

struct Material {
  int law = 0;
  double rho = 0, E = 0, nu = 0, sigma_y = 0, h = 0;
  double lambda = 0, mu = 0;
  int64_t n = 0;          // elements in the filter
  bool contiguous = true; // filter == first..first+n-1
  int64_t first = 0;
  int64_t pos_offset = 0; // assembly position of its first element
  std::vector<int32_t> filter_host;
  DevBuf<int32_t> filter;
  DevBuf<double> lam_el, mu_el, sy_el, h_el;
  DevBuf<float2> lm_f; // FP32 form: (float(lambda), float(mu)) per element
  // FP32 Tet4: distinct nodes of each 256-element block and local corners
  DevBuf<int32_t> bl_nodes;
  DevBuf<uint16_t> lconn;
  int32_t bl_maxu = 0;
  std::vector<int32_t> bl_nodes_host; // [blocks][bl_maxu], -1 padded
  // block-partial assembly (FP32 Tet4): LDS slot of every corner, slot
  // offsets of every distinct node, table entry of every (block, node) partial
  DevBuf<uint16_t> lslot, bl_soff;
  DevBuf<int32_t> bl_pos;
  // partial-sum assembly of the FP64 Tet4 pass: per-corner codes, per-block
  // summed runs (slot ends, table entries)
  DevBuf<int32_t> pcode, poff, ppos;
  DevBuf<uint16_t> pend;
  int64_t t_nblk = 0;
  int32_t t_slots = 0;
  bool per_element = false;
  // materialised internals and plastic history [2][9][n*nqp] / [2][n*nqp]
  DevBuf<double> grad, stress; // materialised by compute_all_stresses
  DevBuf<double> hgrad[2], hsig[2], heps[2], hiso[2];
  DevBuf<double> hpe[2]; // accumulated plastic energy [n*nqp]
  int cur = 0; // index of the current plastic state
  // Material::eigengradu [9][n*nqp] (null: zero); MaterialThermal's sigma_th
  // [n*nqp] (null: -0) and, for the plastic law, its history [2][n*nqp]
  DevBuf<double> eig, sth_now, hsth[2];
  std::vector<double> E_el_host, nu_el_host; // per-element E, nu as given
  bool ext() const { return eig.p != nullptr || sth_now.p != nullptr; }
  // a law computed by the caller (the computeStress plugin)
  bool user() const { return law == AKX_LAW_USER || law == AKX_LAW_USER_FD; }
  // the caller's law compiled for the device, or null (host callback)
  std::unique_ptr<DeviceLaw> dlaw;
};

struct akx_model {
  int device = 0;
  hipStream_t stream = nullptr;
  int etype = -1, nnpe = 0, nqp = 0;
  int64_t nn = 0, nel = 0;
  int precision = AKX_PRECISION_FP64;
  bool initialized = false;
  bool has_fext = false;
  double dt = 0.;
  std::vector<int32_t> conn_host; // AoS copy for filters and the gather table
  DevBuf<double> X, u, v, a, fint, fext, mass;
  DevBuf<uint8_t> blocked;
  DevBuf<float> cu, cv;
  DevBuf<int32_t> conn, pos;
  DevBuf<double2> jw;      // Hex8 plain FP64: J^-1, w det J [5][nel * 8]
  DevBuf<double> xu;       // Tet4: [nn][6] {X, u} node records of the element pass
  // the records' u is kept current (by the predictor and the gathers) only
  // while a pass that reads the records runs: the per-element Tet4 passes
  // (filtered materials, user laws, the FP64 fast form, the pipelined
  // schedule's partial launches); the staged passes read X and u
  bool xu_static = false;
  bool records_live() const { return xu_static || (pipe_parts > 1 && pipe_ok); }
  DevBuf<int64_t> ell_off; // [nn/256 + 1]
  DevBuf<uint8_t> deg;     // [nn]
  int64_t ell_total = 0;   // entries of the node-major table
  DevBuf<unsigned char> fe; // [entries][3] of double or float
  // FP32 Tet4 block-partial assembly: the table holds one entry per (block,
  // distinct node of the block) instead of one per element corner
  bool bpa = false;
  DevBuf<int64_t> bp_off;
  DevBuf<uint8_t> bp_deg;
  int64_t bp_total = 0;
  // FP64 Tet4 partial-sum assembly: prefix partials (same bits, fewer
  // entries: each node's first-block sum, then one entry per later corner)
  // or, opted into, block partials (one entry per (block, node))
  bool ppa = false;
  int assembly = AKX_ASSEMBLY_DEFAULT; // akx_model_set_block_assembly
  DevBuf<int64_t> pp_off;
  DevBuf<uint8_t> pp_deg;
  int64_t pp_total = 0;
  int64_t pp_entries = 0, pp_runs = 0; // table entries, summed runs
  DevBuf<unsigned char> staging;
  DevBuf<unsigned long long> scratch_u64;
  DevBuf<long long> scratch_i64;
  std::vector<std::unique_ptr<Material>> mats;
  // computeStress of the user materials, and host copies of their internals
  akx_compute_stress_fn stress_fn = nullptr;
  void *stress_user = nullptr;
  std::vector<double> user_grad, user_stress;
  bool has_user() const {
    for (auto &mp : mats)
      if (mp->user())
        return true;
    return false;
  }
  // pipelined schedule (akx_model_set_pipeline): element parts and the node
  // groups whose gather they complete, a second stream and the events
  int pipe_parts = 0;
  bool pipe_planned = false, pipe_ok = false;
  std::vector<int64_t> pipe_el;   // [P + 1] element ranges of the parts
  std::vector<int64_t> pipe_node; // [P + 1] node ranges of the gather groups
  std::vector<int> pipe_need;     // [P] last group part p's nodes belong to
  hipStream_t stream2 = nullptr;
  std::vector<hipEvent_t> ev_elem, ev_gather[2];
  // kernel timing
  bool timing = false;
  double t_ms[3] = {0, 0, 0};
  int64_t t_n[3] = {0, 0, 0};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[3];
  std::vector<hipEvent_t> event_pool;

  ~akx_model() {
    drain_timing();
    for (auto e : event_pool)
      (void)hipEventDestroy(e);
    for (auto e : ev_elem)
      (void)hipEventDestroy(e);
    for (auto &v : ev_gather)
      for (auto e : v)
        (void)hipEventDestroy(e);
    if (stream2)
      (void)hipStreamDestroy(stream2);
    if (stream)
      (void)hipStreamDestroy(stream);
  }

  void set_device() { HIPCHECK(hipSetDevice(device)); }

  hipEvent_t get_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }

  void drain_timing() {
    for (int k = 0; k < 3; ++k) {
      for (auto &pr : pending[k]) {
        if (hipEventSynchronize(pr.second) == hipSuccess) {
          float ms = 0.f;
          if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            t_ms[k] += ms;
            t_n[k] += 1;
          }
        }
        event_pool.push_back(pr.first);
        event_pool.push_back(pr.second);
      }
      pending[k].clear();
    }
  }

  template <class F> void timed(int cls, F &&launch) {
    if (!timing) {
      launch();
      return;
    }
    hipEvent_t e0 = get_event(), e1 = get_event();
    HIPCHECK(hipEventRecord(e0, stream));
    launch();
    HIPCHECK(hipEventRecord(e1, stream));
    pending[cls].push_back({e0, e1});
    if (pending[cls].size() > 4096)
      drain_timing();
  }

  size_t fe_elem_size() const {
    return precision == AKX_PRECISION_FP32 ? sizeof(float) : sizeof(double);
  }
};

namespace {

constexpr int TPB = 256;

// device address of the no-allocation energy's scratch array (set up once
// per device)
double *device_scratch(int device);

inline unsigned grid_for(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }
inline int64_t padded_blocks(int64_t n) {
  int64_t nb = (n + TPB - 1) / TPB;
  return (nb + 7) / 8 * 8;
}

void upload_hex8_tables() {
  static bool done_device[64] = {false};
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 64 && done_device[dev])
    return;
  Hex8Tables t;
  fill_hex8_tables(t);
  HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_hex8), &t, sizeof(t)));
  HIPCHECK(akxf::upload_hex8_dnds(&t.dnds[0][0][0]));
  if (dev < 64)
    done_device[dev] = true;
}

template <class T>
void upload_soa(akx_model *m, const T *host_aos, T *dev_soa, int64_t n, int comps) {
  size_t bytes = (size_t)n * comps * sizeof(T);
  m->staging.alloc(std::max(m->staging.n, bytes));
  HIPCHECK(hipMemcpyAsync(m->staging.p, host_aos, bytes, hipMemcpyHostToDevice,
                          m->stream));
  hipLaunchKernelGGL(k_aos_to_soa<T>, dim3(grid_for(n * comps)), dim3(TPB), 0,
                     m->stream, (const T *)m->staging.p, dev_soa, n, comps);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(m->stream));
}

template <class T>
void download_soa(akx_model *m, const T *dev_soa, T *host_aos, int64_t n, int comps) {
  size_t bytes = (size_t)n * comps * sizeof(T);
  m->staging.alloc(std::max(m->staging.n, bytes));
  hipLaunchKernelGGL(k_soa_to_aos<T>, dim3(grid_for(n * comps)), dim3(TPB), 0,
                     m->stream, dev_soa, (T *)m->staging.p, n, comps);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(host_aos, m->staging.p, bytes, hipMemcpyDeviceToHost,
                          m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
}

// Tet4 node records: X and u (init) or u alone (after an upload of u)
void refresh_xu(akx_model *m, bool with_x) {
  if (!m->xu.p || m->nn == 0)
    return;
  hipLaunchKernelGGL(k_fill_xu, dim3(grid_for(m->nn)), dim3(TPB), 0, m->stream,
                     (const double *)m->X.p, (const double *)m->u.p, m->nn, m->xu.p, with_x);
  HIPCHECK(hipGetLastError());
}

// Element passes with block lists: for each block of bs consecutive
// elements of a contiguous material, its distinct nodes in order of first
// use (every block's list padded with -1 to the longest) and each element's
// corners as indices into them (the Tet4 passes stage these nodes' records
// in LDS).  In FP32 also, for block-partial assembly, the LDS slot of every
// corner: the corners of distinct node k of a block take consecutive slots
// in ascending element (then corner) order.
void build_block_nodes(akx_model *m, Material &mat, int64_t bs = 256) {
  const int nnpe = m->nnpe;
  const int64_t n = mat.n, nb = (n + bs - 1) / bs;
  std::vector<int32_t> off(nb + 1, 0), nodes;
  std::vector<uint16_t> lc(n * nnpe);
  std::vector<int64_t> stamp(m->nn, -1);
  std::vector<int32_t> local(m->nn, 0);
  nodes.reserve(n);
  int32_t maxu = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t start = (int64_t)nodes.size();
    for (int64_t i = b * bs; i < std::min<int64_t>(n, b * bs + bs); ++i)
      for (int a = 0; a < nnpe; ++a) {
        const int32_t node = m->conn_host[(mat.first + i) * nnpe + a];
        if (stamp[node] != b) {
          stamp[node] = b;
          local[node] = (int32_t)((int64_t)nodes.size() - start);
          nodes.push_back(node);
        }
        lc[i * nnpe + a] = (uint16_t)local[node];
      }
    require((int64_t)nodes.size() < (int64_t)INT32_MAX, "block node lists too large");
    off[b + 1] = (int32_t)nodes.size();
    maxu = std::max<int32_t>(maxu, off[b + 1] - off[b]);
  }
  std::vector<int32_t> padded((size_t)nb * maxu, -1);
  for (int64_t b = 0; b < nb; ++b)
    std::copy(nodes.begin() + off[b], nodes.begin() + off[b + 1],
              padded.begin() + b * maxu);
  mat.bl_nodes.alloc(padded.size());
  mat.lconn.alloc(n * nnpe);
  HIPCHECK(hipMemcpy(mat.bl_nodes.p, padded.data(), padded.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(mat.lconn.p, lc.data(), lc.size() * sizeof(uint16_t),
                     hipMemcpyHostToDevice));
  mat.bl_maxu = maxu;
  mat.bl_nodes_host = std::move(padded);
  if (m->precision != AKX_PRECISION_FP32)
    return;
  std::vector<uint16_t> soff((size_t)nb * (maxu + 1), 0), slot(n * nnpe);
  std::vector<uint16_t> cnt(maxu + 1);
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t e0 = b * bs, e1 = std::min<int64_t>(n, b * bs + bs);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = e0 * nnpe; i < e1 * nnpe; ++i)
      cnt[lc[i]]++;
    uint16_t *o = &soff[b * (maxu + 1)];
    o[0] = 0;
    for (int k = 0; k < maxu; ++k)
      o[k + 1] = (uint16_t)(o[k] + cnt[k]);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = e0 * nnpe; i < e1 * nnpe; ++i)
      slot[i] = (uint16_t)(o[lc[i]] + cnt[lc[i]]++);
  }
  mat.lslot.alloc(slot.size());
  mat.bl_soff.alloc(soff.size());
  HIPCHECK(hipMemcpy(mat.lslot.p, slot.data(), slot.size() * sizeof(uint16_t),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(mat.bl_soff.p, soff.data(), soff.size() * sizeof(uint16_t),
                     hipMemcpyHostToDevice));
}

// Gather table of the block-partial assembly: node n's entries are the
// partials of the (material, block) pairs whose elements touch it, ranked
// in material order, then ascending block -- a fixed order, so repeated
// runs give the same bits.  Same node-major sliced-ELL layout as the
// per-element table.
void build_bpa_table(akx_model *m) {
  std::vector<int32_t> deg(m->nn, 0);
  for (auto &mp : m->mats)
    for (int32_t node : mp->bl_nodes_host)
      if (node >= 0)
        deg[node]++;
  const int64_t nslices = (m->nn + 255) / 256;
  std::vector<int64_t> off(nslices + 1, 0);
  for (int64_t sl = 0; sl < nslices; ++sl) {
    int w = 0;
    for (int64_t nd = sl * 256; nd < std::min<int64_t>(m->nn, sl * 256 + 256); ++nd)
      w = std::max(w, deg[nd]);
    require(w <= 255, "a node belongs to more than 255 element blocks");
    off[sl + 1] = off[sl] + (int64_t)w * 256;
  }
  m->bp_total = off[nslices];
  require(m->bp_total < (int64_t)INT32_MAX, "block-partial table too large for int32");
  std::vector<int32_t> fill(m->nn, 0);
  for (auto &mp : m->mats) {
    std::vector<int32_t> pos(mp->bl_nodes_host.size(), -1);
    for (size_t k = 0; k < pos.size(); ++k) {
      const int32_t node = mp->bl_nodes_host[k];
      if (node < 0)
        continue;
      const int32_t d = fill[node]++;
      pos[k] = (int32_t)(off[node >> 8] + (int64_t)d * 256 + (node & 255));
    }
    mp->bl_pos.alloc(pos.size());
    HIPCHECK(hipMemcpy(mp->bl_pos.p, pos.data(), pos.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
  }
  std::vector<uint8_t> deg8(m->nn);
  for (int64_t nd = 0; nd < m->nn; ++nd)
    deg8[nd] = (uint8_t)deg[nd];
  m->bp_off.alloc(off.size());
  m->bp_deg.alloc(std::max<int64_t>(m->nn, 1));
  HIPCHECK(hipMemcpy(m->bp_off.p, off.data(), off.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(m->bp_deg.p, deg8.data(), deg8.size(), hipMemcpyHostToDevice));
}

// Prefix- or block-partial assembly of the FP64 Tet4 path (k_tet4_ppa) on
// the 256-element blocks of the staged node records (build_block_nodes).
// Node n's corners, in assembly order (materials in order, elements
// ascending), come in runs: the corners in one block of one material.
// full = false: the node's first run is summed in LDS in that order (the
// gather then adds the same values in the same order: same bits) and every
// later corner keeps an entry of its own.  full = true: every run is summed,
// one entry per (block, node), added by the gather in block order (other
// rounding).  The table keeps the node-major sliced-ELL layout.  Requires
// every material contiguous and its block lists built.
void build_ppa(akx_model *m, bool full) {
  const int64_t TB = PPA_TB;
  const size_t nm = m->mats.size();
  std::vector<int32_t> nent(m->nn, 0), seen(m->nn, 0), base(m->nn, 0), cnt(m->nn, 0);
  std::vector<int64_t> stamp(m->nn, -1), entry0(m->nn, 0);
  std::vector<uint8_t> summed(m->nn, 0);
  std::vector<int64_t> off;
  std::vector<int32_t> nodes;
  int32_t maxslots = 0;
  // one traversal in assembly order; the second also writes the tables
  auto traverse = [&](bool write, std::vector<std::vector<int32_t>> *code,
                      std::vector<std::vector<int32_t>> *poff,
                      std::vector<std::vector<uint16_t>> *pend,
                      std::vector<std::vector<int32_t>> *ppos) {
    std::fill(nent.begin(), nent.end(), 0);
    std::fill(stamp.begin(), stamp.end(), -1);
    int64_t gb = 0;
    for (size_t k = 0; k < nm; ++k) {
      const Material &mat = *m->mats[k];
      const int64_t nb = (mat.n + TB - 1) / TB;
      if (write) {
        (*code)[k].assign(mat.n * 4, 0);
        (*poff)[k].assign(nb + 1, 0);
        (*pend)[k].clear();
        (*ppos)[k].clear();
      }
      for (int64_t b = 0; b < nb; ++b, ++gb) {
        const int64_t e0 = b * TB, e1 = std::min<int64_t>(mat.n, e0 + TB);
        nodes.clear();
        for (int64_t i = e0; i < e1; ++i)
          for (int a = 0; a < 4; ++a) {
            const int32_t node = m->conn_host[(mat.first + i) * 4 + a];
            if (stamp[node] != gb) {
              stamp[node] = gb;
              cnt[node] = 0;
              nodes.push_back(node);
            }
            cnt[node]++;
          }
        int32_t slots = 0;
        for (int32_t node : nodes) {
          summed[node] = full || nent[node] == 0;
          seen[node] = 0;
          entry0[node] = nent[node];
          if (summed[node]) {
            base[node] = slots;
            slots += cnt[node];
            if (write) {
              (*pend)[k].push_back((uint16_t)slots);
              (*ppos)[k].push_back(
                  (int32_t)(off[node >> 8] + (int64_t)nent[node] * 256 + (node & 255)));
            }
            nent[node] += 1;
          } else {
            nent[node] += cnt[node];
          }
        }
        maxslots = std::max(maxslots, slots);
        if (!write)
          continue;
        (*poff)[k][b + 1] = (int32_t)(*pend)[k].size();
        for (int64_t i = e0; i < e1; ++i)
          for (int a = 0; a < 4; ++a) {
            const int32_t node = m->conn_host[(mat.first + i) * 4 + a];
            const int32_t j = seen[node]++;
            (*code)[k][i * 4 + a] =
                summed[node] ? -1 - (base[node] + j)
                             : (int32_t)(off[node >> 8] + (entry0[node] + j) * 256 + (node & 255));
          }
      }
    }
  };
  traverse(false, nullptr, nullptr, nullptr, nullptr);
  const int64_t nslices = (m->nn + 255) / 256;
  off.assign(nslices + 1, 0);
  std::vector<uint8_t> deg8(m->nn, 0);
  for (int64_t sl = 0; sl < nslices; ++sl) {
    int w = 0;
    for (int64_t nd = sl * 256; nd < std::min<int64_t>(m->nn, sl * 256 + 256); ++nd) {
      require(nent[nd] <= 255, "a node has more than 255 table entries");
      deg8[nd] = (uint8_t)nent[nd];
      w = std::max<int>(w, nent[nd]);
    }
    off[sl + 1] = off[sl] + (int64_t)w * 256;
  }
  require(off[nslices] < (int64_t)INT32_MAX, "partial-sum table too large for int32");
  std::vector<std::vector<int32_t>> code(nm), poff(nm), ppos(nm);
  std::vector<std::vector<uint16_t>> pend(nm);
  traverse(true, &code, &poff, &pend, &ppos);
  for (size_t k = 0; k < nm; ++k) {
    Material &mat = *m->mats[k];
    mat.t_nblk = (mat.n + TB - 1) / TB;
    mat.t_slots = maxslots;
    const size_t np = pend[k].size();
    mat.pcode.alloc(mat.n * 4);
    mat.poff.alloc(poff[k].size());
    mat.pend.alloc(std::max<size_t>(np, 1));
    mat.ppos.alloc(std::max<size_t>(np, 1));
    HIPCHECK(hipMemcpy(mat.pcode.p, code[k].data(), mat.n * 4 * sizeof(int32_t),
                       hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(mat.poff.p, poff[k].data(), poff[k].size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
    if (np > 0) {
      HIPCHECK(hipMemcpy(mat.pend.p, pend[k].data(), np * sizeof(uint16_t),
                         hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(mat.ppos.p, ppos[k].data(), np * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    }
  }
  m->pp_total = off[nslices];
  m->pp_entries = 0;
  for (int64_t nd = 0; nd < m->nn; ++nd)
    m->pp_entries += nent[nd];
  m->pp_runs = 0;
  for (size_t k = 0; k < nm; ++k)
    m->pp_runs += (int64_t)pend[k].size();
  m->pp_off.alloc(off.size());
  m->pp_deg.alloc(std::max<int64_t>(m->nn, 1));
  HIPCHECK(hipMemcpy(m->pp_off.p, off.data(), off.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(m->pp_deg.p, deg8.data(), deg8.size(), hipMemcpyHostToDevice));
}

// FP32 form: (float(lambda), float(mu)) pairs of a material with per-element
// constants, the values its element pass rounds the FP64 pair to
void fill_lame_f32(akx_model *m, Material &mat) {
  mat.lm_f.alloc(mat.n);
  hipLaunchKernelGGL(k_lame_f32, dim3(grid_for(mat.n)), dim3(TPB), 0, m->stream,
                     (const double *)mat.lam_el.p, (const double *)mat.mu_el.p,
                     mat.lm_f.p, mat.n);
  HIPCHECK(hipGetLastError());
}

void zero_async(akx_model *m, double *p, int64_t n) {
  if (n <= 0)
    return;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(TPB), 0, m->stream, p, n, 0.);
  HIPCHECK(hipGetLastError());
}

// ------------------------------------------------------------- launches ----
ElemArgs elem_args(akx_model *m, Material &mat, bool materialize) {
  ElemArgs A{};
  A.n = mat.n;
  A.nel = m->nel;
  A.filter = mat.contiguous ? nullptr : mat.filter.p;
  A.first = mat.first;
  A.pos_offset = mat.pos_offset;
  A.conn = m->conn.p;
  A.X = m->X.p;
  A.u = m->u.p;
  A.nn = m->nn;
  A.lambda = mat.lambda;
  A.mu = mat.mu;
  A.sigma_y = mat.sigma_y;
  A.h = mat.h;
  A.lam_el = mat.per_element ? mat.lam_el.p : nullptr;
  A.mu_el = mat.per_element ? mat.mu_el.p : nullptr;
  A.lm_f = mat.per_element ? mat.lm_f.p : nullptr;
  A.bl_nodes = mat.bl_nodes.p;
  A.lconn = mat.lconn.p;
  A.bl_maxu = mat.bl_maxu;
  if (m->bpa) {
    A.lslot = mat.lslot.p;
    A.bl_soff = mat.bl_soff.p;
    A.bl_pos = mat.bl_pos.p;
  }
  if (m->ppa) {
    A.pcode = mat.pcode.p;
    A.poff = mat.poff.p;
    A.pend = mat.pend.p;
    A.ppos = mat.ppos.p;
    A.t_nblk = mat.t_nblk;
    A.t_slots = mat.t_slots;
  }
  A.sy_el = mat.sy_el.p;
  A.h_el = mat.h_el.p;
  A.pos = m->pos.p;
  A.fe = m->fe.p;
  A.jw = m->jw.p;
  A.xu = m->xu.p;
  if (mat.user()) {
    A.stress = mat.stress.p; // read by the force pass
  } else if (materialize && mat.law != AKX_LAW_PLASTIC_LIH) {
    A.grad_u = mat.grad.p;
    A.stress = mat.stress.p;
  }
  if (mat.law == AKX_LAW_PLASTIC_LIH) {
    int p = mat.cur, c = 1 - mat.cur;
    A.p_grad = mat.hgrad[p].p;
    A.p_sig = mat.hsig[p].p;
    A.p_eps = mat.heps[p].p;
    A.p_iso = mat.hiso[p].p;
    A.c_grad = mat.hgrad[c].p;
    A.c_sig = mat.hsig[c].p;
    A.c_eps = mat.heps[c].p;
    A.c_iso = mat.hiso[c].p;
    A.p_pe = mat.hpe[p].p;
    A.c_pe = mat.hpe[c].p;
    if (mat.sth_now.p) {
      A.sth_p = mat.hsth[p].p;
      A.sth_c = mat.hsth[c].p;
    }
  }
  A.eig = mat.eig.p;
  A.sth_now = mat.sth_now.p;
  A.nblocks_padded = padded_blocks(mat.n);
  return A;
}

// Host-side checks that every buffer a kernel variant dereferences exists
// (a kernel must never be launched with a null pointer it will use).
void check_gather_args(const GatherArgs &G, bool update, bool comp) {
  require(G.ell_off && G.deg && G.fe, "gather launched without its table");
  require(G.fint || update, "a gather without update must store f_int");
  if (update)
    require(G.mass && G.blocked && G.u && G.v && G.a,
            "nodal update launched without nodal state");
  if (comp)
    require(G.cu && G.cv, "compensated update launched without compensation terms");
}

void check_elem_args(const ElemArgs &A, int law) {
  require(A.conn && A.X && A.u && A.fe && A.pos,
          "element pass launched without mesh or state");
  require(A.n <= 0 || A.nblocks_padded > 0, "element pass launched with an empty grid");
  if (law == AKX_LAW_PLASTIC_LIH)
    require(A.p_grad && A.p_sig && A.p_eps && A.p_iso && A.c_grad && A.c_sig &&
                A.c_eps && A.c_iso,
            "plastic element pass launched without history buffers");
  if (A.lam_el || A.mu_el)
    require(A.lam_el && A.mu_el, "per-element parameters incomplete");
  if (law == AKX_LAW_USER || law == AKX_LAW_USER_FD)
    require(A.stress != nullptr, "user material without its stress array");
  else if (A.grad_u || A.stress)
    require(A.grad_u && A.stress, "materialised internals incomplete");
}

template <int ET, class CT, class FT, bool EXT>
void launch_elem_law_x(akx_model *m, Material &mat, const ElemArgs &A) {
  dim3 grid((unsigned)A.nblocks_padded), block(TPB);
  if constexpr (ET == ET_TET4) {
    if (A.lconn && TPB == 256) { // block node lists of a whole contiguous material
      const size_t lds = (size_t)A.bl_maxu * 48;
      switch (mat.law) {
      case AKX_LAW_ELASTIC:
        hipLaunchKernelGGL((k_tet4_strict_lds<LAW_ELASTIC, CT, FT, EXT>), grid, block, lds,
                           m->stream, A);
        HIPCHECK(hipGetLastError());
        return;
      case AKX_LAW_ELASTIC_FD:
        hipLaunchKernelGGL((k_tet4_strict_lds<LAW_ELASTIC_FD, CT, FT, EXT>), grid, block, lds,
                           m->stream, A);
        HIPCHECK(hipGetLastError());
        return;
      case AKX_LAW_NEOHOOKEAN:
        hipLaunchKernelGGL((k_tet4_strict_lds<LAW_NEOHOOKEAN, CT, FT, EXT>), grid, block, lds,
                           m->stream, A);
        HIPCHECK(hipGetLastError());
        return;
      case AKX_LAW_PLASTIC_LIH:
        hipLaunchKernelGGL((k_tet4_strict_lds<LAW_PLASTIC, CT, FT, EXT>), grid, block, lds,
                           m->stream, A);
        HIPCHECK(hipGetLastError());
        return;
      default: // user laws: the per-element pass
        break;
      }
    }
  }
  switch (mat.law) {
  case AKX_LAW_ELASTIC:
    hipLaunchKernelGGL((k_element_forces<ET, LAW_ELASTIC, CT, FT, EXT>), grid, block,
                       0, m->stream, A);
    break;
  case AKX_LAW_ELASTIC_FD:
    hipLaunchKernelGGL((k_element_forces<ET, LAW_ELASTIC_FD, CT, FT, EXT>), grid,
                       block, 0, m->stream, A);
    break;
  case AKX_LAW_NEOHOOKEAN:
    hipLaunchKernelGGL((k_element_forces<ET, LAW_NEOHOOKEAN, CT, FT, EXT>), grid,
                       block, 0, m->stream, A);
    break;
  case AKX_LAW_PLASTIC_LIH:
    hipLaunchKernelGGL((k_element_forces<ET, LAW_PLASTIC, CT, FT, EXT>), grid, block,
                       0, m->stream, A);
    break;
  case AKX_LAW_USER:
    hipLaunchKernelGGL((k_element_forces<ET, LAW_USER, CT, FT, EXT>), grid, block, 0,
                       m->stream, A);
    break;
  case AKX_LAW_USER_FD:
    hipLaunchKernelGGL((k_element_forces<ET, LAW_USER_FD, CT, FT, EXT>), grid, block, 0,
                       m->stream, A);
    break;
  default:
    throw Error(AKX_ERR_UNSUPPORTED, "unknown law");
  }
  HIPCHECK(hipGetLastError());
}

template <int ET, class CT, class FT>
void launch_elem_law(akx_model *m, Material &mat, const ElemArgs &A) {
  if (mat.ext())
    launch_elem_law_x<ET, CT, FT, true>(m, mat, A);
  else
    launch_elem_law_x<ET, CT, FT, false>(m, mat, A);
}

template <bool EXT>
void launch_hex8_strict_x(akx_model *m, Material &mat, const ElemArgs &A) {
  require(A.jw != nullptr, "Hex8 element pass without its jacobian table");
  dim3 grid((unsigned)A.nblocks_padded), block(64);
  switch (mat.law) {
  case AKX_LAW_ELASTIC:
    hipLaunchKernelGGL((k_hex8_strict<LAW_ELASTIC, EXT>), grid, block, 0, m->stream, A);
    break;
  case AKX_LAW_ELASTIC_FD:
    hipLaunchKernelGGL((k_hex8_strict<LAW_ELASTIC_FD, EXT>), grid, block, 0, m->stream, A);
    break;
  case AKX_LAW_NEOHOOKEAN:
    hipLaunchKernelGGL((k_hex8_strict<LAW_NEOHOOKEAN, EXT>), grid, block, 0, m->stream, A);
    break;
  case AKX_LAW_PLASTIC_LIH:
    hipLaunchKernelGGL((k_hex8_strict<LAW_PLASTIC, EXT>), grid, block, 0, m->stream, A);
    break;
  case AKX_LAW_USER:
    hipLaunchKernelGGL((k_hex8_strict<LAW_USER, EXT>), grid, block, 0, m->stream, A);
    break;
  case AKX_LAW_USER_FD:
    hipLaunchKernelGGL((k_hex8_strict<LAW_USER_FD, EXT>), grid, block, 0, m->stream, A);
    break;
  default:
    throw Error(AKX_ERR_UNSUPPORTED, "unknown law");
  }
  HIPCHECK(hipGetLastError());
}

void launch_hex8_strict(akx_model *m, Material &mat, const ElemArgs &A) {
  if (mat.ext())
    launch_hex8_strict_x<true>(m, mat, A);
  else
    launch_hex8_strict_x<false>(m, mat, A);
}

template <bool EXT>
void launch_tet4_ppa_x(akx_model *m, Material &mat, const ElemArgs &A) {
  dim3 grid((unsigned)A.nblocks_padded), block(PPA_TB);
  const size_t lds = std::max((size_t)A.bl_maxu * 48, (size_t)A.t_slots * 24);
  switch (mat.law) {
  case AKX_LAW_ELASTIC:
    hipLaunchKernelGGL((k_tet4_ppa<LAW_ELASTIC, EXT>), grid, block, lds, m->stream, A);
    break;
  case AKX_LAW_ELASTIC_FD:
    hipLaunchKernelGGL((k_tet4_ppa<LAW_ELASTIC_FD, EXT>), grid, block, lds, m->stream, A);
    break;
  case AKX_LAW_NEOHOOKEAN:
    hipLaunchKernelGGL((k_tet4_ppa<LAW_NEOHOOKEAN, EXT>), grid, block, lds, m->stream, A);
    break;
  case AKX_LAW_PLASTIC_LIH:
    hipLaunchKernelGGL((k_tet4_ppa<LAW_PLASTIC, EXT>), grid, block, lds, m->stream, A);
    break;
  default: // user laws never take this path
    throw Error(AKX_ERR_UNSUPPORTED, "prefix-partial pass for a user law");
  }
  HIPCHECK(hipGetLastError());
}

void launch_tet4_ppa(akx_model *m, Material &mat, const ElemArgs &A) {
  if (mat.ext())
    launch_tet4_ppa_x<true>(m, mat, A);
  else
    launch_tet4_ppa_x<false>(m, mat, A);
}

void launch_device_law(akx_model *m, Material &mat, ElemArgs A, bool materialize);

void launch_element_pass(akx_model *m, bool materialize) {
  const bool strict = m->precision == AKX_PRECISION_FP64;
  for (auto &mp : m->mats) {
    Material &mat = *mp;
    if (mat.n == 0)
      continue;
    ElemArgs A = elem_args(m, mat, materialize);
    check_elem_args(A, mat.law);
    require(!mat.ext() || strict,
            "eigen_grad_u and thermal stresses run on the plain FP64 path");
    require(!(mat.sth_now.p && mat.law == AKX_LAW_PLASTIC_LIH) || (A.sth_p && A.sth_c),
            "plastic thermal stress without its history");
    require(m->etype != AKX_TET4 || A.xu != nullptr, "Tet4 element pass without its node records");
    require(m->fe.n >= (size_t)(m->ell_total * 3 * m->fe_elem_size()) &&
                m->fe.n >= (size_t)(m->bp_total * 3 * m->fe_elem_size()) &&
                m->pos.n >= (size_t)((mat.pos_offset + mat.n) * m->nnpe),
            "element force table too small for this material");
    if (m->bpa)
      require(A.lslot && A.bl_soff && A.bl_pos && A.bl_nodes && A.lconn,
              "block-partial pass launched without its block tables");
    if (strict && m->etype == AKX_HEX8) {
      int64_t nb = (mat.n + HEX_STRICT_EPB - 1) / HEX_STRICT_EPB;
      A.nblocks_padded = (nb + 7) / 8 * 8;
    } else if (!strict) {
      int64_t epb = akxf::elements_per_block(m->etype, m->bpa);
      int64_t nb = (mat.n + epb - 1) / epb;
      A.nblocks_padded = (nb + 7) / 8 * 8;
    }
    if (strict && m->ppa) {
      require(A.pcode && A.poff && A.pend && A.ppos && A.bl_nodes && A.lconn &&
                  A.t_nblk == (mat.n + PPA_TB - 1) / PPA_TB,
              "prefix-partial pass launched without its block tables");
      A.nblocks_padded = (A.t_nblk + 7) / 8 * 8;
    }
    m->timed(0, [&] {
      if (mat.dlaw) {
        require(strict, "device laws run on the plain FP64 path");
        launch_device_law(m, mat, A, materialize);
      } else if (strict && m->ppa) {
        launch_tet4_ppa(m, mat, A);
      } else if (strict) {
        if (m->etype == AKX_TET4)
          launch_elem_law<ET_TET4, double, double>(m, mat, A);
        else
          launch_hex8_strict(m, mat, A);
      } else {
        HIPCHECK(akxf::launch_elements(m->etype, mat.law,
                                       m->precision == AKX_PRECISION_FP32, A,
                                       m->stream));
      }
    });
  }
}

// The computeStress plugin of the user materials (Material::computeStress,
// material.cc:128-154): grad u at every Gauss point of the material (the
// reference order, as the other laws form it), handed to the caller, whose
// stress goes back to the device for the force pass.
void user_stress_phase(akx_model *m) {
  if (!m->has_user())
    return;
  for (int k = 0; k < (int)m->mats.size(); ++k) {
    Material &mat = *m->mats[k];
    if (!mat.user() || mat.n == 0)
      continue;
    if (mat.dlaw) // the law runs inside the element pass
      continue;
    require(m->stress_fn != nullptr,
            "a user material needs a computeStress callback (akx_model_set_stress_callback) "
            "or a device law (akx_model_set_device_stress)");
    const int64_t nq = mat.n * m->nqp;
    ElemArgs A = elem_args(m, mat, false);
    A.grad_u = mat.grad.p;
    A.stress = nullptr;
    require(A.grad_u && A.conn && A.X && A.u, "user material gradient pass without buffers");
    require(m->etype != AKX_TET4 || A.xu != nullptr, "Tet4 gradient pass without its node records");
    if (m->etype == AKX_HEX8) {
      require(A.jw != nullptr, "Hex8 gradient pass without its jacobian table");
      const int64_t nb = (mat.n + HEX_STRICT_EPB - 1) / HEX_STRICT_EPB;
      A.nblocks_padded = (nb + 7) / 8 * 8;
      if (mat.ext())
        hipLaunchKernelGGL((k_hex8_strict<LAW_USER_GRAD, true>),
                           dim3((unsigned)A.nblocks_padded), dim3(64), 0, m->stream, A);
      else
        hipLaunchKernelGGL((k_hex8_strict<LAW_USER_GRAD, false>),
                           dim3((unsigned)A.nblocks_padded), dim3(64), 0, m->stream, A);
    } else {
      if (mat.ext())
        hipLaunchKernelGGL((k_element_forces<ET_TET4, LAW_USER_GRAD, double, double, true>),
                           dim3((unsigned)A.nblocks_padded), dim3(TPB), 0, m->stream, A);
      else
        hipLaunchKernelGGL((k_element_forces<ET_TET4, LAW_USER_GRAD, double, double, false>),
                           dim3((unsigned)A.nblocks_padded), dim3(TPB), 0, m->stream, A);
    }
    HIPCHECK(hipGetLastError());
    m->user_grad.resize(9 * nq);
    m->user_stress.assign(9 * nq, 0.);
    download_soa(m, mat.grad.p, m->user_grad.data(), nq, 9);
    const int rc = m->stress_fn(m->stress_user, k, nq, m->user_grad.data(),
                                m->user_stress.data());
    if (rc != 0)
      throw Error(AKX_ERR_INVALID,
                  "computeStress of material " + std::to_string(k) + " failed");
    upload_soa(m, m->user_stress.data(), mat.stress.p, nq, 9);
  }
}

#include "akx_rtc_sources.inc"

// Compiles the caller's law into the material's element pass with hiprtc for
// the model's GPU -- the library's own kernel headers, the plain path's
// floating-point rules (-O3, no contraction: every multiply and add rounds on
// its own) -- for the pass without and with the eigen-gradient terms, and
// loads it.  The parameters become constants of the program (hexadecimal
// literals: exact).
std::unique_ptr<DeviceLaw> compile_device_law(akx_model *m, const Material &mat,
                                              const char *source, const double *params,
                                              int nparams) {
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, m->device));
  std::string arch(prop.gcnArchName);
  arch = "--offload-arch=" + arch.substr(0, arch.find(':'));
  const int law = mat.law == AKX_LAW_USER_FD ? LAW_DEVICE_FD : LAW_DEVICE;
  std::string src =
      "typedef __hip_internal::int8_t int8_t;\n"
      "typedef __hip_internal::uint8_t uint8_t;\n"
      "typedef __hip_internal::int16_t int16_t;\n"
      "typedef __hip_internal::uint16_t uint16_t;\n"
      "typedef __hip_internal::int32_t int32_t;\n"
      "typedef __hip_internal::uint32_t uint32_t;\n"
      "typedef __hip_internal::int64_t int64_t;\n"
      "typedef __hip_internal::uint64_t uint64_t;\n"
      "#define AKX_DEVICE_LAW 1\n";
  src += "#line 1 \"device law\"\n";
  src += source;
  src += "\n#line 1 \"akx device law program\"\n";
  src += "__constant__ double akx_law_params[" + std::to_string(std::max(nparams, 1)) + "] = {";
  for (int k = 0; k < std::max(nparams, 1); ++k) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", nparams > 0 ? params[k] : 0.);
    src += (k ? ", " : "") + std::string(buf);
  }
  src += "};\n#include \"akx_kernels.hpp\"\nnamespace akx {\n__constant__ Hex8Tables c_hex8;\n}\n";
  std::string names[2];
  for (int ext = 0; ext < 2; ++ext)
    names[ext] = m->etype == AKX_HEX8
                     ? "akx::k_hex8_strict<" + std::to_string(law) + ", " +
                           (ext ? "true" : "false") + ">"
                     : "akx::k_element_forces<0, " + std::to_string(law) + ", double, double, " +
                           (ext ? "true" : "false") + ">";
  const std::string table = "&akx::c_hex8";
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "akx_device_law.hip", 3, k_rtc_headers,
                          k_rtc_header_names) != HIPRTC_SUCCESS)
    throw Error(AKX_ERR_HIP, "hiprtcCreateProgram failed");
  for (auto &nm : names)
    (void)hiprtcAddNameExpression(prog, nm.c_str());
  (void)hiprtcAddNameExpression(prog, table.c_str());
  const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 8, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 0) {
      log.resize(n);
      (void)hiprtcGetProgramLog(prog, &log[0]);
    }
    (void)hiprtcDestroyProgram(&prog);
    throw Error(AKX_ERR_INVALID, std::string("the device law did not compile (") +
                                     hiprtcGetErrorString(rc) + "):\n" + log);
  }
  std::string lowered[3];
  for (int k = 0; k < 3; ++k) {
    const char *low = nullptr;
    if (hiprtcGetLoweredName(prog, k < 2 ? names[k].c_str() : table.c_str(), &low) !=
            HIPRTC_SUCCESS ||
        !low) {
      (void)hiprtcDestroyProgram(&prog);
      throw Error(AKX_ERR_HIP, "the device law program lost an instantiation");
    }
    lowered[k] = low;
  }
  size_t size = 0;
  std::vector<char> code;
  if (hiprtcGetCodeSize(prog, &size) == HIPRTC_SUCCESS && size > 0) {
    code.resize(size);
    if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS)
      code.clear();
  }
  (void)hiprtcDestroyProgram(&prog);
  require(!code.empty(), "the device law compiled to no code");
  auto dl = std::make_unique<DeviceLaw>();
  HIPCHECK(hipModuleLoadData(&dl->module, code.data()));
  for (int ext = 0; ext < 2; ++ext)
    HIPCHECK(hipModuleGetFunction(&dl->kernel[ext], dl->module, lowered[ext].c_str()));
  hipDeviceptr_t tab = nullptr;
  size_t tab_bytes = 0;
  HIPCHECK(hipModuleGetGlobal(&tab, &tab_bytes, dl->module, lowered[2].c_str()));
  Hex8Tables t;
  fill_hex8_tables(t);
  require(tab_bytes == sizeof(t), "the device law program's Hex8 table has another size");
  HIPCHECK(hipMemcpy((void *)tab, &t, sizeof(t), hipMemcpyHostToDevice));
  return dl;
}

// The element pass of a material whose law was compiled in
void launch_device_law(akx_model *m, Material &mat, ElemArgs A, bool materialize) {
  A.grad_u = materialize ? mat.grad.p : nullptr; // grad u and the stress, when asked
  A.stress = mat.stress.p;
  void *args[] = {&A};
  const unsigned grid = (unsigned)A.nblocks_padded;
  const unsigned block = m->etype == AKX_HEX8 ? 64 : TPB;
  HIPCHECK(hipModuleLaunchKernel(mat.dlaw->kernel[mat.ext() ? 1 : 0], grid, 1, 1, block, 1, 1,
                                 0, m->stream, args, nullptr));
}

void flip_plastic_state(akx_model *m) {
  for (auto &mp : m->mats)
    if (mp->law == AKX_LAW_PLASTIC_LIH)
      mp->cur = 1 - mp->cur;
}

// per_corner: the per-corner table whatever the model's assembly (the
// pipelined schedule's partial launches write it)
GatherArgs gather_args(akx_model *m, bool per_corner = false) {
  GatherArgs G{};
  G.nn = m->nn;
  G.node0 = 0;
  G.node_end = m->nn;
  G.ell_off = m->ell_off.p;
  G.deg = m->deg.p;
  if (m->bpa && !per_corner) {
    G.ell_off = m->bp_off.p;
    G.deg = m->bp_deg.p;
  } else if (m->ppa && !per_corner) {
    G.ell_off = m->pp_off.p;
    G.deg = m->pp_deg.p;
  }
  G.fe = m->fe.p;
  G.fint = m->fint.p;
  G.fext = m->has_fext ? m->fext.p : nullptr;
  G.mass = m->mass.p;
  G.blocked = m->blocked.p;
  G.u = m->u.p;
  G.xu = m->records_live() ? m->xu.p : nullptr;
  G.v = m->v.p;
  G.a = m->a.p;
  G.cu = m->cu.p;
  G.cv = m->cv.p;
  G.dt = m->dt;
  G.nblocks_padded = padded_blocks(m->nn);
  return G;
}

// f_int is stored by the last step of a solve_step call (and by the gathers
// without update): the steps in between use it only inside the gather, and
// nothing reads the stored array before the call returns.
void launch_gather(akx_model *m, bool update, bool predict) {
  GatherArgs G = gather_args(m);
  if (update && predict)
    G.fint = nullptr;
  dim3 grid((unsigned)G.nblocks_padded), block(TPB);
  const bool f64 = m->precision != AKX_PRECISION_FP32;
  check_gather_args(G, update, update && !f64);
  m->timed(1, [&] {
    if (!update) {
      if (f64)
        hipLaunchKernelGGL((k_gather_update<double, false, false, false>), grid,
                           block, 0, m->stream, G);
      else
        hipLaunchKernelGGL((k_gather_update<float, false, false, false>), grid,
                           block, 0, m->stream, G);
    } else if (f64) {
      if (predict)
        hipLaunchKernelGGL((k_gather_update<double, true, true, false>), grid,
                           block, 0, m->stream, G);
      else
        hipLaunchKernelGGL((k_gather_update<double, true, false, false>), grid,
                           block, 0, m->stream, G);
    } else {
      if (predict)
        hipLaunchKernelGGL((k_gather_update<float, true, true, true>), grid,
                           block, 0, m->stream, G);
      else
        hipLaunchKernelGGL((k_gather_update<float, true, false, true>), grid,
                           block, 0, m->stream, G);
    }
  });
  HIPCHECK(hipGetLastError());
}

void launch_predictor(akx_model *m) {
  GatherArgs G = gather_args(m);
  dim3 grid((unsigned)G.nblocks_padded), block(TPB);
  check_gather_args(G, true, m->precision == AKX_PRECISION_FP32);
  m->timed(2, [&] {
    if (m->precision != AKX_PRECISION_FP32)
      hipLaunchKernelGGL((k_predictor<false>), grid, block, 0, m->stream, G);
    else
      hipLaunchKernelGGL((k_predictor<true>), grid, block, 0, m->stream, G);
  });
  HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Pipelined step (opt-in): the elements are cut in P consecutive parts and the
// nodes in P groups, group q = the 256-node slices whose last element (in
// assembly order) lies in part q.  Then
//   gather of group q (step k)        needs the element pass of parts <= q
//   element part p (step k + 1)       needs the gather of the groups its nodes
//                                     belong to (the last is pipe_need[p])
// so the memory-bound gather of group q runs on a second stream while the
// compute-bound element pass of part q + 1 runs.  Values and summation
// orders are those of the plain schedule: it reorders launches, not
// arithmetic.  Planned only for one contiguous non-plastic material whose
// slice groups are monotone (otherwise the plain schedule is used).
// ---------------------------------------------------------------------------
// a change of the materials' inputs re-decides whether the pipelined
// schedule applies
void refresh_pipeline_plan(akx_model *m) { m->pipe_planned = false; }

void plan_pipeline(akx_model *m) {
  m->pipe_planned = true;
  m->pipe_ok = false;
  const int P = m->pipe_parts;
  if (P < 2 || m->mats.size() != 1 || m->bpa || m->mats[0]->ext() || m->mats[0]->dlaw)
    return;
  Material &mat = *m->mats[0];
  if (!mat.contiguous || mat.law == AKX_LAW_PLASTIC_LIH || mat.first != 0 ||
      mat.n != m->nel || mat.pos_offset != 0 || m->nel < P)
    return;
  const int64_t nel = m->nel, nn = m->nn, nnpe = m->nnpe;
  auto part_of = [&](int64_t e) { return (int)(e * P / nel); };
  std::vector<int64_t> maxel(nn, -1);
  for (int64_t e = 0; e < nel; ++e)
    for (int a = 0; a < nnpe; ++a) {
      const int32_t nd = m->conn_host[e * nnpe + a];
      maxel[nd] = std::max<int64_t>(maxel[nd], e);
    }
  const int64_t nsl = (nn + 255) / 256;
  std::vector<int> lp(nsl, 0);
  int last = 0;
  for (int64_t s = 0; s < nsl; ++s) {
    int v = -1;
    for (int64_t n = s * 256; n < std::min<int64_t>(nn, s * 256 + 256); ++n)
      if (maxel[n] >= 0)
        v = std::max(v, part_of(maxel[n]));
    if (v < 0)
      v = last;
    if (v < last)
      return; // groups would not be contiguous node ranges
    lp[s] = last = v;
  }
  m->pipe_el.assign(P + 1, 0);
  m->pipe_node.assign(P + 1, nn);
  for (int p = 0; p <= P; ++p)
    m->pipe_el[p] = (int64_t)p * nel / P;
  // part boundaries must agree with part_of
  for (int p = 0; p < P; ++p)
    if (m->pipe_el[p] < m->pipe_el[p + 1] && part_of(m->pipe_el[p]) != p)
      return;
  m->pipe_node[0] = 0;
  for (int p = 1; p < P; ++p) {
    int64_t s = 0;
    while (s < nsl && lp[s] < p)
      ++s;
    m->pipe_node[p] = std::min<int64_t>(nn, s * 256);
  }
  m->pipe_node[P] = nn;
  m->pipe_need.assign(P, 0);
  for (int p = 0; p < P; ++p) {
    int64_t smax = 0;
    for (int64_t e = m->pipe_el[p]; e < m->pipe_el[p + 1]; ++e)
      for (int a = 0; a < nnpe; ++a)
        smax = std::max<int64_t>(smax, m->conn_host[e * nnpe + a] / 256);
    m->pipe_need[p] = lp[smax];
  }
  if (!m->stream2)
    HIPCHECK(hipStreamCreateWithFlags(&m->stream2, hipStreamNonBlocking));
  auto make = [&](std::vector<hipEvent_t> &v) {
    while ((int)v.size() < P) {
      hipEvent_t e;
      HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      v.push_back(e);
    }
  };
  make(m->ev_elem);
  make(m->ev_gather[0]);
  make(m->ev_gather[1]);
  m->pipe_ok = true;
  if (!m->xu_static) // the partial launches read the records: bring u in
    refresh_xu(m, false);
}

void launch_element_part(akx_model *m, int p) {
  Material &mat = *m->mats[0];
  const int64_t e0 = m->pipe_el[p], cnt = m->pipe_el[p + 1] - e0;
  if (cnt <= 0)
    return;
  const bool strict = m->precision == AKX_PRECISION_FP64;
  ElemArgs A = elem_args(m, mat, false);
  A.n = cnt;
  A.first = mat.first + e0;
  A.pos_offset = mat.pos_offset + e0;
  if (A.lm_f)
    A.lm_f += e0;
  A.lconn = nullptr; // the block node lists follow the whole range's blocks
  if (A.lam_el) {
    A.lam_el += e0;
    A.mu_el += e0;
  }
  if (A.sy_el)
    A.sy_el += e0;
  if (A.h_el)
    A.h_el += e0;
  int64_t epb = strict ? (m->etype == AKX_HEX8 ? HEX_STRICT_EPB : TPB)
                       : akxf::elements_per_block(m->etype);
  A.nblocks_padded = ((cnt + epb - 1) / epb + 7) / 8 * 8;
  check_elem_args(A, mat.law);
  require(m->etype != AKX_TET4 || A.xu != nullptr, "Tet4 element pass without its node records");
  if (strict) {
    if (m->etype == AKX_TET4)
      launch_elem_law<ET_TET4, double, double>(m, mat, A);
    else
      launch_hex8_strict(m, mat, A);
  } else {
    HIPCHECK(akxf::launch_elements(m->etype, mat.law, m->precision == AKX_PRECISION_FP32,
                                   A, m->stream));
  }
}

void launch_gather_group(akx_model *m, int p, bool predict) {
  GatherArgs G = gather_args(m, true);
  if (predict)
    G.fint = nullptr; // stored by the last step only (launch_gather)
  G.node0 = m->pipe_node[p];
  G.node_end = m->pipe_node[p + 1];
  if (G.node_end <= G.node0)
    return;
  G.nblocks_padded = padded_blocks(G.node_end - G.node0);
  const bool f64 = m->precision != AKX_PRECISION_FP32;
  check_gather_args(G, true, !f64);
  dim3 grid((unsigned)G.nblocks_padded), block(TPB);
  if (f64) {
    if (predict)
      hipLaunchKernelGGL((k_gather_update<double, true, true, false>), grid, block, 0,
                         m->stream2, G);
    else
      hipLaunchKernelGGL((k_gather_update<double, true, false, false>), grid, block, 0,
                         m->stream2, G);
  } else {
    if (predict)
      hipLaunchKernelGGL((k_gather_update<float, true, true, true>), grid, block, 0,
                         m->stream2, G);
    else
      hipLaunchKernelGGL((k_gather_update<float, true, false, true>), grid, block, 0,
                         m->stream2, G);
  }
  HIPCHECK(hipGetLastError());
}

void solve_pipelined(akx_model *m, int64_t nb_steps) {
  const int P = m->pipe_parts;
  launch_predictor(m);
  for (int64_t s = 0; s < nb_steps; ++s) {
    const int cur = (int)(s & 1), prev = cur ^ 1;
    for (int p = 0; p < P; ++p) {
      if (s > 0) // the gather of the previous step on this part's nodes
        HIPCHECK(hipStreamWaitEvent(m->stream, m->ev_gather[prev][m->pipe_need[p]], 0));
      launch_element_part(m, p);
      HIPCHECK(hipEventRecord(m->ev_elem[p], m->stream));
      HIPCHECK(hipStreamWaitEvent(m->stream2, m->ev_elem[p], 0));
      launch_gather_group(m, p, s + 1 < nb_steps);
      HIPCHECK(hipEventRecord(m->ev_gather[cur][p], m->stream2));
    }
  }
  // join: the model's stream ends after the last gather
  HIPCHECK(hipStreamWaitEvent(m->stream, m->ev_gather[(nb_steps - 1) & 1][P - 1], 0));
}

InitArgs init_args(akx_model *m, Material *mat) {
  InitArgs A{};
  A.nel = m->nel;
  A.nn = m->nn;
  A.conn = m->conn.p;
  A.X = m->X.p;
  A.u = m->u.p;
  if (mat) {
    A.n = mat->n;
    A.first = mat->first;
    A.pos_offset = mat->pos_offset;
    A.filter = mat->contiguous ? nullptr : mat->filter.p;
    A.rho = mat->rho;
    A.lambda = mat->lambda;
    A.mu = mat->mu;
    A.lam_el = mat->per_element ? mat->lam_el.p : nullptr;
    A.mu_el = mat->per_element ? mat->mu_el.p : nullptr;
  }
  return A;
}

void check_ready(akx_model *m) {
  require(m != nullptr, "null model");
  require(m->initialized, "initFull has not been called");
}

// Gather table.  Every (element position p, local node a) gets an entry of a
// node-major sliced-ELL table: nodes are cut in slices of 256, slice s has
// width w_s = its largest node degree and starts at ell_off[s]; the entry of
// the d-th element of node n is ell_off[s] + d*256 + n%256.  Elements are
// ranked per node in ascending assembly order -- materials in creation
// order, each material's filter in order -- which is the order of the
// reference's sequential scatter (dof_manager.cc:92-114, called per material
// by material.cc:98-120).  MeshUtils::buildNode2Elements
// (mesh_utils.cc:37-82) builds the same node -> element lists.
void build_gather_table(akx_model *m) {
  int64_t nslots = 0;
  for (auto &mp : m->mats)
    nslots += mp->n * m->nnpe;
  require(nslots < (int64_t)INT32_MAX, "too many element slots for int32 indices");
  std::vector<int32_t> deg(m->nn, 0);
  for (auto &mp : m->mats)
    for (int64_t i = 0; i < mp->n; ++i) {
      int64_t e = mp->contiguous ? mp->first + i : mp->filter_host[i];
      for (int a = 0; a < m->nnpe; ++a)
        deg[m->conn_host[e * m->nnpe + a]]++;
    }
  const int64_t nslices = (m->nn + 255) / 256;
  std::vector<int64_t> off(nslices + 1, 0);
  for (int64_t sl = 0; sl < nslices; ++sl) {
    int w = 0;
    for (int64_t n = sl * 256; n < std::min<int64_t>(m->nn, sl * 256 + 256); ++n)
      w = std::max(w, deg[n]);
    require(w <= 255, "a node belongs to more than 255 elements");
    off[sl + 1] = off[sl] + (int64_t)w * 256;
  }
  m->ell_total = off[nslices];
  require(m->ell_total < (int64_t)INT32_MAX, "element force table too large for int32");
  std::vector<int32_t> fill(m->nn, 0), pos(std::max<int64_t>(nslots, 1));
  for (auto &mp : m->mats)
    for (int64_t i = 0; i < mp->n; ++i) {
      int64_t e = mp->contiguous ? mp->first + i : mp->filter_host[i];
      int64_t p = mp->pos_offset + i;
      for (int a = 0; a < m->nnpe; ++a) {
        int32_t node = m->conn_host[e * m->nnpe + a];
        int32_t d = fill[node]++;
        pos[p * m->nnpe + a] = (int32_t)(off[node >> 8] + (int64_t)d * 256 + (node & 255));
      }
    }
  std::vector<uint8_t> deg8(m->nn);
  for (int64_t n = 0; n < m->nn; ++n)
    deg8[n] = (uint8_t)deg[n];
  m->pos.alloc(pos.size());
  m->ell_off.alloc(off.size());
  m->deg.alloc(m->nn);
  HIPCHECK(hipMemcpy(m->pos.p, pos.data(), pos.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(m->ell_off.p, off.data(), off.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(m->deg.p, deg8.data(), m->nn, hipMemcpyHostToDevice));
}

void alloc_internals(akx_model *m, Material &mat) {
  int64_t nq = mat.n * m->nqp;
  if (mat.law == AKX_LAW_PLASTIC_LIH) {
    for (int s = 0; s < 2; ++s) {
      mat.hgrad[s].alloc(9 * nq);
      mat.hsig[s].alloc(9 * nq);
      mat.heps[s].alloc(9 * nq);
      mat.hiso[s].alloc(nq);
      mat.hpe[s].alloc(nq);
      zero_async(m, mat.hgrad[s].p, 9 * nq);
      zero_async(m, mat.hsig[s].p, 9 * nq);
      zero_async(m, mat.heps[s].p, 9 * nq);
      zero_async(m, mat.hiso[s].p, nq);
      zero_async(m, mat.hpe[s].p, nq);
    }
    mat.cur = 0;
  }
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *akx_last_error(void) { return g_last_error.c_str(); }

int akx_version(void) { return 100; }

int akx_device_count(int *count) {
  return guard([&] {
    require(count != nullptr, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
      n = 0;
    int good = 0;
    for (int d = 0; d < n; ++d) {
      hipDeviceProp_t p;
      if (hipGetDeviceProperties(&p, d) == hipSuccess &&
          std::strncmp(p.gcnArchName, "gfx950", 6) == 0)
        ++good;
    }
    *count = good;
  });
}

int akx_model_create(int device, akx_model **out) {
  return guard([&] {
    require(out != nullptr, "null output pointer");
    auto m = std::make_unique<akx_model>();
    m->device = device;
    m->set_device();
    hipDeviceProp_t p;
    HIPCHECK(hipGetDeviceProperties(&p, device));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
      throw Error(AKX_ERR_UNSUPPORTED,
                  std::string("device is not gfx950: ") + p.gcnArchName);
    HIPCHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    (void)device_scratch(device); // the no-allocation energy's array on this device
    m->scratch_u64.alloc(4);
    m->scratch_i64.alloc(4);
    *out = m.release();
  });
}

int akx_model_destroy(akx_model *m) {
  return guard([&] {
    if (!m)
      return;
    (void)hipSetDevice(m->device);
    if (m->stream)
      (void)hipStreamSynchronize(m->stream);
    delete m;
  });
}

int akx_model_set_mesh(akx_model *m, int element_type, int64_t nb_nodes,
                       const double *nodes, int64_t nb_elements,
                       const int32_t *connectivity) {
  return guard([&] {
    require(m != nullptr, "null model");
    require(!m->initialized, "mesh already initialised");
    require(element_type == AKX_TET4 || element_type == AKX_HEX8,
            "element type must be _tetrahedron_4 or _hexahedron_8");
    require(nb_nodes > 0 && nb_elements > 0, "empty mesh");
    require(nb_nodes < INT32_MAX, "too many nodes for int32 connectivity");
    require(nodes && connectivity, "null mesh arrays");
    m->set_device();
    m->etype = element_type;
    m->nnpe = element_type == AKX_TET4 ? 4 : 8;
    m->nqp = element_type == AKX_TET4 ? 1 : 8;
    m->nn = nb_nodes;
    m->nel = nb_elements;
    m->conn_host.assign(connectivity, connectivity + nb_elements * m->nnpe);
    for (int32_t c : m->conn_host)
      require(c >= 0 && c < nb_nodes, "connectivity refers to a missing node");
    m->X.alloc(3 * nb_nodes);
    m->u.alloc(3 * nb_nodes);
    m->v.alloc(3 * nb_nodes);
    m->a.alloc(3 * nb_nodes);
    m->fint.alloc(3 * nb_nodes);
    m->mass.alloc(nb_nodes);
    m->blocked.alloc(nb_nodes);
    m->conn.alloc(nb_elements * m->nnpe);
    upload_soa(m, nodes, m->X.p, nb_nodes, 3);
    HIPCHECK(hipMemcpy(m->conn.p, connectivity,
                       nb_elements * m->nnpe * sizeof(int32_t),
                       hipMemcpyHostToDevice));
    zero_async(m, m->u.p, 3 * nb_nodes);
    zero_async(m, m->v.p, 3 * nb_nodes);
    zero_async(m, m->a.p, 3 * nb_nodes);
    zero_async(m, m->fint.p, 3 * nb_nodes);
    zero_async(m, m->mass.p, nb_nodes);
    HIPCHECK(hipMemsetAsync(m->blocked.p, 0, nb_nodes, m->stream));
    if (m->etype == AKX_HEX8)
      upload_hex8_tables();
    HIPCHECK(hipStreamSynchronize(m->stream));
  });
}

int akx_material_law_by_name(const char *name, int finite_deformation, int *law) {
  return guard([&] {
    require(name && law, "null argument");
    std::string s(name);
    if (s == "elastic")
      *law = finite_deformation ? AKX_LAW_ELASTIC_FD : AKX_LAW_ELASTIC;
    else if (s == "neohookean")
      *law = AKX_LAW_NEOHOOKEAN;
    else if (s == "plastic_linear_isotropic_hardening") {
      // The reference's finite-deformation branch (material_linear_isotropic
      // _hardening.cc:36-52, _inline_impl.hh:88-186) is not provided: refuse
      // it rather than run the small-strain law in its place.
      if (finite_deformation)
        throw Error(AKX_ERR_UNSUPPORTED,
                    "plastic_linear_isotropic_hardening with finite_deformation = true "
                    "is not provided (small-strain plasticity only)");
      *law = AKX_LAW_PLASTIC_LIH;
    } else
      throw Error(AKX_ERR_UNSUPPORTED, "unknown material law: " + s);
  });
}

int akx_model_add_material(akx_model *m, int law, const double *params,
                           int nparams, const int32_t *filter,
                           int64_t nb_filter, int *material_id) {
  return guard([&] {
    require(m != nullptr && m->etype >= 0, "set the mesh first");
    require(!m->initialized, "materials must be added before initFull");
    require(law >= AKX_LAW_ELASTIC && law <= AKX_LAW_USER_FD, "unknown law");
    require(params && nparams >= 3, "params = {rho, E, nu[, sigma_y, h]}");
    m->set_device();
    auto mat = std::make_unique<Material>();
    mat->law = law;
    mat->rho = params[0];
    mat->E = params[1];
    mat->nu = params[2];
    mat->sigma_y = nparams > 3 ? params[3] : 0.;
    mat->h = nparams > 4 ? params[4] : 0.;
    require(mat->rho > 0, "rho must be positive");
    lame(mat->E, mat->nu, mat->lambda, mat->mu);
    int64_t pos = 0;
    for (auto &o : m->mats)
      pos += o->n;
    mat->pos_offset = pos;
    if (filter == nullptr) {
      mat->n = m->nel;
      mat->first = 0;
      mat->contiguous = true;
    } else {
      require(nb_filter >= 0, "negative filter size");
      mat->n = nb_filter;
      mat->filter_host.assign(filter, filter + nb_filter);
      bool contig = true;
      for (int64_t i = 0; i < nb_filter; ++i) {
        require(filter[i] >= 0 && filter[i] < m->nel, "filter refers to a missing element");
        if (filter[i] != filter[0] + i)
          contig = false;
      }
      mat->contiguous = contig;
      mat->first = nb_filter ? filter[0] : 0;
      if (!contig && nb_filter) {
        mat->filter.alloc(nb_filter);
        HIPCHECK(hipMemcpy(mat->filter.p, filter, nb_filter * sizeof(int32_t),
                           hipMemcpyHostToDevice));
      }
    }
    if (material_id)
      *material_id = (int)m->mats.size();
    m->mats.push_back(std::move(mat));
  });
}

int akx_model_set_element_parameters(akx_model *m, int material_id,
                                     const double *E, const double *nu,
                                     const double *sigma_y, const double *h) {
  return guard([&] {
    require(m != nullptr, "null model");
    require(material_id >= 0 && material_id < (int)m->mats.size(), "bad material id");
    m->set_device();
    Material &mat = *m->mats[material_id];
    int64_t n = mat.n;
    if (n == 0)
      return;
    auto up = [&](DevBuf<double> &b, const double *src) {
      b.alloc(n);
      HIPCHECK(hipMemcpy(b.p, src, n * sizeof(double), hipMemcpyHostToDevice));
    };
    if (E || nu) {
      mat.E_el_host.assign(n, mat.E);
      mat.nu_el_host.assign(n, mat.nu);
      if (E)
        mat.E_el_host.assign(E, E + n);
      if (nu)
        mat.nu_el_host.assign(nu, nu + n);
      DevBuf<double> dE, dnu;
      if (E)
        up(dE, E);
      if (nu)
        up(dnu, nu);
      mat.lam_el.alloc(n);
      mat.mu_el.alloc(n);
      hipLaunchKernelGGL(k_lame, dim3(grid_for(n)), dim3(TPB), 0, m->stream,
                         (const double *)dE.p, (const double *)dnu.p, mat.E,
                         mat.nu, mat.lam_el.p, mat.mu_el.p, n);
      HIPCHECK(hipGetLastError());
      mat.per_element = true;
      if (m->precision == AKX_PRECISION_FP32)
        fill_lame_f32(m, mat);
      HIPCHECK(hipStreamSynchronize(m->stream));
    }
    if (sigma_y)
      up(mat.sy_el, sigma_y);
    if (h)
      up(mat.h_el, h);
  });
}

int akx_model_set_eigen_grad_u(akx_model *m, int material_id, const double *eigen_grad_u) {
  return guard([&] {
    require(m != nullptr && m->etype >= 0, "set the mesh first");
    require(material_id >= 0 && material_id < (int)m->mats.size(), "bad material id");
    if (m->precision != AKX_PRECISION_FP64)
      throw Error(AKX_ERR_UNSUPPORTED, "eigen_grad_u runs on the plain FP64 path");
    m->set_device();
    Material &mat = *m->mats[material_id];
    if (!eigen_grad_u || mat.n == 0) {
      mat.eig.release();
      return;
    }
    const int64_t nq = mat.n * m->nqp;
    mat.eig.alloc(9 * nq);
    upload_soa(m, eigen_grad_u, mat.eig.p, nq, 9);
    refresh_pipeline_plan(m);
  });
}

int akx_model_set_thermal(akx_model *m, int material_id, double alpha, const double *delta_T) {
  return guard([&] {
    require(m != nullptr && m->etype >= 0, "set the mesh first");
    require(material_id >= 0 && material_id < (int)m->mats.size(), "bad material id");
    if (m->precision != AKX_PRECISION_FP64)
      throw Error(AKX_ERR_UNSUPPORTED, "thermal stresses run on the plain FP64 path");
    m->set_device();
    Material &mat = *m->mats[material_id];
    if (!delta_T || mat.n == 0) {
      mat.sth_now.release();
      return;
    }
    require(std::isfinite(alpha), "non-finite alpha");
    const int64_t nq = mat.n * m->nqp;
    // sigma_th = -E / (1 - 2 nu) * alpha * delta_T (material_thermal.hh:91-95)
    std::vector<double> sth(nq);
    for (int64_t i = 0; i < mat.n; ++i) {
      const double E = mat.E_el_host.empty() ? mat.E : mat.E_el_host[i];
      const double nu = mat.nu_el_host.empty() ? mat.nu : mat.nu_el_host[i];
      for (int q = 0; q < m->nqp; ++q)
        sth[i * m->nqp + q] = -E / (1. - 2. * nu) * alpha * delta_T[i * m->nqp + q];
    }
    mat.sth_now.alloc(nq);
    HIPCHECK(hipMemcpy(mat.sth_now.p, sth.data(), nq * sizeof(double), hipMemcpyHostToDevice));
    if (mat.law == AKX_LAW_PLASTIC_LIH && !mat.hsth[0].p) {
      // the history starts at zero (material_thermal.cc:34)
      for (int k = 0; k < 2; ++k) {
        mat.hsth[k].alloc(nq);
        zero_async(m, mat.hsth[k].p, nq);
      }
      HIPCHECK(hipStreamSynchronize(m->stream));
    }
    refresh_pipeline_plan(m);
  });
}

int akx_model_set_precision(akx_model *m, int precision) {
  return guard([&] {
    require(m != nullptr, "null model");
    require(!m->initialized, "precision must be set before initFull");
    for (auto &mp : m->mats)
      require(precision == AKX_PRECISION_FP64 || !mp->ext(),
              "eigen_grad_u and thermal stresses run on the plain FP64 path");
    require(precision == AKX_PRECISION_FP64 || precision == AKX_PRECISION_FP32 ||
                precision == AKX_PRECISION_FP64_FAST,
            "unknown precision");
    m->precision = precision;
  });
}

int akx_model_set_block_assembly(akx_model *m, int enable) {
  return guard([&] {
    require(m != nullptr, "null model");
    require(!m->initialized, "the assembly must be chosen before initFull");
    require(enable >= AKX_ASSEMBLY_PER_CORNER && enable <= AKX_ASSEMBLY_BLOCK_PARTIALS,
            "assembly must be AKX_ASSEMBLY_PER_CORNER, _DEFAULT or _BLOCK_PARTIALS");
    m->assembly = enable;
  });
}

int akx_model_get_block_assembly(akx_model *m, int *active) {
  return guard([&] {
    check_ready(m);
    require(active != nullptr, "null output");
    *active = m->bpa ? AKX_ASSEMBLY_BLOCK_PARTIALS
              : m->ppa ? (m->assembly == AKX_ASSEMBLY_BLOCK_PARTIALS ? AKX_ASSEMBLY_BLOCK_PARTIALS
                                                                     : AKX_ASSEMBLY_PREFIX_PARTIALS)
                       : AKX_ASSEMBLY_PER_CORNER;
  });
}

int akx_model_init_full(akx_model *m) {
  return guard([&] {
    require(m != nullptr && m->etype >= 0, "set the mesh first");
    require(!m->initialized, "initFull called twice");
    require(!m->mats.empty(), "no material");
    require(!m->has_user() || m->precision == AKX_PRECISION_FP64,
            "user materials run on the plain FP64 path");
    m->set_device();
    for (auto &mp : m->mats)
      if (mp->user()) { // the internals the plugin reads and writes
        mp->grad.alloc(9 * mp->n * m->nqp);
        mp->stress.alloc(9 * mp->n * m->nqp);
      }
    // IntegratorGauss::checkJacobians
    {
      InitArgs A = init_args(m, nullptr);
      long long init = INT64_MAX;
      HIPCHECK(hipMemcpy(m->scratch_i64.p, &init, sizeof(init), hipMemcpyHostToDevice));
      A.first_bad = m->scratch_i64.p;
      if (m->etype == AKX_TET4)
        hipLaunchKernelGGL(k_check_jacobians<ET_TET4>, dim3(grid_for(m->nel)),
                           dim3(TPB), 0, m->stream, A);
      else
        hipLaunchKernelGGL(k_check_jacobians<ET_HEX8>, dim3(grid_for(m->nel)),
                           dim3(TPB), 0, m->stream, A);
      HIPCHECK(hipGetLastError());
      long long bad = 0;
      HIPCHECK(hipMemcpyAsync(&bad, m->scratch_i64.p, sizeof(bad),
                              hipMemcpyDeviceToHost, m->stream));
      HIPCHECK(hipStreamSynchronize(m->stream));
      if (bad != INT64_MAX) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "Negative jacobian computed, possible problem in the "
                      "element node ordering (Quadrature Point %lld:%lld)",
                      bad % m->nqp, bad / m->nqp);
        throw Error(AKX_ERR_JACOBIAN, msg);
      }
    }
    // FEEngine precompute of the plain FP64 Hex8 pass: J^-1 and w det J per
    // Gauss point (fe_engine_template_tmpl.hh:98-120)
    if (m->etype == AKX_HEX8 && m->precision == AKX_PRECISION_FP64 && m->nel > 0) {
      const int64_t nq = m->nel * 8;
      m->jw.alloc(5 * nq);
      hipLaunchKernelGGL(k_hex8_jacobian_table, dim3(grid_for(nq)), dim3(TPB), 0,
                         m->stream, (const int32_t *)m->conn.p, (const double *)m->X.p,
                         m->nn, m->nel, m->jw.p);
      HIPCHECK(hipGetLastError());
    }
    if (m->etype == AKX_TET4 && m->nn > 0) {
      m->xu.alloc(6 * m->nn);
      refresh_xu(m, true);
      for (auto &mp : m->mats) // Tet4 passes with staged node records
        if (mp->contiguous && mp->n > 0 && m->precision != AKX_PRECISION_FP64_FAST)
          build_block_nodes(m, *mp);
      m->xu_static = m->precision == AKX_PRECISION_FP64_FAST;
      for (auto &mp : m->mats)
        m->xu_static = m->xu_static || (mp->n > 0 && (!mp->contiguous || mp->user()));
      bool ppa = m->assembly != AKX_ASSEMBLY_PER_CORNER &&
                 m->precision == AKX_PRECISION_FP64 && !m->has_user();
      for (auto &mp : m->mats)
        ppa = ppa && mp->contiguous && mp->n > 0 && mp->lconn.p != nullptr;
      if (ppa)
        build_ppa(m, m->assembly == AKX_ASSEMBLY_BLOCK_PARTIALS);
      m->ppa = ppa;
    }
    if (m->etype == AKX_HEX8 && m->precision == AKX_PRECISION_FP32 && m->nn > 0 &&
        m->assembly != AKX_ASSEMBLY_PER_CORNER)
      for (auto &mp : m->mats) // FP32 Hex8 block partials
        if (mp->contiguous && mp->n > 0)
          build_block_nodes(m, *mp, akxf::elements_per_block(AKX_HEX8, true));
    build_gather_table(m);
    m->bpa = m->assembly != AKX_ASSEMBLY_PER_CORNER && m->precision == AKX_PRECISION_FP32 &&
             m->nn > 0;
    for (auto &mp : m->mats)
      m->bpa = m->bpa && mp->contiguous && mp->n > 0 && mp->lslot.p != nullptr;
    if (m->bpa)
      build_bpa_table(m);
    m->fe.alloc(std::max<int64_t>(std::max(m->ell_total, m->bp_total), 1) * 3 *
                m->fe_elem_size());
    // lumped mass
    {
      DevBuf<double> me;
      me.alloc(std::max<int64_t>(m->ell_total, 1));
      for (auto &mp : m->mats) {
        if (mp->n == 0)
          continue;
        InitArgs A = init_args(m, mp.get());
        A.me = me.p;
        A.pos = m->pos.p;
        if (m->etype == AKX_TET4)
          hipLaunchKernelGGL(k_mass_elements<ET_TET4>, dim3(grid_for(mp->n)),
                             dim3(TPB), 0, m->stream, A);
        else
          hipLaunchKernelGGL(k_mass_elements<ET_HEX8>, dim3(grid_for(mp->n)),
                             dim3(TPB), 0, m->stream, A);
        HIPCHECK(hipGetLastError());
      }
      hipLaunchKernelGGL(k_gather_mass, dim3(grid_for(m->nn)), dim3(TPB), 0,
                         m->stream, m->nn, (const int64_t *)m->ell_off.p,
                         (const uint8_t *)m->deg.p, (const double *)me.p,
                         m->mass.p);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(m->stream));
    }
    for (auto &mp : m->mats)
      alloc_internals(m, *mp);
    if (m->precision == AKX_PRECISION_FP32) {
      for (auto &mp : m->mats)
        if (mp->per_element && mp->n > 0)
          fill_lame_f32(m, *mp);
      m->cu.alloc(3 * m->nn);
      m->cv.alloc(3 * m->nn);
      hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(3 * m->nn)), dim3(TPB), 0,
                         m->stream, m->cu.p, 3 * m->nn, 0.f);
      hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(3 * m->nn)), dim3(TPB), 0,
                         m->stream, m->cv.p, 3 * m->nn, 0.f);
      HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipStreamSynchronize(m->stream));
    m->initialized = true;
  });
}

int akx_model_get_stable_time_step(akx_model *m, double *dt) {
  return guard([&] {
    check_ready(m);
    require(dt != nullptr, "null output");
    m->set_device();
    unsigned long long init = 0x7FEFFFFFFFFFFFFFull; // DBL_MAX bits
    HIPCHECK(hipMemcpyAsync(m->scratch_u64.p, &init, sizeof(init),
                            hipMemcpyHostToDevice, m->stream));
    for (auto &mp : m->mats) {
      if (mp->n == 0)
        continue;
      InitArgs A = init_args(m, mp.get());
      A.min_bits = m->scratch_u64.p;
      if (m->etype == AKX_TET4)
        hipLaunchKernelGGL(k_stable_dt<ET_TET4>, dim3(grid_for(mp->n)),
                           dim3(TPB), 0, m->stream, A);
      else
        hipLaunchKernelGGL(k_stable_dt<ET_HEX8>, dim3(grid_for(mp->n)),
                           dim3(TPB), 0, m->stream, A);
      HIPCHECK(hipGetLastError());
    }
    unsigned long long bits = 0;
    HIPCHECK(hipMemcpyAsync(&bits, m->scratch_u64.p, sizeof(bits),
                            hipMemcpyDeviceToHost, m->stream));
    HIPCHECK(hipStreamSynchronize(m->stream));
    double r;
    std::memcpy(&r, &bits, sizeof(r));
    *dt = r;
  });
}

int akx_model_set_time_step(akx_model *m, double dt) {
  return guard([&] {
    require(m != nullptr, "null model");
    require(dt > 0 && std::isfinite(dt), "time step must be positive");
    m->dt = dt;
  });
}

int akx_model_solve_steps(akx_model *m, int64_t nb_steps) {
  return guard([&] {
    check_ready(m);
    require(m->dt > 0, "setTimeStep has not been called");
    require(nb_steps >= 0, "negative step count");
    if (nb_steps == 0)
      return;
    m->set_device();
    if (m->has_user()) { // the plugin runs on the host between the passes
      launch_predictor(m);
      for (int64_t s = 0; s < nb_steps; ++s) {
        user_stress_phase(m);
        launch_element_pass(m, false);
        flip_plastic_state(m);
        launch_gather(m, true, s + 1 < nb_steps);
      }
      HIPCHECK(hipStreamSynchronize(m->stream));
      return;
    }
    if (m->pipe_parts > 1 && !m->pipe_planned)
      plan_pipeline(m);
    if (m->pipe_parts > 1 && m->pipe_ok && !m->timing) {
      solve_pipelined(m, nb_steps);
      return;
    }
    launch_predictor(m);
    for (int64_t s = 0; s < nb_steps; ++s) {
      launch_element_pass(m, false);
      flip_plastic_state(m);
      launch_gather(m, true, s + 1 < nb_steps);
    }
  });
}

// ModelSolver::solveStep(solver_id) (model_solver.cc:267-280): "" names the
// default solver, which initFull(_explicit_lumped_mass) registers as
// "explicit_lumped" (solid_mechanics_model.cc:165-169); any other id is not a
// solver of this model and fails as DOFManager::getTimeStepSolver does
// (dof_manager.cc:585-596).
int akx_model_solve_step(akx_model *m, const char *solver_id) {
  const int st = guard([&] {
    const std::string id = solver_id ? solver_id : "";
    if (!id.empty() && id != "explicit_lumped")
      throw Error(AKX_ERR_INVALID, "The non linear solver dof_manager:tss:" + id +
                                       " does not exists in dof_manager");
  });
  return st != AKX_OK ? st : akx_model_solve_steps(m, 1);
}

int akx_model_set_pipeline(akx_model *m, int parts, int *active) {
  return guard([&] {
    check_ready(m);
    require(parts >= 0 && parts <= 64, "parts must be in [0, 64]");
    m->set_device();
    HIPCHECK(hipStreamSynchronize(m->stream));
    m->pipe_parts = parts;
    m->pipe_planned = false;
    if (parts > 1)
      plan_pipeline(m);
    if (active)
      *active = parts > 1 && m->pipe_ok;
  });
}

int akx_model_assemble_internal_forces(akx_model *m) {
  return guard([&] {
    check_ready(m);
    m->set_device();
    // recompute from the previous plastic state so that calling it twice is
    // idempotent (the reference recomputes from the saved previous state)
    for (auto &mp : m->mats)
      if (mp->law == AKX_LAW_PLASTIC_LIH)
        mp->cur = 1 - mp->cur;
    user_stress_phase(m);
    launch_element_pass(m, false);
    flip_plastic_state(m);
    launch_gather(m, false, false);
    HIPCHECK(hipStreamSynchronize(m->stream));
  });
}

int akx_model_get_nb_nodes(akx_model *m, int64_t *nb_nodes) {
  return guard([&] {
    require(m && nb_nodes, "null argument");
    *nb_nodes = m->nn;
  });
}

int akx_model_get_field(akx_model *m, int field, void *host) {
  return guard([&] {
    require(m != nullptr && m->etype >= 0, "set the mesh first");
    require(host != nullptr, "null host array");
    m->set_device();
    HIPCHECK(hipStreamSynchronize(m->stream));
    double *src = nullptr;
    switch (field) {
    case AKX_FIELD_DISPLACEMENT: src = m->u.p; break;
    case AKX_FIELD_VELOCITY: src = m->v.p; break;
    case AKX_FIELD_ACCELERATION: src = m->a.p; break;
    case AKX_FIELD_INTERNAL_FORCE: src = m->fint.p; break;
    case AKX_FIELD_NODES: src = m->X.p; break;
    case AKX_FIELD_EXTERNAL_FORCE:
      if (!m->has_fext) {
        std::memset(host, 0, sizeof(double) * 3 * m->nn);
        return;
      }
      src = m->fext.p;
      break;
    case AKX_FIELD_MASS: {
      size_t bytes = sizeof(double) * 3 * m->nn;
      m->staging.alloc(std::max(m->staging.n, bytes));
      hipLaunchKernelGGL(k_mass_to3, dim3(grid_for(m->nn)), dim3(TPB), 0,
                         m->stream, (const double *)m->mass.p,
                         (double *)m->staging.p, m->nn);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(host, m->staging.p, bytes, hipMemcpyDeviceToHost,
                              m->stream));
      HIPCHECK(hipStreamSynchronize(m->stream));
      return;
    }
    case AKX_FIELD_BLOCKED_DOFS: {
      size_t bytes = 3 * m->nn;
      m->staging.alloc(std::max(m->staging.n, bytes));
      hipLaunchKernelGGL(k_bool3_from_bits, dim3(grid_for(m->nn)), dim3(TPB), 0,
                         m->stream, (const uint8_t *)m->blocked.p,
                         (uint8_t *)m->staging.p, m->nn);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(host, m->staging.p, bytes, hipMemcpyDeviceToHost,
                              m->stream));
      HIPCHECK(hipStreamSynchronize(m->stream));
      return;
    }
    default:
      throw Error(AKX_ERR_INVALID, "unknown field");
    }
    download_soa(m, (const double *)src, (double *)host, m->nn, 3);
  });
}

} // extern "C"

namespace {

// BC::Neumann::FromTraction on facets (boundary_condition_tmpl.hh:100-160,
// boundary_condition_functor_inline_impl.hh:109-113): per facet
// intf(i, a) = sum_q N_a(q) t_i wJ_q, then f_ext += intf in facet order
// (dof_manager.cc:65-120).  Facet classes: _triangle_3 (one Gauss point
// (1/3, 1/3), weight 1/2; element_class_triangle_3_inline_impl.hh:52-90,
// gauss_integration.cc:56-57) and _quadrangle_4 (2 x 2 Gauss points, xi
// fastest; element_class_quadrangle_4_inline_impl.hh:64-110); surface
// jacobian |J_xi x J_eta| (element_class_tmpl.hh:352-371).  Host code: it runs
// once per boundary condition, on the facets only.
void traction_on_facets(int facet_type, int64_t nb_facets, const int32_t *conn,
                        const double *X /* [nn][3] */, int64_t nn, const double *t,
                        double *fext /* [nn][3] */) {
  const int nfn = facet_type == AKX_TRI3 ? 3 : 4, nq = facet_type == AKX_TRI3 ? 1 : 4;
  double xi[4][2], w[4];
  if (facet_type == AKX_TRI3) {
    xi[0][0] = 1. / 3.;
    xi[0][1] = 1. / 3.;
    w[0] = 1. / 2.;
  } else {
    const double g = 1. / std::sqrt(3.), pos[2] = {-g, g};
    for (int q = 0; q < 4; ++q) {
      xi[q][0] = pos[q & 1];
      xi[q][1] = pos[q >> 1];
      w[q] = 1. * 1.;
    }
  }
  for (int64_t f = 0; f < nb_facets; ++f) {
    double Xe[4][3];
    for (int a = 0; a < nfn; ++a) {
      const int32_t nd = conn[f * nfn + a];
      require(nd >= 0 && nd < nn, "facet refers to a missing node");
      for (int c = 0; c < 3; ++c)
        Xe[a][c] = X[3 * (int64_t)nd + c];
    }
    double intf[12];
    for (int q = 0; q < nq; ++q) {
      double N[4], d[2][4];
      const double x = xi[q][0], y = xi[q][1];
      if (facet_type == AKX_TRI3) {
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
          double sum = d[r][0] * Xe[0][c];
          for (int a = 1; a < nfn; ++a)
            sum += d[r][a] * Xe[a][c];
          J[r][c] = sum;
        }
      const double cx = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      const double cy = J[0][2] * J[1][0] - J[0][0] * J[1][2];
      const double cz = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double wJ = std::sqrt(cx * cx + cy * cy + cz * cz) * w[q];
      for (int a = 0; a < nfn; ++a)
        for (int i = 0; i < 3; ++i) {
          const double v = N[a] * t[i] * wJ;
          intf[i + 3 * a] = (q == 0) ? v : intf[i + 3 * a] + v;
        }
    }
    for (int a = 0; a < nfn; ++a)
      for (int i = 0; i < 3; ++i)
        fext[3 * (int64_t)conn[f * nfn + a] + i] += 1. * intf[i + 3 * a];
  }
}

} // namespace

extern "C" {

int akx_model_apply_traction(akx_model *m, int facet_type, int64_t nb_facets,
                             const int32_t *facet_conn, const double *traction) {
  return guard([&] {
    require(m != nullptr && m->etype >= 0, "set the mesh first");
    require(facet_type == AKX_TRI3 || facet_type == AKX_QUAD4, "facet type must be "
                                                               "AKX_TRI3 or AKX_QUAD4");
    require(nb_facets >= 0 && (nb_facets == 0 || facet_conn) && traction,
            "null argument");
    for (int i = 0; i < 3; ++i)
      require(std::isfinite(traction[i]), "non-finite traction");
    m->set_device();
    HIPCHECK(hipStreamSynchronize(m->stream));
    std::vector<double> X(3 * m->nn), f(3 * m->nn, 0.);
    download_soa(m, (const double *)m->X.p, X.data(), m->nn, 3);
    if (m->has_fext)
      download_soa(m, (const double *)m->fext.p, f.data(), m->nn, 3);
    traction_on_facets(facet_type, nb_facets, facet_conn, X.data(), m->nn, traction,
                       f.data());
    m->fext.alloc(3 * m->nn);
    m->has_fext = true;
    upload_soa(m, (const double *)f.data(), m->fext.p, m->nn, 3);
  });
}

int akx_model_set_field(akx_model *m, int field, const void *host) {
  return guard([&] {
    require(m != nullptr && m->etype >= 0, "set the mesh first");
    require(host != nullptr, "null host array");
    m->set_device();
    HIPCHECK(hipStreamSynchronize(m->stream));
    double *dst = nullptr;
    switch (field) {
    case AKX_FIELD_DISPLACEMENT: dst = m->u.p; break;
    case AKX_FIELD_VELOCITY: dst = m->v.p; break;
    case AKX_FIELD_ACCELERATION: dst = m->a.p; break;
    case AKX_FIELD_EXTERNAL_FORCE: {
      const double *h = (const double *)host;
      bool any = false;
      for (int64_t i = 0; i < 3 * m->nn && !any; ++i)
        any = h[i] != 0.;
      if (!any && !m->has_fext)
        return;
      m->fext.alloc(3 * m->nn);
      m->has_fext = true;
      dst = m->fext.p;
      break;
    }
    case AKX_FIELD_BLOCKED_DOFS: {
      size_t bytes = 3 * m->nn;
      m->staging.alloc(std::max(m->staging.n, bytes));
      HIPCHECK(hipMemcpyAsync(m->staging.p, host, bytes, hipMemcpyHostToDevice,
                              m->stream));
      hipLaunchKernelGGL(k_bits_from_bool3, dim3(grid_for(m->nn)), dim3(TPB), 0,
                         m->stream, (const uint8_t *)m->staging.p, m->blocked.p,
                         m->nn);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(m->stream));
      return;
    }
    default:
      throw Error(AKX_ERR_INVALID, "field is read only or unknown");
    }
    upload_soa(m, (const double *)host, dst, m->nn, 3);
    if (field == AKX_FIELD_DISPLACEMENT)
      refresh_xu(m, false);
  });
}

int akx_model_compute_all_stresses(akx_model *m) {
  return guard([&] {
    check_ready(m);
    m->set_device();
    for (auto &mp : m->mats) {
      int64_t nq = mp->n * m->nqp;
      if (mp->user())
        continue; // allocated at initFull, filled by the plugin
      if (mp->law != AKX_LAW_PLASTIC_LIH) {
        mp->grad.alloc(9 * nq);
        mp->stress.alloc(9 * nq);
      } else {
        mp->cur = 1 - mp->cur; // recompute the current state from the previous
      }
    }
    user_stress_phase(m);
    launch_element_pass(m, true);
    flip_plastic_state(m);
    HIPCHECK(hipStreamSynchronize(m->stream));
  });
}

int akx_model_set_device_stress(akx_model *m, int material_id, const char *source,
                                const double *params, int nparams) {
  return guard([&] {
    require(m != nullptr, "null model");
    require(material_id >= 0 && material_id < (int)m->mats.size(), "no such material");
    Material &mat = *m->mats[material_id];
    require(mat.user(), "a device law replaces the computeStress of a user material");
    require(nparams >= 0 && nparams <= 4096 && (nparams == 0 || params != nullptr),
            "device law parameters: 0 to 4096 values");
    m->set_device();
    require(m->etype >= 0, "set the mesh first");
    std::unique_ptr<DeviceLaw> law;
    if (source != nullptr)
      law = compile_device_law(m, mat, source, params, nparams);
    HIPCHECK(hipStreamSynchronize(m->stream)); // the law it replaces may be running
    mat.dlaw = std::move(law);
  });
}

int akx_model_set_stress_callback(akx_model *m, akx_compute_stress_fn fn, void *user) {
  return guard([&] {
    require(m != nullptr, "null model");
    m->stress_fn = fn;
    m->stress_user = user;
  });
}

} // extern "C"

namespace {

// Tree sum of the n values at in[] down to one (the deterministic shape of
// k_tree_sum); in and out are overwritten, out holds >= ceil(n / 2048).
// Returns the pointer to the final value.
const double *tree_reduce(akx_model *m, double *in, double *out, int64_t n) {
  while (n > 1) {
    const int64_t nb = (n + 2047) / 2048;
    hipLaunchKernelGGL(k_tree_sum, dim3((unsigned)nb), dim3(256), 0, m->stream,
                       (const double *)in, n, out);
    HIPCHECK(hipGetLastError());
    std::swap(in, out);
    n = nb;
  }
  return in;
}

double read_back(akx_model *m, const double *p) {
  double r = 0.;
  HIPCHECK(hipMemcpyAsync(&r, p, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return r;
}

// deterministic sum of n device values (buffer a is overwritten)
double device_sum(akx_model *m, DevBuf<double> &a, DevBuf<double> &b, int64_t n) {
  if (n == 0)
    return 0.;
  return read_back(m, tree_reduce(m, a.p, b.p, n));
}

NodeEnergyArgs node_energy_args(akx_model *m, int kind) {
  NodeEnergyArgs A{};
  A.nn = m->nn;
  A.kind = kind;
  A.v = m->v.p;
  A.mass = m->mass.p;
  A.fint = m->fint.p;
  A.fext = m->has_fext ? m->fext.p : nullptr;
  A.blocked = m->blocked.p;
  require(A.v && A.mass && (kind == 0 || (A.fint && A.blocked)), "energy inputs missing");
  return A;
}

template <int ET>
void launch_energy_elements(akx_model *m, int law, const EnergyArgs &A) {
  dim3 grid(grid_for(A.n)), block(TPB);
  switch (law) {
  case AKX_LAW_ELASTIC:
    hipLaunchKernelGGL((k_energy_elements<ET, LAW_ELASTIC>), grid, block, 0, m->stream, A);
    break;
  case AKX_LAW_ELASTIC_FD:
    hipLaunchKernelGGL((k_energy_elements<ET, LAW_ELASTIC_FD>), grid, block, 0, m->stream,
                       A);
    break;
  case AKX_LAW_NEOHOOKEAN:
    hipLaunchKernelGGL((k_energy_elements<ET, LAW_NEOHOOKEAN>), grid, block, 0, m->stream,
                       A);
    break;
  default:
    hipLaunchKernelGGL((k_energy_elements<ET, LAW_PLASTIC>), grid, block, 0, m->stream, A);
  }
  HIPCHECK(hipGetLastError());
}

// the fused first tree level: ceil(n / 2048) partial sums into out
template <int ET>
void launch_energy_elements_sum(akx_model *m, int law, const EnergyArgs &A, double *out) {
  dim3 grid((unsigned)((A.n + 2047) / 2048)), block(256);
  switch (law) {
  case AKX_LAW_ELASTIC:
    hipLaunchKernelGGL((k_energy_elements_sum<ET, LAW_ELASTIC>), grid, block, 0, m->stream,
                       A, out);
    break;
  case AKX_LAW_ELASTIC_FD:
    hipLaunchKernelGGL((k_energy_elements_sum<ET, LAW_ELASTIC_FD>), grid, block, 0,
                       m->stream, A, out);
    break;
  case AKX_LAW_NEOHOOKEAN:
    hipLaunchKernelGGL((k_energy_elements_sum<ET, LAW_NEOHOOKEAN>), grid, block, 0,
                       m->stream, A, out);
    break;
  default:
    hipLaunchKernelGGL((k_energy_elements_sum<ET, LAW_PLASTIC>), grid, block, 0, m->stream,
                       A, out);
  }
  HIPCHECK(hipGetLastError());
}

// inputs of Material::getEnergy("potential" / "plastic") of one material
EnergyArgs energy_args(akx_model *m, Material &mat, bool plastic_energy) {
  const bool plastic = mat.law == AKX_LAW_PLASTIC_LIH;
  EnergyArgs A{};
  A.n = mat.n;
  A.first = mat.first;
  A.nn = m->nn;
  A.nq_all = mat.n * m->nqp;
  A.filter = mat.filter.p;
  A.conn = m->conn.p;
  A.X = m->X.p;
  A.lambda = mat.lambda;
  A.mu = mat.mu;
  A.lam_el = mat.per_element ? mat.lam_el.p : nullptr;
  A.mu_el = mat.per_element ? mat.mu_el.p : nullptr;
  if (plastic) {
    A.g = mat.hgrad[mat.cur].p;
    A.s = mat.hsig[mat.cur].p;
    A.eps = mat.heps[mat.cur].p;
    if (plastic_energy)
      A.pe = mat.hpe[mat.cur].p;
  } else {
    A.g = mat.grad.p;
    A.s = mat.stress.p;
  }
  require(A.conn && A.X && (A.pe || (A.g && A.s)), "energy inputs missing");
  return A;
}

// Material::getEnergy gives 0 for ids it does not know: "plastic" of a
// material that is not plastic
bool material_has_energy(const Material &mat, bool plastic_energy) {
  return mat.n > 0 && (!plastic_energy || mat.law == AKX_LAW_PLASTIC_LIH);
}

// Material::getEnergy("potential" / "plastic") of one material
double material_energy(akx_model *m, Material &mat, bool plastic_energy,
                       DevBuf<double> &a, DevBuf<double> &b) {
  if (!material_has_energy(mat, plastic_energy))
    return 0.;
  EnergyArgs A = energy_args(m, mat, plastic_energy);
  a.reserve(mat.n);
  b.reserve((mat.n + 2047) / 2048 + 1);
  A.out = a.p;
  if (m->etype == AKX_TET4)
    launch_energy_elements<ET_TET4>(m, mat.law, A);
  else
    launch_energy_elements<ET_HEX8>(m, mat.law, A);
  return device_sum(m, a, b, mat.n);
}

void refuse_user_potential(akx_model *m, int material_id) {
  for (int k = 0; k < (int)m->mats.size(); ++k)
    if ((material_id == -1 || material_id == k) && m->mats[k]->user())
      throw Error(AKX_ERR_UNSUPPORTED,
                  "the potential energy of a user material is not computed here");
}

// The potential energy reads the stress of the current displacement: one
// element pass that materialises grad u and sigma of the non-plastic
// materials and recomputes the plastic state from the previous one.
// allocate = false: the internals must exist already (no device allocation).
void evaluate_stresses_for_energy(akx_model *m, bool allocate) {
  bool need = false;
  for (auto &mp : m->mats)
    need = need || mp->law != AKX_LAW_PLASTIC_LIH;
  if (!need)
    return;
  for (auto &mp : m->mats) {
    if (mp->user() || mp->law == AKX_LAW_PLASTIC_LIH)
      continue;
    const size_t nq9 = (size_t)(9 * mp->n * m->nqp);
    if (allocate) {
      mp->grad.alloc(nq9);
      mp->stress.alloc(nq9);
    } else if (nq9 > 0) {
      require(mp->grad.p && mp->stress.p && mp->grad.n == nq9 && mp->stress.n == nq9,
              "the stress arrays of the materials do not exist yet: call "
              "akx_model_get_energy(\"potential\") or akx_model_compute_all_stresses once "
              "before akx_model_get_energy_noalloc(\"potential\")");
    }
  }
  user_stress_phase(m);
  for (auto &mp : m->mats)
    if (mp->law == AKX_LAW_PLASTIC_LIH)
      mp->cur = 1 - mp->cur;
  launch_element_pass(m, true);
  flip_plastic_state(m);
}

// ---------------------------------------------------------------------------
// No-allocation energy (akx_model_get_energy_noalloc): the partial sums live
// in one file-scope device array.  Each device holds its own copy of it; the
// host reaches it only through the address hipGetSymbolAddress gives on that
// device (the array's name is a host-side stand-in, not a device address).
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_PARTIALS = AKX_NOALLOC_MAX_ITEMS / 2048; // first level
constexpr int64_t NOALLOC_SCRATCH = NOALLOC_PARTIALS + 256;        // + second level

} // namespace

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[MAX_DEVICES] = {};

// Set up once per device, with that device current (akx_model_create).
double *device_scratch(int device) {
  require(device >= 0 && device < MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    HIPCHECK(hipGetDevice(&cur));
    require(cur == device, "scratch set-up needs its device current");
    void *p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "no device address for the energy scratch array");
    g_scratch_ptr[device] = static_cast<double *>(p);
  }
  return g_scratch_ptr[device];
}

// partials of the first level at sc[0 ..), the second level at
// sc[NOALLOC_PARTIALS ..); returns the final value
double noalloc_reduce(akx_model *m, double *sc, int64_t first_level_sums) {
  return read_back(m, tree_reduce(m, sc, sc + NOALLOC_PARTIALS, first_level_sums));
}

void require_noalloc_size(int64_t n, const char *what) {
  if (n > AKX_NOALLOC_MAX_ITEMS)
    throw Error(AKX_ERR_INVALID,
                std::string("akx_model_get_energy_noalloc: ") + what + " (" +
                    std::to_string(n) + ") exceed the scratch limit of " +
                    std::to_string((long long)AKX_NOALLOC_MAX_ITEMS));
}

} // namespace

extern "C" {

int akx_model_get_energy(akx_model *m, const char *energy_id, int material_id,
                         double *energy) {
  return guard([&] {
    check_ready(m);
    require(energy_id != nullptr && energy != nullptr, "null argument");
    require(material_id >= -1 && material_id < (int)m->mats.size(), "bad material id");
    m->set_device();
    const std::string id(energy_id);
    DevBuf<double> a, b;
    if (id == "kinetic" || id == "external work") {
      require(material_id == -1, "kinetic energy and external work are model energies");
      const int kind = id == "kinetic" ? 0 : 1;
      a.reserve(m->nn);
      b.reserve((m->nn + 2047) / 2048 + 1);
      hipLaunchKernelGGL(k_node_energy, dim3(grid_for(m->nn)), dim3(TPB), 0, m->stream,
                         node_energy_args(m, kind), a.p);
      HIPCHECK(hipGetLastError());
      const double s = device_sum(m, a, b, m->nn);
      *energy = kind == 0 ? s / 2. : s * m->dt;
      return;
    }
    if (id != "potential" && id != "plastic") {
      *energy = 0.; // the reference's Material::getEnergy returns 0 for other ids
      return;
    }
    const bool plastic_energy = id == "plastic";
    if (!plastic_energy) {
      refuse_user_potential(m, material_id);
      evaluate_stresses_for_energy(m, true);
    }
    double e = 0.;
    for (int k = 0; k < (int)m->mats.size(); ++k)
      if (material_id == -1 || material_id == k)
        e += material_energy(m, *m->mats[k], plastic_energy, a, b);
    *energy = e;
  });
}

int akx_model_get_energy_noalloc(akx_model *m, const char *energy_id, int material_id,
                                 double *energy) {
  return guard([&] {
    check_ready(m);
    require(energy_id != nullptr && energy != nullptr, "null argument");
    require(material_id >= -1 && material_id < (int)m->mats.size(), "bad material id");
    m->set_device();
    double *sc = device_scratch(m->device);
    const std::string id(energy_id);
    if (id == "kinetic" || id == "external work") {
      require(material_id == -1, "kinetic energy and external work are model energies");
      require_noalloc_size(m->nn, "nodes");
      const int kind = id == "kinetic" ? 0 : 1;
      double s = 0.;
      if (m->nn > 0) {
        const int64_t nb = (m->nn + 2047) / 2048;
        hipLaunchKernelGGL(k_node_energy_sum, dim3((unsigned)nb), dim3(256), 0, m->stream,
                           node_energy_args(m, kind), sc);
        HIPCHECK(hipGetLastError());
        s = noalloc_reduce(m, sc, nb);
      }
      *energy = kind == 0 ? s / 2. : s * m->dt;
      return;
    }
    if (id != "potential" && id != "plastic") {
      *energy = 0.;
      return;
    }
    const bool plastic_energy = id == "plastic";
    for (int k = 0; k < (int)m->mats.size(); ++k)
      if (material_id == -1 || material_id == k)
        require_noalloc_size(m->mats[k]->n, "elements of a material");
    if (!plastic_energy) {
      refuse_user_potential(m, material_id);
      if (m->has_user())
        throw Error(AKX_ERR_UNSUPPORTED,
                    "akx_model_get_energy_noalloc(\"potential\") runs no computeStress "
                    "plugin: the model has a user material");
      evaluate_stresses_for_energy(m, false);
    }
    double e = 0.;
    for (int k = 0; k < (int)m->mats.size(); ++k) {
      if (material_id != -1 && material_id != k)
        continue;
      Material &mat = *m->mats[k];
      if (!material_has_energy(mat, plastic_energy))
        continue;
      EnergyArgs A = energy_args(m, mat, plastic_energy);
      if (m->etype == AKX_TET4)
        launch_energy_elements_sum<ET_TET4>(m, mat.law, A, sc);
      else
        launch_energy_elements_sum<ET_HEX8>(m, mat.law, A, sc);
      e += noalloc_reduce(m, sc, (mat.n + 2047) / 2048);
    }
    *energy = e;
  });
}

int akx_model_get_internal(akx_model *m, int material_id, int internal,
                           double *host) {
  return guard([&] {
    check_ready(m);
    require(host != nullptr, "null host array");
    require(material_id >= 0 && material_id < (int)m->mats.size(), "bad material id");
    m->set_device();
    Material &mat = *m->mats[material_id];
    int64_t nq = mat.n * m->nqp;
    const double *src = nullptr;
    int comps = 9;
    bool plastic = mat.law == AKX_LAW_PLASTIC_LIH;
    switch (internal) {
    case AKX_INTERNAL_GRAD_U: src = plastic ? mat.hgrad[mat.cur].p : mat.grad.p; break;
    case AKX_INTERNAL_STRESS: src = plastic ? mat.hsig[mat.cur].p : mat.stress.p; break;
    case AKX_INTERNAL_INELASTIC_STRAIN:
      require(plastic, "inelastic_strain exists only for plastic materials");
      src = mat.heps[mat.cur].p;
      break;
    case AKX_INTERNAL_ISO_HARDENING:
      require(plastic, "iso_hardening exists only for plastic materials");
      src = mat.hiso[mat.cur].p;
      comps = 1;
      break;
    default:
      throw Error(AKX_ERR_INVALID, "unknown internal");
    }
    require(src != nullptr, "call computeAllStresses first");
    if (nq == 0)
      return;
    download_soa(m, src, host, nq, comps);
  });
}

} // extern "C"

namespace {

// ---------------------------------------------------------------------------
// FEEngine / DOFManager calls on arbitrary fields, in the operation order of
// the element passes (the reference's):
//   gradientOnIntegrationPoints  g(k, j) = sum_n u(k, n) B(j, n)
//                                (shape_functions_inline_impl.hh:296-310)
//   computeBtD                   BtD(k, n) = sum_j D(k, j) B(j, n)
//                                (shape_lagrange_inline_impl.hh:373-410)
//   integrate                    I(k) = sum_q f(k, q) wJ_q
//                                (integrator_gauss_inline_impl.hh:401-441)
//   assembleElementalArrayLocalArray
//                                out(conn(e, n), k) += s * r_e(k + d n), the
//                                elements in filter order
//                                (dof_manager.cc:65-120)
// Per-point arrays are [point][comps] with the points of element i at rows
// i * nqp ..; matrices per point are column-major (rows k, columns j).
// ---------------------------------------------------------------------------
struct FeArgs {
  int64_t n;             // elements of the call (filter size)
  const int32_t *filter; // null: element i is i
  const int32_t *conn;
  const double *X; // [3][nn]
  int64_t nn;
  const double *in;
  double *out;
  int d; // rows of the per-point matrix (degrees of freedom)
};

template <int ET>
__device__ __forceinline__ void fe_point(const FeArgs &A, int64_t i, int q, int32_t (&nodes)[Elem<ET>::nnpe],
                                         double (&b)[Elem<ET>::nnpe][3], double &wJ) {
  constexpr int NN = Elem<ET>::nnpe;
  const int64_t e = A.filter ? (int64_t)A.filter[i] : i;
  double X[NN][3];
#pragma unroll
  for (int n = 0; n < NN; ++n) {
    nodes[n] = A.conn[e * NN + n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      X[n][c] = A.X[(int64_t)c * A.nn + nodes[n]];
  }
  double Ji[3][3], detJ;
  wJ = Shape<ET, double>::jacobian_at(X, q, Ji, detJ);
#pragma unroll
  for (int n = 0; n < NN; ++n)
    Shape<ET, double>::bcol(Ji, q, n, b[n]);
}

// one thread per (element, Gauss point); in = field [nn][d]
template <int ET> __global__ void k_fe_gradient(FeArgs A) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.n * NQ)
    return;
  int32_t nodes[NN];
  double b[NN][3], wJ;
  fe_point<ET>(A, t / NQ, (int)(t % NQ), nodes, b, wJ);
  double *o = A.out + t * 3 * A.d;
  for (int k = 0; k < A.d; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = A.in[(int64_t)nodes[0] * A.d + k] * b[0][j];
#pragma unroll
      for (int n = 1; n < NN; ++n)
        s = s + A.in[(int64_t)nodes[n] * A.d + k] * b[n][j];
      o[k + A.d * j] = s;
    }
}

// one thread per (element, Gauss point); in = D [point][d * 3]
template <int ET> __global__ void k_fe_btd(FeArgs A) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.n * NQ)
    return;
  int32_t nodes[NN];
  double b[NN][3], wJ;
  fe_point<ET>(A, t / NQ, (int)(t % NQ), nodes, b, wJ);
  const double *D = A.in + t * 3 * A.d;
  double *o = A.out + t * NN * A.d;
  for (int k = 0; k < A.d; ++k)
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      double s = D[k] * b[n][0];
      s += D[k + A.d] * b[n][1];
      s += D[k + 2 * A.d] * b[n][2];
      o[k + A.d * n] = s;
    }
}

// one thread per element; in = f [point][d]
template <int ET> __global__ void k_fe_integrate(FeArgs A) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n)
    return;
  double wJ[NQ];
  for (int q = 0; q < NQ; ++q) {
    int32_t nodes[NN];
    double b[NN][3];
    fe_point<ET>(A, i, q, nodes, b, wJ[q]);
  }
  for (int k = 0; k < A.d; ++k) {
    double s = A.in[(i * NQ) * A.d + k] * wJ[0];
    for (int q = 1; q < NQ; ++q)
      s = s + A.in[(i * NQ + q) * A.d + k] * wJ[q];
    A.out[i * A.d + k] = s;
  }
}

// one thread per node: its contributions (offsets into the elemental
// array, in filter order) added in that order
__global__ void k_dof_assemble(int64_t nn, const int64_t *rowptr, const int64_t *entry,
                               const double *elem, int d, double scale, double *out) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn)
    return;
  for (int k = 0; k < d; ++k) {
    double acc = out[node * d + k];
    for (int64_t r = rowptr[node]; r < rowptr[node + 1]; ++r)
      acc += scale * elem[entry[r] + k];
    out[node * d + k] = acc;
  }
}

struct FeCall {
  akx_model *m;
  FeArgs A{};
  DevBuf<int32_t> filter;
  DevBuf<double> in, out;
  FeCall(akx_model *mm, const int32_t *filt, int64_t nb_filter, int d) : m(mm) {
    check_ready(m);
    require(d >= 1 && d <= 64, "nb_degree_of_freedom must be in [1, 64]");
    m->set_device();
    A.n = filt ? nb_filter : m->nel;
    require(A.n >= 0, "negative filter size");
    if (filt && nb_filter > 0) {
      for (int64_t i = 0; i < nb_filter; ++i)
        require(filt[i] >= 0 && filt[i] < m->nel, "filter refers to a missing element");
      filter.alloc(nb_filter);
      HIPCHECK(hipMemcpy(filter.p, filt, nb_filter * 4, hipMemcpyHostToDevice));
    }
    A.filter = filter.p;
    A.conn = m->conn.p;
    A.X = m->X.p;
    A.nn = m->nn;
    A.d = d;
  }
  void put(const double *host, int64_t count) {
    in.alloc(count > 0 ? count : 1);
    if (count > 0)
      HIPCHECK(hipMemcpy(in.p, host, count * 8, hipMemcpyHostToDevice));
    A.in = in.p;
  }
  void make_out(int64_t count) {
    out.alloc(count > 0 ? count : 1);
    A.out = out.p;
  }
  void get(double *host, int64_t count) {
    HIPCHECK(hipStreamSynchronize(m->stream));
    if (count > 0)
      HIPCHECK(hipMemcpy(host, out.p, count * 8, hipMemcpyDeviceToHost));
  }
};

template <template <int> class K> void fe_launch(akx_model *m, const FeArgs &A, int64_t threads) {
  if (threads == 0)
    return;
  const dim3 grid(grid_for(threads)), block(TPB);
  if (m->etype == AKX_HEX8)
    hipLaunchKernelGGL(K<ET_HEX8>::fn, grid, block, 0, m->stream, A);
  else
    hipLaunchKernelGGL(K<ET_TET4>::fn, grid, block, 0, m->stream, A);
  HIPCHECK(hipGetLastError());
}

template <int ET> struct KGrad { static constexpr auto fn = k_fe_gradient<ET>; };
template <int ET> struct KBtd { static constexpr auto fn = k_fe_btd<ET>; };
template <int ET> struct KInt { static constexpr auto fn = k_fe_integrate<ET>; };

} // namespace

extern "C" {

int akx_fe_gradient_on_integration_points(akx_model *m, const double *field, int nb_dof,
                                          const int32_t *filter, int64_t nb_filter,
                                          double *out) {
  return guard([&] {
    require(field && out, "null array");
    FeCall c(m, filter, nb_filter, nb_dof);
    const int64_t npt = c.A.n * m->nqp;
    c.put(field, m->nn * nb_dof);
    c.make_out(npt * 3 * nb_dof);
    fe_launch<KGrad>(m, c.A, npt);
    c.get(out, npt * 3 * nb_dof);
  });
}

int akx_fe_compute_btd(akx_model *m, const double *Ds, int nb_comp, const int32_t *filter,
                       int64_t nb_filter, double *BtDs) {
  return guard([&] {
    require(Ds && BtDs, "null array");
    require(nb_comp % 3 == 0, "D must have 3 columns per point (nb_component = 3 d)");
    FeCall c(m, filter, nb_filter, nb_comp / 3);
    const int64_t npt = c.A.n * m->nqp;
    c.put(Ds, npt * nb_comp);
    c.make_out(npt * m->nnpe * c.A.d);
    fe_launch<KBtd>(m, c.A, npt);
    c.get(BtDs, npt * m->nnpe * c.A.d);
  });
}

int akx_fe_integrate(akx_model *m, const double *f, int nb_dof, const int32_t *filter,
                     int64_t nb_filter, double *intf) {
  return guard([&] {
    require(f && intf, "null array");
    FeCall c(m, filter, nb_filter, nb_dof);
    c.put(f, c.A.n * m->nqp * nb_dof);
    c.make_out(c.A.n * nb_dof);
    fe_launch<KInt>(m, c.A, c.A.n);
    c.get(intf, c.A.n * nb_dof);
  });
}

int akx_dof_assemble_elemental_array(akx_model *m, const double *elementary, int nb_dof,
                                     const int32_t *filter, int64_t nb_filter,
                                     double scale, double *assembled) {
  return guard([&] {
    require(elementary && assembled, "null array");
    FeCall c(m, filter, nb_filter, nb_dof);
    const int nnpe = m->nnpe;
    // node -> contributions, in filter order (a counting sort)
    std::vector<int64_t> rowptr(m->nn + 1, 0), entry(c.A.n * nnpe);
    auto elem_of = [&](int64_t i) { return filter ? (int64_t)filter[i] : i; };
    for (int64_t i = 0; i < c.A.n; ++i)
      for (int a = 0; a < nnpe; ++a)
        rowptr[m->conn_host[elem_of(i) * nnpe + a] + 1]++;
    for (int64_t k = 0; k < m->nn; ++k)
      rowptr[k + 1] += rowptr[k];
    std::vector<int64_t> fill(rowptr.begin(), rowptr.end() - 1);
    for (int64_t i = 0; i < c.A.n; ++i)
      for (int a = 0; a < nnpe; ++a)
        entry[fill[m->conn_host[elem_of(i) * nnpe + a]]++] = (i * nnpe + a) * nb_dof;
    DevBuf<int64_t> drow, dent;
    drow.alloc(rowptr.size());
    dent.alloc(entry.empty() ? 1 : entry.size());
    HIPCHECK(hipMemcpy(drow.p, rowptr.data(), rowptr.size() * 8, hipMemcpyHostToDevice));
    if (!entry.empty())
      HIPCHECK(hipMemcpy(dent.p, entry.data(), entry.size() * 8, hipMemcpyHostToDevice));
    c.put(elementary, c.A.n * nnpe * nb_dof);
    c.make_out(m->nn * nb_dof);
    HIPCHECK(hipMemcpy(c.out.p, assembled, m->nn * nb_dof * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_dof_assemble, dim3(grid_for(m->nn)), dim3(TPB), 0, m->stream, m->nn,
                       (const int64_t *)drow.p, (const int64_t *)dent.p, (const double *)c.in.p,
                       nb_dof, scale, c.out.p);
    HIPCHECK(hipGetLastError());
    c.get(assembled, m->nn * nb_dof);
  });
}

int akx_calibration_copy(int device, int64_t bytes, int repeats, double *ms) {
  return guard([&] {
    require(bytes >= 32 && repeats > 0 && ms, "bad calibration arguments");
    HIPCHECK(hipSetDevice(device));
    const int64_t n = bytes / 32;
    DevBuf<double4> a, b;
    a.alloc(n);
    b.alloc(n);
    hipStream_t st;
    HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    require((n + 511) / 512 < (int64_t)UINT32_MAX, "calibration copy too large");
    const unsigned grid = (unsigned)((n + 511) / 512);
    hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, st,
                       (const double4 *)a.p, b.p, n); // warm-up
    HIPCHECK(hipEventRecord(e0, st));
    for (int r = 0; r < repeats; ++r)
      hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, st,
                         (const double4 *)a.p, b.p, n);
    HIPCHECK(hipEventRecord(e1, st));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    *ms = (double)t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
  });
}

int akx_model_synchronize(akx_model *m) {
  return guard([&] {
    require(m != nullptr, "null model");
    m->set_device();
    HIPCHECK(hipStreamSynchronize(m->stream));
  });
}

int akx_model_enable_kernel_timing(akx_model *m, int enable) {
  return guard([&] {
    require(m != nullptr, "null model");
    m->drain_timing();
    m->timing = enable != 0;
    for (int k = 0; k < 3; ++k) {
      m->t_ms[k] = 0;
      m->t_n[k] = 0;
    }
  });
}

int akx_model_get_kernel_timing(akx_model *m, int kernel_class, double *total_ms,
                                int64_t *launches) {
  return guard([&] {
    require(m != nullptr, "null model");
    require(kernel_class >= 0 && kernel_class < 3, "bad kernel class");
    m->set_device();
    m->drain_timing();
    if (total_ms)
      *total_ms = m->t_ms[kernel_class];
    if (launches)
      *launches = m->t_n[kernel_class];
  });
}

int akx_model_get_kernel_bytes(akx_model *m, int kernel_class, double *bytes) {
  return guard([&] {
    check_ready(m);
    require(bytes != nullptr, "null output");
    const double s_node = 8.0; // FP64 nodal state
    const double s_f = (double)m->fe_elem_size();
    const double s_c = m->precision == AKX_PRECISION_FP32 ? 4.0 : 8.0;
    int64_t nel_all = 0;
    double extra = 0.;
    for (auto &mp : m->mats) {
      nel_all += mp->n;
      if (mp->per_element) // lambda, mu per element: FP64, or the FP32 pair
        extra += mp->n * (mp->lm_f.p ? 2.0 * 4.0 : 2.0 * 8.0);
      if (mp->lconn.p && m->etype == AKX_TET4) // staged records: 2-byte local
        // corners instead of 4-byte node ids, plus the blocks' node lists
        extra += -2.0 * m->nnpe * mp->n + 4.0 * (double)mp->bl_nodes.n;
      if (mp->law == AKX_LAW_PLASTIC_LIH)
        extra += (double)mp->n * m->nqp * (2 * (9 + 9 + 9 + 1) * 8.0); // read + write history
      if (!mp->contiguous)
        extra += mp->n * 4.0;
    }
    (void)s_c;
    // table entries the element pass writes and the gather reads: one per
    // element corner, or one per (block, distinct node) partial
    double entries = (double)nel_all * m->nnpe;
    if (m->bpa) {
      double partials = 0.;
      for (auto &mp : m->mats)
        for (int32_t node : mp->bl_nodes_host)
          partials += node >= 0 ? 1. : 0.;
      // corner slots (2 B) instead of table positions (4 B); per partial its
      // slot offset (2 B) and table position (4 B)
      extra += -2.0 * m->nnpe * nel_all + 6.0 * partials;
      entries = partials;
    } else if (m->ppa) {
      // corner codes (4 B, as the positions); per summed run its slot end
      // (2 B) and table position (4 B)
      extra += 6.0 * (double)m->pp_runs;
      entries = (double)m->pp_entries;
    }
    switch (kernel_class) {
    case 0: // element pass: connectivity + table entries, X and u once per
            // node, r_e write; with the Hex8 jacobian table, the table (80 B
            // per Gauss point) instead of X
      if (m->jw.p)
        *bytes = nel_all * (8.0 * m->nnpe + 80.0 * m->nqp) + entries * 3.0 * s_f +
                 m->nn * 3.0 * s_node + extra;
      else
        *bytes = nel_all * 8.0 * m->nnpe + entries * 3.0 * s_f +
                 m->nn * (3.0 * s_node + 3.0 * s_node) + extra;
      break;
    case 1: { // gather + update of a step followed by another: degree, r_e
              // read, mass, blocked, a/v read+write, u read+write (fused
              // predictor); the last step of a call also writes f_int.  The
              // FP32 form reads and writes no a between the steps of a call
              // (its velocity array holds the half-step velocity), and its
              // compensation terms
      const bool f32 = m->precision == AKX_PRECISION_FP32;
      double per_node = 4.0 + s_node /*M*/ + 1.0 + (m->records_live() ? 3 * s_node : 0.) /*u records*/ +
                        (f32 ? 0. : 2 * 3 * s_node) /*a*/ + 2 * 3 * s_node /*v*/ +
                        2 * 3 * s_node /*u*/ + (m->has_fext ? 3 * s_node : 0.);
      if (f32)
        per_node += 2 * 2 * 3 * 4.0; // compensation terms read + write
      *bytes = entries * (3.0 * s_f) + m->nn * (per_node - 3.0);
      break;
    }
    case 2:
      *bytes = m->nn * (1.0 + 3 * s_node + 2 * 2 * 3 * s_node);
      break;
    default:
      throw Error(AKX_ERR_INVALID, "bad kernel class");
    }
  });
}

} // extern "C"
