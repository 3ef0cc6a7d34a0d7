// akx_batch.hip -- many independent meshes advanced together on one GPU: the
// batching form (one persistent launch per substep over all instances,
// captured in a hipGraph, launch shape chosen by a project-written
// autotuner) and the substrate of the serving form (instances admitted and
// retired between chunks of steps while the others stay resident).
//
// Memory: arenas sized at creation.
//   nodes    one 48-byte record {X, u} per node (the element pass reads a node
//            with three 16-byte loads); SoA [3][node_cap] v, a, f_int;
//            M [node_cap]; blocked bits; an instance owns a node range
//            aligned to 256 (whole ELL slices)
//   elements AoS connectivity and table positions per element type
//   table    node-major sliced-ELL elemental forces (FP64 or FP32)
//   history  plastic ping-pong: FP64 form grad u, stress, plastic strain
//            [2][9][qp_cap] and hardening; FP32 form the compact state,
//            plastic strain [2][6][qp_cap] and hardening, FP32
// Free lists (first fit) hand out and take back ranges.
//
// Per step: element pass (one persistent launch per (element type, law)
// group present, over tiles that never span an instance, so the material
// parameters and the plastic parity are uniform per block) and the nodal
// gather + update (one persistent launch over the 256-node slices of the
// live instances, dt uniform per block).  The arithmetic of the FP64 form is
// the reference's operation order, element by element identical to the
// single-model plain path, so an instance gives the same bits alone and in
// the mix.
#include "../../include/akantu_mi355x.h"
#include "akx_batch.hpp"
#include "akx_fast.hpp"
#include "akx_host.hpp"
#include "akx_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace akx {

// ---------------------------------------------------------------------------
// Element pass of the FP64 form: the single-model code (reference operation
// order) over the tiles of one (element type, law) group, one tile per
// workgroup (persistent workgroups walking chunks of tiles measured 7-20%
// slower: their load latencies ran back to back).
// ---------------------------------------------------------------------------
template <int LAW>
__global__ void __launch_bounds__(256) k_batch_tet4(BatchElemArgs A) {
  {
    const int64_t tt = blockIdx.x; // one tile per workgroup
    if (tt >= A.ntiles)
      return;
    const BTile T = A.tiles[tt];
    const BInst P = A.inst[T.inst];
    const QpView V = batch_view<double>(A, (P.parity_base ^ A.parity) & 1);
    for (int j = threadIdx.x; j < T.count; j += blockDim.x) {
      const int64_t e = T.first + j;
      element_forces_one<ET_TET4, LAW, double, double, double, NodesXU, FE_W>(
          A.conn + e * 4, NodesXU{A.xu}, P.lambda, P.mu, P.sigma_y, P.h, V, A.hoff + e,
          A.pos + e * 4, A.fe);
    }
  }
}

template <int LAW, int EPB>
__global__ void __launch_bounds__(EPB * 8) k_batch_hex8(BatchElemArgs A) {
  __shared__ double sN[EPB][8][6];
  __shared__ double sD[8][3][8];
  const int t = threadIdx.x, el = t >> 3, q = t & 7;
  for (int k = t; k < 192; k += EPB * 8)
    (&sD[0][0][0])[k] = A.hex_tab[k];
  const double wq = A.hex_tab[HEX_TAB_W + q];
  __syncthreads();
  {
    const int64_t tt = blockIdx.x; // one tile per workgroup
    if (tt >= A.ntiles)
      return;
    const BTile T = A.tiles[tt];
    const BInst P = A.inst[T.inst];
    const QpView V = batch_view<double>(A, (P.parity_base ^ A.parity) & 1);
    for (int base = 0; base < T.count; base += EPB) {
      const bool active = base + el < T.count;
      const int64_t e = T.first + (active ? base + el : 0);
      hex8_strict_group<LAW, EPB, double, NodesXU, FE_W>(sN, sD, wq, el, q, active,
                                                         A.conn + e * 8,
                                          NodesXU{A.xu}, P.lambda, P.mu, P.sigma_y, P.h,
                                          V, A.hoff + e * 8 + q, A.pos + e * 8,
                                          reinterpret_cast<double *>(A.fe));
    }
  }
}

// ---------------------------------------------------------------------------
// Gather + residual + lumped solve + corrector (+ next predictor), persistent
// over the 256-node slices of the live instances: the single-model
// k_gather_update arithmetic (zero external force) with the instance's dt.
// ---------------------------------------------------------------------------
struct BatchGatherArgs {
  const BSlice *slices;
  int64_t nslices;
  const BInst *inst;
  const int64_t *ell_off; // [ncap / 256] first entry of each slice
  const uint8_t *deg;
  const void *fe;
  double *fint, *v, *a; // [3][ncap]
  double *xu;           // [ncap][6] node records {X, u}
  double *u;            // [3][ncap] displacement of the instances with urec = 0
  const double *mass;
  const uint8_t *blocked;
  float *cu, *cv; // compensation terms (FP32 form) or null
  int64_t ncap;
  int64_t nblocks;
};

// One 256-node slice per workgroup (a persistent loop over slices measured
// 8% slower: each wave walked its slices' dependent loads one after another).
template <class FT, bool PREDICT, bool COMP>
__global__ void __launch_bounds__(256) k_batch_gather(BatchGatherArgs G) {
  {
    const int64_t s = blockIdx.x;
    if (s >= G.nslices)
      return;
    const BSlice SL = G.slices[s];
    if ((int)threadIdx.x >= SL.count)
      return;
    const double dt = G.inst[SL.inst].dt;
    const bool urec = G.inst[SL.inst].urec != 0;
    const int64_t n = SL.node0 + threadIdx.x;
    const FT *p = reinterpret_cast<const FT *>(G.fe) + (G.ell_off[n >> 8] + (n & 255)) * FE_W;
    double f[3] = {0., 0., 0.};
    const int deg = G.deg[n];
    for (int d = 0; d < deg; ++d, p += 256 * FE_W) {
      double e0, e1, e2;
      if constexpr (FE_W == 3) { // read once per step: non-temporal
        e0 = __builtin_nontemporal_load(p);
        e1 = __builtin_nontemporal_load(p + 1);
        e2 = __builtin_nontemporal_load(p + 2);
      } else if constexpr (sizeof(FT) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        e0 = v.x;
        e1 = v.y;
        e2 = v.z;
      } else {
        const double2 v0 = reinterpret_cast<const double2 *>(p)[0];
        const double2 v1 = reinterpret_cast<const double2 *>(p)[1];
        e0 = v0.x;
        e1 = v0.y;
        e2 = v1.x;
      }
      f[0] = f[0] + -1. * e0;
      f[1] = f[1] + -1. * e1;
      f[2] = f[2] + -1. * e2;
    }
    const double M = G.mass[n];
    const uint8_t blk = G.blocked[n];
    // u of the node's record (u0 at 48 n + 24 bytes, then an aligned pair)
    // or of the [3][ncap] array, read and written whole; blocked components
    // keep their value
    double us[3] = {0., 0., 0.};
    if constexpr (PREDICT) {
      if (urec) {
        const double2 u12 = *reinterpret_cast<const double2 *>(G.xu + 6 * n + 4);
        us[0] = G.xu[6 * n + 3];
        us[1] = u12.x;
        us[2] = u12.y;
      } else {
        us[0] = G.u[n];
        us[1] = G.u[G.ncap + n];
        us[2] = G.u[2 * G.ncap + n];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t o = (int64_t)c * G.ncap + n;
      if (G.fint) // null while more steps follow in the same call
        G.fint[o] = f[c];
      if ((blk >> c) & 1)
        continue;
      if constexpr (COMP) { // FP32 form: half-step velocity (akx_math.hpp)
        double w = G.v[o], a;
        float cw = G.cv[o], cu = 0.f;
        if constexpr (PREDICT)
          cu = G.cu[o];
        half_step_gather<PREDICT>(f[c], M, dt, w, cw, us[c], cu, a);
        if constexpr (PREDICT)
          G.cu[o] = cu;
        else
          G.a[o] = a;
        G.v[o] = w;
        G.cv[o] = cw;
      } else { // reference order, zero external force (akx_math.hpp)
        double a = G.a[o], v = G.v[o];
        reference_gather<PREDICT>(0., f[c], M, dt, a, v, us[c]);
        G.a[o] = a;
        G.v[o] = v;
      }
    }
    if constexpr (PREDICT) {
      if (urec) {
        G.xu[6 * n + 3] = us[0];
        *reinterpret_cast<double2 *>(G.xu + 6 * n + 4) = make_double2(us[1], us[2]);
      } else {
        G.u[n] = us[0];
        G.u[G.ncap + n] = us[1];
        G.u[2 * G.ncap + n] = us[2];
      }
    }
  }
}

// predictor alone (first step of a solve call): newmark-beta.cc:47-77
template <bool COMP>
__global__ void __launch_bounds__(256) k_batch_predictor(BatchGatherArgs G) {
  {
    const int64_t s = blockIdx.x;
    if (s >= G.nslices)
      return;
    const BSlice SL = G.slices[s];
    if ((int)threadIdx.x >= SL.count)
      return;
    const double dt = G.inst[SL.inst].dt;
    const bool urec = G.inst[SL.inst].urec != 0;
    const int64_t n = SL.node0 + threadIdx.x;
    const uint8_t blk = G.blocked[n];
    double us[3];
    if (urec) {
      const double2 u12 = *reinterpret_cast<const double2 *>(G.xu + 6 * n + 4);
      us[0] = G.xu[6 * n + 3];
      us[1] = u12.x;
      us[2] = u12.y;
    } else {
      us[0] = G.u[n];
      us[1] = G.u[G.ncap + n];
      us[2] = G.u[2 * G.ncap + n];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if ((blk >> c) & 1)
        continue;
      const int64_t o = (int64_t)c * G.ncap + n;
      const double a = G.a[o];
      double v = G.v[o], u = us[c];
      if constexpr (COMP) { // the velocity array takes the half-step velocity
        float cu = G.cu[o], cv = G.cv[o];
        half_step_predict(a, dt, v, cv, u, cu);
        G.cu[o] = cu;
        G.cv[o] = cv;
      } else {
        reference_predict(a, dt, v, u);
      }
      us[c] = u;
      G.v[o] = v;
    }
    if (urec) {
      G.xu[6 * n + 3] = us[0];
      *reinterpret_cast<double2 *>(G.xu + 6 * n + 4) = make_double2(us[1], us[2]);
    } else {
      G.u[n] = us[0];
      G.u[G.ncap + n] = us[1];
      G.u[2 * G.ncap + n] = us[2];
    }
  }
}

// ---------------------------------------------------------------------------
// Init kernels of one instance (the single-model init arithmetic, Hex8
// tables from the device table): jacobian check, lumped-mass contributions
// m_e(a) = sum_q (rho N_a(q)) wJ_q (fe_engine_template_tmpl_field.hh:112-138),
// min_e h_e / c_e (solid_mechanics_model.cc:423-473).
// ---------------------------------------------------------------------------
struct BatchInitArgs {
  int64_t first, n;
  const int32_t *conn, *pos;
  const double *xu; // [ncap][6] node records
  int64_t ncap;
  double rho, lambda, mu;
  const double *hex_tab;
  double *me; // mass contribution of table entry ell0 + k at me[k]
  int64_t ell0;
  unsigned long long *min_bits;
  long long *first_bad;
};

template <int ET>
__device__ __forceinline__ void batch_load_X(const BatchInitArgs &A, int64_t e,
                                             double (&X)[Elem<ET>::nnpe][3]) {
#pragma unroll
  for (int n = 0; n < Elem<ET>::nnpe; ++n) {
    const int32_t node = A.conn[e * Elem<ET>::nnpe + n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      X[n][c] = A.xu[6 * (int64_t)node + c];
  }
}

// w * detJ at Gauss point q (the sums of Shape<ET, double>::jacobian_at)
template <int ET>
__device__ __forceinline__ double batch_wj(const BatchInitArgs &A,
                                           const double (&X)[Elem<ET>::nnpe][3], int q,
                                           double &detJ) {
  double Ji[3][3];
  if constexpr (ET == ET_TET4) {
    return Shape<ET_TET4, double>::jacobian_at(X, q, Ji, detJ);
  } else {
    const double *d = A.hex_tab + q * 24;
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = d[i * 8] * X[0][j];
#pragma unroll
        for (int n = 1; n < 8; ++n)
          s += d[i * 8 + n] * X[n][j];
        J[i][j] = s;
      }
    detJ = jacobian_det_inv(J, Ji);
    return detJ * A.hex_tab[HEX_TAB_W + q];
  }
}

template <int ET> __global__ void k_batch_mass(BatchInitArgs A) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n)
    return;
  const int64_t e = A.first + i;
  double X[NN][3];
  batch_load_X<ET>(A, e, X);
  double m[NN];
#pragma unroll
  for (int n = 0; n < NN; ++n)
    m[n] = 0.;
  for (int q = 0; q < NQ; ++q) {
    double detJ;
    const double wJ = batch_wj<ET>(A, X, q, detJ);
    if (detJ < 0.)
      atomicMin(A.first_bad, (long long)(i * NQ + q));
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      const double N = (ET == ET_TET4) ? .25 : A.hex_tab[HEX_TAB_N + q * 8 + n];
      m[n] += (A.rho * N) * wJ;
    }
  }
#pragma unroll
  for (int n = 0; n < NN; ++n)
    A.me[A.pos[e * NN + n] - A.ell0] = m[n];
}

// Lumped-mass contributions of one instance under block-partial assembly
// (FP32 batches): block b (BPA_TET4 / BPA_HEX8 elements) puts each corner's m_e(a) (the
// sums of k_batch_mass) into an LDS slot (the corners of one distinct node
// take consecutive slots in ascending element order), then one thread per
// distinct node adds its slots in that order and writes the block's partial
// to the node's table entry; k_batch_gather_mass adds the partials in
// ascending block order.
template <int ET>
__global__ void __launch_bounds__(256) k_batch_mass_bpa(BatchInitArgs A, const BBlk *blks,
                                                        const uint16_t *ls,
                                                        const uint16_t *bsend,
                                                        const int32_t *bpos) {
  constexpr int NN = Elem<ET>::nnpe, NQ = Elem<ET>::nqp;
  constexpr int BS = ET == ET_TET4 ? BPA_TET4 : BPA_HEX8; // elements per block
  __shared__ double s_m[BS * NN];
  const BBlk B = blks[blockIdx.x];
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if ((int)threadIdx.x < BS && i < A.n) {
    const int64_t e = A.first + i;
    double X[NN][3];
    batch_load_X<ET>(A, e, X);
    double m[NN];
#pragma unroll
    for (int n = 0; n < NN; ++n)
      m[n] = 0.;
    for (int q = 0; q < NQ; ++q) {
      double detJ;
      const double wJ = batch_wj<ET>(A, X, q, detJ);
      if (detJ < 0.)
        atomicMin(A.first_bad, (long long)(i * NQ + q));
#pragma unroll
      for (int n = 0; n < NN; ++n) {
        const double N = (ET == ET_TET4) ? .25 : A.hex_tab[HEX_TAB_N + q * 8 + n];
        m[n] += (A.rho * N) * wJ;
      }
    }
#pragma unroll
    for (int n = 0; n < NN; ++n)
      s_m[ls[e * NN + n]] = m[n];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < B.count; k += blockDim.x) {
    const int s0 = k > 0 ? (int)bsend[B.off + k - 1] : 0, s1 = bsend[B.off + k];
    double m = 0.;
    for (int q = s0; q < s1; ++q)
      m += s_m[q];
    A.me[bpos[B.off + k] - A.ell0] = m;
  }
}

template <int ET> __global__ void __launch_bounds__(256) k_batch_dt(BatchInitArgs A) {
  constexpr int NN = Elem<ET>::nnpe;
  __shared__ double red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dt = 1.7976931348623157e308;
  if (i < A.n) {
    double X[NN][3];
    batch_load_X<ET>(A, A.first + i, X);
    const double h = inradius<ET>(X);
    const double c = sqrt((A.lambda + 2 * A.mu) / A.rho);
    dt = h / c;
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) // positive doubles order like their bit patterns
    atomicMin(A.min_bits, (unsigned long long)__double_as_longlong(red[0]));
}

// M = 0 + sum 1 * m_e over the node's entries in order
__global__ void k_batch_gather_mass(int64_t node0, int64_t n, const int64_t *ell_off,
                                    const uint8_t *deg, const double *me, int64_t ell0,
                                    double *mass) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int64_t nd = node0 + i;
  double m = 0.;
  const double *p = me + (ell_off[nd >> 8] - ell0) + (nd & 255);
  for (int d = 0; d < deg[nd]; ++d)
    m += 1. * p[d * 256];
  mass[nd] = m;
}

// host [n][3] <-> arena [3][ncap] (stride ncap, offset 0) or the node
// records [ncap][6] (stride 1 per component, 6 per node, offset 0 for X and
// 3 for u), at node0
__global__ void k_batch_put3(const double *aos, double *dst, int64_t comp_stride,
                             int64_t node_stride, int64_t node0, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n)
    return;
  const int64_t r = i / 3, c = i % 3;
  dst[c * comp_stride + (node0 + r) * node_stride] = aos[i];
}

__global__ void k_batch_get3(const double *src, double *aos, int64_t comp_stride,
                             int64_t node_stride, int64_t node0, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n)
    return;
  const int64_t r = i / 3, c = i % 3;
  aos[i] = src[c * comp_stride + (node0 + r) * node_stride];
}

template <class T> __global__ void k_batch_fill(T *p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

} // namespace akx

using namespace akx;

// ---------------------------------------------------------------------------
// Host engine
// ---------------------------------------------------------------------------
namespace {

// first-fit free list of [offset, offset + size) ranges
struct FreeList {
  std::map<int64_t, int64_t> free; // offset -> size
  void reset(int64_t c) {
    free.clear();
    if (c > 0)
      free[0] = c;
  }
  int64_t alloc(int64_t n, int64_t align = 1) {
    for (auto it = free.begin(); it != free.end(); ++it) {
      const int64_t off = (it->first + align - 1) / align * align;
      const int64_t end = it->first + it->second;
      if (off + n <= end) {
        const int64_t o0 = it->first;
        free.erase(it);
        if (off > o0)
          free[o0] = off - o0;
        if (off + n < end)
          free[off + n] = end - (off + n);
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    if (n <= 0)
      return;
    auto it = free.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free.erase(nx);
    }
    if (it != free.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free.erase(it);
      }
    }
  }
  int64_t largest() const {
    int64_t m = 0;
    for (auto &f : free)
      m = std::max(m, f.second);
    return m;
  }
};

struct HostInst {
  bool live = false;
  int etype = 0, law = 0;
  double rho = 0, lambda = 0, mu = 0, sy = 0, h = 0;
  double dt = 0, dt_crit = 0;
  int64_t node0 = 0, nn = 0, nn_pad = 0; // nodes [node0, node0 + nn_pad)
  int64_t elem0 = 0, nel = 0;            // in the arena of its element type
  int64_t ell0 = 0, nell = 0;            // force-table entries
  int64_t steps = 0;                     // steps done
  bool urec = true;                      // u in the node records (BInst::urec)
  // block-partial assembly (Tet4, FP32): the instance's block-entry range
  // and the distinct-node count of each of its blocks
  int64_t bent0 = 0, nbent = 0;
  std::vector<int32_t> blk_cnt;
};

constexpr int NGROUP = 8; // (element type, law)
constexpr int MAX_INSTANCES = 4096;

} // namespace

struct akx_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = AKX_PRECISION_FP64;
  bool plastic = false;
  int64_t ncap = 0, tcap = 0, hexcap = 0, ellcap = 0;
  // arenas
  DevBuf<double> xu; // [ncap][6] node records {X, u}
  DevBuf<double> bx, bu; // [3][ncap] X and u (FP32 batches)
  DevBuf<double> v, a, fint, mass;
  DevBuf<uint8_t> blocked, deg;
  DevBuf<float> cu, cv;
  DevBuf<int32_t> conn_t, pos_t, conn_h, pos_h;
  DevBuf<int64_t> ell_off;
  DevBuf<unsigned char> fe;
  DevBuf<unsigned char> hist[2][4]; // grad u, stress, plastic strain, hardening
  DevBuf<double> hex_tab;
  DevBuf<unsigned long long> dt_bits;
  DevBuf<long long> bad;
  DevBuf<double> staging;
  FreeList fl_nodes, fl_tet, fl_hex, fl_ell;
  // FP32 Tet4 block-partial assembly: per element its local corners and LDS
  // slots, per (block, distinct node) the arena node id, the end of its slot
  // range and its table entry
  bool bpa = false;
  int64_t bcap = 0;
  DevBuf<uint16_t> lc_t, ls_t, ls_h; // local corners (Tet4), slots (Tet4, Hex8)
  DevBuf<int32_t> bnode, bpos;
  DevBuf<uint16_t> bsend;
  FreeList fl_bent;
  DevBuf<BBlk> d_blks; // parallel to d_tiles
  int32_t bmaxu_live = 0; // most distinct nodes of a live Tet4 block (staged)
  std::vector<HostInst> inst;
  // device tables
  DevBuf<BInst> d_inst; // MAX_INSTANCES entries, address fixed
  DevBuf<BTile> d_tiles;
  DevBuf<BSlice> d_slices;
  int64_t group_first[NGROUP + 1] = {0};
  int64_t nslices_live = 0;
  bool lists_dirty = true, inst_dirty = true;
  // launch shape
  int block = 256, tile = 2048, blocks_per_cu = 0, cus = 256;
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  // timing
  bool timing = false;
  double t_ms[2] = {0, 0};
  int64_t t_n[2] = {0, 0};

  ~akx_batch() {
    if (graph)
      (void)hipGraphExecDestroy(graph);
    if (stream)
      (void)hipStreamDestroy(stream);
  }
  bool fp32() const { return precision == AKX_PRECISION_FP32; }
  size_t fe_size() const { return fp32() ? 4 : 8; }
  size_t hist_size() const { return fp32() ? 4 : 8; }
  int64_t hist_stride() const { return tcap + 8 * hexcap; }
  // components of history slot k (grad u, stress, plastic strain, hardening):
  // the FP32 form keeps only the plastic strain (Voigt) and the hardening
  int hist_comps(int k) const {
    if (fp32())
      return k == 2 ? 6 : (k == 3 ? 1 : 0);
    return k == 3 ? 1 : 9;
  }
};

namespace {

constexpr int TPB = 256;
unsigned gridn(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

void sync(akx_batch *b) { HIPCHECK(hipStreamSynchronize(b->stream)); }

void drop_graph(akx_batch *b) {
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
}

HostInst &live_inst(akx_batch *b, int id) {
  require(b != nullptr, "null batch");
  require(id >= 0 && id < (int)b->inst.size() && b->inst[id].live, "no such instance");
  return b->inst[id];
}

// instance table: material parameters, dt and the history parity base
void upload_inst(akx_batch *b) {
  std::vector<BInst> ti(std::max<size_t>(b->inst.size(), 1));
  for (size_t id = 0; id < b->inst.size(); ++id) {
    const HostInst &I = b->inst[id];
    ti[id] = BInst{I.lambda, I.mu, I.sy, I.h, I.rho, I.dt, (int32_t)(I.steps & 1), I.law,
                   I.urec ? 1 : 0, 0};
  }
  sync(b);
  HIPCHECK(hipMemcpy(b->d_inst.p, ti.data(), ti.size() * sizeof(BInst),
                     hipMemcpyHostToDevice));
  b->inst_dirty = false;
}

// work lists of the live instances: element tiles per group, node slices
void rebuild_lists(akx_batch *b) {
  std::vector<std::vector<BTile>> groups(NGROUP);
  std::vector<std::vector<BBlk>> gblks(NGROUP);
  std::vector<BSlice> slices;
  b->bmaxu_live = 0;
  for (size_t id = 0; id < b->inst.size(); ++id) {
    const HostInst &I = b->inst[id];
    if (!I.live)
      continue;
    auto &G = groups[I.etype * 4 + I.law];
    auto &GB = gblks[I.etype * 4 + I.law];
    if (!I.blk_cnt.empty()) { // block-partial assembly: one tile per block
      const int64_t bs = I.etype == AKX_TET4 ? BPA_TET4 : BPA_HEX8;
      int64_t off = I.bent0;
      for (size_t k = 0; k < I.blk_cnt.size(); ++k) {
        const int64_t f = (int64_t)k * bs;
        G.push_back({I.elem0 + f, (int32_t)std::min<int64_t>(bs, I.nel - f), (int32_t)id});
        GB.push_back({off, I.blk_cnt[k], 0});
        off += I.blk_cnt[k];
        if (I.etype == AKX_TET4)
          b->bmaxu_live = std::max(b->bmaxu_live, I.blk_cnt[k]);
      }
    } else {
      for (int64_t f = 0; f < I.nel; f += b->tile) {
        G.push_back({I.elem0 + f, (int32_t)std::min<int64_t>(b->tile, I.nel - f), (int32_t)id});
        GB.push_back({0, 0, 0});
      }
    }
    for (int64_t s = 0; s < I.nn; s += 256)
      slices.push_back({I.node0 + s, (int32_t)std::min<int64_t>(256, I.nn - s), (int32_t)id});
  }
  std::vector<BTile> tiles;
  std::vector<BBlk> blks;
  b->group_first[0] = 0;
  for (int g = 0; g < NGROUP; ++g) {
    tiles.insert(tiles.end(), groups[g].begin(), groups[g].end());
    blks.insert(blks.end(), gblks[g].begin(), gblks[g].end());
    b->group_first[g + 1] = (int64_t)tiles.size();
  }
  sync(b);
  drop_graph(b);
  b->d_tiles.reserve(tiles.size());
  b->d_blks.reserve(blks.size());
  b->d_slices.reserve(slices.size());
  if (!tiles.empty()) {
    HIPCHECK(hipMemcpy(b->d_tiles.p, tiles.data(), tiles.size() * sizeof(BTile),
                       hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b->d_blks.p, blks.data(), blks.size() * sizeof(BBlk),
                       hipMemcpyHostToDevice));
  }
  if (!slices.empty())
    HIPCHECK(hipMemcpy(b->d_slices.p, slices.data(), slices.size() * sizeof(BSlice),
                       hipMemcpyHostToDevice));
  b->nslices_live = (int64_t)slices.size();
  b->lists_dirty = false;
}

BatchElemArgs elem_args(akx_batch *b, int etype, int64_t t0, int64_t t1, int parity) {
  BatchElemArgs A{};
  A.tiles = b->d_tiles.p + t0;
  A.ntiles = t1 - t0;
  A.inst = b->d_inst.p;
  A.conn = etype == AKX_TET4 ? b->conn_t.p : b->conn_h.p;
  A.pos = etype == AKX_TET4 ? b->pos_t.p : b->pos_h.p;
  A.xu = b->xu.p;
  A.X = b->bx.p;
  A.u = b->bu.p;
  A.ncap = b->ncap;
  A.fe = b->fe.p;
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < 4; ++k)
      A.hist[s][k] = b->hist[s][k].p;
  A.hcap = b->hist_stride();
  A.hoff = etype == AKX_TET4 ? 0 : b->tcap;
  A.hex_tab = b->hex_tab.p;
  A.parity = parity;
  if (b->bpa && etype == AKX_TET4) {
    A.blks = b->d_blks.p + t0;
    A.bnode = b->bnode.p;
    A.bsend = b->bsend.p;
    A.bpos = b->bpos.p;
    A.lc = b->lc_t.p;
    A.ls = b->ls_t.p;
    A.bmaxu = b->bmaxu_live;
    require(A.blks && A.bnode && A.bsend && A.bpos && A.lc && A.ls && A.bmaxu > 0 &&
                A.bmaxu <= 1024,
            "block-partial pass launched without its block tables");
  } else if (b->bpa) { // Hex8: corner slots only (no staged records)
    A.blks = b->d_blks.p + t0;
    A.bsend = b->bsend.p;
    A.bpos = b->bpos.p;
    A.ls = b->ls_h.p;
    require(A.blks && A.bsend && A.bpos && A.ls,
            "block-partial pass launched without its block tables");
  }
  return A;
}

template <int LAW> void launch_strict(akx_batch *b, int etype, const BatchElemArgs &A) {
  dim3 grid((unsigned)A.ntiles);
  if (etype == AKX_TET4)
    hipLaunchKernelGGL(k_batch_tet4<LAW>, grid, dim3(b->block), 0, b->stream, A);
  else if (b->block == 64)
    hipLaunchKernelGGL((k_batch_hex8<LAW, 8>), grid, dim3(64), 0, b->stream, A);
  else if (b->block == 128)
    hipLaunchKernelGGL((k_batch_hex8<LAW, 16>), grid, dim3(128), 0, b->stream, A);
  else
    hipLaunchKernelGGL((k_batch_hex8<LAW, 32>), grid, dim3(256), 0, b->stream, A);
  HIPCHECK(hipGetLastError());
}

void element_pass(akx_batch *b, int parity) {
  for (int g = 0; g < NGROUP; ++g) {
    const int64_t t0 = b->group_first[g], t1 = b->group_first[g + 1];
    if (t1 == t0)
      continue;
    const int etype = g / 4, law = g % 4;
    BatchElemArgs A = elem_args(b, etype, t0, t1, parity);
    require(A.conn && A.pos && A.xu && A.fe && A.inst && A.tiles && A.hex_tab,
            "batch element arena not allocated");
    if (law == AKX_LAW_PLASTIC_LIH)
      for (int s = 0; s < 2; ++s)
        for (int k = 0; k < 4; ++k)
          require(b->hist_comps(k) == 0 || A.hist[s][k] != nullptr,
                  "plastic history not allocated");
    if (b->fp32()) {
      HIPCHECK(akxf::launch_batch_elements_f32(etype, law, b->block, A, b->stream));
    } else {
      switch (law) {
      case AKX_LAW_ELASTIC: launch_strict<LAW_ELASTIC>(b, etype, A); break;
      case AKX_LAW_ELASTIC_FD: launch_strict<LAW_ELASTIC_FD>(b, etype, A); break;
      case AKX_LAW_NEOHOOKEAN: launch_strict<LAW_NEOHOOKEAN>(b, etype, A); break;
      default: launch_strict<LAW_PLASTIC>(b, etype, A); break;
      }
    }
  }
}

BatchGatherArgs gather_args(akx_batch *b) {
  BatchGatherArgs G{};
  G.slices = b->d_slices.p;
  G.nslices = b->nslices_live;
  G.inst = b->d_inst.p;
  G.ell_off = b->ell_off.p;
  G.deg = b->deg.p;
  G.fe = b->fe.p;
  G.fint = b->fint.p;
  G.xu = b->xu.p;
  G.u = b->bu.p;
  G.v = b->v.p;
  G.a = b->a.p;
  G.mass = b->mass.p;
  G.blocked = b->blocked.p;
  G.cu = b->cu.p;
  G.cv = b->cv.p;
  G.ncap = b->ncap;
  G.nblocks = G.nslices; // one slice per workgroup
  require(G.slices && G.ell_off && G.deg && G.fe && G.fint && G.xu && G.v && G.a &&
              G.mass && G.blocked,
          "batch node arena not allocated");
  require(!b->fp32() || (G.cu && G.cv), "compensation terms not allocated");
  return G;
}

// f_int is stored by the last step of a call only: the steps before it use
// it inside the gather, and nothing reads the array before the call returns.
void gather_pass(akx_batch *b, bool predict) {
  if (b->nslices_live == 0)
    return;
  BatchGatherArgs G = gather_args(b);
  if (predict)
    G.fint = nullptr;
  dim3 grid((unsigned)G.nblocks), block(256);
  if (b->fp32()) {
    if (predict)
      hipLaunchKernelGGL((k_batch_gather<float, true, true>), grid, block, 0, b->stream, G);
    else
      hipLaunchKernelGGL((k_batch_gather<float, false, true>), grid, block, 0, b->stream, G);
  } else {
    if (predict)
      hipLaunchKernelGGL((k_batch_gather<double, true, false>), grid, block, 0, b->stream, G);
    else
      hipLaunchKernelGGL((k_batch_gather<double, false, false>), grid, block, 0, b->stream, G);
  }
  HIPCHECK(hipGetLastError());
}

void predictor_pass(akx_batch *b) {
  if (b->nslices_live == 0)
    return;
  const BatchGatherArgs G = gather_args(b);
  dim3 grid((unsigned)G.nblocks), block(256);
  if (b->fp32())
    hipLaunchKernelGGL((k_batch_predictor<true>), grid, block, 0, b->stream, G);
  else
    hipLaunchKernelGGL((k_batch_predictor<false>), grid, block, 0, b->stream, G);
  HIPCHECK(hipGetLastError());
}

template <class F> void timed(akx_batch *b, int cls, F &&f) {
  if (!b->timing) {
    f();
    return;
  }
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  HIPCHECK(hipEventRecord(e0, b->stream));
  f();
  HIPCHECK(hipEventRecord(e1, b->stream));
  HIPCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
  b->t_ms[cls] += ms;
  b->t_n[cls] += 1;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// Two steps (step parities 0 and 1, each element pass + gather with the next
// predictor fused) captured once; the graph holds kernels only.
void ensure_graph(akx_batch *b) {
  if (b->graph)
    return;
  hipGraph_t g = nullptr;
  HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int p = 0; p < 2; ++p) {
      element_pass(b, p);
      gather_pass(b, true);
    }
  } catch (...) {
    hipGraph_t tmp = nullptr;
    (void)hipStreamEndCapture(b->stream, &tmp);
    if (tmp)
      (void)hipGraphDestroy(tmp);
    throw;
  }
  HIPCHECK(hipStreamEndCapture(b->stream, &g));
  hipGraphExec_t ge = nullptr;
  const hipError_t st = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHECK(st);
  b->graph = ge;
}

// n steps of every live instance: predictor, then per step element pass +
// gather (the gather of step k carries the predictor of step k + 1), the
// last gather without it.  Step k reads history buffer (steps_i + k) & 1.
void do_steps(akx_batch *b, int64_t n) {
  if (n <= 0)
    return;
  if (b->lists_dirty)
    rebuild_lists(b);
  if (b->inst_dirty)
    upload_inst(b);
  if (b->nslices_live == 0)
    return;
  predictor_pass(b);
  int64_t k = 0;
  if (b->use_graph && !b->timing && n > 2) {
    ensure_graph(b);
    const int64_t pairs = (n - 1) / 2;
    for (int64_t p = 0; p < pairs; ++p)
      HIPCHECK(hipGraphLaunch(b->graph, b->stream));
    k = 2 * pairs;
  }
  for (; k < n; ++k) {
    timed(b, 0, [&] { element_pass(b, (int)(k & 1)); });
    timed(b, 1, [&] { gather_pass(b, k + 1 < n); });
  }
  for (auto &I : b->inst)
    if (I.live)
      I.steps += n;
  if (n & 1)
    b->inst_dirty = true; // parity bases moved
}

// device bytes an instance takes in a batch (node range padded to 256,
// force-table entries of FE_W values estimated with a 25 % slice padding)
int64_t footprint_bytes(int precision, int etype, int law, int64_t nn, int64_t nel) {
  const int nnpe = etype == AKX_TET4 ? 4 : 8, nqp = etype == AKX_TET4 ? 1 : 8;
  const bool f32 = precision == AKX_PRECISION_FP32;
  const double sf = f32 ? 4 : 8;
  const int64_t nn_pad = (nn + 255) / 256 * 256;
  // node records, v, a, f_int, mass, blocked bits, degree; FP32: the X and
  // u arrays and the two compensation terms
  double bytes = (double)nn_pad * (5 * 3 * 8 + 8 + 1 + 1 + (f32 ? 2 * 3 * 8 + 2 * 3 * 4 : 0)) +
                 (double)nn_pad / 256 * 8;
  bytes += (double)nel * nnpe * (4 + 4 + 1.25 * FE_W * sf);
  if (f32 && etype == AKX_TET4) // block partials: corners and slots, and at most
                                // two distinct-node entries per element
    bytes += (double)nel * (4 * 2 * 2 + 2 * (4 + 4 + 2));
  else if (f32) // Hex8: slots, at most five distinct-node entries per element
    bytes += (double)nel * (8 * 2 + 5 * (4 + 4 + 2));
  if (law == AKX_LAW_PLASTIC_LIH) // two buffers of the history
    bytes += (double)nel * nqp * 2 * (f32 ? 7 * 4 : 28 * 8);
  return (int64_t)bytes;
}

void release_ranges(akx_batch *b, const HostInst &I) {
  b->fl_nodes.release(I.node0, I.nn_pad);
  (I.etype == AKX_TET4 ? b->fl_tet : b->fl_hex).release(I.elem0, I.nel);
  b->fl_ell.release(I.ell0, I.nell);
  if (I.nbent > 0)
    b->fl_bent.release(I.bent0, I.nbent);
}

// Hex8 device table: dN/dxi [8][3][8], N [8][8], w [8] (the numbers the
// single-model path uploads)
std::vector<double> hex8_table() {
  Hex8Tables t;
  fill_hex8_tables(t);
  std::vector<double> tab(HEX_TAB_SIZE);
  std::memcpy(tab.data(), &t.dnds[0][0][0], 192 * sizeof(double));
  std::memcpy(tab.data() + HEX_TAB_N, &t.N[0][0], 64 * sizeof(double));
  std::memcpy(tab.data() + HEX_TAB_W, &t.w[0], 8 * sizeof(double));
  return tab;
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int akx_batch_create(int device, int precision, int64_t node_capacity,
                     int64_t tet_capacity, int64_t hex_capacity, int with_plasticity,
                     akx_batch **out) {
  return guard([&] {
    require(out != nullptr, "null output");
    require(precision == AKX_PRECISION_FP64 || precision == AKX_PRECISION_FP32,
            "a batch runs the FP64 plain form or the FP32 form");
    require(node_capacity > 0 && tet_capacity >= 0 && hex_capacity >= 0 &&
                tet_capacity + hex_capacity > 0,
            "bad capacities");
    auto b = std::make_unique<akx_batch>();
    b->device = device;
    HIPCHECK(hipSetDevice(device));
    hipDeviceProp_t p;
    HIPCHECK(hipGetDeviceProperties(&p, device));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
      throw Error(AKX_ERR_UNSUPPORTED, std::string("device is not gfx950: ") + p.gcnArchName);
    b->cus = p.multiProcessorCount;
    HIPCHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->precision = precision;
    b->plastic = with_plasticity != 0;
    b->ncap = (node_capacity + 255) / 256 * 256;
    b->tcap = tet_capacity;
    b->hexcap = hex_capacity;
    // force-table entries: one per element corner plus slice padding
    b->ellcap = ((4 * b->tcap + 8 * b->hexcap) * 5 / 4 / 256 + 64) * 256;
    require(b->ncap < INT32_MAX && b->ellcap < INT32_MAX && 8 * b->hexcap < INT32_MAX &&
                4 * b->tcap < INT32_MAX,
            "capacities exceed the int32 indices of the arenas");
    b->xu.alloc(6 * b->ncap);
    for (auto *f : {&b->v, &b->a, &b->fint})
      f->alloc(3 * b->ncap);
    b->mass.alloc(b->ncap);
    b->blocked.alloc(b->ncap);
    b->deg.alloc(b->ncap);
    b->ell_off.alloc(b->ncap / 256);
    if (b->fp32()) {
      b->cu.alloc(3 * b->ncap);
      b->cv.alloc(3 * b->ncap);
    }
    if (b->tcap) {
      b->conn_t.alloc(4 * b->tcap);
      b->pos_t.alloc(4 * b->tcap);
    }
    // FP32 instances assemble by block partials; a block of BPA_TET4 / BPA_HEX8 box
    // elements has about 0.7 (Tet4) and at most 4.03 (Hex8: part of one row
    // of cells) distinct nodes per element, so twice the Tet4 and five times
    // the Hex8 capacity hold the distinct-node entries of any box mesh
    b->bpa = b->fp32() && (b->tcap > 0 || b->hexcap > 0);
    if (b->bpa) {
      b->bcap = 2 * b->tcap + 5 * b->hexcap + 2048;
      if (b->tcap) {
        b->lc_t.alloc(4 * b->tcap);
        b->ls_t.alloc(4 * b->tcap);
      }
      b->bx.alloc(3 * b->ncap);
      b->bu.alloc(3 * b->ncap);
      if (b->hexcap)
        b->ls_h.alloc(8 * b->hexcap);
      b->bnode.alloc(b->bcap);
      b->bpos.alloc(b->bcap);
      b->bsend.alloc(b->bcap);
      b->fl_bent.reset(b->bcap);
    }
    if (b->hexcap) {
      b->conn_h.alloc(8 * b->hexcap);
      b->pos_h.alloc(8 * b->hexcap);
    }
    b->fe.alloc(b->ellcap * FE_W * b->fe_size());
    if (b->plastic)
      for (int s = 0; s < 2; ++s)
        for (int k = 0; k < 4; ++k)
          if (b->hist_comps(k) > 0)
            b->hist[s][k].alloc(b->hist_comps(k) * b->hist_stride() * b->hist_size());
    const std::vector<double> tab = hex8_table();
    b->hex_tab.alloc(tab.size());
    HIPCHECK(hipMemcpy(b->hex_tab.p, tab.data(), tab.size() * sizeof(double),
                       hipMemcpyHostToDevice));
    b->dt_bits.alloc(1);
    b->bad.alloc(1);
    b->d_inst.alloc(MAX_INSTANCES);
    b->fl_nodes.reset(b->ncap);
    b->fl_tet.reset(b->tcap);
    b->fl_hex.reset(b->hexcap);
    b->fl_ell.reset(b->ellcap);
    *out = b.release();
  });
}

int akx_batch_destroy(akx_batch *b) {
  return guard([&] {
    if (!b)
      return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    delete b;
  });
}

int akx_batch_footprint(int precision, int element_type, int law, int64_t nb_nodes,
                        int64_t nb_elements, int64_t *bytes) {
  return guard([&] {
    require(bytes != nullptr, "null output");
    require(element_type == AKX_TET4 || element_type == AKX_HEX8, "bad element type");
    require(nb_nodes >= 0 && nb_elements >= 0, "negative size");
    *bytes = footprint_bytes(precision, element_type, law, nb_nodes, nb_elements);
  });
}

int akx_batch_add_instance(akx_batch *b, int element_type, int law, const double *params,
                           int64_t nb_nodes, const double *nodes, int64_t nb_elements,
                           const int32_t *conn, const uint8_t *blocked_dofs,
                           const double *velocity, double dt_factor, int *instance_id) {
  return guard([&] {
    require(b && params && nodes && conn && instance_id, "null argument");
    require(element_type == AKX_TET4 || element_type == AKX_HEX8, "bad element type");
    require(law >= AKX_LAW_ELASTIC && law <= AKX_LAW_PLASTIC_LIH, "bad law");
    require(law != AKX_LAW_PLASTIC_LIH || b->plastic,
            "the batch was created without plastic history");
    require(nb_nodes > 0 && nb_elements > 0, "empty instance");
    require(dt_factor > 0 && std::isfinite(dt_factor), "dt_factor must be positive");
    require(params[0] > 0 && params[1] > 0 && params[2] > -1 && params[2] < .5,
            "need rho > 0, E > 0, -1 < nu < 0.5");
    for (int64_t i = 0; i < 3 * nb_nodes; ++i)
      require(std::isfinite(nodes[i]), "non-finite node coordinate");
    if (velocity)
      for (int64_t i = 0; i < 3 * nb_nodes; ++i)
        require(std::isfinite(velocity[i]), "non-finite initial velocity");
    HIPCHECK(hipSetDevice(b->device));
    const int nnpe = element_type == AKX_TET4 ? 4 : 8;
    const int nqp = element_type == AKX_TET4 ? 1 : 8;
    HostInst I;
    I.etype = element_type;
    I.law = law;
    I.urec = !b->fp32();
    require(I.urec || (b->bx.p && b->bu.p), "FP32 instance without its X and u arrays");
    I.rho = params[0];
    lame(params[1], params[2], I.lambda, I.mu);
    I.sy = params[3];
    I.h = params[4];
    I.nn = nb_nodes;
    I.nn_pad = (nb_nodes + 255) / 256 * 256;
    I.nel = nb_elements;
    for (int64_t i = 0; i < nb_elements * nnpe; ++i)
      require(conn[i] >= 0 && conn[i] < nb_nodes, "connectivity refers to a missing node");
    // block-partial assembly (FP32): per block of bs elements its distinct
    // nodes in order of first use, each corner's index among them and its
    // LDS slot (the corners of one distinct node take consecutive slots in
    // ascending element order); the table then holds one entry per (block,
    // distinct node) and a node's degree is the number of blocks touching it
    const bool bpa = b->bpa;
    const int64_t bs = element_type == AKX_TET4 ? BPA_TET4 : BPA_HEX8; // per block
    const int64_t nblk = bpa ? (nb_elements + bs - 1) / bs : 0;
    std::vector<int32_t> bl_nodes; // instance-local node ids, blocks in order
    std::vector<uint16_t> lc, ls, send;
    if (bpa) {
      lc.resize(nb_elements * nnpe);
      ls.resize(nb_elements * nnpe);
      bl_nodes.reserve(nb_elements);
      I.blk_cnt.resize(nblk);
      std::vector<int64_t> stamp(nb_nodes, -1);
      std::vector<int32_t> local(nb_nodes, 0);
      std::vector<uint16_t> cnt(bs * nnpe + 1);
      for (int64_t bk = 0; bk < nblk; ++bk) {
        const int64_t e0 = bk * bs, e1 = std::min<int64_t>(nb_elements, e0 + bs);
        const int64_t start = (int64_t)bl_nodes.size();
        for (int64_t i = e0 * nnpe; i < e1 * nnpe; ++i) {
          const int32_t nd = conn[i];
          if (stamp[nd] != bk) {
            stamp[nd] = bk;
            local[nd] = (int32_t)((int64_t)bl_nodes.size() - start);
            bl_nodes.push_back(nd);
          }
          lc[i] = (uint16_t)local[nd];
        }
        const int32_t nloc = (int32_t)((int64_t)bl_nodes.size() - start);
        I.blk_cnt[bk] = nloc;
        std::fill(cnt.begin(), cnt.begin() + nloc + 1, 0);
        for (int64_t i = e0 * nnpe; i < e1 * nnpe; ++i)
          cnt[lc[i] + 1]++;
        for (int32_t k = 0; k < nloc; ++k) { // slot ends (inclusive prefix)
          cnt[k + 1] = (uint16_t)(cnt[k + 1] + cnt[k]);
          send.push_back(cnt[k + 1]);
        }
        for (int64_t i = e0 * nnpe; i < e1 * nnpe; ++i) // cnt[k]: start of k's slots
          ls[i] = cnt[lc[i]]++;
      }
      I.nbent = (int64_t)bl_nodes.size();
    }
    // node degrees and the sliced-ELL layout of this instance
    std::vector<int32_t> deg(I.nn_pad, 0);
    if (bpa) {
      for (int32_t nd : bl_nodes)
        deg[nd]++;
    } else {
      for (int64_t i = 0; i < nb_elements * nnpe; ++i)
        deg[conn[i]]++;
    }
    const int64_t nsl = I.nn_pad / 256;
    std::vector<int64_t> off(nsl + 1, 0);
    for (int64_t s = 0; s < nsl; ++s) {
      int w = 0;
      for (int64_t k = 0; k < 256; ++k)
        w = std::max(w, deg[s * 256 + k]);
      require(w <= 255, "a node belongs to more than 255 elements");
      off[s + 1] = off[s] + (int64_t)w * 256;
    }
    I.nell = off[nsl];
    // ranges in the arenas
    if (b->inst.size() >= (size_t)MAX_INSTANCES) {
      bool slot = false;
      for (auto &J : b->inst)
        slot = slot || (!J.live);
      if (!slot)
        throw Error(AKX_ERR_MEMORY, "instance table full");
    }
    FreeList &fl_el = element_type == AKX_TET4 ? b->fl_tet : b->fl_hex;
    I.node0 = b->fl_nodes.alloc(I.nn_pad, 256);
    if (I.node0 < 0)
      throw Error(AKX_ERR_MEMORY, "node arena full");
    I.elem0 = fl_el.alloc(I.nel);
    if (I.elem0 < 0) {
      b->fl_nodes.release(I.node0, I.nn_pad);
      throw Error(AKX_ERR_MEMORY, "element arena full");
    }
    I.ell0 = b->fl_ell.alloc(I.nell, 256); // nell is a multiple of 256
    if (I.ell0 < 0) {
      b->fl_nodes.release(I.node0, I.nn_pad);
      fl_el.release(I.elem0, I.nel);
      throw Error(AKX_ERR_MEMORY, "force table full");
    }
    if (bpa) {
      I.bent0 = b->fl_bent.alloc(I.nbent);
      if (I.bent0 < 0) {
        b->fl_nodes.release(I.node0, I.nn_pad);
        fl_el.release(I.elem0, I.nel);
        b->fl_ell.release(I.ell0, I.nell);
        throw Error(AKX_ERR_MEMORY, "block arena full");
      }
    }
    try {
      // arena connectivity, table entries, slice offsets, degrees, blocked bits
      std::vector<int32_t> gconn(nb_elements * nnpe), gpos(bpa ? 0 : nb_elements * nnpe);
      std::vector<int32_t> fill(I.nn_pad, 0);
      for (int64_t e = 0; e < nb_elements; ++e)
        for (int a = 0; a < nnpe; ++a) {
          const int32_t nd = conn[e * nnpe + a];
          gconn[e * nnpe + a] = (int32_t)(I.node0 + nd);
          if (bpa)
            continue;
          const int32_t d = fill[nd]++; // elements in ascending order
          gpos[e * nnpe + a] =
              (int32_t)(I.ell0 + off[nd >> 8] + (int64_t)d * 256 + (nd & 255));
        }
      // block partials: entry of (block, distinct node) in ascending block order
      std::vector<int32_t> gbnode(I.nbent), gbpos(I.nbent);
      for (int64_t k = 0; k < I.nbent; ++k) {
        const int32_t nd = bl_nodes[k];
        gbnode[k] = (int32_t)(I.node0 + nd);
        const int32_t d = fill[nd]++;
        gbpos[k] = (int32_t)(I.ell0 + off[nd >> 8] + (int64_t)d * 256 + (nd & 255));
      }
      std::vector<int64_t> goff(nsl);
      for (int64_t s = 0; s < nsl; ++s)
        goff[s] = I.ell0 + off[s];
      std::vector<uint8_t> deg8(I.nn_pad), bits(I.nn_pad, 7);
      for (int64_t k = 0; k < I.nn_pad; ++k)
        deg8[k] = (uint8_t)deg[k];
      for (int64_t k = 0; k < nb_nodes; ++k)
        bits[k] = blocked_dofs ? (uint8_t)((blocked_dofs[3 * k] ? 1 : 0) |
                                           (blocked_dofs[3 * k + 1] ? 2 : 0) |
                                           (blocked_dofs[3 * k + 2] ? 4 : 0))
                               : 0;
      sync(b);
      int32_t *dconn = element_type == AKX_TET4 ? b->conn_t.p : b->conn_h.p;
      int32_t *dpos = element_type == AKX_TET4 ? b->pos_t.p : b->pos_h.p;
      HIPCHECK(hipMemcpy(dconn + I.elem0 * nnpe, gconn.data(), gconn.size() * 4,
                         hipMemcpyHostToDevice));
      if (!bpa)
        HIPCHECK(hipMemcpy(dpos + I.elem0 * nnpe, gpos.data(), gpos.size() * 4,
                           hipMemcpyHostToDevice));
      if (bpa && element_type == AKX_TET4) {
        HIPCHECK(hipMemcpy(b->lc_t.p + I.elem0 * 4, lc.data(), lc.size() * 2,
                           hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(b->ls_t.p + I.elem0 * 4, ls.data(), ls.size() * 2,
                           hipMemcpyHostToDevice));
      } else if (bpa) {
        HIPCHECK(hipMemcpy(b->ls_h.p + I.elem0 * 8, ls.data(), ls.size() * 2,
                           hipMemcpyHostToDevice));
      }
      if (bpa) {
        HIPCHECK(hipMemcpy(b->bnode.p + I.bent0, gbnode.data(), gbnode.size() * 4,
                           hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(b->bpos.p + I.bent0, gbpos.data(), gbpos.size() * 4,
                           hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(b->bsend.p + I.bent0, send.data(), send.size() * 2,
                           hipMemcpyHostToDevice));
      }
      HIPCHECK(hipMemcpy(b->ell_off.p + I.node0 / 256, goff.data(), nsl * 8,
                         hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(b->deg.p + I.node0, deg8.data(), I.nn_pad, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(b->blocked.p + I.node0, bits.data(), I.nn_pad,
                         hipMemcpyHostToDevice));
      // nodal state: X and v from the host, u = a = f_int = 0 (the records
      // of the padding nodes are zeroed too)
      hipLaunchKernelGGL((k_batch_fill<double>), dim3(gridn(6 * I.nn_pad)), dim3(TPB), 0,
                         b->stream, b->xu.p + 6 * I.node0, 6 * I.nn_pad, 0.);
      b->staging.reserve(3 * nb_nodes);
      HIPCHECK(hipStreamSynchronize(b->stream));
      HIPCHECK(hipMemcpy(b->staging.p, nodes, 3 * nb_nodes * sizeof(double),
                         hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_batch_put3, dim3(gridn(3 * nb_nodes)), dim3(TPB), 0, b->stream,
                         (const double *)b->staging.p, b->xu.p, (int64_t)1, (int64_t)6,
                         I.node0, nb_nodes);
      HIPCHECK(hipGetLastError());
      if (!I.urec) { // X and u = 0 in the [3][ncap] arrays
        for (int c = 0; c < 3; ++c)
          hipLaunchKernelGGL((k_batch_fill<double>), dim3(gridn(I.nn_pad)), dim3(TPB), 0,
                             b->stream, b->bx.p + c * b->ncap + I.node0, I.nn_pad, 0.);
        hipLaunchKernelGGL(k_batch_put3, dim3(gridn(3 * nb_nodes)), dim3(TPB), 0, b->stream,
                           (const double *)b->staging.p, b->bx.p, b->ncap, (int64_t)1,
                           I.node0, nb_nodes);
        HIPCHECK(hipGetLastError());
      }
      for (auto *f : {&b->v, &b->a, &b->fint, &b->bu})
        for (int c = 0; c < 3 && f->p; ++c)
          hipLaunchKernelGGL((k_batch_fill<double>), dim3(gridn(I.nn_pad)), dim3(TPB), 0,
                             b->stream, f->p + c * b->ncap + I.node0, I.nn_pad, 0.);
      if (b->fp32())
        for (auto *f : {&b->cu, &b->cv})
          for (int c = 0; c < 3; ++c)
            hipLaunchKernelGGL((k_batch_fill<float>), dim3(gridn(I.nn_pad)), dim3(TPB), 0,
                               b->stream, f->p + c * b->ncap + I.node0, I.nn_pad, 0.f);
      HIPCHECK(hipGetLastError());
      sync(b);
      if (velocity) {
        HIPCHECK(hipMemcpy(b->staging.p, velocity, 3 * nb_nodes * sizeof(double),
                           hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_batch_put3, dim3(gridn(3 * nb_nodes)), dim3(TPB), 0,
                           b->stream, (const double *)b->staging.p, b->v.p, b->ncap,
                           (int64_t)1, I.node0, nb_nodes);
        HIPCHECK(hipGetLastError());
      }
      if (law == AKX_LAW_PLASTIC_LIH) {
        const int64_t stride = b->hist_stride();
        const int64_t q0 = (element_type == AKX_TET4 ? 0 : b->tcap) + I.elem0 * nqp;
        const int64_t nq = I.nel * nqp;
        for (int s = 0; s < 2; ++s)
          for (int k = 0; k < 4; ++k)
            for (int c = 0; c < b->hist_comps(k); ++c) {
              if (b->fp32())
                hipLaunchKernelGGL((k_batch_fill<float>), dim3(gridn(nq)), dim3(TPB), 0,
                                   b->stream, (float *)b->hist[s][k].p + c * stride + q0,
                                   nq, 0.f);
              else
                hipLaunchKernelGGL((k_batch_fill<double>), dim3(gridn(nq)), dim3(TPB), 0,
                                   b->stream, (double *)b->hist[s][k].p + c * stride + q0,
                                   nq, 0.);
            }
        HIPCHECK(hipGetLastError());
      }
      // jacobians, lumped mass, stable time step; the mass contributions go
      // to the instance's own (not yet used) force-table range
      BatchInitArgs A{};
      A.first = I.elem0;
      A.n = I.nel;
      A.conn = dconn;
      A.pos = dpos;
      A.xu = b->xu.p;
      A.ncap = b->ncap;
      A.rho = I.rho;
      A.lambda = I.lambda;
      A.mu = I.mu;
      A.hex_tab = b->hex_tab.p;
      // 8 bytes per entry fit in the FE_W * 4 or FE_W * 8 bytes of the
      // entry's force; ell0 is a multiple of 256, so the address is aligned
      A.me = reinterpret_cast<double *>(b->fe.p + I.ell0 * FE_W * b->fe_size());
      A.ell0 = I.ell0;
      const long long big = INT64_MAX;
      const unsigned long long dmax = 0x7FEFFFFFFFFFFFFFull;
      HIPCHECK(hipMemcpy(b->bad.p, &big, 8, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(b->dt_bits.p, &dmax, 8, hipMemcpyHostToDevice));
      A.first_bad = b->bad.p;
      A.min_bits = b->dt_bits.p;
      DevBuf<BBlk> iblk; // this instance's blocks, for the mass kernel
      if (bpa) {
        std::vector<BBlk> hb(nblk);
        int64_t o = I.bent0;
        for (int64_t k = 0; k < nblk; ++k) {
          hb[k] = BBlk{o, I.blk_cnt[k], 0};
          o += I.blk_cnt[k];
        }
        iblk.alloc(nblk);
        HIPCHECK(hipMemcpy(iblk.p, hb.data(), hb.size() * sizeof(BBlk), hipMemcpyHostToDevice));
        if (element_type == AKX_TET4) {
          hipLaunchKernelGGL(k_batch_mass_bpa<ET_TET4>, dim3((unsigned)nblk), dim3(256), 0,
                             b->stream, A, (const BBlk *)iblk.p, (const uint16_t *)b->ls_t.p,
                             (const uint16_t *)b->bsend.p, (const int32_t *)b->bpos.p);
          hipLaunchKernelGGL(k_batch_dt<ET_TET4>, dim3(gridn(I.nel)), dim3(TPB), 0,
                             b->stream, A);
        } else {
          hipLaunchKernelGGL(k_batch_mass_bpa<ET_HEX8>, dim3((unsigned)nblk), dim3(256), 0,
                             b->stream, A, (const BBlk *)iblk.p, (const uint16_t *)b->ls_h.p,
                             (const uint16_t *)b->bsend.p, (const int32_t *)b->bpos.p);
          hipLaunchKernelGGL(k_batch_dt<ET_HEX8>, dim3(gridn(I.nel)), dim3(TPB), 0,
                             b->stream, A);
        }
      } else if (element_type == AKX_TET4) {
        hipLaunchKernelGGL(k_batch_mass<ET_TET4>, dim3(gridn(I.nel)), dim3(TPB), 0,
                           b->stream, A);
        hipLaunchKernelGGL(k_batch_dt<ET_TET4>, dim3(gridn(I.nel)), dim3(TPB), 0,
                           b->stream, A);
      } else {
        hipLaunchKernelGGL(k_batch_mass<ET_HEX8>, dim3(gridn(I.nel)), dim3(TPB), 0,
                           b->stream, A);
        hipLaunchKernelGGL(k_batch_dt<ET_HEX8>, dim3(gridn(I.nel)), dim3(TPB), 0,
                           b->stream, A);
      }
      HIPCHECK(hipGetLastError());
      hipLaunchKernelGGL(k_batch_gather_mass, dim3(gridn(I.nn)), dim3(TPB), 0, b->stream,
                         I.node0, I.nn, (const int64_t *)b->ell_off.p,
                         (const uint8_t *)b->deg.p, (const double *)A.me, I.ell0,
                         b->mass.p);
      HIPCHECK(hipGetLastError());
      long long bad = 0;
      unsigned long long bits_dt = 0;
      HIPCHECK(hipMemcpyAsync(&bad, b->bad.p, 8, hipMemcpyDeviceToHost, b->stream));
      HIPCHECK(hipMemcpyAsync(&bits_dt, b->dt_bits.p, 8, hipMemcpyDeviceToHost, b->stream));
      sync(b);
      if (bad != INT64_MAX)
        throw Error(AKX_ERR_JACOBIAN, "Negative jacobian computed at Gauss point " +
                                          std::to_string(bad) + " of the instance");
      std::memcpy(&I.dt_crit, &bits_dt, 8);
      I.dt = dt_factor * I.dt_crit;
    } catch (...) {
      release_ranges(b, I);
      throw;
    }
    I.live = true;
    I.steps = 0;
    int id = -1;
    for (size_t k = 0; k < b->inst.size(); ++k)
      if (!b->inst[k].live) {
        id = (int)k;
        break;
      }
    if (id < 0) {
      id = (int)b->inst.size();
      b->inst.push_back(I);
    } else {
      b->inst[id] = I;
    }
    b->lists_dirty = b->inst_dirty = true;
    *instance_id = id;
  });
}

int akx_batch_remove_instance(akx_batch *b, int id) {
  return guard([&] {
    HostInst &I = live_inst(b, id);
    HIPCHECK(hipSetDevice(b->device));
    sync(b);
    release_ranges(b, I);
    I = HostInst();
    b->lists_dirty = b->inst_dirty = true;
  });
}

int akx_batch_get_time_step(akx_batch *b, int id, double *dt, double *dt_crit) {
  return guard([&] {
    const HostInst &I = live_inst(b, id);
    if (dt)
      *dt = I.dt;
    if (dt_crit)
      *dt_crit = I.dt_crit;
  });
}

int akx_batch_set_time_step(akx_batch *b, int id, double dt) {
  return guard([&] {
    HostInst &I = live_inst(b, id);
    require(dt > 0 && std::isfinite(dt), "time step must be positive");
    I.dt = dt;
    b->inst_dirty = true;
  });
}

int akx_batch_get_steps(akx_batch *b, int id, int64_t *steps) {
  return guard([&] {
    const HostInst &I = live_inst(b, id);
    require(steps != nullptr, "null output");
    *steps = I.steps;
  });
}

int akx_batch_solve_step(akx_batch *b, int64_t nb_steps) {
  return guard([&] {
    require(b != nullptr, "null batch");
    require(nb_steps >= 0, "negative step count");
    HIPCHECK(hipSetDevice(b->device));
    do_steps(b, nb_steps);
  });
}

int akx_batch_synchronize(akx_batch *b) {
  return guard([&] {
    require(b != nullptr, "null batch");
    HIPCHECK(hipSetDevice(b->device));
    sync(b);
  });
}

int akx_batch_get_field(akx_batch *b, int id, int field, double *host) {
  return guard([&] {
    const HostInst &I = live_inst(b, id);
    require(host != nullptr, "null output");
    HIPCHECK(hipSetDevice(b->device));
    const double *src = nullptr;
    int64_t comp_stride = b->ncap, node_stride = 1;
    switch (field) {
    case AKX_FIELD_DISPLACEMENT:
      if (I.urec) {
        src = b->xu.p + 3;
        comp_stride = 1;
        node_stride = 6;
      } else {
        src = b->bu.p;
      }
      break;
    case AKX_FIELD_VELOCITY: src = b->v.p; break;
    case AKX_FIELD_ACCELERATION: src = b->a.p; break;
    case AKX_FIELD_INTERNAL_FORCE: src = b->fint.p; break;
    case AKX_FIELD_NODES:
      src = b->xu.p;
      comp_stride = 1;
      node_stride = 6;
      break;
    default: throw Error(AKX_ERR_INVALID, "field not available for batch instances");
    }
    sync(b);
    b->staging.reserve(3 * I.nn);
    hipLaunchKernelGGL(k_batch_get3, dim3(gridn(3 * I.nn)), dim3(TPB), 0, b->stream, src,
                       b->staging.p, comp_stride, node_stride, I.node0, I.nn);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(host, b->staging.p, 3 * I.nn * sizeof(double),
                            hipMemcpyDeviceToHost, b->stream));
    sync(b);
  });
}

int akx_batch_set_launch(akx_batch *b, int block, int tile, int blocks_per_cu,
                         int use_graph) {
  return guard([&] {
    require(b != nullptr, "null batch");
    require(block == 64 || block == 128 || block == 256, "block must be 64, 128 or 256");
    require(tile >= 64 && tile <= (1 << 20), "tile must be in [64, 2^20]");
    require(blocks_per_cu >= 0 && blocks_per_cu <= 32, "blocks_per_cu must be in [0, 32]");
    HIPCHECK(hipSetDevice(b->device));
    sync(b);
    drop_graph(b);
    b->block = block;
    b->tile = tile;
    b->blocks_per_cu = blocks_per_cu;
    b->use_graph = use_graph != 0;
    b->lists_dirty = true;
  });
}

int akx_batch_get_launch(akx_batch *b, int *block, int *tile, int *blocks_per_cu,
                         int *use_graph) {
  return guard([&] {
    require(b != nullptr, "null batch");
    if (block)
      *block = b->block;
    if (tile)
      *tile = b->tile;
    if (blocks_per_cu)
      *blocks_per_cu = b->blocks_per_cu;
    if (use_graph)
      *use_graph = b->use_graph;
  });
}

// Autotuner: times the element pass of the live instances for every
// (threads per block, elements per tile, persistent blocks per CU) candidate
// and keeps the fastest.  The trials leave the state untouched: the element
// pass writes only the force table and the history buffer that the next
// real step writes again from the same inputs.
int akx_batch_autotune(akx_batch *b, int repeats, int *block, int *tile,
                       int *blocks_per_cu, double *best_ms) {
  return guard([&] {
    require(b != nullptr, "null batch");
    require(repeats >= 1 && repeats <= 100, "repeats must be in [1, 100]");
    HIPCHECK(hipSetDevice(b->device));
    if (b->inst_dirty)
      upload_inst(b);
    const int blocks[3] = {64, 128, 256};
    const int tiles[5] = {256, 512, 1024, 2048, 4096};
    const int bpcs[1] = {0}; // one tile per workgroup: no persistent blocks
    double best = 1e300;
    int bb = b->block, bt = b->tile, bc = b->blocks_per_cu;
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    try {
      for (int blk : blocks)
        for (int tl : tiles)
          for (int bpc : bpcs) {
            b->block = blk;
            b->tile = tl;
            b->blocks_per_cu = bpc;
            rebuild_lists(b);
            element_pass(b, 0); // warm-up
            HIPCHECK(hipEventRecord(e0, b->stream));
            for (int r = 0; r < repeats; ++r)
              element_pass(b, 0);
            HIPCHECK(hipEventRecord(e1, b->stream));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms / repeats < best) {
              best = ms / repeats;
              bb = blk;
              bt = tl;
              bc = bpc;
            }
          }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->block = bb;
    b->tile = bt;
    b->blocks_per_cu = bc;
    rebuild_lists(b);
    if (block)
      *block = bb;
    if (tile)
      *tile = bt;
    if (blocks_per_cu)
      *blocks_per_cu = bc;
    if (best_ms)
      *best_ms = best;
  });
}

int akx_batch_get_stats(akx_batch *b, int64_t *live, int64_t *elements,
                        int64_t *quadrature_points, int64_t *nodes) {
  return guard([&] {
    require(b != nullptr, "null batch");
    int64_t nl = 0, ne = 0, nq = 0, nn = 0;
    for (auto &I : b->inst)
      if (I.live) {
        ++nl;
        ne += I.nel;
        nq += I.nel * (I.etype == AKX_TET4 ? 1 : 8);
        nn += I.nn;
      }
    if (live)
      *live = nl;
    if (elements)
      *elements = ne;
    if (quadrature_points)
      *quadrature_points = nq;
    if (nodes)
      *nodes = nn;
  });
}

int akx_batch_capacity_left(akx_batch *b, int64_t *nodes, int64_t *tets, int64_t *hexes,
                            int64_t *entries) {
  return guard([&] {
    require(b != nullptr, "null batch");
    if (nodes)
      *nodes = b->fl_nodes.largest();
    if (tets)
      *tets = b->fl_tet.largest();
    if (hexes)
      *hexes = b->fl_hex.largest();
    if (entries)
      *entries = b->fl_ell.largest();
  });
}

// Unique bytes one launch of each kernel class moves over the live instances:
// element pass = connectivity + table positions + the elemental forces
// written + X and u of every node once (+ plastic history read and written);
// gather = the forces read (one entry per element corner, or per block
// partial) + per node degree,
// blocked bits, mass, f_int written and u, v, a (+ compensation) read and
// written.
int akx_batch_get_kernel_bytes(akx_batch *b, int kernel_class, double *bytes) {
  return guard([&] {
    require(b != nullptr && bytes != nullptr && kernel_class >= 0 && kernel_class < 2,
            "bad arguments");
    const double fs = (double)b->fe_size(), hs = (double)b->hist_size();
    double total = 0;
    for (auto &I : b->inst) {
      if (!I.live)
        continue;
      const double nnpe = I.etype == AKX_TET4 ? 4 : 8, nqp = I.etype == AKX_TET4 ? 1 : 8;
      const double corners = (double)I.nel * nnpe, nn = (double)I.nn;
      // table entries written and read: one per element corner, or one per
      // (block, distinct node) under block-partial assembly
      const bool bpa = !I.blk_cnt.empty();
      const double entries = bpa ? (double)I.nbent : corners;
      if (kernel_class == 0) {
        if (bpa) // per corner the local corner (Tet4, 2 B) or the node id
                 // (Hex8, 4 B) and the slot (2 B); per entry the node id
                 // (Tet4), slot end and table position, the partial written
          total += corners * (I.etype == AKX_TET4 ? 4 : 6) +
                   entries * ((I.etype == AKX_TET4 ? 4 : 0) + 2 + 4 + 3 * fs) + nn * 2 * 3 * 8;
        else
          total += corners * (4 + 4 + 3 * fs) + nn * 2 * 3 * 8;
        if (I.law == AKX_LAW_PLASTIC_LIH) // history read and written
          total += (double)I.nel * nqp * 2 * (b->fp32() ? 7 : 28) * hs;
      } else {
        // a step followed by another (f_int is written by the last one
        // only): u, v and a read and written; FP32: u, the half-step
        // velocity and their compensation terms
        total += entries * 3 * fs + nn * (1 + 1 + 8 + (b->fp32() ? 2 * 3 * 2 * (8 + 4)
                                                                  : 3 * 3 * 2 * 8));
      }
    }
    *bytes = total;
  });
}

int akx_batch_enable_kernel_timing(akx_batch *b, int enable) {
  return guard([&] {
    require(b != nullptr, "null batch");
    b->timing = enable != 0;
    b->t_ms[0] = b->t_ms[1] = 0;
    b->t_n[0] = b->t_n[1] = 0;
  });
}

int akx_batch_get_kernel_timing(akx_batch *b, int kernel_class, double *total_ms,
                                int64_t *launches) {
  return guard([&] {
    require(b != nullptr && kernel_class >= 0 && kernel_class < 2, "bad arguments");
    if (total_ms)
      *total_ms = b->t_ms[kernel_class];
    if (launches)
      *launches = b->t_n[kernel_class];
  });
}

} // extern "C"
