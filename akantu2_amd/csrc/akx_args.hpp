// akx_args.hpp -- plain argument structs shared by the element kernels of
// the strict (reference operation order) and fast translation units.
#pragma once
#ifndef __HIPCC_RTC__ // run-time compilation provides these
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace akx {

// ---------------------------------------------------------------------------
// Element pass arguments.  One launch covers the elements of one material.
// ---------------------------------------------------------------------------
struct ElemArgs {
  int64_t n;              // elements in this launch
  int64_t nel;            // elements in the mesh (connectivity stride)
  const int32_t *filter;  // element ids, or nullptr (ids first + i)
  int64_t first;
  int64_t pos_offset;     // assembly position of the first element
  const int32_t *conn;    // [nel][nnpe]
  const double *X;        // [3][nn]
  const double *u;        // [3][nn]
  int64_t nn;
  double lambda, mu, sigma_y, h;
  const double *lam_el, *mu_el, *sy_el, *h_el; // per launch element, or null
  const float2 *lm_f; // FP32 form: (float(lambda), float(mu)) per element, or null
  // FP32 Tet4: per 256-element block, the block's distinct nodes (bl_nodes
  // [b * bl_maxu, (b + 1) * bl_maxu), padded with -1) and each element's
  // corners as indices into them (lconn, [n][4]); null when not built
  const int32_t *bl_nodes;
  const uint16_t *lconn;
  int32_t bl_maxu; // most distinct nodes of one block
  // FP32 Tet4 with block-partial assembly (null otherwise): each block sums
  // the forces its elements put on each of its distinct nodes and writes one
  // partial per (block, distinct node).  lslot [n][4]: the LDS slot of each
  // element corner (the corners of distinct node k of the block take slots
  // bl_soff[b * (bl_maxu + 1) + k] .. [k + 1] in ascending element order);
  // bl_pos [b * bl_maxu + k]: the table entry of the partial (-1: padding).
  const uint16_t *lslot;
  const uint16_t *bl_soff;
  const int32_t *bl_pos;
  const int32_t *pos;     // [(pos_offset + i) * nnpe + a] -> entry of the
                          // node-major elemental-force table
  // FP64 Tet4 with partial-sum assembly (null otherwise), on the 256-element
  // blocks of bl_nodes / lconn.  pcode [n][4]: >= 0 the table entry of the
  // corner, < 0 the LDS slot -1 - pcode of a corner whose run (its node's
  // corners in this block) is summed in the block.  The summed runs of block
  // b are poff[b] .. poff[b + 1] - 1; run k sums slots pend[k - 1] (0 for
  // the block's first) .. pend[k] - 1 into table entry ppos[k].  t_nblk:
  // blocks of the launch; t_slots: LDS slots of the largest block.
  const int32_t *pcode, *poff, *ppos;
  const uint16_t *pend;
  int64_t t_nblk;
  int32_t t_slots;
  // Hex8, plain FP64: J^-1 and w det J of every Gauss point of the mesh,
  // [5][nel * 8] double2 {Ji00, Ji01} {Ji02, Ji10} {Ji11, Ji12} {Ji20, Ji21}
  // {Ji22, wJ}, or null
  const double2 *jw;
  // Tet4: one 48-byte record {X, u} per node, [nn][6], kept equal to X and u
  // by every kernel that writes u; the element pass reads it instead of the
  // six [3][nn] arrays (three 16-byte loads per node instead of six 8-byte
  // loads from six cache lines)
  const double *xu;
  void *fe;               // [entry][3] of FT
  // materialised internals [9][n*nqp] (or null), column-major 3x3 index
  double *grad_u, *stress;
  // plastic history, [9][n*nqp] and [n*nqp]: previous (read) / current (write)
  const double *p_grad, *p_sig, *p_eps, *p_iso;
  double *c_grad, *c_sig, *c_eps, *c_iso;
  // plastic energy [n*nqp]: previous (read) / current (write), or null
  const double *p_pe;
  double *c_pe;
  // Material::eigengradu, [9][n*nqp], subtracted from grad u (null: zero);
  // thermal stress sigma_th of every Gauss point [n*nqp] (null: -0), and for
  // the plastic law its previous (read) / current (write) values
  const double *eig, *sth_now, *sth_p;
  double *sth_c;
  int64_t nblocks_padded;
};

// ---------------------------------------------------------------------------
// Where the Gauss-point state of one launch lives: previous (read) and
// current (written) plastic history, [9][nq_all] column-major 3x3 and
// [nq_all] for the hardening, stored in HT; and the optional grad u /
// stress output (FP64, same indexing).
// ---------------------------------------------------------------------------
template <class HT> struct QpViewT {
  const HT *p_grad, *p_sig, *p_eps, *p_iso;
  HT *c_grad, *c_sig, *c_eps, *c_iso;
  double *grad_u, *stress;
  int64_t nq_all;
  const HT *p_pe; // accumulated plastic energy, or null (not tracked)
  HT *c_pe;
  // eigen grad u [9][nq_all] and thermal stress [nq_all] (plastic: previous /
  // current), read by the element passes instantiated with them
  const double *eig = nullptr, *sth_now = nullptr, *sth_p = nullptr;
  double *sth_c = nullptr;
};
using QpView = QpViewT<double>;

__host__ __device__ inline QpView qp_view(const ElemArgs &A, int nqp) {
  return QpView{A.p_grad, A.p_sig,  A.p_eps,  A.p_iso,  A.c_grad, A.c_sig,
                A.c_eps,  A.c_iso,  A.grad_u, A.stress, A.n * nqp, A.p_pe,
                A.c_pe,   A.eig,    A.sth_now, A.sth_p, A.sth_c};
}

// ---------------------------------------------------------------------------
// Where the element pass finds the coordinates X and displacement u of a
// node: structure-of-arrays [3][stride] (single model) or one 48-byte record
// per node {X0, X1, X2, u0, u1, u2} read with three 16-byte loads (batches).
// ---------------------------------------------------------------------------
struct NodesSoA {
  const double *X, *u;
  int64_t stride;
  __device__ __forceinline__ void load(int32_t n, double (&x)[3], double (&d)[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c] = X[(int64_t)c * stride + n];
      d[c] = u[(int64_t)c * stride + n];
    }
  }
  __device__ __forceinline__ void load_u(int32_t n, double (&d)[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      d[c] = u[(int64_t)c * stride + n];
  }
};

// The four corners' {X, u} records, already read into registers (corner
// a is node index a)
struct NodesRegs4 {
  double r[4][6];
  __device__ __forceinline__ void load(int32_t n, double (&x)[3], double (&d)[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c] = r[n][c];
      d[c] = r[n][3 + c];
    }
  }
};

// Write one elemental-force entry: W = 3 values, or W = 4 with a zero pad so
// that the entry is one aligned vector store (16 B in FP32, 2 x 16 B in FP64).
template <int W, class FT>
__device__ __forceinline__ void put_entry(FT *fe, int64_t entry, FT a, FT b, FT c) {
  if constexpr (W == 3) {
    FT *o = fe + entry * 3;
    o[0] = a;
    o[1] = b;
    o[2] = c;
  } else if constexpr (sizeof(FT) == 4) {
    *reinterpret_cast<float4 *>(fe + entry * 4) = make_float4(a, b, c, 0.f);
  } else {
    double2 *o = reinterpret_cast<double2 *>(fe + entry * 4);
    o[0] = make_double2(a, b);
    o[1] = make_double2(c, 0.);
  }
}

// Tet4 passes with staged records: builds the {X, u} records of block blk's
// distinct nodes in LDS from the [3][nn] arrays X and u (six 8-byte loads
// per node; a block's list holds runs of consecutive node ids, so they
// coalesce), three 16-byte LDS stores per node.  Reading X and u rather than
// the global records lets the gathers skip the records' u, whose 24-byte
// half-record stores cost the FP32 gather of 60 M Tet4 a quarter of its
// time.  The caller synchronises the block before reading them.
__device__ __forceinline__ void stage_block_records(const ElemArgs &A, int64_t blk,
                                                    double2 *s) {
  const int32_t *list = A.bl_nodes + blk * A.bl_maxu;
  for (int k = threadIdx.x; k < A.bl_maxu; k += blockDim.x) {
    const int64_t node = list[k];
    if (node >= 0) {
      s[3 * k] = make_double2(A.X[node], A.X[A.nn + node]);
      s[3 * k + 1] = make_double2(A.X[2 * A.nn + node], A.u[node]);
      s[3 * k + 2] = make_double2(A.u[A.nn + node], A.u[2 * A.nn + node]);
    }
  }
}

// {X, u} records staged in LDS, indexed by a block's distinct-node number
struct NodesLDS {
  const double *s; // [distinct nodes][6]
  __device__ __forceinline__ void load(int32_t n, double (&x)[3], double (&d)[3]) const {
    const double2 *p = reinterpret_cast<const double2 *>(s + 6 * n);
    const double2 a = p[0], b = p[1], c = p[2];
    x[0] = a.x;
    x[1] = a.y;
    x[2] = b.x;
    d[0] = b.y;
    d[1] = c.x;
    d[2] = c.y;
  }
};

struct NodesXU {
  const double *xu; // [n][6]
  __device__ __forceinline__ void load(int32_t n, double (&x)[3], double (&d)[3]) const {
    const double2 *p = reinterpret_cast<const double2 *>(xu + 6 * (int64_t)n);
    const double2 a = p[0], b = p[1], c = p[2];
    x[0] = a.x;
    x[1] = a.y;
    x[2] = b.x;
    d[0] = b.y;
    d[1] = c.x;
    d[2] = c.y;
  }
};

} // namespace akx
