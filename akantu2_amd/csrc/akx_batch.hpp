// akx_batch.hpp -- device-side tables and launch arguments of the batched
// multi-instance engine (akx_batch.hip, strict FP64 kernels) shared with the
// FP32-form batched element kernels (akx_fast.hip).
//
// Work lists:
//   tiles  : runs of at most `tile` consecutive elements of ONE instance, so
//            material parameters and the history parity are uniform over a
//            tile; one list per (element type, law) group.
//   slices : 256-node slices of the live instances (an instance owns whole
//            slices: its node range starts on a multiple of 256).
// Every launch gives each workgroup one tile (element passes) or one slice
// (gathers): persistent workgroups walking chunks of the lists measured
// slower, their dependent load latencies running back to back.
#pragma once
#include "akx_args.hpp"
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace akx {

struct BInst { // one per instance id
  double lambda, mu, sigma_y, h, rho, dt;
  int32_t parity_base; // steps done before the current solve call, mod 2
  int32_t law;
  // where the instance's displacement lives: 1 in the {X, u} node records
  // (FP64: its passes read corners from them), 0 in the [3][ncap] array u
  // (FP32: the Tet4 block pass builds the records in LDS from X and u, the
  // Hex8 pass reads its corners from them, and the gathers store u whole
  // instead of half-records)
  int32_t urec;
  int32_t pad;
};

struct BTile {
  int64_t first; // first element (index in the element arena of its type)
  int32_t count;
  int32_t inst;
};

struct BSlice {
  int64_t node0; // multiple of 256
  int32_t count; // nodes of the instance in this slice
  int32_t inst;
};

// FP32 Tet4 block-partial assembly: the distinct nodes of one 256-element
// block of an instance (its tile) are entries off .. off + count - 1 of the
// block arenas (node ids, slot ends, table positions)
struct BBlk {
  int64_t off;
  int32_t count;
  int32_t pad;
};

struct BatchElemArgs {
  const BTile *tiles;
  int64_t ntiles;
  // block-partial assembly (FP32 Tet4 groups; null otherwise): per tile (one
  // 256-element block) its distinct-node range, the block arenas, and each
  // element's local corners and LDS slots ([elem][4] uint16)
  const BBlk *blks;
  const int32_t *bnode;   // arena node id of each distinct node
  const uint16_t *bsend;  // end of each distinct node's slot range
  const int32_t *bpos;    // table entry of each (block, distinct node) partial
  const uint16_t *lc, *ls;
  int32_t bmaxu;          // most distinct nodes of one live block
  const BInst *inst;
  const int32_t *conn; // [elem][nnpe] arena node ids
  const int32_t *pos;  // [elem][nnpe] table entries
  const double *xu;    // [ncap][6] node records {X, u}
  const double *X, *u; // [3][ncap] coordinates and displacement (BInst::urec = 0)
  int64_t ncap;
  void *fe;            // [entry][FE_W] elemental forces (FP64 or FP32)
  // plastic history, two buffers (ping-pong by step parity).  FP64 form:
  // grad u, stress, plastic strain ([9][hcap]) and hardening ([hcap]);
  // FP32 form: plastic strain in Voigt order ([6][hcap], in the third
  // slot) and hardening ([hcap]).  Gauss point q of element e is index
  // hoff + e * nqp + q.
  void *hist[2][4];
  int64_t hcap, hoff;
  const double *hex_tab; // Hex8: dN/dxi [8][3][8], N [8][8], w [8]
  int32_t parity;        // step index within the solve call, mod 2
};

constexpr int HEX_TAB_N = 192, HEX_TAB_W = 256, HEX_TAB_SIZE = 264;

// elements per block of the block-partial assembly: 256 Tet4; 192 Hex8, whose
// 1,536 corner slots (18 KB of LDS) leave room for eight 256-lane
// workgroups per CU (256 elements' 24 KB allowed six)
constexpr int BPA_TET4 = 256, BPA_HEX8 = 192;

// values per elemental-force entry in a batch's table.  3: a zero-padded
// 4-wide entry (one vector store per corner) measured slower, its extra
// bytes cost more than the saved cache accesses.
constexpr int FE_W = 3;

// the Gauss-point view of one step: read buffer `par`, write buffer par ^ 1
template <class HT>
__device__ __forceinline__ QpViewT<HT> batch_view(const BatchElemArgs &A, int par) {
  return QpViewT<HT>{(const HT *)A.hist[par][0], (const HT *)A.hist[par][1],
                     (const HT *)A.hist[par][2], (const HT *)A.hist[par][3],
                     (HT *)A.hist[par ^ 1][0],   (HT *)A.hist[par ^ 1][1],
                     (HT *)A.hist[par ^ 1][2],   (HT *)A.hist[par ^ 1][3],
                     nullptr,                    nullptr,
                     A.hcap,                     nullptr,
                     nullptr};
}

// the compact FP32 plastic state of one step (read `par`, write par ^ 1)
struct QpCompact {
  const float *p_eps, *p_iso;
  float *c_eps, *c_iso;
  double *grad_u, *stress; // not materialised: always null
  int64_t nq_all;
};

__device__ __forceinline__ QpCompact batch_compact_view(const BatchElemArgs &A, int par) {
  return QpCompact{(const float *)A.hist[par][2], (const float *)A.hist[par][3],
                   (float *)A.hist[par ^ 1][2], (float *)A.hist[par ^ 1][3], nullptr,
                   nullptr, A.hcap};
}

} // namespace akx
