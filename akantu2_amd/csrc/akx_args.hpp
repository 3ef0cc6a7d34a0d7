// akx_args.hpp -- plain argument structs shared by the element kernels of
// the strict (reference operation order) and fast translation units.
#pragma once
#include <stdint.h>

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
  const int32_t *pos;     // [(pos_offset + i) * nnpe + a] -> entry of the
                          // node-major elemental-force table
  void *fe;               // [entry][3] of FT
  // materialised internals [9][n*nqp] (or null), column-major 3x3 index
  double *grad_u, *stress;
  // plastic history, [9][n*nqp] and [n*nqp]: previous (read) / current (write)
  const double *p_grad, *p_sig, *p_eps, *p_iso;
  double *c_grad, *c_sig, *c_eps, *c_iso;
  int64_t nblocks_padded;
};

} // namespace akx
