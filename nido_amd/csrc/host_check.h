// host_check.h -- host-side CSR validation and per-graph sums (OpenMP),
// shared by the device handle's constructors.
#pragma once
#include <stdint.h>
#include "fixed.h"

namespace nido {
// Offsets non-decreasing; every arc of graph g (vertices [vptr[g],
// vptr[g+1])) points inside graph g; weights finite with |w| < 2^62 and the
// bounds on their totals.  off/col/w are indexed from off[0].  *nonneg:
// every weight >= 0; twom (if not null, G entries): every graph's exact 2m.
int check_csr(int64_t nv, const int64_t* off, const int64_t* col, const double* w, int G,
              const int64_t* vptr, bool* nonneg, fx_t* twom = nullptr);
// Its first part: offsets non-decreasing, col and w present when there are
// arcs (the rest is checked while large inputs are uploaded).
int check_offsets(int64_t nv, const int64_t* off, const int64_t* col, const double* w);
// The greedy colouring of Graph::coloring run by one thread (coloring.hip):
// colors[i] = smallest colour not held by a neighbour j < i with
// r[j] >= r[i]; off[nv + 1], col[off[nv]] (ids in [0, nv), narrowed to
// int32 on the way to the device), host arrays.
int device_coloring(int64_t nv, const int64_t* off, const int64_t* col, const int64_t* r, int* colors,
                    int64_t* rounds);
}  // namespace nido
