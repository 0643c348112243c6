// Plain-data types of the discrete adjoint shared by its kernels
// (chns_adjoint.hpp) and the host API (chns_adjoint.cpp).
#pragma once
#include "chns_types.hpp"

namespace chns {

// partial sums: fac, G1, G2, G3, g1i, g2i, g3i (adj_rowB), mc1, mc2 (adj_colBC),
// A11, A12, A21, A22 (adj_rowD2), lg1, lg2 (adj_colBC)
constexpr int ADJ_NQ = 15;
enum {
  AQ_FAC = 0, AQ_G1 = 1, AQ_G2 = 2, AQ_G3 = 3, AQ_G1I = 4, AQ_G2I = 5, AQ_G3I = 6, AQ_MC1 = 7, AQ_MC2 = 8,
  AQ_A11 = 9, AQ_A12 = 10, AQ_A21 = 11, AQ_A22 = 12, AQ_LG1 = 13, AQ_LG2 = 14
};

struct AdjBufs {
  const cd *W, *C1, *C2;                           // linearisation point, column-major [Ny2][N]
  const cd *Hc1, *Hc2, *Hu, *Hv, *Hwx, *Hwy, *Hmu1, *Hmu2, *Hdx1, *Hdx2;   // passes A, C at it (TJ, ldJ)
  const cd *Mh1, *Mh2;                             // pass B at it (TF, ldF)
  cd *Ub[3], *Ab[3], *Gi[3];                       // ubar, abar, product input (column-major)
  cd *Hxi, *Hxix, *Het1, *Het2;                    // column iFFTs of the input (TJ, ldJ)
  cd* O[14];                                       // adjoint row-pass outputs (TF, ldJ, j < Jt)
  cd *Hmb1, *Hmb2;                                 // adj_colBC outputs (TJ, ldJ)
  cd *Ocb1, *Ocb2;                                 // adj_rowB outputs (TF, ldJ)
  double* part;                                    // [ADJ_NQ][mb] workgroup partials of this product
  int mb;                                          // slots per quantity
  const cd* tw;
  const cd* tabw1;
  const cd* tabc1;
  const double* tabw2;
  const double* tabc2;
  const int* amask;
};

struct AdjArgs {
  AdjBufs A;
  KParams P;
};

}  // namespace chns
