// Device-side descriptors shared by the RCWA kernels and the plan layer.
#pragma once
#include <stdint.h>
#include "orf_common.h"

#define ORF_MAX_SLOTS 8  // distinct materials per patterned layer

struct DevShape {
  int type;    // 0 circle, 1 rect, 2 polygon
  int slot;    // material slot painted by this shape
  double cx, cy, a, b;
  int vert0;   // polygon: offset of x-coordinates in the vertex array (y follow after nvert)
  int nvert;
};

struct DevGrid {
  int shape0, nshape;
  double Lx, Ly;  // period of the structure this grid belongs to
};

// One (structure, context) page for a given layer; a context is one
// (wavelength, period) pair with its own k-vectors and half-space S-matrices.
struct DevPage {
  int b, lam;   // lam: context index
  int grid;    // label grid of this layer (-1: homogeneous)
  int nslot;
};

// One solve page: a mode page in one parity sector (sector 0 in the full basis).
struct DevSolve {
  int mode;    // mode page index within the chunk
  int sector;
};

// Orthonormal parity-sector basis (symmetryBasis.m:2-15) as sparse lists.
struct DevSector {
  const int* ue_row;       // [4 * nr] full row of each nonzero of UE column a (-1 = none)
  const double* ue_coef;   // [4 * nr]
  const int* uh_row;
  const double* uh_coef;
  const int* ue_rowmap;    // [nf] reduced column of full row i (-1 = outside the sector)
  const double* ue_rowcoef;
};

struct DevSource {
  int pol;          // output slot
  int sector;       // solve sector carrying this source
  cplx ex, ey;      // transverse incident E at the zero order
  int srcCol;       // reduced basis: the one nonzero of UE' c
  cplx srcVal;
};
