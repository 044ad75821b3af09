/* oracle_internal.h -- preconditioner internals shared by the oracle's
 * translation units.  TEST INFRASTRUCTURE ONLY (see parcel_oracle.h). */
#ifndef PARCEL_ORACLE_INTERNAL_H
#define PARCEL_ORACLE_INTERNAL_H

#include "parcel_oracle.h"

struct oracle_precon {
  int ipreflag, ilu_method, nblocks, independ_nvec, n;
  int* blo; /* 1-based first row of each block */
  int* bhi; /* 1-based last row of each block */
  /* point Jacobi and D-ILU */
  double* dinv;
  /* D-ILU uses the original matrix entries and per-row L/U ranges */
  const int* rowptr;
  const int* col;
  const double* val;
  int* lstart; int* lend; int* ustart; int* uend; /* 0-based entry ids */
  /* ILU(0): per block split factors, 0-based local CSR */
  int** lptr; int** lcol; double** lval;
  int** uptr; int** ucol; double** uval;
  /* ORACLE_PRECON_FACTOR: M = L L^T on a five-point grid (batched.c) */
  int nx, ny;
  double* F64; /* ls, 1/Ld, a_fwd, a_bwd (4n) */
  float* F32;  /* the FP32 apply arrays (4n) */
  int apply32in64; /* FP64 vectors through the FP32 apply */
  int nstrip;      /* independent strips of grid lines */
  /* ORACLE_PRECON_RAS: block D-ILU of the expanded matrix, gather / scatter */
  oracle_precon* inner;
  int ras_n;
  int* ras_src;
  int* ras_dst;
  double* ras_re;
  double* ras_ze;
};

/* factor solves of batched.c */
void oracle_lscan_apply(int nx, int ny, const double* F, const double* r, double* z);
void oracle_lscan_apply32(int nx, int ny, const float* F, const float* r, float* z);
void oracle_lscan_apply32in64(int nx, int ny, int nstrip, const float* F, const double* r, double* z);
void oracle_lscan_apply_strips(int nx, int ny, int nstrip, const double* F, const double* r, double* z);
void oracle_lscan_apply32_strips(int nx, int ny, int nstrip, const float* F, const float* r, float* z);

/* D-ILU L/U entry ranges of rows lo..hi (1-based), init_solve_ILU_crs order */
void oracle_dilu_ranges(oracle_precon* p, int lo, int hi);

#endif
