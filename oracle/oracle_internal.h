/* oracle_internal.h -- preconditioner internals shared by the oracle's
 * translation units.  TEST INFRASTRUCTURE ONLY (see parcel_oracle.h). */
#ifndef PARCEL_ORACLE_INTERNAL_H
#define PARCEL_ORACLE_INTERNAL_H

#include "parcel_oracle.h"

struct oracle_precon {
  int ipreflag, ilu_method, nblocks, independ_nvec, n;
  int base; /* D-ILU arrays and vectors hold row k (1-based) at k - 1 - base
               (0; an additive-Schwarz region: its first row - 1) */
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
  /* ORACLE_PRECON_AS: additive Schwarz over divide_region blocks extended by
     addL rows; region t is a D-ILU precon (base = its first row - 1) */
  int addL, iflagAS;
  oracle_precon** as_reg;
  int* olo; /* own rows of block t (1-based, inclusive) */
  int* ohi;
  double** as_in;
  double** as_out;
};

/* D-ILU factor and L/U ranges of one block (rows lo..hi, 1-based) */
int oracle_dilu_block(oracle_precon* p, int lo, int hi);
void oracle_dilu_apply_block(const oracle_precon* p, int lo, int hi, const double* r, double* s);
/* additive Schwarz (as.c) */
void oracle_as_apply(const oracle_precon* p, const double* r, double* s);
void oracle_as_free(oracle_precon* p);

/* factor solves of batched.c */
void oracle_lscan_apply(int nx, int ny, const double* F, const double* r, double* z);
void oracle_lscan_apply32(int nx, int ny, const float* F, const float* r, float* z);
void oracle_lscan_apply32in64(int nx, int ny, int nstrip, const float* F, const double* r, double* z);
void oracle_lscan_apply_strips(int nx, int ny, int nstrip, const double* F, const double* r, double* z);
void oracle_lscan_apply32_strips(int nx, int ny, int nstrip, const float* F, const float* r, float* z);

/* D-ILU L/U entry ranges of rows lo..hi (1-based), init_solve_ILU_crs order */
void oracle_dilu_ranges(oracle_precon* p, int lo, int hi);

#endif
