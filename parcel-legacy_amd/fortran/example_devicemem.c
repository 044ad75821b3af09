/* example_devicemem.c -- a C driver in the form of the reference's GPU-build
 * C driver (src/example/main_direct_c.c:195-243 with CUDA_SOLVER): the
 * system goes to device memory, every array is passed as the address of a
 * devicemem handle, the solver is called through the by-reference entry
 * points (libparcel_devicemem.so).  Prints niter, iret, the residual
 * history and the first and last entries of x, like example_poisson.
 *
 * usage: example_devicemem solver gnx ipreflag ILU_method nblocks itrmax rtol [p1 p2 p3 p4]
 *   solver 1 dcg, 2 dbicgstab, 3 dgmres (p1 m, p2 GS flag), 4 dcagmres
 *   (p1 s, p2 t, p3 basis, p4 QR flag), 5 dcbcg (p1 k, p2 eigenvalue flag),
 *   6 ddcg (lo parts zero), 7 dcg_ddm (margin 1, p1 addL) */
#include <stdio.h>
#include <stdlib.h>

#include "parcel_devicemem.h"
#include "parcel_hip.h"

static parcel_devicemem to_device(const void* host, size_t bytes) {
  parcel_devicemem h = {NULL, NULL, NULL};
  if (parcel_hip_malloc(&h.mem, bytes) || parcel_hip_memcpy(h.mem, host, bytes, 1)) {
    fprintf(stderr, "device copy failed: %s\n", parcel_hip_last_error());
    exit(2);
  }
  return h;
}

int main(int argc, char** argv) {
  if (argc < 8) {
    fprintf(stderr, "usage: %s solver gnx ipreflag ILU_method nblocks itrmax rtol [p1..p4]\n", argv[0]);
    return 2;
  }
  int solver = atoi(argv[1]), g = atoi(argv[2]), ipre = atoi(argv[3]), ilu = atoi(argv[4]);
  int nblk = atoi(argv[5]), itrmax = atoi(argv[6]);
  double rtol = atof(argv[7]);
  int p[4] = {4, 1, 0, 4};
  if (solver == 5) { p[0] = 4; p[1] = 1; p[2] = 0; p[3] = 0; }
  for (int k = 0; k < 4 && 8 + k < argc; ++k) p[k] = atoi(argv[8 + k]);
  const int zero = 0, one = 1, four = 4, nb2000 = 2000, two = 2, m1 = -1, hundred = 100, ten = 10, three = 3;
  int n = g * g * g, nnz = 0, iret = -99, niter = itrmax, x1 = 0, xn = n - 1;
  double* hist = calloc((size_t)itrmax + 64, sizeof(double));
  parcel_init_c_(&zero);
  double* x = NULL;
  parcel_devicemem dx = {NULL, NULL, NULL};
  if (solver == 7) {
    const int m = (g + 2) * (g + 2) * (g + 2), nnd = 7;
    double* vd = calloc((size_t)7 * m, sizeof(double));
    double* b = calloc((size_t)m, sizeof(double));
    double* xe = calloc((size_t)m, sizeof(double));
    int ioff[7], iog[21], nbr[1] = {0}, ndiv[3] = {1, 1, 1}, ng[3] = {g, g, g};
    parcel_gen_poisson3d_ddm(g, g, g, 1, vd, ioff, iog, b, xe);
    x = calloc((size_t)m, sizeof(double));
    dx = to_device(x, sizeof(double) * m);
    parcel_devicemem db = to_device(b, sizeof(double) * m), dv = to_device(vd, sizeof(double) * 7 * m);
    parcel_devicemem doff = to_device(ioff, sizeof(int) * 7);
    parcel_dcg_ddm_c_(&zero, &dx, &db, &m, &nnd, &doff, &dv, &zero, nbr, &one, &three, ndiv, ng, iog, &one, &ipre,
                      &p[0], &zero, &niter, &rtol, hist, &zero, &m1, &one, &one, &nb2000, &iret);
    parcel_hip_memcpy(x, dx.mem, sizeof(double) * m, 2);
    x1 = 1 + (g + 2) + (g + 2) * (g + 2);
    xn = g + (g + 2) * g + (g + 2) * (g + 2) * g;
  } else {
    parcel_gen_poisson3d_parcel(g, g, g, &nnz, NULL, NULL, NULL);
    int* rp = malloc(sizeof(int) * (n + 1));
    int* col = malloc(sizeof(int) * nnz);
    double* val = malloc(sizeof(double) * nnz);
    double* b = malloc(sizeof(double) * n);
    double* xe = malloc(sizeof(double) * n);
    double* zn = calloc((size_t)(n > nnz ? n : nnz), sizeof(double));
    parcel_gen_poisson3d_parcel(g, g, g, &nnz, rp, col, val);
    parcel_gen_poisson3d_rhs_parcel(g, g, g, b, xe);
    x = calloc((size_t)n, sizeof(double));
    dx = to_device(x, sizeof(double) * n);
    parcel_devicemem db = to_device(b, sizeof(double) * n), da = to_device(val, sizeof(double) * nnz);
    parcel_devicemem drp = to_device(rp, sizeof(int) * (n + 1)), dcol = to_device(col, sizeof(int) * nnz);
    parcel_devicemem dxl = to_device(zn, sizeof(double) * n), dbl = to_device(zn, sizeof(double) * n);
    parcel_devicemem dal = to_device(zn, sizeof(double) * nnz);
    switch (solver) {
      case 1:
        parcel_dcg_c_(&zero, &dx, &db, &n, &n, &nnz, &one, &da, &drp, &dcol, &ipre, &ilu, &zero, &four, &niter,
                      &rtol, hist, &zero, &nblk, &one, &one, &nb2000, &iret);
        break;
      case 2:
        parcel_dbicgstab_c_(&zero, &dx, &db, &n, &n, &nnz, &one, &da, &drp, &dcol, &ipre, &ilu, &zero, &four,
                            &niter, &rtol, hist, &zero, &nblk, &one, &one, &nb2000, &iret);
        break;
      case 3:
        parcel_dgmres_c_(&zero, &dx, &db, &n, &n, &nnz, &one, &da, &drp, &dcol, &ipre, &ilu, &zero, &four, &niter,
                         &rtol, hist, &zero, &nblk, &one, &one, &nb2000, &p[0], &p[1], &iret);
        break;
      case 4:
        parcel_dcagmres_c_(&zero, &dx, &db, &n, &n, &nnz, &one, &da, &drp, &dcol, &ipre, &ilu, &zero, &four,
                           &niter, &rtol, hist, &zero, &nblk, &one, &one, &nb2000, &p[0], &p[1], &p[2], &p[3],
                           &iret);
        break;
      case 5:
        parcel_dcbcg_c_(&zero, &dx, &db, &n, &n, &nnz, &one, &da, &drp, &dcol, &ipre, &ilu, &zero, &four, &niter,
                        &rtol, hist, &zero, &nblk, &one, &one, &nb2000, &p[0], &p[1], &hundred, &four, &ten, &one,
                        &three, &iret);
        break;
      case 6:
        parcel_ddcg_c_(&zero, &dx, &dxl, &db, &dbl, &n, &n, &nnz, &one, &da, &dal, &drp, &dcol, &ipre, &ilu, &zero,
                       &four, &niter, &rtol, hist, &zero, &nblk, &one, &one, &nb2000, &two, &two, &two, &two, &iret);
        break;
      default:
        fprintf(stderr, "unknown solver %d\n", solver);
        return 2;
    }
    parcel_hip_memcpy(x, dx.mem, sizeof(double) * n, 2);
  }
  parcel_free_c_();
  printf("niter %d iret %d\n", niter, iret);
  for (int i = 0; i < niter; ++i) printf("%25.17e\n", hist[i]);
  printf("x1 %25.17e\n", x[x1]);
  printf("xn %25.17e\n", x[xn]);
  return 0;
}
