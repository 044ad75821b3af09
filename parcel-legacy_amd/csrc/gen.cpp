// gen.cpp -- synthetic problem generators (host code).
//
// The counterpart of the reference's example generators
// (src/example/problem.f90) plus the seeded generators of the synthetic
// workloads.  Matrices come out as 1-based CSR with sorted columns, the
// layout parcel_dcg takes (src/wrapper.f90:1995-2010).
//
// Random numbers: SplitMix64 (Steele, Lea, Flood 2014).  "Seed s" means the
// stream z_i = mix(s + (i+1) * 0x9E3779B97F4A7C15), i = 0, 1, ...; the top
// 53 bits give u_i in [0, 1).  Each value depends only on (s, i), so the
// stream can be generated in parallel.
#include <cmath>
#include <cstdint>
#include <vector>

#include "parcel_hip.h"

namespace {

inline uint64_t splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline double u01(uint64_t seed, uint64_t i) {
  return (double)(splitmix64(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

// 7-point stencil in the reference's entry order b, s, w, c, e, n, t
// (make_poisson3D_CRS, problem.f90:448-618).
int stencil3d(int gnx, int gny, int gnz, double off, double diag, int* nnz, int* rowptr, int* col,
              double* val) {
  if (gnx < 1 || gny < 1 || gnz < 1 || !nnz) return PARCEL_HIP_ERR_ARG;
  long n = (long)gnx * gny * gnz;
  long cnt = n * 7 - 2L * gny * gnz - 2L * gnx * gnz - 2L * gnx * gny;
  if (n > 0x7fffffffL || cnt > 0x7fffffffL) return PARCEL_HIP_ERR_ARG;
  *nnz = (int)cnt;
  if (!rowptr || !col || !val) return PARCEL_HIP_OK;
  const long pxy = (long)gnx * gny;
#pragma omp parallel for schedule(static)
  for (int iz = 1; iz <= gnz; ++iz) {
    // entries of planes 1..iz-1: each holds the 5-point count plus one z
    // neighbour per row below (planes > 1) and one above (planes < gnz)
    const long plane5 = 5 * pxy - 2L * gny - 2L * gnx;
    long e = (long)(iz - 1) * plane5 + pxy * (long)(iz - 1) + pxy * (long)(iz > 2 ? iz - 2 : 0);
    for (int iy = 1; iy <= gny; ++iy) {
      for (int ix = 1; ix <= gnx; ++ix) {
        long row = ix + (long)(iy - 1) * gnx + (long)(iz - 1) * pxy;  // 1-based
        rowptr[row - 1] = (int)(e + 1);
        if (iz > 1) { col[e] = (int)(row - pxy); val[e] = off; ++e; }
        if (iy > 1) { col[e] = (int)(row - gnx); val[e] = off; ++e; }
        if (ix > 1) { col[e] = (int)(row - 1); val[e] = off; ++e; }
        col[e] = (int)row; val[e] = diag; ++e;
        if (ix < gnx) { col[e] = (int)(row + 1); val[e] = off; ++e; }
        if (iy < gny) { col[e] = (int)(row + gnx); val[e] = off; ++e; }
        if (iz < gnz) { col[e] = (int)(row + pxy); val[e] = off; ++e; }
      }
    }
  }
  rowptr[n] = (int)(cnt + 1);
  return PARCEL_HIP_OK;
}

}  // namespace

extern "C" {

int parcel_gen_poisson3d_parcel(int gnx, int gny, int gnz, int* nnz, int* rowptr, int* col,
                                double* val) {
  // cc = -2.0*(dxi2+dyi2+dzi2) with dx = dy = dz = 1, off-diagonals +1
  return stencil3d(gnx, gny, gnz, 1.0, -6.0, nnz, rowptr, col, val);
}

int parcel_gen_poisson3d_spd(int gnx, int gny, int gnz, int* nnz, int* rowptr, int* col,
                             double* val) {
  return stencil3d(gnx, gny, gnz, -1.0, 6.0, nnz, rowptr, col, val);
}

// make_poisson3D_right_vector (problem.f90:772-875) for one rank holding the
// whole grid.  real(i)**2 of the reference is exact in single precision for
// these sizes; the boundary updates `b - a*a*(c*c)` are contracted to one
// fused multiply-add as -march=native builds of the reference do.
int parcel_gen_poisson3d_rhs_parcel(int gnx, int gny, int gnz, double* vecb, double* vecx) {
  if (gnx < 1 || gny < 1 || gnz < 1 || !vecb) return PARCEL_HIP_ERR_ARG;
  const double hx = 1.0 / (gnx + 1), hy = 1.0 / (gny + 1), hz = 1.0 / (gnz + 1);
  const long nx = gnx, ny = gny, nz = gnz;
  const double z2 = 2.0 * (hx * hx) * (hy * hy) * (hz * hz);
#pragma omp parallel for schedule(static)
  for (long iz = 1; iz <= nz; ++iz)
    for (long iy = 1; iy <= ny; ++iy)
      for (long ix = 1; ix <= nx; ++ix) {
        const double cx = (double)(ix * ix), cy = (double)(iy * iy), cz = (double)(iz * iz);
        const long i = ix + (iy - 1) * nx + (iz - 1) * nx * ny;
        vecb[i - 1] = z2 * (cx * cy + cy * cz + cz * cx);
      }
  // x = gnx face
  for (long iz = 1; iz <= nz; ++iz)
    for (long iy = 1; iy <= ny; ++iy) {
      const long i = nx + (iy - 1) * nx + (iz - 1) * nx * ny;
      const double a = hy * (double)iy, c = hz * (double)iz;
      vecb[i - 1] = std::fma(-(a * a), c * c, vecb[i - 1]);
    }
  // y = gny face
  for (long iz = 1; iz <= nz; ++iz)
    for (long ix = 1; ix <= nx; ++ix) {
      const long i = ix + (ny - 1) * nx + (iz - 1) * nx * ny;
      const double a = hx * (double)ix, c = hz * (double)iz;
      vecb[i - 1] = std::fma(-(a * a), c * c, vecb[i - 1]);
    }
  // z = gnz face
  for (long iy = 1; iy <= ny; ++iy)
    for (long ix = 1; ix <= nx; ++ix) {
      const long i = ix + (iy - 1) * nx + (nz - 1) * nx * ny;
      const double a = hx * (double)ix, c = hy * (double)iy;
      vecb[i - 1] = std::fma(-(a * a), c * c, vecb[i - 1]);
    }
  if (vecx) {
    const double z1 = (hx * hx) * (hy * hy) * (hz * hz);
#pragma omp parallel for schedule(static)
    for (long iz = 1; iz <= nz; ++iz)
      for (long iy = 1; iy <= ny; ++iy)
        for (long ix = 1; ix <= nx; ++ix) {
          const double cx = (double)(ix * ix), cy = (double)(iy * iy), cz = (double)(iz * iz);
          const long i = ix + (iy - 1) * nx + (iz - 1) * nx * ny;
          vecx[i - 1] = z1 * (cx * cy * cz);
        }
  }
  return PARCEL_HIP_OK;
}

// make_poisson2D_CRS (problem.f90:210-340): entry order s, w, c, e, n;
// diagonal -4, off-diagonals +1, all multiplied by `scale`.
int parcel_gen_poisson2d(int gnx, int gny, double scale, int* nnz, int* rowptr, int* col,
                         double* val) {
  if (gnx < 1 || gny < 1 || !nnz) return PARCEL_HIP_ERR_ARG;
  long n = (long)gnx * gny;
  long cnt = 5 * n - 2L * gnx - 2L * gny;
  if (cnt > 0x7fffffffL) return PARCEL_HIP_ERR_ARG;
  *nnz = (int)cnt;
  if (!rowptr || !col || !val) return PARCEL_HIP_OK;
  const double off = 1.0 * scale, diag = -4.0 * scale;
  long e = 0;
  for (int iy = 1; iy <= gny; ++iy)
    for (int ix = 1; ix <= gnx; ++ix) {
      long row = ix + (long)(iy - 1) * gnx;
      rowptr[row - 1] = (int)(e + 1);
      if (iy > 1) { col[e] = (int)(row - gnx); val[e] = off; ++e; }
      if (ix > 1) { col[e] = (int)(row - 1); val[e] = off; ++e; }
      col[e] = (int)row; val[e] = diag; ++e;
      if (ix < gnx) { col[e] = (int)(row + 1); val[e] = off; ++e; }
      if (iy < gny) { col[e] = (int)(row + gnx); val[e] = off; ++e; }
    }
  rowptr[n] = (int)(e + 1);
  return PARCEL_HIP_OK;
}

// make_poisson_3D_DDM (src/example/ddm_setup.f90:133-349), one rank: the
// right-hand side is make_poisson3D_right_vector's (same formula), placed in
// the padded layout.
int parcel_gen_poisson3d_ddm(int gnx, int gny, int gnz, int margin, double* val_dia, int* ioff_dia,
                             int* ioff_grids, double* vecb, double* vecx_exact) {
  if (gnx < 1 || gny < 1 || gnz < 1 || margin < 1 || !val_dia || !ioff_dia || !ioff_grids || !vecb)
    return PARCEL_HIP_ERR_ARG;
  const long mx = gnx + 2L * margin, my = gny + 2L * margin, mz = gnz + 2L * margin, m = mx * my * mz;
  const long n = (long)gnx * gny * gnz;
  for (long k = 0; k < 7 * m; ++k) val_dia[k] = 0.0;
  for (long k = 0; k < m; ++k) { vecb[k] = 0.0; if (vecx_exact) vecx_exact[k] = 0.0; }
  static const int od[7][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int ip = 0; ip < 7; ++ip) {
    ioff_dia[ip] = (int)(mx * my * od[ip][2] + mx * od[ip][1] + od[ip][0]);
    for (int d = 0; d < 3; ++d) ioff_grids[ip + 7 * d] = od[ip][d];
  }
  std::vector<double> b(n), xe(n);
  int rc = parcel_gen_poisson3d_rhs_parcel(gnx, gny, gnz, b.data(), xe.data());
  if (rc) return rc;
  for (int jz = 0; jz < gnz; ++jz)
    for (int jy = 0; jy < gny; ++jy)
      for (int jx = 0; jx < gnx; ++jx) {
        const long j = (jx + margin) + mx * (jy + margin) + mx * my * (jz + margin);
        const long i = jx + (long)gnx * jy + (long)gnx * gny * jz;
        if (jz > 0) val_dia[j + 0 * m] = 1.0;
        if (jy > 0) val_dia[j + 1 * m] = 1.0;
        if (jx > 0) val_dia[j + 2 * m] = 1.0;
        val_dia[j + 3 * m] = -6.0;
        if (jx < gnx - 1) val_dia[j + 4 * m] = 1.0;
        if (jy < gny - 1) val_dia[j + 5 * m] = 1.0;
        if (jz < gnz - 1) val_dia[j + 6 * m] = 1.0;
        vecb[j] = b[i];
        if (vecx_exact) vecx_exact[j] = xe[i];
      }
  return PARCEL_HIP_OK;
}

int parcel_gen_uniform(long n, unsigned long long seed, double lo, double hi, double* x) {
  if (n < 0 || !x) return PARCEL_HIP_ERR_ARG;
  const double w = hi - lo;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) x[i] = lo + w * u01(seed, (uint64_t)i);
  return PARCEL_HIP_OK;
}

}  // extern "C"
