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

#include "dd.hpp"
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

// 3D 7-point first-order upwind convection-diffusion, finite volumes on
// cubic cells, divided by D*h so that a diffusion face is -1.  The velocity
// direction (vx, vy, vz) is scaled so that the cell Peclet number
// Pe = max|v_d| h / (2 D) equals `peclet`; the convective number of
// direction d is then c_d = 2 Pe v_d / max|v|.  Upwind: for v_d > 0 the
// neighbour at -d gets -(1 + c_d) and the one at +d gets -1 (mirrored for
// v_d < 0); the diagonal is 6 + sum |c_d| on every row (Dirichlet walls:
// couplings to cells outside the grid are dropped, the diagonal is kept).
// Entry order b, s, w, c, e, n, t as make_poisson3D_CRS
// (problem.f90:448-618).
int parcel_gen_convdiff3d(int gnx, int gny, int gnz, double peclet, double vx, double vy, double vz, int* nnz,
                          int* rowptr, int* col, double* val) {
  const double vm = std::fmax(std::fabs(vx), std::fmax(std::fabs(vy), std::fabs(vz)));
  if (!(vm > 0.0) || !(peclet >= 0.0)) return PARCEL_HIP_ERR_ARG;
  int rc = stencil3d(gnx, gny, gnz, -1.0, 6.0, nnz, rowptr, col, val);
  if (rc || !rowptr || !col || !val) return rc;
  const double c[3] = {2.0 * peclet * vx / vm, 2.0 * peclet * vy / vm, 2.0 * peclet * vz / vm};
  // coefficient of the neighbour at -d and +d, and the diagonal
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = c[d] > 0.0 ? -(1.0 + c[d]) : -1.0;
    hi[d] = c[d] < 0.0 ? -(1.0 - c[d]) : -1.0;
  }
  const double diag = 6.0 + ((std::fabs(c[0]) + std::fabs(c[1])) + std::fabs(c[2]));
  const long n = (long)gnx * gny * gnz, pxy = (long)gnx * gny;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < n; ++r) {
    const long row = r + 1;
    for (int e = rowptr[r] - 1; e < rowptr[r + 1] - 1; ++e) {
      const long dc = (long)col[e] - row;
      double v = diag;
      if (dc == -pxy) v = lo[2];
      else if (dc == -gnx) v = lo[1];
      else if (dc == -1) v = lo[0];
      else if (dc == 1) v = hi[0];
      else if (dc == gnx) v = hi[1];
      else if (dc == pxy) v = hi[2];
      val[e] = v;
    }
  }
  return PARCEL_HIP_OK;
}

// kappa_i = exp(ln kmin + (ln kmax - ln kmin) u_i), u_i the SplitMix64 stream
// `seed`: log-uniform in [kmin, kmax).
int parcel_gen_loguniform(long n, unsigned long long seed, double kmin, double kmax, double* kappa) {
  if (n < 0 || !kappa || !(kmin > 0.0) || !(kmax >= kmin)) return PARCEL_HIP_ERR_ARG;
  const double a = std::log(kmin), w = std::log(kmax) - a;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) kappa[i] = std::exp(a + w * u01(seed, (uint64_t)i));
  return PARCEL_HIP_OK;
}

// make_poisson3D_adv_CRS (problem.f90:620-770) for one rank holding the
// whole grid: dx = 1/gnx (likewise y, z), face coefficient
// -dy*dz / (dx*0.5*(1/k_c + 1/k_nb)), diagonal -(cw + ce + cs + cn + cb + ct)
// summed in that order, entries b, s, w, c, e, n, t.  The reference leaves
// its ghost layer ramda(0,..), ramda(gnx+1,..) to the caller; here a ghost
// carries the boundary cell's own conductivity (homogeneous Dirichlet).
int parcel_gen_diffusion3d(int gnx, int gny, int gnz, const double* kappa, int* nnz, int* rowptr, int* col,
                           double* val) {
  if (!kappa && rowptr) return PARCEL_HIP_ERR_ARG;
  int rc = stencil3d(gnx, gny, gnz, -1.0, 6.0, nnz, rowptr, col, val);
  if (rc || !rowptr || !col || !val) return rc;
  const double dx = 1.0 / gnx, dy = 1.0 / gny, dz = 1.0 / gnz;
  const long pxy = (long)gnx * gny;
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < gnz; ++iz)
    for (int iy = 0; iy < gny; ++iy)
      for (int ix = 0; ix < gnx; ++ix) {
        const long r = ix + (long)iy * gnx + (long)iz * pxy;
        const double kc = kappa[r];
        const double kw = ix > 0 ? kappa[r - 1] : kc, ke = ix < gnx - 1 ? kappa[r + 1] : kc;
        const double ks = iy > 0 ? kappa[r - gnx] : kc, kn = iy < gny - 1 ? kappa[r + gnx] : kc;
        const double kb = iz > 0 ? kappa[r - pxy] : kc, kt = iz < gnz - 1 ? kappa[r + pxy] : kc;
        const double cw = -(dy * dz / (dx * 0.5 * (1.0 / kc + 1.0 / kw)));
        const double ce = -(dy * dz / (dx * 0.5 * (1.0 / kc + 1.0 / ke)));
        const double cs = -(dx * dz / (dy * 0.5 * (1.0 / kc + 1.0 / ks)));
        const double cn = -(dx * dz / (dy * 0.5 * (1.0 / kc + 1.0 / kn)));
        const double cb = -(dx * dy / (dz * 0.5 * (1.0 / kc + 1.0 / kb)));
        const double ct = -(dx * dy / (dz * 0.5 * (1.0 / kc + 1.0 / kt)));
        const double cc = -(((((cw + ce) + cs) + cn) + cb) + ct);
        long e = rowptr[r] - 1;
        if (iz > 0) val[e++] = cb;
        if (iy > 0) val[e++] = cs;
        if (ix > 0) val[e++] = cw;
        val[e++] = cc;
        if (ix < gnx - 1) val[e++] = ce;
        if (iy < gny - 1) val[e++] = cn;
        if (iz < gnz - 1) val[e++] = ct;
      }
  return PARCEL_HIP_OK;
}

int parcel_gen_crs_to_dia(int n, const int* rowptr, const int* col, const double* val, int* nnd, int* offset,
                          double* diaA) {
  if (n < 1 || !rowptr || !col || !nnd) return PARCEL_HIP_ERR_ARG;
  // the offsets present, ascending (make_poisson2D_matrix_DIA_get_nnd)
  std::vector<char> seen(2 * (size_t)n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int e = rowptr[i] - 1; e < rowptr[i + 1] - 1; ++e) {
      const long d = (long)(col[e] - 1) - i;
      if (d <= -n || d >= n) return PARCEL_HIP_ERR_ARG;
      seen[(size_t)(d + n)] = 1;
    }
  std::vector<int> offs;
  for (long d = -(long)n; d <= n; ++d)
    if (seen[(size_t)(d + n)]) offs.push_back((int)d);
  *nnd = (int)offs.size();
  if (!offset) return PARCEL_HIP_OK;
  if (!val || !diaA) return PARCEL_HIP_ERR_ARG;
  std::vector<int> slot(2 * (size_t)n + 1, -1);
  for (size_t j = 0; j < offs.size(); ++j) {
    offset[j] = offs[j];
    slot[(size_t)(offs[j] + n)] = (int)j;
  }
  for (long k = 0; k < (long)offs.size() * n; ++k) diaA[k] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int e = rowptr[i] - 1; e < rowptr[i + 1] - 1; ++e) {
      const long d = (long)(col[e] - 1) - i;
      diaA[(long)slot[(size_t)(d + n)] * n + i] = val[e];
    }
  return PARCEL_HIP_OK;
}

// b = A x in double-double: every product val*x exact (Dekker's product of
// two doubles), the row's products added in CRS order with dd_add.
int parcel_gen_spmv_dd(int n, const int* rowptr, const int* col, const double* val, const double* x, double* b_hi,
                       double* b_lo) {
  if (n < 0 || !rowptr || !col || !val || !x || !b_hi || !b_lo) return PARCEL_HIP_ERR_ARG;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < n; ++r) {
    parcel::dd acc = parcel::mkdd(0.0, 0.0);
    for (int e = rowptr[r] - 1; e < rowptr[r + 1] - 1; ++e)
      acc = parcel::dd_add(acc, parcel::dd_mul(parcel::mkdd(val[e], 0.0), parcel::mkdd(x[col[e] - 1], 0.0)));
    b_hi[r] = acc.hi;
    b_lo[r] = acc.lo;
  }
  return PARCEL_HIP_OK;
}

}  // extern "C"
