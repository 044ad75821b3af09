// grf.cpp -- random-coefficient diffusion systems (host code).
//
// The learned-preconditioner workload: a 2D five-point diffusion operator on
// an nx*ny cell grid whose per-cell conductivity is kappa = exp(g), g a
// stationary Gaussian random field with squared-exponential covariance
//     C(dx, dy) = sigma^2 exp(-(dx^2 + dy^2) / (2 ell^2)).
// The reference has no generator for such fields; the matrix assembly
// follows make_poisson3D_adv_CRS (src/example/problem.f90:620-770) reduced to
// two dimensions: harmonic-mean face coefficients, entry order s, w, c, e, n.
//
// Field synthesis (circulant embedding on the nx*ny torus, FP64):
//   g = Re IFFT2(sqrt(lambda) * FFT2(w)),  w ~ N(0, 1) white noise,
// lambda the eigenvalues of the circulant covariance, i.e. the DFT of the
// periodised C.  C is separable, so lambda(kx, ky) = sigma^2 S_nx(kx) S_ny(ky)
// with, by Poisson summation,
//   S_N(k) = sum_i exp(-i^2 / (2 ell^2)) e^{-2 pi i i k / N}
//          = sqrt(2 pi) ell sum_m exp(-2 pi^2 ell^2 (k/N - m)^2),
// which is non-negative (it underflows to 0 only below 1e-300), so no
// eigenvalue is clamped and the modes far below an FFT's rounding level keep
// their true (tiny) weight.  For N comparable to ell the periodised C is the
// sum over all images, not the minimum-image truncation.  Cov(g) equals C to
// exp(-N^2 / (8 ell^2)) (exp(-128) for N = 512, ell = 16).  w_k is the k-th
// normal of seed `seed` (Box-Muller on consecutive SplitMix64 uniforms,
// below), k = ix + nx*iy.
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "parcel_hip.h"

namespace {

typedef std::complex<double> cplx;

inline uint64_t splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline double u01(uint64_t seed, uint64_t i) {
  return (double)(splitmix64(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

// Normal k of stream `seed`: pair p = k/2 takes u1 = u_{2p}, u2 = u_{2p+1};
// radius sqrt(-2 ln(1 - u1)) (1 - u1 is in (0, 1]), angle 2 pi u2; even k
// takes the cosine, odd k the sine.
inline double normal(uint64_t seed, uint64_t k) {
  const uint64_t p = k >> 1;
  const double u1 = u01(seed, 2 * p), u2 = u01(seed, 2 * p + 1);
  const double rad = std::sqrt(-2.0 * std::log(1.0 - u1));
  const double ang = 6.283185307179586 * u2;
  return (k & 1) ? rad * std::sin(ang) : rad * std::cos(ang);
}

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Iterative radix-2 FFT of length n over a strided sequence; sign -1 forward,
// +1 inverse (unnormalised).
struct FFT {
  int n = 0;
  std::vector<cplx> tw;   // exp(-2 pi i k / n), k < n/2
  std::vector<int> rev;
  explicit FFT(int n_) : n(n_), tw(n_ / 2), rev(n_) {
    for (int k = 0; k < n / 2; ++k) {
      const double a = -2.0 * M_PI * (double)k / (double)n;
      tw[k] = cplx(std::cos(a), std::sin(a));
    }
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    for (int i = 0; i < n; ++i) {
      int r = 0;
      for (int b = 0; b < lg; ++b)
        if (i & (1 << b)) r |= 1 << (lg - 1 - b);
      rev[i] = r;
    }
  }
  void run(cplx* a, long stride, int sign, cplx* tmp) const {
    for (int i = 0; i < n; ++i) tmp[rev[i]] = a[i * stride];
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len >> 1, step = n / len;
      for (int s = 0; s < n; s += len)
        for (int k = 0; k < half; ++k) {
          cplx w = tw[k * step];
          if (sign > 0) w = std::conj(w);
          const cplx u = tmp[s + k], v = tmp[s + k + half] * w;
          tmp[s + k] = u + v;
          tmp[s + k + half] = u - v;
        }
    }
    for (int i = 0; i < n; ++i) a[i * stride] = tmp[i];
  }
};

// In-place 2D FFT of an nx*ny array stored x fastest.
void fft2(cplx* a, int nx, int ny, int sign, const FFT& fx, const FFT& fy) {
  std::vector<cplx> tmp(nx > ny ? nx : ny);
  for (int j = 0; j < ny; ++j) fx.run(a + (long)j * nx, 1, sign, tmp.data());
  for (int i = 0; i < nx; ++i) fy.run(a + i, nx, sign, tmp.data());
}

// S_N(k) of the comment above; m runs over the four images nearest k/N.
std::vector<double> spectrum1d(int N, double ell) {
  std::vector<double> S(N);
  const double a = 2.0 * M_PI * M_PI * ell * ell, pre = std::sqrt(2.0 * M_PI) * ell;
  for (int k = 0; k < N; ++k) {
    const double f = (double)k / (double)N;
    double t = 0.0;
    for (int m = -1; m <= 2; ++m) t += std::exp(-a * (f - m) * (f - m));
    S[k] = pre * t;
  }
  return S;
}

void spectrum(int nx, int ny, double ell, double sigma, std::vector<double>& sq) {
  const std::vector<double> sx = spectrum1d(nx, ell), sy = spectrum1d(ny, ell);
  sq.resize((size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) sq[(size_t)j * nx + i] = std::sqrt(sigma * sigma * sx[i] * sy[j]);
}

void field(int nx, int ny, uint64_t seed, const std::vector<double>& sq, const FFT& fx, const FFT& fy,
           double* g) {
  const size_t n = (size_t)nx * ny;
  std::vector<cplx> w(n);
  for (size_t k = 0; k < n; ++k) w[k] = cplx(normal(seed, k), 0.0);
  fft2(w.data(), nx, ny, -1, fx, fy);
  for (size_t k = 0; k < n; ++k) w[k] *= sq[k];
  fft2(w.data(), nx, ny, +1, fx, fy);
  const double s = 1.0 / (double)n;
  for (size_t k = 0; k < n; ++k) g[k] = w[k].real() * s;
}

// Face coefficient between cells of conductivity ka, kb (problem.f90:684-689
// with dx = dy): 1 / (0.5 (1/ka + 1/kb)).
inline double face(double ka, double kb) { return 1.0 / (0.5 * (1.0 / ka + 1.0 / kb)); }

// Five-point rows; the ghost cell beyond the boundary carries the boundary
// cell's own conductivity, so a boundary face contributes kappa to the
// diagonal (homogeneous Dirichlet).  Only values are written when pattern
// arrays are NULL.
void assemble(int nx, int ny, const double* kappa, int* rowptr, int* col, double* val) {
  long e = 0;
  for (int iy = 0; iy < ny; ++iy)
    for (int ix = 0; ix < nx; ++ix) {
      const long r = ix + (long)iy * nx;
      const double kc = kappa[r];
      const double fs = face(kc, iy > 0 ? kappa[r - nx] : kc);
      const double fw = face(kc, ix > 0 ? kappa[r - 1] : kc);
      const double fe = face(kc, ix < nx - 1 ? kappa[r + 1] : kc);
      const double fn = face(kc, iy < ny - 1 ? kappa[r + nx] : kc);
      const double diag = ((fs + fw) + fe) + fn;
      if (rowptr) rowptr[r] = (int)(e + 1);
      if (iy > 0) { if (col) col[e] = (int)(r - nx + 1); val[e++] = -fs; }
      if (ix > 0) { if (col) col[e] = (int)(r - 1 + 1); val[e++] = -fw; }
      if (col) col[e] = (int)(r + 1);
      val[e++] = diag;
      if (ix < nx - 1) { if (col) col[e] = (int)(r + 1 + 1); val[e++] = -fe; }
      if (iy < ny - 1) { if (col) col[e] = (int)(r + nx + 1); val[e++] = -fn; }
    }
  if (rowptr) rowptr[(long)nx * ny] = (int)(e + 1);
}

}  // namespace

extern "C" {

int parcel_gen_normal(long n, unsigned long long seed, double* x) {
  if (n < 0 || !x) return PARCEL_HIP_ERR_ARG;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) x[i] = normal(seed, (uint64_t)i);
  return PARCEL_HIP_OK;
}

int parcel_gen_grf2d(int nx, int ny, double ell, double sigma, unsigned long long seed, double* g) {
  if (!pow2(nx) || !pow2(ny) || !(ell > 0.0) || !(sigma >= 0.0) || !g) return PARCEL_HIP_ERR_ARG;
  FFT fx(nx), fy(ny);
  std::vector<double> sq;
  spectrum(nx, ny, ell, sigma, sq);
  field(nx, ny, seed, sq, fx, fy, g);
  return PARCEL_HIP_OK;
}

int parcel_gen_diffusion2d(int nx, int ny, const double* kappa, int* nnz, int* rowptr, int* col,
                           double* val) {
  if (nx < 1 || ny < 1 || !kappa || !nnz) return PARCEL_HIP_ERR_ARG;
  const long n = (long)nx * ny, cnt = 5 * n - 2L * nx - 2L * ny;
  if (cnt > 0x7fffffffL) return PARCEL_HIP_ERR_ARG;
  *nnz = (int)cnt;
  if (!rowptr || !col || !val) return PARCEL_HIP_OK;
  assemble(nx, ny, kappa, rowptr, col, val);
  return PARCEL_HIP_OK;
}

// nsys systems of seeds seed0, seed0+1, ...: kappa = exp(g) for each, the
// shared pattern once (rowptr/col may be NULL), values of system s at
// vals + s*nnz.  Optional outputs: the fields g (nsys*nx*ny).
int parcel_gen_grf_diffusion2d(int nx, int ny, double ell, double sigma, unsigned long long seed0,
                               int nsys, int* nnz, int* rowptr, int* col, double* vals, double* g_out) {
  if (!pow2(nx) || !pow2(ny) || !(ell > 0.0) || !(sigma >= 0.0) || nsys < 0 || !nnz)
    return PARCEL_HIP_ERR_ARG;
  const long n = (long)nx * ny, cnt = 5 * n - 2L * nx - 2L * ny;
  if (cnt > 0x7fffffffL) return PARCEL_HIP_ERR_ARG;
  *nnz = (int)cnt;
  if (!vals) return PARCEL_HIP_OK;
  FFT fx(nx), fy(ny);
  std::vector<double> sq;
  spectrum(nx, ny, ell, sigma, sq);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int s = 0; s < nsys; ++s) {
    std::vector<double> g(n), kap(n);
    field(nx, ny, seed0 + (unsigned long long)s, sq, fx, fy, g.data());
    for (long k = 0; k < n; ++k) kap[k] = std::exp(g[k]);
    assemble(nx, ny, kap.data(), s == 0 ? rowptr : nullptr, s == 0 ? col : nullptr, vals + (long)s * cnt);
    if (g_out)
      for (long k = 0; k < n; ++k) g_out[(long)s * n + k] = g[k];
    for (long k = 0; k < n; ++k) bad |= !std::isfinite(kap[k]);
  }
  return bad ? PARCEL_HIP_ERR_ARG : PARCEL_HIP_OK;
}

}  // extern "C"
