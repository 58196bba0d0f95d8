// TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the SyN stage with the ANTS neighbourhood cross-correlation
// metric, as antsRegistration drives it through RegistrationHelper
// (Examples/itkantsRegistrationHelper.hxx:1983-2220).  It is the parity oracle
// for the HIP path and the timed CPU baseline ("port"): only tests/, the smoke
// check and bench.py's cpu_baseline leg load it.  The shipped library never
// links or calls it.
//
// The arithmetic of the path lives in ITK (pinned v6 18d89fd6 / v5.4.6, absent
// from the reference tree), so each function below restates the ITK filter the
// reference calls and cites the ANTs call site plus the in-tree restatements
// of the same algorithm.  Every routine is templated on the value type:
// double is the FP64 oracle; float models the reference's `--float 1` runs
// and is the CPU baseline.  Coordinates and continuous indices are double.
//
// Layout: images are dense x-fastest arrays; displacement fields are D-vector
// AoS (x fastest), in physical units (mm), on an identity-direction grid.
//
// Parity status per function is stated in the comment above it.

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
typedef struct {
  int dim;             // 2 or 3
  int64_t n[3];        // size, x fastest; n[2] == 1 when dim == 2
  double spacing[3];
  double origin[3];
  double direction[9]; // row-major dim x dim, orthonormal; all zero = identity
} og_grid;

typedef struct {
  int n_levels;
  int shrink[8];
  double sigma[8];
  int sigma_in_mm;
  int iterations[8];
  double learning_rate;
  double update_variance;
  double total_variance;
  int radius;
  double conv_threshold;
  int conv_window;
  int preprocess;        // window both images to [0,1] as PreprocessImage does
  int force_iterations;  // run every iteration (no convergence stop)
  int threads;           // 0 = OpenMP default
  int inv_tie_rule;      // inversion max test: 0 guarded (max > 0.1 (1 + 1e-5)), 1 strict (max > 0.1 in R)
  int model;             // 0 SyN (Gaussian field smoothing), 1 BSplineSyN (B-spline fits)
  int update_mesh[3];    // BSplineSyN: spans of the update field at the first level (doubled per level)
  int total_mesh[3];     // BSplineSyN: spans of the total field at the first level (0: no fit)
  int spline_order;      // BSplineSyN: 1..3
} og_params;

// Geometry and transforms of a composite: antsApplyTransforms and the SyN
// stage's initial transforms.
struct og_geom {
  int dim;
  int64_t n[3];
  double spacing[3];
  double origin[3];
  double direction[9];  // row-major dim x dim
};
struct og_xform {
  int kind;  // 0 identity, 1 affine, 2 displacement field
  double matrix[9];
  double translation[3];
  double center[3];
  og_geom field;
  const float* f;
};

typedef struct {
  int max_records;
  int n_records;
  int* level;
  int* iteration;
  double* energy;
  double* convergence;
  int* inv_iters;        // 4 per record: F2M inv, F2M fwd, M2M inv, M2M fwd
  double* seconds;       // optional: wall time of each iteration
} og_trace;
}

namespace {

// Image geometry (itk::ImageBase): point = org + dir * (index .* sp),
// index = (dir^T (point - org)) ./ sp.  `oblique` marks a non-identity
// direction; identity grids keep the plain per-axis arithmetic.
struct Grid {
  int D;
  int64_t n[3];
  double sp[3], org[3];
  double dir[9];  // 3x3 row-major, identity-padded
  bool oblique;
  int64_t count() const { return n[0] * n[1] * n[2]; }
};

Grid to_grid(const og_grid* g) {
  Grid r;
  r.D = g->dim;
  for (int d = 0; d < 3; ++d) {
    r.n[d] = d < g->dim ? g->n[d] : 1;
    r.sp[d] = d < g->dim ? g->spacing[d] : 1.0;
    r.org[d] = d < g->dim ? g->origin[d] : 0.0;
  }
  bool all_zero = true;
  for (int k = 0; k < g->dim * g->dim; ++k)
    if (g->direction[k] != 0.0) all_zero = false;
  r.oblique = false;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const double id = a == b ? 1.0 : 0.0;
      const double v = (a < g->dim && b < g->dim && !all_zero) ? g->direction[a * g->dim + b] : id;
      r.dir[a * 3 + b] = v;
      if (v != id) r.oblique = true;
    }
  return r;
}

void set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

// ---------------------------------------------------------------------------
// Modified Bessel functions, Numerical Recipes in C (2nd ed.) sec. 6.6, with
// the polynomial fits of Abramowitz & Stegun 9.8.1-9.8.4.  ITK's
// GaussianOperator builds its coefficients from this method.
// Pinned by tests against scipy.special.ive (relative 1e-6).
double bessel_i0(double x) {
  double ax = std::fabs(x);
  if (ax < 3.75) {
    double y = x / 3.75;
    y *= y;
    return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  double y = 3.75 / ax;
  return (std::exp(ax) / std::sqrt(ax)) *
         (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

double bessel_i1(double x) {
  double ax = std::fabs(x), ans;
  if (ax < 3.75) {
    double y = x / 3.75;
    y *= y;
    ans = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    double y = 3.75 / ax;
    ans = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    ans = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y * (-0.1031555e-1 + y * ans))));
    ans *= std::exp(ax) / std::sqrt(ax);
  }
  return x < 0.0 ? -ans : ans;
}

// Miller's downward recurrence (NR bessi), normalised by I0.
double bessel_in(int n, double x) {
  if (n == 0) return bessel_i0(x);
  if (n == 1) return bessel_i1(x);
  if (x == 0.0) return 0.0;
  const double acc = 40.0, bigno = 1.0e10, bigni = 1.0e-10;
  double tox = 2.0 / std::fabs(x), bip = 0.0, ans = 0.0, bi = 1.0;
  for (int j = 2 * (n + static_cast<int>(std::sqrt(acc * n))); j > 0; --j) {
    double bim = bip + j * tox * bi;
    bip = bi;
    bi = bim;
    if (std::fabs(bi) > bigno) {
      ans *= bigni;
      bi *= bigni;
      bip *= bigni;
    }
    if (j == n) ans = bip;
  }
  ans *= bessel_i0(x) / bi;
  return (x < 0.0 && (n & 1)) ? -ans : ans;
}

// ITK GaussianOperator::GenerateCoefficients: c_i = e^-t I_i(t), grown while
// 1 + 2*sum(c_1..c_i) < 1 - maxError, stopped early when a coefficient is
// negligible or the half-width passes maxWidth; normalised to sum 1 and
// mirrored.  Returns the half kernel c_0..c_R.  Variance t is in voxels
// (SyN passes it unscaled; DiscreteGaussian divides by spacing^2 for mm).
// Pinned: radius/values against scipy.special.ive (variances 3, 4, 1, 0.5).
std::vector<double> gaussian_half_kernel(double variance, double max_error, int64_t max_width) {
  std::vector<double> c;
  const double et = std::exp(-variance);
  const double cap = 1.0 - max_error;
  double sum = 0.0;
  c.push_back(et * bessel_i0(variance));
  sum += c[0];
  c.push_back(et * bessel_i1(variance));
  sum += 2.0 * c[1];
  for (int i = 2; sum < cap; ++i) {
    c.push_back(et * bessel_in(i, variance));
    sum += 2.0 * c[i];
    if (c[i] < sum * std::numeric_limits<double>::epsilon()) break;
    if (static_cast<int64_t>(c.size()) > max_width) break;
  }
  for (double& v : c) v /= sum;
  return c;
}

// ---------------------------------------------------------------------------
// Separable 1-D correlation along axis `ax` with edge replication (ITK's
// ZeroFluxNeumannBoundaryCondition), `ncomp` interleaved components.
template <class R>
void convolve_axis(const Grid& g, int ncomp, const R* in, R* out, const std::vector<double>& half, int ax) {
  const int64_t rad = static_cast<int64_t>(half.size()) - 1;
  const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int64_t stride = ax == 0 ? 1 : (ax == 1 ? nx : nx * ny);
  const int64_t len = g.n[ax];
  std::vector<R> k(2 * rad + 1);
  for (int64_t i = -rad; i <= rad; ++i) k[i + rad] = static_cast<R>(half[std::abs(i)]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        const int64_t idx = x + nx * (y + ny * z);
        const int64_t pos = ax == 0 ? x : (ax == 1 ? y : z);
        const int64_t base = idx - pos * stride;
        for (int c = 0; c < ncomp; ++c) {
          R acc = 0;
          for (int64_t t = -rad; t <= rad; ++t) {
            int64_t p = pos + t;
            p = p < 0 ? 0 : (p >= len ? len - 1 : p);
            acc += k[t + rad] * in[(base + p * stride) * ncomp + c];
          }
          out[idx * ncomp + c] = acc;
        }
      }
}

// DiscreteGaussianImageFilter as ImageRegistrationMethodv4 configures it per
// level (variance sigma^2, MaximumError 0.01, MaximumKernelWidth 32).  sigma
// is in voxels unless in_mm (then variance / spacing^2 per axis).
// Helper call site: Examples/itkantsRegistrationHelper.hxx:1275-1289, 2141-2150.
// Pinned: kernel values (above), impulse response and scipy correlate1d.
template <class R>
void smooth_image(const Grid& g, const R* in, R* out, double sigma, bool in_mm) {
  const int64_t N = g.count();
  std::vector<R> a(in, in + N), b(N);
  if (sigma > 0.0) {
    for (int d = 0; d < g.D; ++d) {
      double var = sigma * sigma;
      if (in_mm) var /= g.sp[d] * g.sp[d];
      std::vector<double> half = gaussian_half_kernel(var, 0.01, 32);
      convolve_axis<R>(g, 1, a.data(), b.data(), half, d);
      std::swap(a, b);
    }
  }
  std::memcpy(out, a.data(), sizeof(R) * N);
}

// SyN GaussianSmoothDisplacementField: per-axis GaussianOperator(variance,
// MaxError 0.001, MaxWidth size[d]), then boundary voxels zeroed and the
// interior blended with the input when variance < 0.5; variance <= 0 is a copy.
// Restatements in tree: Utilities/itkSimulatedExponentialDisplacementFieldSource.hxx:140-238,
// Utilities/itkDiReCTImageFilter.hxx:892-971.
// Pinned: kernel values, impulse response, boundary-zero/blend closed forms.
template <class R>
void smooth_field(const Grid& g, const R* in, R* out, double variance) {
  const int D = g.D;
  const int64_t N = g.count();
  if (variance <= 0.0) {
    std::memcpy(out, in, sizeof(R) * N * D);
    return;
  }
  std::vector<R> a(in, in + N * D), b(N * D);
  for (int d = 0; d < D; ++d) {
    std::vector<double> half = gaussian_half_kernel(variance, 0.001, g.n[d]);
    convolve_axis<R>(g, D, a.data(), b.data(), half, d);
    std::swap(a, b);
  }
  const double w1 = variance < 0.5 ? 1.0 - variance / 0.5 : 1.0;
  const double w2 = 1.0 - w1;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < g.n[2]; ++z)
    for (int64_t y = 0; y < g.n[1]; ++y)
      for (int64_t x = 0; x < g.n[0]; ++x) {
        const int64_t i = x + g.n[0] * (y + g.n[1] * z);
        bool bnd = x == 0 || x == g.n[0] - 1 || y == 0 || y == g.n[1] - 1;
        if (D == 3) bnd = bnd || z == 0 || z == g.n[2] - 1;
        for (int c = 0; c < D; ++c)
          out[i * D + c] = bnd ? R(0) : static_cast<R>(a[i * D + c] * w1 + in[i * D + c] * w2);
      }
}

// ---------------------------------------------------------------------------
// B-spline fit of a displacement field: BSplineSyN's BSplineSmoothDisplacement-
// Field, i.e. ITK's DisplacementFieldToBSplineImageFilter over
// BSplineScatteredDataPointSetToImageFilter (ITK v5.4.6, absent from the
// reference tree) with one fitting level, stationary boundary and no
// confidence image, as the helper configures the transform
// (Examples/itkantsRegistrationHelper.hxx:2222-2340: mesh size << level,
// spline order) and as the reference's own callers of the filter set it up
// (Utilities/itkDiReCTImageFilter.hxx:975-1000,
// Examples/SmoothDisplacementField.cxx:134-165).
//  * every voxel is a data point at its parametric coordinate
//    p_d = i_d * spans_d / (n_d - 1) (the B-spline domain is the field's own);
//    p_d = spans_d is moved to spans_d - eps, eps = 100 FLT_EPSILON, times 10
//    while float(S) == float(S) - eps for S the largest span count;
//  * voxels on the image boundary carry a zero vector with weight 1e10, the
//    others their vector with weight 1;
//  * one level of multilevel B-spline approximation with confidences (Lee,
//    Wolberg & Shin, IEEE TVCG 3(3) 1997; Tustison & Gee, Insight Journal
//    2005): for point c with basis products B_kc over its (q+1)^D control
//    points, delta_k += w_c B_kc^2 (z_c B_kc / sum_l B_lc^2) and
//    omega_k += w_c B_kc^2; phi_k = delta_k / omega_k (0 where omega_k = 0);
//  * the output is the B-spline with control points phi evaluated at every
//    voxel; the basis is ITK's centred BSplineKernelFunction of order q at
//    t - j + (q - 1) / 2, j = 0..q.
// A mesh of 0 spans means no fit (the total field of BSplineSyN[., ., 0]).
// Pinned: an independent MBA built from scipy's B-spline basis on dense
// matrices, tensor-product separability, mirror symmetry, linearity.
double bsp_kernel(int q, double u) {
  u = std::fabs(u);
  switch (q) {
    case 1:
      return u < 1.0 ? 1.0 - u : 0.0;
    case 2:
      return u < 0.5 ? 0.75 - u * u : (u < 1.5 ? 0.5 * (1.5 - u) * (1.5 - u) : 0.0);
    default:
      return u < 1.0 ? (4.0 - 6.0 * u * u + 3.0 * u * u * u) / 6.0 : (u < 2.0 ? (2.0 - u) * (2.0 - u) * (2.0 - u) / 6.0 : 0.0);
  }
}

double bsp_epsilon(int max_spans) {
  volatile float eps = 100.0f * FLT_EPSILON;
  const volatile float S = static_cast<float>(max_spans);
  while (static_cast<float>(S) == static_cast<float>(S - eps)) eps = eps * 10.0f;
  return static_cast<double>(eps);
}

// span index and basis values of voxel i of an axis of n voxels and `spans` spans
void bsp_point(int64_t i, int64_t n, int spans, int q, double eps, int* s, double b[4]) {
  double p = n > 1 ? static_cast<double>(i) * spans / static_cast<double>(n - 1) : 0.0;
  if (std::fabs(p - spans) <= eps) p = spans - eps;
  const int si = static_cast<int>(std::floor(p));
  const double t = p - si;
  for (int j = 0; j < 4; ++j) b[j] = j <= q ? bsp_kernel(q, t - j + 0.5 * (q - 1)) : 0.0;
  *s = si;
}

template <class R>
void bspline_fit_field(const Grid& g, const R* in, R* out, const int mesh[3], int q) {
  const int D = g.D;
  const int64_t N = g.count();
  bool fit = true;
  for (int d = 0; d < D; ++d) fit = fit && mesh[d] > 0;
  if (!fit) {
    std::memcpy(out, in, sizeof(R) * N * D);
    return;
  }
  int maxs = 0;
  for (int d = 0; d < D; ++d) maxs = std::max(maxs, mesh[d]);
  const double eps = bsp_epsilon(maxs);
  int64_t m[3] = {1, 1, 1};
  for (int d = 0; d < D; ++d) m[d] = mesh[d] + q;
  const int64_t NC = m[0] * m[1] * m[2];
  std::vector<double> delta(NC * D, 0.0), omega(NC, 0.0);
  const int taps[3] = {q + 1, D >= 2 ? q + 1 : 1, D == 3 ? q + 1 : 1};
  const double kBoundaryWeight = 1.0e10;
  for (int64_t z = 0; z < g.n[2]; ++z)
    for (int64_t y = 0; y < g.n[1]; ++y)
      for (int64_t x = 0; x < g.n[0]; ++x) {
        const int64_t c[3] = {x, y, z};
        int s[3] = {0, 0, 0};
        double b[3][4] = {{1, 0, 0, 0}, {1, 0, 0, 0}, {1, 0, 0, 0}};
        bool bnd = false;
        for (int d = 0; d < D; ++d) {
          bsp_point(c[d], g.n[d], mesh[d], q, eps, &s[d], b[d]);
          bnd = bnd || c[d] == 0 || c[d] == g.n[d] - 1;
        }
        const int64_t i = x + g.n[0] * (y + g.n[1] * z);
        const double w = bnd ? kBoundaryWeight : 1.0;
        double w2 = 0.0;
        for (int kz = 0; kz < taps[2]; ++kz)
          for (int ky = 0; ky < taps[1]; ++ky)
            for (int kx = 0; kx < taps[0]; ++kx) {
              const double B = b[0][kx] * b[1][ky] * b[2][kz];
              w2 += B * B;
            }
        for (int kz = 0; kz < taps[2]; ++kz)
          for (int ky = 0; ky < taps[1]; ++ky)
            for (int kx = 0; kx < taps[0]; ++kx) {
              const double B = b[0][kx] * b[1][ky] * b[2][kz];
              const int64_t k = (s[0] + kx) + m[0] * ((s[1] + ky) + m[1] * (s[2] + kz));
              omega[k] += w * B * B;
              for (int cc = 0; cc < D; ++cc) {
                const double zc = bnd ? 0.0 : static_cast<double>(in[i * D + cc]);
                delta[k * D + cc] += w * B * B * (zc * B / w2);
              }
            }
      }
  std::vector<double> phi(NC * D, 0.0);
  for (int64_t k = 0; k < NC; ++k)
    if (omega[k] != 0.0)
      for (int cc = 0; cc < D; ++cc) phi[k * D + cc] = delta[k * D + cc] / omega[k];
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < g.n[2]; ++z)
    for (int64_t y = 0; y < g.n[1]; ++y)
      for (int64_t x = 0; x < g.n[0]; ++x) {
        const int64_t c[3] = {x, y, z};
        int s[3] = {0, 0, 0};
        double b[3][4] = {{1, 0, 0, 0}, {1, 0, 0, 0}, {1, 0, 0, 0}};
        for (int d = 0; d < D; ++d) bsp_point(c[d], g.n[d], mesh[d], q, eps, &s[d], b[d]);
        double v[3] = {0.0, 0.0, 0.0};
        for (int kz = 0; kz < taps[2]; ++kz)
          for (int ky = 0; ky < taps[1]; ++ky)
            for (int kx = 0; kx < taps[0]; ++kx) {
              const double B = b[0][kx] * b[1][ky] * b[2][kz];
              const int64_t k = (s[0] + kx) + m[0] * ((s[1] + ky) + m[1] * (s[2] + kz));
              for (int cc = 0; cc < D; ++cc) v[cc] += B * phi[k * D + cc];
            }
        const int64_t i = x + g.n[0] * (y + g.n[1] * z);
        for (int cc = 0; cc < D; ++cc) out[i * D + cc] = static_cast<R>(v[cc]);
      }
}

// ---------------------------------------------------------------------------
// Buffer test and linear interpolation with ITK semantics (ImageFunction::
// IsInsideBuffer: -0.5 <= cidx < n-0.5 per axis; Linear/VectorLinear
// interpolate functions clamp neighbours to the buffer, which inside the
// buffer equals edge-replicated trilinear interpolation).
// Pinned: scipy.ndimage.map_coordinates(order=1, mode='nearest') on inside points.
inline bool inside(const Grid& g, const double c[3]) {
  for (int d = 0; d < g.D; ++d)
    if (!(c[d] >= -0.5 && c[d] < static_cast<double>(g.n[d]) - 0.5)) return false;
  return true;
}

inline void to_cidx(const Grid& g, const double p[3], double c[3]) {
  if (g.oblique) {
    double q[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < g.D; ++k) q[k] = p[k] - g.org[k];
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int k = 0; k < g.D; ++k) s += g.dir[k * 3 + d] * q[k];
      c[d] = d < g.D ? s / g.sp[d] : 0.0;
    }
    return;
  }
  for (int d = 0; d < 3; ++d) c[d] = d < g.D ? (p[d] - g.org[d]) / g.sp[d] : 0.0;
}

template <class R>
void interp(const Grid& g, int ncomp, const R* img, const double c[3], double* outv) {
  int64_t lo[3], hi[3];
  double w[3];
  for (int d = 0; d < 3; ++d) {
    if (d >= g.D) {
      lo[d] = hi[d] = 0;
      w[d] = 0.0;
      continue;
    }
    double f = std::floor(c[d]);
    int64_t b = static_cast<int64_t>(f);
    double t = c[d] - f;
    if (b < 0) {  // c in [-0.5, 0): clamped to the first voxel
      b = 0;
      t = 0.0;
    }
    lo[d] = b;
    hi[d] = b + 1 < g.n[d] ? b + 1 : b;
    w[d] = t;
  }
  for (int k = 0; k < ncomp; ++k) outv[k] = 0.0;
  for (int corner = 0; corner < 8; ++corner) {
    double wt = 1.0;
    int64_t ix[3];
    for (int d = 0; d < 3; ++d) {
      bool up = (corner >> d) & 1;
      if (d >= g.D && up) {
        wt = 0.0;
        break;
      }
      wt *= up ? w[d] : 1.0 - w[d];
      ix[d] = up ? hi[d] : lo[d];
    }
    if (wt == 0.0) continue;
    const int64_t idx = ix[0] + g.n[0] * (ix[1] + g.n[1] * ix[2]);
    for (int k = 0; k < ncomp; ++k) outv[k] += wt * static_cast<double>(img[idx * ncomp + k]);
  }
}

inline void grid_point(const Grid& g, int64_t x, int64_t y, int64_t z, double p[3]) {
  if (g.oblique) {
    const double v[3] = {x * g.sp[0], y * g.sp[1], g.D == 3 ? z * g.sp[2] : 0.0};
    for (int r = 0; r < 3; ++r) {
      double s = g.org[r];
      for (int k = 0; k < g.D; ++k) s += g.dir[r * 3 + k] * v[k];
      p[r] = r < g.D ? s : 0.0;
    }
    return;
  }
  p[0] = g.org[0] + x * g.sp[0];
  p[1] = g.org[1] + y * g.sp[1];
  p[2] = g.D == 3 ? g.org[2] + z * g.sp[2] : 0.0;
}

// ResampleImageFilter(image on gi, transform x -> x + u(x), output grid go,
// linear interpolator, default 0).  u is sampled at the output voxel itself
// (the fields live on the output grid), as in SyN ComputeMetricGradientField
// with DownsampleImagesForMetricDerivatives on.  u == nullptr: identity.
// Helper call sites: Examples/itkantsRegistrationHelper.hxx:719-763; in-tree
// mirror of the SyN resampling: Examples/antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate.h:342-381.
void xform_point(const og_xform& T, int D, double p[3]);

// With initial transforms xf[0..nt) (SyNImageRegistrationMethod's
// fixed/moving composite, Examples/itkantsRegistrationHelper.hxx:2119-2137):
// the displaced point then passes through xf[0], xf[1], ... before the image
// is sampled, as the composite [initial, middle^-1] applies the middle
// field's inverse first.
template <class R>
void warp_image(const Grid& gi, const R* img, const Grid& go, const R* u, R* out, const og_xform* xf = nullptr,
                int nt = 0) {
  const int D = go.D;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < go.n[2]; ++z)
    for (int64_t y = 0; y < go.n[1]; ++y)
      for (int64_t x = 0; x < go.n[0]; ++x) {
        const int64_t i = x + go.n[0] * (y + go.n[1] * z);
        double p[3], c[3];
        grid_point(go, x, y, z, p);
        if (u)
          for (int d = 0; d < D; ++d) p[d] += static_cast<double>(u[i * D + d]);
        for (int t = 0; t < nt; ++t) xform_point(xf[t], D, p);
        to_cidx(gi, p, c);
        double v = 0.0;
        if (inside(gi, c)) interp<R>(gi, 1, img, c, &v);
        out[i] = static_cast<R>(v);
      }
}

// ComposeDisplacementFieldsImageFilter(displacement d, warping w):
// out(x) = w(x) + d(x + w(x)), d linearly interpolated, 0 outside its buffer.
// Output on w's grid.  Restatement: Utilities/itkComposeDiffeomorphismsImageFilter.hxx:76-104.
// Pinned: constant fields compose additively; affine fields exactly.
template <class R>
void compose(const Grid& g, const R* disp, const R* warp, R* out) {
  const int D = g.D;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < g.n[2]; ++z)
    for (int64_t y = 0; y < g.n[1]; ++y)
      for (int64_t x = 0; x < g.n[0]; ++x) {
        const int64_t i = x + g.n[0] * (y + g.n[1] * z);
        double p[3], c[3], v[3] = {0, 0, 0};
        grid_point(g, x, y, z, p);
        for (int d = 0; d < D; ++d) p[d] += static_cast<double>(warp[i * D + d]);
        to_cidx(g, p, c);
        if (inside(g, c)) interp<R>(g, D, disp, c, v);
        for (int d = 0; d < D; ++d) out[i * D + d] = static_cast<R>(static_cast<double>(warp[i * D + d]) + v[d]);
      }
}

// DisplacementFieldTransformParametersAdaptor: resample a field onto a new
// grid with the identity transform, vector linear interpolation, 0 outside.
// SyN level set-up, Examples/itkantsRegistrationHelper.hxx:2033-2055.
template <class R>
void resample_field(const Grid& gi, const R* in, const Grid& go, R* out) {
  const int D = go.D;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < go.n[2]; ++z)
    for (int64_t y = 0; y < go.n[1]; ++y)
      for (int64_t x = 0; x < go.n[0]; ++x) {
        const int64_t i = x + go.n[0] * (y + go.n[1] * z);
        double p[3], c[3], v[3] = {0, 0, 0};
        grid_point(go, x, y, z, p);
        to_cidx(gi, p, c);
        if (inside(gi, c)) interp<R>(gi, D, in, c, v);
        for (int d = 0; d < D; ++d) out[i * D + d] = static_cast<R>(v[d]);
      }
}

// ---------------------------------------------------------------------------
// ANTSNeighborhoodCorrelationImageToImageMetricv4::GetValueAndDerivative with
// identity transforms on images already resampled to the level grid.
// Box sums over the in-image part of the (2r+1)^D window of f, m, f^2, m^2,
// f*m and the count; sFF, sMM, sFM and the derivative
//   2 sFM/(sFF sMM) (f~ - sFM/sMM m~) grad m
// as in the in-tree restatement ImageRegistration/itkCrossCorrelationRegistrationFunction.cxx:178-306, 441-503.
// A window with sFF or sMM not above eps keeps localCC = 1 and a zero
// derivative.  Value = -sum(localCC)/N.  grad m is the
// CentralDifferenceImageFunction rule (half-voxel linear samples; a component
// is 0 where p+h/2 leaves the buffer, a half difference at index 0).
// Both SyN calls are evaluated from one set of box sums: call 1 (fixed=F,
// moving=M) gives the fixed-to-middle update `g_f2m`, call 2 (fixed=M,
// moving=F) the moving-to-middle update `g_m2m`.
// Pinned: brute-force window loops, CC=1 for M=aF+b, degenerate windows, and
// a finite-difference check of d localCC / d m_centre.
template <class R>
inline R grad_component(const R* img, int64_t i, int64_t pos, int64_t len, int64_t stride, double h) {
  if (pos == len - 1) return R(0);
  const double up = 0.5 * (static_cast<double>(img[i]) + static_cast<double>(img[i + stride]));
  const double dn = pos == 0 ? static_cast<double>(img[i]) : 0.5 * (static_cast<double>(img[i - stride]) + static_cast<double>(img[i]));
  return static_cast<R>((up - dn) / h);
}

// gext (optional): the moving image's gradient at each voxel's mapped point,
// already physical (ImageToImageMetricv4 with a field transform: the
// CentralDifferenceImageFunction of the moving image at the mapped point);
// then g_f2m = c1 * gext and g_m2m is not written.
template <class R>
double cc_metric(const Grid& g, const R* F, const R* M, int radius, double eps, R* g_f2m, R* g_m2m,
                 const R* gext = nullptr) {
  const int D = g.D;
  const int64_t N = g.count();
  const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
  // 6 quantities: 1, f, m, ff, mm, fm.  Window sums and the per-voxel
  // formula in A: double for the FP64 oracle (the GPU's sums are double
  // too); float in the FP32 model, as ITK's metric computes in its float
  // InternalComputationValueType under --float 1 (with separable box sums
  // here, not ITK's sliding queue, so the rounding differs)
  using A = typename std::conditional<std::is_same<R, float>::value, float, double>::type;
  std::vector<A> q(N * 6), t(N * 6);
  for (int64_t i = 0; i < N; ++i) {
    const A f = F[i], m = M[i];
    A* o = &q[i * 6];
    o[0] = 1;
    o[1] = f;
    o[2] = m;
    o[3] = f * f;
    o[4] = m * m;
    o[5] = f * m;
  }
  for (int ax = 0; ax < D; ++ax) {
    const int64_t stride = ax == 0 ? 1 : (ax == 1 ? nx : nx * ny);
    const int64_t len = g.n[ax];
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 0; z < nz; ++z)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
          const int64_t idx = x + nx * (y + ny * z);
          const int64_t pos = ax == 0 ? x : (ax == 1 ? y : z);
          const int64_t lo = std::max<int64_t>(0, pos - radius), hi = std::min<int64_t>(len - 1, pos + radius);
          A acc[6] = {0, 0, 0, 0, 0, 0};
          for (int64_t p = lo; p <= hi; ++p) {
            const A* s = &q[(idx + (p - pos) * stride) * 6];
            for (int k = 0; k < 6; ++k) acc[k] += s[k];
          }
          for (int k = 0; k < 6; ++k) t[idx * 6 + k] = acc[k];
        }
    std::swap(q, t);
  }
  // per-slab partial sums, summed in slab order: deterministic for any thread count
  std::vector<double> part(nz, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < nz; ++z) {
    double local = 0.0;
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        const int64_t i = x + nx * (y + ny * z);
        const A* s = &q[i * 6];
        const A cnt = s[0], sf = s[1], sm = s[2], sff = s[3], smm = s[4], sfm = s[5];
        const A mf = sf / cnt, mm = sm / cnt;
        const A sFF = sff - mf * sf - mf * sf + cnt * mf * mf;
        const A sMM = smm - mm * sm - mm * sm + cnt * mm * mm;
        const A sFM = sfm - mm * sf - mf * sm + cnt * mm * mf;
        const A fA = F[i] - mf, mA = M[i] - mm;
        A cc = 1;
        R gm[3] = {0, 0, 0}, gf[3] = {0, 0, 0};
        if (sFF > eps && sMM > eps) {
          cc = sFM * sFM / (sFF * sMM);
          const int64_t pos[3] = {x, y, z};
          for (int d = 0; d < D; ++d) {
            const int64_t stride = d == 0 ? 1 : (d == 1 ? nx : nx * ny);
            gm[d] = grad_component<R>(M, i, pos[d], g.n[d], stride, g.sp[d]);
            gf[d] = grad_component<R>(F, i, pos[d], g.n[d], stride, g.sp[d]);
          }
          const A a = A(2) * sFM / (sFF * sMM);
          const R c1 = static_cast<R>(a * (fA - sFM / sMM * mA));
          const R c2 = static_cast<R>(a * (mA - sFM / sFF * fA));
          if (gext)
            for (int d = 0; d < D; ++d) gm[d] = gext[i * D + d];
          for (int d = 0; d < D; ++d) {
            gm[d] *= c1;
            gf[d] *= c2;
          }
          // CentralDifferenceImageFunction with UseImageDirection: the
          // index-axis derivative is turned into a physical vector
          if (g.oblique && !gext) {
            R rm[3] = {0, 0, 0}, rf[3] = {0, 0, 0};
            for (int r = 0; r < D; ++r)
              for (int k = 0; k < D; ++k) {
                rm[r] += static_cast<R>(g.dir[r * 3 + k]) * gm[k];
                rf[r] += static_cast<R>(g.dir[r * 3 + k]) * gf[k];
              }
            for (int d = 0; d < D; ++d) {
              gm[d] = rm[d];
              gf[d] = rf[d];
            }
          }
        }
        for (int d = 0; d < D; ++d) {
          if (g_f2m) g_f2m[i * D + d] = gm[d];
          if (g_m2m && !gext) g_m2m[i * D + d] = gf[d];
        }
        local += static_cast<double>(cc);
      }
    part[z] = local;
  }
  double total = 0.0;
  for (int64_t z = 0; z < nz; ++z) total += part[z];
  return -total / static_cast<double>(N);
}

// SyN ScaleUpdateField: u *= learningRate / max_i |u_i / spacing| (when > 0).
template <class R>
double scale_update(const Grid& g, R* u, double lr) {
  const int D = g.D;
  const int64_t N = g.count();
  double mx = 0.0;
#pragma omp parallel for reduction(max : mx) schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
      double v = static_cast<double>(u[i * D + d]) / g.sp[d];
      s += v * v;
    }
    s = std::sqrt(s);
    if (s > mx) mx = s;
  }
  double scale = lr;
  if (mx > 0.0) scale /= mx;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N * D; ++i) u[i] = static_cast<R>(static_cast<double>(u[i]) * scale);
  return mx;
}

// InvertDisplacementFieldImageFilter as SyN configures it (MaximumNumberOfIterations 20,
// MeanErrorToleranceThreshold 0.001, MaxErrorToleranceThreshold 0.1,
// EnforceBoundaryCondition on): inv starts at the initial estimate; each step
// e = compose(field, inv), norms |e/spacing| (mean over all voxels, max),
// eps = 0.75 on the first step then 0.5, e clamped to eps*max, inv += -eps*e,
// boundary voxels of inv zeroed.  The loop runs while step < 20 and the last
// max > 0.1 and mean > 0.001.  Restatement with the same constants:
// ImageRegistration/itkANTSImageRegistrationOptimizer.h:1845-1988.
// Pinned: a constant shift inverts exactly in the interior; the residual at
// exit respects the tolerances; constants cross-checked with the restatement.
//
// Tie rule (`strict` false, the default): the max test is
// `max > 0.1 * (1 + 1e-5)`.  When the initial inverse is zero the first
// residual is the scaled update itself, whose max norm equals the learning
// rate, so SyN[0.1,...] lands exactly on the 0.1 tolerance and the
// reference's decision is set by rounding; the guard makes that tie resolve
// the same way in every precision.  `strict` follows ITK's test as written:
// the max norm rounded to R, compared with 0.1 in R (`> 0.1f` for --float 1).
constexpr double kInvMaxTol = 0.1 * (1.0 + 1e-5);
template <class R>
inline bool inv_max_above(double maxn, bool strict) {
  return strict ? static_cast<R>(maxn) > static_cast<R>(0.1) : maxn > kInvMaxTol;
}
template <class R>
int invert_field(const Grid& g, const R* field, const R* init, R* inv, double* out_max, double* out_mean,
                 bool strict = false) {
  const int D = g.D;
  const int64_t N = g.count();
  const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
  if (init)
    std::memcpy(inv, init, sizeof(R) * N * D);
  else
    std::fill(inv, inv + N * D, R(0));
  std::vector<R> e(N * D), norm(N);
  double maxn = std::numeric_limits<double>::max(), meann = std::numeric_limits<double>::max();
  int it = 0;
  while (it++ < 20 && inv_max_above<R>(maxn, strict) && meann > 0.001) {
    compose<R>(g, field, inv, e.data());
    std::vector<double> psum(nz, 0.0), pmax(nz, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < nz; ++z) {
      double s = 0.0, m = 0.0;
      for (int64_t j = 0; j < nx * ny; ++j) {
        const int64_t i = z * nx * ny + j;
        double q = 0.0;
        for (int d = 0; d < D; ++d) {
          double v = static_cast<double>(e[i * D + d]) / g.sp[d];
          q += v * v;
        }
        q = std::sqrt(q);
        norm[i] = static_cast<R>(q);
        s += q;
        if (q > m) m = q;
        for (int d = 0; d < D; ++d) e[i * D + d] = -e[i * D + d];
      }
      psum[z] = s;
      pmax[z] = m;
    }
    meann = 0.0;
    maxn = 0.0;
    for (int64_t z = 0; z < nz; ++z) {
      meann += psum[z];
      maxn = std::max(maxn, pmax[z]);
    }
    meann /= static_cast<double>(N);
    const double eps = it == 1 ? 0.75 : 0.5;
    const double lim = eps * maxn;
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < nz; ++z)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
          const int64_t i = x + nx * (y + ny * z);
          double sc = 1.0;
          const double q = static_cast<double>(norm[i]);
          if (q > lim) sc = lim / q;
          bool bnd = x == 0 || x == nx - 1 || y == 0 || y == ny - 1;
          if (D == 3) bnd = bnd || z == 0 || z == nz - 1;
          for (int d = 0; d < D; ++d)
            inv[i * D + d] = bnd ? R(0) : static_cast<R>(static_cast<double>(inv[i * D + d]) + static_cast<double>(e[i * D + d]) * sc * eps);
        }
  }
  if (out_max) *out_max = maxn;
  if (out_mean) *out_mean = meann;
  return it - 1;
}

// WindowConvergenceMonitoringFunction (window W): energies pushed into a
// window of the last W values, a running total of |E| over the whole level;
// fewer than W values -> max(double); otherwise an order-1 B-spline with 2
// control points, one level, is fitted by the multilevel B-spline
// approximation (BSplineScatteredDataPointSetToImageFilter) to
// (t_n = n/(W-1), E_n/total) and the convergence value is -dphi/dt at t=1.
// Described at Examples/antsRegistration.cxx:397-404.
// Parity unpinned beyond invariants: a constant window gives 0, a falling
// energy a positive value (the exact normalisation lives in ITK only).
struct ConvMonitor {
  int window;
  std::vector<double> vals;
  double total = 0.0;
  explicit ConvMonitor(int w) : window(w) {}
  void add(double e) {
    vals.push_back(e);
    if (static_cast<int>(vals.size()) > window) vals.erase(vals.begin());
    total += std::fabs(e);
  }
  double value() const {
    if (static_cast<int>(vals.size()) < window || window < 2) return std::numeric_limits<double>::max();
    double num[2] = {0, 0}, den[2] = {0, 0};
    for (int n = 0; n < window; ++n) {
      const double t = static_cast<double>(n) / static_cast<double>(window - 1);
      const double v = vals[n] / total;
      const double w[2] = {1.0 - t, t};
      const double w2 = w[0] * w[0] + w[1] * w[1];
      for (int k = 0; k < 2; ++k) {
        num[k] += w[k] * w[k] * (w[k] * v / w2);
        den[k] += w[k] * w[k];
      }
    }
    const double p0 = den[0] > 0 ? num[0] / den[0] : 0.0;
    const double p1 = den[1] > 0 ? num[1] / den[1] : 0.0;
    return -(p1 - p0);
  }
};

// RegistrationHelper::CalculateShrinkFactorsPerDimension (Examples/itkantsRegistrationHelper.hxx:593-641)
// followed by ShrinkImageFilter output information (hxx:16-39): size floor(n/s)
// (at least 1), spacing*s, origin moved so the physical centres coincide.
Grid shrink_grid(const Grid& g, int factor) {
  int minIndex = 0;
  double minSp = g.sp[0];
  for (int d = 1; d < g.D; ++d)
    if (minSp > g.sp[d]) {
      minSp = g.sp[d];
      minIndex = d;
    }
  int f[3] = {1, 1, 1};
  f[minIndex] = factor;
  const double target = g.sp[minIndex] * factor;
  for (int d = 0; d < g.D; ++d) {
    if (d == minIndex) continue;
    double best = std::fabs(g.sp[d] * factor - target);
    int bf = factor;
    for (int k = factor - 1; k > 0; --k) {
      double diff = std::fabs(g.sp[d] * k - target);
      if (diff <= best) {
        best = diff;
        bf = k;
      }
    }
    f[d] = bf;
  }
  Grid o = g;
  double delta[3] = {0.0, 0.0, 0.0};  // index-aligned offset of the output origin
  for (int d = 0; d < g.D; ++d) {
    o.sp[d] = g.sp[d] * f[d];
    o.n[d] = std::max<int64_t>(1, static_cast<int64_t>(std::floor(static_cast<double>(g.n[d]) / f[d])));
    const double in_c = (g.n[d] - 1) / 2.0, out_c = (o.n[d] - 1) / 2.0;
    o.org[d] = g.org[d] + (g.org[d] + in_c * g.sp[d]) - (g.org[d] + out_c * o.sp[d]);
    delta[d] = in_c * g.sp[d] - out_c * o.sp[d];
  }
  // with a direction the centres coincide in physical space: origin + dir * delta
  if (g.oblique)
    for (int r = 0; r < g.D; ++r) {
      double s = g.org[r];
      for (int k = 0; k < g.D; ++k) s += g.dir[r * 3 + k] * delta[k];
      o.org[r] = s;
    }
  return o;
}

// PreprocessImage (Examples/itkantsRegistrationHelper.hxx:74-136) with the
// CLI defaults (no winsorizing, no histogram matching): lo = min,
// hi = max + (max-min)/2560 (256-bin histogram, marginal scale 10), both
// stored as float; IntensityWindowing to [0,1] with clamping.
template <class R>
void preprocess(int64_t N, const R* in, R* out) {
  double mn = std::numeric_limits<double>::max(), mx = -std::numeric_limits<double>::max();
  for (int64_t i = 0; i < N; ++i) {
    mn = std::min(mn, static_cast<double>(in[i]));
    mx = std::max(mx, static_cast<double>(in[i]));
  }
  const float lo = static_cast<float>(mn);
  const float hi = static_cast<float>(mx + (mx - mn) / 2560.0);
  const double scale = (hi > lo) ? 1.0 / (static_cast<double>(hi) - static_cast<double>(lo)) : 0.0;
  const double shift = -static_cast<double>(lo) * scale;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const double x = static_cast<double>(in[i]);
    double v;
    if (x < lo)
      v = 0.0;
    else if (x > hi)
      v = 1.0;
    else
      v = x * scale + shift;
    out[i] = static_cast<R>(v);
  }
}

// The SyN stage: SyNImageRegistrationMethod level loop (InitializeRegistrationAtEachLevel,
// StartOptimization, final composition) as configured by
// Examples/itkantsRegistrationHelper.hxx:1983-2220.  Per iteration (both
// updates from the same state): warp F and M through the inverse middle
// fields, CC value and both derivatives, smooth at the update variance,
// scale to the learning rate, compose onto the forward middle fields,
// smooth at the total variance, invert (inverse from the previous inverse,
// then forward as the inverse of the new inverse), E = mean of both values,
// convergence monitor.  Final: fwd = M2M.inv o F2M.fwd, inv = F2M.inv o M2M.fwd
// (same as Examples/itkantsRegistrationHelper.hxx:3698-3711).
template <class R>
int syn_register(const og_params* P, const Grid& gF, const R* fixed_in, const Grid& gM, const R* moving_in, R* out_fwd,
                 R* out_inv, og_trace* trace, R* state4, double eps, const og_xform* fx = nullptr, int nfx = 0,
                 const og_xform* mx = nullptr, int nmx = 0, const Grid* gS = nullptr, const R* state_in = nullptr,
                 R* lvl_fwd = nullptr, R* lvl_inv = nullptr) {
  set_threads(P->threads);
  const bool strict = P->inv_tie_rule != 0;
  // per-level snapshots: the four middle fields at the end of each level
  // before the last, with their grid
  std::vector<std::vector<R>> snap(lvl_fwd || lvl_inv ? 4 * P->n_levels : 0);
  std::vector<Grid> snap_grid(P->n_levels);
  const int D = gF.D;
  const int64_t NF = gF.count(), NM = gM.count();
  std::vector<R> F(fixed_in, fixed_in + NF), M(moving_in, moving_in + NM);
  if (P->preprocess) {
    preprocess<R>(NF, fixed_in, F.data());
    preprocess<R>(NM, moving_in, M.data());
  }
  std::vector<R> Fs(NF), Ms(NM);
  std::vector<R> f2m_f, f2m_i, m2m_f, m2m_i;
  Grid gPrev = gF;
  if (trace) trace->n_records = 0;
  for (int l = 0; l < P->n_levels; ++l) {
    const Grid gl = shrink_grid(gF, P->shrink[l]);
    const int64_t N = gl.count();
    if (l == 0 && state_in) {
      // restored state (FixedToMiddle fwd/inv, MovingToMiddle fwd/inv on grid
      // gS): the level-0 parameter adaptor resamples all four fields onto the
      // level grid, as it does between levels
      const int64_t NS = gS->count();
      std::vector<R>* dst[4] = {&f2m_f, &f2m_i, &m2m_f, &m2m_i};
      for (int k = 0; k < 4; ++k) {
        dst[k]->assign(N * D, R(0));
        resample_field<R>(*gS, state_in + k * NS * D, gl, dst[k]->data());
      }
    } else if (l == 0) {
      f2m_f.assign(N * D, R(0));
      f2m_i.assign(N * D, R(0));
      m2m_f.assign(N * D, R(0));
      m2m_i.assign(N * D, R(0));
    } else {
      std::vector<R> tmp(N * D);
      resample_field<R>(gPrev, f2m_f.data(), gl, tmp.data());
      f2m_f = tmp;
      resample_field<R>(gPrev, f2m_i.data(), gl, tmp.data());
      f2m_i = tmp;
      resample_field<R>(gPrev, m2m_f.data(), gl, tmp.data());
      m2m_f = tmp;
      resample_field<R>(gPrev, m2m_i.data(), gl, tmp.data());
      m2m_i = tmp;
    }
    gPrev = gl;
    snap_grid[l] = gl;
    if (P->iterations[l] <= 0) {
      if (!snap.empty()) {
        snap[4 * l + 0] = f2m_f;
        snap[4 * l + 1] = f2m_i;
        snap[4 * l + 2] = m2m_f;
        snap[4 * l + 3] = m2m_i;
      }
      continue;
    }
    smooth_image<R>(gF, F.data(), Fs.data(), P->sigma[l], P->sigma_in_mm != 0);
    smooth_image<R>(gM, M.data(), Ms.data(), P->sigma[l], P->sigma_in_mm != 0);
    std::vector<R> Fw(N), Mw(N), g1(N * D), g2(N * D), u1(N * D), u2(N * D), t1(N * D), t2(N * D), s1(N * D), s2(N * D);
    ConvMonitor mon(P->conv_window);
    bool converged = false;
    int iter = 0;
    // BSplineSyN: mesh sizes of this level (the helper's << level)
    const bool bsp = P->model == 1;
    int um[3] = {0, 0, 0}, tm[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
      um[d] = P->update_mesh[d] << l;
      tm[d] = P->total_mesh[d] << l;
    }
    while (iter++ < P->iterations[l] && !converged) {
      const auto t_iter = std::chrono::steady_clock::now();
      warp_image<R>(gF, Fs.data(), gl, f2m_i.data(), Fw.data(), fx, nfx);
      warp_image<R>(gM, Ms.data(), gl, m2m_i.data(), Mw.data(), mx, nmx);
      const double value = cc_metric<R>(gl, Fw.data(), Mw.data(), P->radius, eps, g1.data(), g2.data());
      if (bsp) {
        bspline_fit_field<R>(gl, g1.data(), u1.data(), um, P->spline_order);
        bspline_fit_field<R>(gl, g2.data(), u2.data(), um, P->spline_order);
      } else {
        smooth_field<R>(gl, g1.data(), u1.data(), P->update_variance);
        smooth_field<R>(gl, g2.data(), u2.data(), P->update_variance);
      }
      scale_update<R>(gl, u1.data(), P->learning_rate);
      scale_update<R>(gl, u2.data(), P->learning_rate);
      compose<R>(gl, u1.data(), f2m_f.data(), t1.data());
      compose<R>(gl, u2.data(), m2m_f.data(), t2.data());
      if (bsp) {
        bspline_fit_field<R>(gl, t1.data(), s1.data(), tm, P->spline_order);
        bspline_fit_field<R>(gl, t2.data(), s2.data(), tm, P->spline_order);
      } else {
        smooth_field<R>(gl, t1.data(), s1.data(), P->total_variance);
        smooth_field<R>(gl, t2.data(), s2.data(), P->total_variance);
      }
      int k[4];
      std::vector<R> ninv(N * D), nfwd(N * D);
      k[0] = invert_field<R>(gl, s1.data(), f2m_i.data(), ninv.data(), nullptr, nullptr, strict);
      k[1] = invert_field<R>(gl, ninv.data(), s1.data(), nfwd.data(), nullptr, nullptr, strict);
      f2m_i.swap(ninv);
      f2m_f.swap(nfwd);
      k[2] = invert_field<R>(gl, s2.data(), m2m_i.data(), ninv.data(), nullptr, nullptr, strict);
      k[3] = invert_field<R>(gl, ninv.data(), s2.data(), nfwd.data(), nullptr, nullptr, strict);
      m2m_i.swap(ninv);
      m2m_f.swap(nfwd);
      const double E = 0.5 * (value + value);
      mon.add(E);
      const double conv = mon.value();
      if (!P->force_iterations && conv < P->conv_threshold) converged = true;
      if (trace && trace->n_records < trace->max_records) {
        const int r = trace->n_records++;
        trace->level[r] = l;
        trace->iteration[r] = iter;
        trace->energy[r] = E;
        trace->convergence[r] = conv;
        for (int j = 0; j < 4; ++j) trace->inv_iters[r * 4 + j] = k[j];
        if (trace->seconds)
          trace->seconds[r] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_iter).count();
      }
    }
    if (!snap.empty()) {
      snap[4 * l + 0] = f2m_f;
      snap[4 * l + 1] = f2m_i;
      snap[4 * l + 2] = m2m_f;
      snap[4 * l + 3] = m2m_i;
    }
  }
  // final composition on the last level grid (the full-resolution grid when the last shrink is 1)
  const int64_t N = gPrev.count();
  // Per-level results: the state at the end of level l carried through the
  // later levels' grids by the parameter adaptor (as a schedule that runs no
  // iterations after level l does) and composed on the last grid.
  for (int l = 0; l < static_cast<int>(snap.size() / 4); ++l) {
    std::vector<R> f[4];
    for (int k = 0; k < 4; ++k) f[k] = snap[4 * l + k];
    for (int m = l + 1; m < P->n_levels; ++m) {
      const int64_t Nm = snap_grid[m].count();
      for (int k = 0; k < 4; ++k) {
        std::vector<R> tmp(Nm * D);
        resample_field<R>(snap_grid[m - 1], f[k].data(), snap_grid[m], tmp.data());
        f[k].swap(tmp);
      }
    }
    if (lvl_fwd) compose<R>(gPrev, f[3].data(), f[0].data(), lvl_fwd + l * N * D);
    if (lvl_inv) compose<R>(gPrev, f[1].data(), f[2].data(), lvl_inv + l * N * D);
  }
  if (out_fwd) compose<R>(gPrev, m2m_i.data(), f2m_f.data(), out_fwd);
  if (out_inv) compose<R>(gPrev, f2m_i.data(), m2m_f.data(), out_inv);
  if (state4) {
    std::memcpy(state4 + 0 * N * D, f2m_f.data(), sizeof(R) * N * D);
    std::memcpy(state4 + 1 * N * D, f2m_i.data(), sizeof(R) * N * D);
    std::memcpy(state4 + 2 * N * D, m2m_f.data(), sizeof(R) * N * D);
    std::memcpy(state4 + 3 * N * D, m2m_i.data(), sizeof(R) * N * D);
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// antsApplyTransforms subset (Examples/antsApplyTransforms.cxx:546-640):
// itk::ResampleImageFilter over the reference grid, with a composite of
// MatrixOffsetTransformBase (y = A (x - c) + t + c) and
// DisplacementFieldTransform (y = x + u(x), u by VectorLinearInterpolate-
// ImageFunction, zero outside its buffer) applied in command-line order
// (itkantsRegistrationHelper.h:1262-1455 adds them from the last -t option
// to the first; itk::CompositeTransform applies the last added first).
// Interpolation: LinearInterpolateImageFunction, NearestNeighbor (half-up
// rounding) or GenericLabel[Linear] (Schaerer et al., "A generic interpolator
// for multi-label images", Insight Journal 2014: the label whose binary
// indicator image interpolates highest; labels iterate in ascending order
// with a strict comparison, so ties go to the smaller label).  Points whose
// continuous index lies outside [-0.5, n-0.5) get the default value.

namespace {

// point = origin + sum_c (direction[r][c] * spacing[c]) * index[c]
void geom_index_to_point(const og_geom& g, const double idx[3], double p[3]) {
  const int D = g.dim;
  for (int r = 0; r < 3; ++r) {
    if (r >= D) {
      p[r] = 0.0;
      continue;
    }
    double s = g.origin[r];
    for (int c = 0; c < D; ++c) {
      const double m = g.direction[r * D + c] * g.spacing[c];
      s = s + m * idx[c];
    }
    p[r] = s;
  }
}

// index = diag(1/spacing) direction^T (point - origin)
void geom_point_to_index(const og_geom& g, const double p[3], double c[3]) {
  const int D = g.dim;
  double q[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < D; ++k) q[k] = p[k] - g.origin[k];
  for (int r = 0; r < 3; ++r) {
    if (r >= D) {
      c[r] = 0.0;
      continue;
    }
    double s = 0.0;
    for (int k = 0; k < D; ++k) {
      const double mi = g.direction[k * D + r] / g.spacing[r];
      s = s + mi * q[k];
    }
    c[r] = s;
  }
}

bool geom_inside(const og_geom& g, const double c[3]) {
  for (int d = 0; d < g.dim; ++d)
    if (!(c[d] >= -0.5 && c[d] < static_cast<double>(g.n[d]) - 0.5)) return false;
  return true;
}

// trilinear (bilinear) cell: corner flat indices and weights, upper neighbour
// clamped to the buffer; corner k has bit 0 = x, bit 1 = y, bit 2 = z
void geom_cell(const og_geom& g, const double c[3], int64_t idx[8], double w[3]) {
  const int D = g.dim;
  int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    w[d] = 0.0;
    if (d >= D) continue;
    const double fl = std::floor(c[d]);
    int64_t b = static_cast<int64_t>(fl);
    double t = c[d] - fl;
    if (b < 0) {
      b = 0;
      t = 0.0;
    }
    lo[d] = b;
    hi[d] = b + 1 < g.n[d] ? b + 1 : b;
    w[d] = t;
  }
  for (int k = 0; k < (1 << D); ++k) {
    const int64_t x = (k & 1) ? hi[0] : lo[0], y = (k & 2) ? hi[1] : lo[1], z = (k & 4) ? hi[2] : lo[2];
    idx[k] = x + g.n[0] * (y + (D == 3 ? g.n[1] * z : 0));
  }
}

// interpolate corner values: along x, then y, then z
double cell_lerp(int D, const double v[8], const double w[3]) {
  const double a00 = v[0] + (v[1] - v[0]) * w[0];
  const double a10 = v[2] + (v[3] - v[2]) * w[0];
  const double b0 = a00 + (a10 - a00) * w[1];
  if (D == 2) return b0;
  const double a01 = v[4] + (v[5] - v[4]) * w[0];
  const double a11 = v[6] + (v[7] - v[6]) * w[0];
  const double b1 = a01 + (a11 - a01) * w[1];
  return b0 + (b1 - b0) * w[2];
}

void xform_point(const og_xform& T, int D, double p[3]) {
  if (T.kind == 1) {
    double q[3] = {0.0, 0.0, 0.0};
    for (int r = 0; r < D; ++r) {
      double ac = 0.0;
      for (int k = 0; k < D; ++k) ac = ac + T.matrix[r * D + k] * T.center[k];
      const double offset = T.translation[r] + T.center[r] - ac;
      double s = 0.0;
      for (int k = 0; k < D; ++k) s = s + T.matrix[r * D + k] * p[k];
      q[r] = s + offset;
    }
    for (int r = 0; r < D; ++r) p[r] = q[r];
  } else if (T.kind == 2) {
    double c[3];
    geom_point_to_index(T.field, p, c);
    if (!geom_inside(T.field, c)) return;
    int64_t idx[8];
    double w[3];
    geom_cell(T.field, c, idx, w);
    for (int d = 0; d < D; ++d) {
      double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int k = 0; k < (1 << D); ++k) v[k] = static_cast<double>(T.f[idx[k] * D + d]);
      p[d] = p[d] + cell_lerp(D, v, w);
    }
  }
}

// BSplineDecompositionImageFilter (ITK, absent from the reference tree; used
// by BSplineInterpolateImageFunction, which antsApplyTransforms -n BSpline[q]
// builds at Examples/make_interpolator_snip.tmpl:30-40): along each axis in
// turn, every line is multiplied by the gain prod_k (1 - z_k)(1 - 1/z_k) of
// the order's poles, then per pole filtered causally, c+[k] = c[k] +
// z c+[k - 1], and anticausally, c-[k] = z (c-[k + 1] - c+[k]), with the
// mirror-boundary initial values of Unser, Aldroubi & Eden (IEEE TSP 41(2),
// 1993) as restated in Unser (IEEE SPM 16(6), 1999): the causal start is the
// geometric sum over the mirrored line, truncated after ceil(ln 1e-10 /
// ln|z|) terms when that is shorter than the line, and the anticausal start
// c-[N-1] = z / (z^2 - 1) (z c+[N-2] + c+[N-1]).  Lines of length 1 are left
// as they are.  Parity: pinned by test_oracle_apply.py against scipy's
// mirror-mode spline_filter / map_coordinates (same recursion, independent
// code), and by the interpolation property (the spline passes through the
// samples).
void bspline_poles(int q, double z[2], int* np) {
  *np = 0;
  if (q == 2) {
    z[0] = std::sqrt(8.0) - 3.0;
    *np = 1;
  } else if (q == 3) {
    z[0] = std::sqrt(3.0) - 2.0;
    *np = 1;
  } else if (q == 4) {
    z[0] = std::sqrt(664.0 - std::sqrt(438976.0)) + std::sqrt(304.0) - 19.0;
    z[1] = std::sqrt(664.0 + std::sqrt(438976.0)) - std::sqrt(304.0) - 19.0;
    *np = 2;
  } else if (q == 5) {
    z[0] = std::sqrt(67.5 - std::sqrt(4436.25)) + std::sqrt(26.25) - 6.5;
    z[1] = std::sqrt(67.5 + std::sqrt(4436.25)) - std::sqrt(26.25) - 6.5;
    *np = 2;
  }
}

void bspline_line(std::vector<double>& c, const double* z, int np) {
  const int64_t N = static_cast<int64_t>(c.size());
  if (N < 2 || np == 0) return;
  double gain = 1.0;
  for (int k = 0; k < np; ++k) gain *= (1.0 - z[k]) * (1.0 - 1.0 / z[k]);
  for (double& v : c) v *= gain;
  for (int k = 0; k < np; ++k) {
    const double zk = z[k];
    const int64_t horizon = static_cast<int64_t>(std::ceil(std::log(1e-10) / std::log(std::fabs(zk))));
    double c0 = 0.0;
    if (horizon < N) {
      for (int64_t n = 0; n < horizon; ++n) c0 += std::pow(zk, static_cast<double>(n)) * c[n];
    } else {
      // full mirrored sum: sum_{n=0}^{2N-3} z^n c[mirror(n)] / (1 - z^(2N-2))
      double acc = 0.0;
      for (int64_t n = 0; n <= 2 * N - 3; ++n) {
        const int64_t m = n < N ? n : 2 * (N - 1) - n;
        acc += std::pow(zk, static_cast<double>(n)) * c[m];
      }
      c0 = acc / (1.0 - std::pow(zk, static_cast<double>(2 * N - 2)));
    }
    c[0] = c0;
    for (int64_t n = 1; n < N; ++n) c[n] += zk * c[n - 1];
    c[N - 1] = zk / (zk * zk - 1.0) * (zk * c[N - 2] + c[N - 1]);
    for (int64_t n = N - 2; n >= 0; --n) c[n] = zk * (c[n + 1] - c[n]);
  }
}

void bspline_coefficients(const og_geom& g, const float* img, int q, std::vector<double>& coef) {
  const int D = g.dim;
  const int64_t n[3] = {g.n[0], g.n[1], D == 3 ? g.n[2] : 1};
  const int64_t N = n[0] * n[1] * n[2];
  coef.assign(img, img + N);
  double z[2];
  int np;
  bspline_poles(q, z, &np);
  for (int d = 0; d < D; ++d) {
    const int64_t len = n[d];
    const int64_t stride = d == 0 ? 1 : (d == 1 ? n[0] : n[0] * n[1]);
    std::vector<double> line(len);
    for (int64_t b = 0; b < N / len; ++b) {
      // the line through the voxel whose axis-d index is 0 and whose other
      // indices enumerate b
      const int64_t lo = b % stride, hi = b / stride;
      const int64_t base = hi * stride * len + lo;
      for (int64_t k = 0; k < len; ++k) line[k] = coef[base + k * stride];
      bspline_line(line, z, np);
      for (int64_t k = 0; k < len; ++k) coef[base + k * stride] = line[k];
    }
  }
}

// Centred B-spline of degree q at t by its truncated-power closed form,
// beta_q(t) = 1/q! sum_{k=0}^{q+1} (-1)^k C(q+1, k) (t + (q+1)/2 - k)_+^q,
// with degree 0 the indicator of [-1/2, 1/2).
double bspline_beta(int q, double t) {
  if (q == 0) return (t >= -0.5 && t < 0.5) ? 1.0 : 0.0;
  double s = 0.0, binom = 1.0, fact = 1.0;
  for (int k = 2; k <= q; ++k) fact *= k;
  for (int k = 0; k <= q + 1; ++k) {
    const double u = t + 0.5 * (q + 1) - k;
    if (u > 0.0) s += ((k & 1) ? -binom : binom) * std::pow(u, q);
    binom = binom * (q + 1 - k) / (k + 1);
  }
  return s / fact;
}

// BSplineInterpolateImageFunction::Evaluate at continuous index c: the
// q + 1 coefficients per axis from floor(c) - q/2 (odd q) or floor(c + 1/2)
// - q/2 (even q), mirror-reflected into the buffer (period 2(n - 1)),
// weighted by beta_q(c - j).
double bspline_evaluate(const og_geom& g, const std::vector<double>& coef, int q, const double c[3]) {
  const int D = g.dim;
  int64_t idx[3][6] = {};
  double w[3][6] = {};
  for (int d = 0; d < 3; ++d) {
    if (d >= D) {
      w[d][0] = 1.0;
      continue;
    }
    const int64_t len = g.n[d];
    const int64_t start = static_cast<int64_t>(std::floor((q % 2) ? c[d] : c[d] + 0.5)) - q / 2;
    for (int k = 0; k <= q; ++k) {
      int64_t j = start + k;
      w[d][k] = bspline_beta(q, c[d] - static_cast<double>(j));
      if (len == 1) {
        j = 0;
      } else {
        while (j < 0 || j >= len) j = j < 0 ? -j : 2 * (len - 1) - j;
      }
      idx[d][k] = j;
    }
  }
  double acc = 0.0;
  const int qz = D == 3 ? q : 0;
  for (int kz = 0; kz <= qz; ++kz)
    for (int ky = 0; ky <= q; ++ky)
      for (int kx = 0; kx <= q; ++kx)
        acc += w[2][kz] * w[1][ky] * w[0][kx] * coef[idx[0][kx] + g.n[0] * (idx[1][ky] + g.n[1] * idx[2][kz])];
  return acc;
}

// WindowedSincInterpolateImageFunction<.., 3, window, ConstantBoundaryCondition>
// (ITK, absent from the reference tree; selected at
// Examples/make_interpolator_snip.tmpl:77-110): per axis the six samples
// floor(c) - 2 .. floor(c) + 3 with weights window(x) sinc(x) at their
// offsets x = c - j (a delta on floor(c) when c is whole), zero outside the
// buffer; window 0 cosine cos(pi x / 6), 1 Hamming 0.54 + 0.46 cos(pi x / 3),
// 2 Lanczos sinc(x / 3), 3 Blackman 0.42 + 0.5 cos(pi x / 3) +
// 0.08 cos(2 pi x / 3), 4 Welch 1 - x^2 / 9 (Meijering et al. 1999).
// Parity: pinned by test_oracle_apply.py against a numpy evaluation of
// the definition at scattered points near the edges (np.sinc), exact
// samples at whole indices and the reconstruction of a band-limited
// signal; the window constants and the offset convention are ITK's as
// documented, not read from ITK's source (unpinned against ITK itself).
double sinc_window(int w, double x) {
  const double pi = 3.14159265358979323846;
  const double sinc = x == 0.0 ? 1.0 : std::sin(pi * x) / (pi * x);
  double win = 1.0;
  switch (w) {
    case 0: win = std::cos(pi * x / 6.0); break;
    case 1: win = 0.54 + 0.46 * std::cos(pi * x / 3.0); break;
    case 2: win = x == 0.0 ? 1.0 : std::sin(pi * x / 3.0) / (pi * x / 3.0); break;
    case 3: win = 0.42 + 0.5 * std::cos(pi * x / 3.0) + 0.08 * std::cos(2.0 * pi * x / 3.0); break;
    default: win = 1.0 - x * x / 9.0; break;
  }
  return win * sinc;
}

double sinc_evaluate(const og_geom& g, const float* img, int w, const double c[3]) {
  const int D = g.dim;
  double wt[3][6];
  int64_t base[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    for (int k = 0; k < 6; ++k) wt[d][k] = (d >= D && k == 0) ? 1.0 : 0.0;
    if (d >= D) continue;
    const double fl = std::floor(c[d]);
    base[d] = static_cast<int64_t>(fl) - 2;
    const double dist = c[d] - fl;
    for (int k = 0; k < 6; ++k) {
      const double off = dist + 2.0 - k;  // c - (floor(c) - 2 + k)
      wt[d][k] = dist == 0.0 ? (k == 2 ? 1.0 : 0.0) : sinc_window(w, off);
    }
  }
  double acc = 0.0;
  const int KZ = D == 3 ? 6 : 1;
  for (int kz = 0; kz < KZ; ++kz)
    for (int ky = 0; ky < 6; ++ky)
      for (int kx = 0; kx < 6; ++kx) {
        const int64_t x = base[0] + kx, y = base[1] + ky, z = base[2] + kz;
        const bool in = x >= 0 && x < g.n[0] && y >= 0 && y < g.n[1] && (D == 2 || (z >= 0 && z < g.n[2]));
        double v = in ? static_cast<double>(img[x + g.n[0] * (y + g.n[1] * z)]) : 0.0;
        v *= wt[0][kx];
        v *= wt[1][ky];
        if (D == 3) v *= wt[2][kz];
        acc += v;
      }
  return acc;
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points (ctypes).  `_f64` is the FP64 oracle, `_f32` the float model.
extern "C" {

// interp: 0 linear, 1 nearest neighbour, 2 generic label, 16 + q B-spline
// of order q (0..5), 32 + w windowed sinc (w: cosine, Hamming, Lanczos,
// Blackman, Welch)
int og_apply_transforms_f64(const og_geom* ref, const og_geom* in, const float* img, const og_xform* xf, int nt,
                            int interp, double defv, float* out) {
  const int D = ref->dim;
  if (in->dim != D) return -1;
  const int64_t N = ref->n[0] * ref->n[1] * (D == 3 ? ref->n[2] : 1);
  const int q = interp - 16;
  std::vector<double> coef;
  if (q >= 0 && q <= 5) bspline_coefficients(*in, img, q, coef);
  else if ((interp < 0 || interp > 2) && (interp < 32 || interp > 36)) return -2;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int64_t x = i % ref->n[0], y = (i / ref->n[0]) % ref->n[1], z = i / (ref->n[0] * ref->n[1]);
    const double id[3] = {static_cast<double>(x), static_cast<double>(y), static_cast<double>(z)};
    double p[3];
    geom_index_to_point(*ref, id, p);
    for (int t = 0; t < nt; ++t) xform_point(xf[t], D, p);
    double c[3];
    geom_point_to_index(*in, p, c);
    if (!geom_inside(*in, c)) {
      out[i] = static_cast<float>(defv);
      continue;
    }
    if (interp >= 32) {
      out[i] = static_cast<float>(sinc_evaluate(*in, img, interp - 32, c));
      continue;
    }
    if (interp >= 16) {
      out[i] = static_cast<float>(bspline_evaluate(*in, coef, q, c));
      continue;
    }
    if (interp == 1) {
      int64_t r[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) r[d] = static_cast<int64_t>(std::floor(c[d] + 0.5));
      out[i] = img[r[0] + in->n[0] * (r[1] + in->n[1] * r[2])];
      continue;
    }
    int64_t idx[8];
    double w[3];
    geom_cell(*in, c, idx, w);
    float lab[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < (1 << D); ++k) lab[k] = img[idx[k]];
    if (interp == 0) {
      double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int k = 0; k < (1 << D); ++k) v[k] = static_cast<double>(lab[k]);
      out[i] = static_cast<float>(cell_lerp(D, v, w));
      continue;
    }
    // generic label: labels in ascending order, strict comparison
    std::vector<float> labels(lab, lab + (1 << D));
    std::sort(labels.begin(), labels.end());
    float best = labels[0];
    double bestv = -1.0;
    for (float l : labels) {
      double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int k = 0; k < (1 << D); ++k) v[k] = lab[k] == l ? 1.0 : 0.0;
      const double val = cell_lerp(D, v, w);
      if (val > bestv) {
        bestv = val;
        best = l;
      }
    }
    out[i] = best;
  }
  return 0;
}

// CreateJacobianDeterminantImage (Examples/CreateJacobianDeterminantImage.cxx:60-160):
// DeformationFieldGradientTensorImageFilter (Utilities/itkDeformationField
// GradientTensorImageFilter.hxx:113-185: the five-point difference
// (-u(x+2) + 8u(x+1) - 8u(x-1) + u(x-2)) / 12 per axis over the zero-flux
// Neumann neighbourhood, weight 1/spacing, rows rotated by the direction,
// + I), DeterminantTensorImageFilter, then max(det, 0) or
// log(max(det, 0.001)).  `field`: dim components per voxel; grad (nullable)
// receives the matrices row-major.  Parity: pinned by test_oracle_apply.py
// (linear and cubic fields on oblique grids against their analytic
// Jacobians, the clamped-stencil edge value by hand, the clamps).
void og_jacobian_f64(const og_geom* g, const float* field, int do_log, double* out, double* grad) {
  const int D = g->dim;
  const int64_t n[3] = {g->n[0], g->n[1], D == 3 ? g->n[2] : 1};
  const int64_t N = n[0] * n[1] * n[2];
  bool all_zero = true;
  for (int k = 0; k < D * D; ++k)
    if (g->direction[k] != 0.0) all_zero = false;
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < N; ++v) {
    const int64_t ix[3] = {v % n[0], (v / n[0]) % n[1], v / (n[0] * n[1])};
    double F[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int j = 0; j < D; ++j) {
      auto at = [&](int64_t step, int c) {
        int64_t q[3] = {ix[0], ix[1], ix[2]};
        q[j] = std::min(std::max(q[j] + step, int64_t{0}), n[j] - 1);
        return static_cast<double>(field[((q[2] * n[1] + q[1]) * n[0] + q[0]) * D + c]);
      };
      for (int c = 0; c < D; ++c)
        F[c][j] = (1.0 / g->spacing[j]) * (-at(2, c) + 8.0 * at(1, c) - 8.0 * at(-1, c) + at(-2, c)) / 12.0;
    }
    double G[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c) {
        double acc = 0.0;
        for (int k = 0; k < D; ++k) {
          const double Rck = all_zero ? (c == k ? 1.0 : 0.0) : g->direction[c * D + k];
          acc += Rck * F[r][k];
        }
        G[r][c] = acc + (r == c ? 1.0 : 0.0);
      }
    if (grad)
      for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) grad[v * D * D + r * D + c] = G[r][c];
    // determinant by Gaussian elimination with partial pivoting
    double M[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) M[r][c] = G[r][c];
    double det = 1.0;
    for (int c = 0; c < D; ++c) {
      int piv = c;
      for (int r = c + 1; r < D; ++r)
        if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
      if (M[piv][c] == 0.0) {
        det = 0.0;
        break;
      }
      if (piv != c) {
        for (int k = 0; k < 3; ++k) std::swap(M[c][k], M[piv][k]);
        det = -det;
      }
      det *= M[c][c];
      for (int r = c + 1; r < D; ++r) {
        const double f = M[r][c] / M[c][c];
        for (int k = c; k < D; ++k) M[r][k] -= f * M[c][k];
      }
    }
    out[v] = do_log ? std::log(std::max(std::max(det, 0.0), 0.001)) : std::max(det, 0.0);
  }
}

// GeometricJacobianDeterminantImageFilter
// (Utilities/itkGeometricJacobianDeterminantImageFilter.hxx:88-335): at each
// voxel's physical point p the simplex p + M o_k (M = direction * spacing;
// o_k the triangle (0, r3), (-1/2, -r3), (1/2, -r3), r3 = sqrt(3)/4, or the
// tetrahedron (+-1/2, 0, -1/(2 sqrt 2)), (0, +-1/2, 1/(2 sqrt 2))), each
// vertex displaced by the field at it (VectorLinearInterpolateImageFunction:
// corners clamped to the buffer), |moved volume| / |unmoved volume|;
// log(max(., 0.001)) with do_log.  Parity: pinned by test_oracle_apply.py
// (zero and uniform fields give 1, a linear field |det A| away from the
// faces, agreement with the gradient-tensor form on a smooth field).
void og_jacobian_geometric_f64(const og_geom* g, const float* field, int do_log, double* out) {
  const int D = g->dim;
  const int64_t n[3] = {g->n[0], g->n[1], D == 3 ? g->n[2] : 1};
  const int64_t N = n[0] * n[1] * n[2];
  const double r3 = 0.25 * std::sqrt(3.0), t2 = 0.5 / std::sqrt(2.0);
  const double tri[3][3] = {{0.0, r3, 0.0}, {-0.5, -r3, 0.0}, {0.5, -r3, 0.0}};
  const double tet[4][3] = {{0.5, 0.0, -t2}, {-0.5, 0.0, -t2}, {0.0, 0.5, t2}, {0.0, -0.5, t2}};
  auto volume = [&](const double P[4][3]) {
    if (D == 2) {
      const double ab0 = P[0][0] - P[1][0], ab1 = P[0][1] - P[1][1];
      const double ac0 = P[0][0] - P[2][0], ac1 = P[0][1] - P[2][1];
      return 0.5 * std::fabs(ab0 * ac1 - ac0 * ab1);
    }
    double a[3], b[3], c[3];
    for (int r = 0; r < 3; ++r) {
      a[r] = P[0][r] - P[3][r];
      b[r] = P[1][r] - P[3][r];
      c[r] = P[2][r] - P[3][r];
    }
    const double x = b[1] * c[2] - b[2] * c[1], y = b[2] * c[0] - b[0] * c[2], z = b[0] * c[1] - b[1] * c[0];
    return std::fabs(a[0] * x + a[1] * y + a[2] * z) / 6.0;
  };
  // the unmoved simplex at the origin index
  double base[4][3] = {};
  {
    const double zero[3] = {0.0, 0.0, 0.0};
    double p0[3];
    geom_index_to_point(*g, zero, p0);
    for (int v = 0; v <= D; ++v) {
      double pv[3];
      geom_index_to_point(*g, D == 2 ? tri[v] : tet[v], pv);
      for (int r = 0; r < 3; ++r) base[v][r] = pv[r] - p0[r];
    }
  }
  const double undisplaced = volume(base);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const double id[3] = {static_cast<double>(i % n[0]), static_cast<double>((i / n[0]) % n[1]),
                          static_cast<double>(i / (n[0] * n[1]))};
    double p[3];
    geom_index_to_point(*g, id, p);
    double P[4][3] = {};
    for (int v = 0; v <= D; ++v) {
      double q[3] = {0.0, 0.0, 0.0};
      for (int r = 0; r < D; ++r) q[r] = p[r] + base[v][r];
      double c[3];
      geom_point_to_index(*g, q, c);
      double disp[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < (1 << D); ++k) {
        double w = 1.0;
        int64_t j[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) {
          const double fl = std::floor(c[d]);
          const double t = c[d] - fl;
          const int64_t b = static_cast<int64_t>(fl) + ((k >> d) & 1);
          j[d] = std::min(std::max(b, int64_t{0}), n[d] - 1);
          w *= ((k >> d) & 1) ? t : 1.0 - t;
        }
        const int64_t lin = (j[2] * n[1] + j[1]) * n[0] + j[0];
        for (int r = 0; r < D; ++r) disp[r] += w * static_cast<double>(field[lin * D + r]);
      }
      for (int r = 0; r < D; ++r) P[v][r] = q[r] + disp[r];
    }
    const double ratio = volume(P) / undisplaced;
    out[i] = do_log ? std::log(std::max(ratio, 0.001)) : ratio;
  }
}

// CompositeTransform::TransformPoint of n points (dim doubles each), as
// antsApplyTransformsToPoints calls it (Examples/antsApplyTransformsToPoints.cxx:166-180).
void og_transform_points_f64(int dim, const double* in, int64_t n, const og_xform* xf, int nt, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double p[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d) p[d] = in[i * dim + d];
    for (int t = 0; t < nt; ++t) xform_point(xf[t], dim, p);
    for (int d = 0; d < dim; ++d) out[i * dim + d] = p[d];
  }
}

// TransformToDisplacementFieldFilter (ITK; called at
// Examples/antsApplyTransforms.cxx:716-724): T(p) - p at every point p of
// `ref`, T the transform list applied in order; dim interleaved components.
int og_compose_field_f64(const og_geom* ref, const og_xform* xf, int nt, float* out) {
  const int D = ref->dim;
  const int64_t N = ref->n[0] * ref->n[1] * (D == 3 ? ref->n[2] : 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int64_t x = i % ref->n[0], y = (i / ref->n[0]) % ref->n[1], z = i / (ref->n[0] * ref->n[1]);
    const double id[3] = {static_cast<double>(x), static_cast<double>(y), static_cast<double>(z)};
    double p0[3], p[3];
    geom_index_to_point(*ref, id, p0);
    for (int d = 0; d < 3; ++d) p[d] = p0[d];
    for (int t = 0; t < nt; ++t) xform_point(xf[t], D, p);
    for (int d = 0; d < D; ++d) out[i * D + d] = static_cast<float>(p[d] - p0[d]);
  }
  return 0;
}

int og_gaussian_kernel(double variance, double max_error, int64_t max_width, double* out, int max_out) {
  std::vector<double> h = gaussian_half_kernel(variance, max_error, max_width);
  const int n = static_cast<int>(h.size());
  for (int i = 0; i < n && i < max_out; ++i) out[i] = h[i];
  return n - 1;  // radius
}

double og_bessel_in(int n, double x) { return bessel_in(n, x); }

void og_shrink_grid(const og_grid* g, int factor, og_grid* out) {
  Grid r = shrink_grid(to_grid(g), factor);
  out->dim = r.D;
  for (int d = 0; d < 3; ++d) {
    out->n[d] = r.n[d];
    out->spacing[d] = r.sp[d];
    out->origin[d] = r.org[d];
  }
  for (int k = 0; k < 9; ++k) out->direction[k] = 0.0;
  for (int a = 0; a < r.D; ++a)
    for (int b = 0; b < r.D; ++b) out->direction[a * r.D + b] = r.dir[a * 3 + b];
}

void og_smooth_image_f64(const og_grid* g, const double* in, double* out, double sigma, int in_mm) {
  smooth_image<double>(to_grid(g), in, out, sigma, in_mm != 0);
}
void og_smooth_field_f64(const og_grid* g, const double* in, double* out, double variance) {
  smooth_field<double>(to_grid(g), in, out, variance);
}
void og_warp_image_f64(const og_grid* gi, const double* img, const og_grid* go, const double* u, double* out) {
  warp_image<double>(to_grid(gi), img, to_grid(go), u, out);
}
void og_compose_f64(const og_grid* g, const double* disp, const double* warp, double* out) {
  compose<double>(to_grid(g), disp, warp, out);
}
void og_resample_field_f64(const og_grid* gi, const double* in, const og_grid* go, double* out) {
  resample_field<double>(to_grid(gi), in, to_grid(go), out);
}
double og_cc_f64(const og_grid* g, const double* F, const double* M, int radius, double eps, double* g_f2m, double* g_m2m) {
  return cc_metric<double>(to_grid(g), F, M, radius, eps, g_f2m, g_m2m);
}
double og_cc_ext_f64(const og_grid* g, const double* F, const double* M, int radius, double eps, const double* gext,
                     double* out) {
  return cc_metric<double>(to_grid(g), F, M, radius, eps, out, nullptr, gext);
}
double og_scale_update_f64(const og_grid* g, double* u, double lr) { return scale_update<double>(to_grid(g), u, lr); }
void og_bspline_fit_field_f64(const og_grid* g, const double* in, double* out, const int* mesh, int order) {
  const int m3[3] = {mesh[0], mesh[1], g->dim == 3 ? mesh[2] : 1};
  bspline_fit_field<double>(to_grid(g), in, out, m3, order);
}
void og_bspline_point(int64_t i, int64_t n, int spans, int order, int max_spans, int* span, double* b) {
  bsp_point(i, n, spans, order, bsp_epsilon(max_spans), span, b);
}
int og_invert_f64(const og_grid* g, const double* field, const double* init, double* inv, double* out_max, double* out_mean) {
  return invert_field<double>(to_grid(g), field, init, inv, out_max, out_mean);
}
void og_preprocess_f64(int64_t n, const double* in, double* out) { preprocess<double>(n, in, out); }

double og_conv_value(const double* energies, int n, int window) {
  ConvMonitor m(window);
  for (int i = 0; i < n; ++i) m.add(energies[i]);
  return m.value();
}

// Nearest-neighbour warp of an integer label map (antsApplyTransforms -n
// NearestNeighbor, Examples/antsApplyTransforms.cxx:593-630): point x + u(x)
// in double, continuous index in the label grid, outside -> 0, else the label
// at floor(c + 0.5) per axis (ITK rounds half-integers up).
void og_warp_labels_nn(const og_grid* gout, const float* u, const og_grid* glab, const int32_t* labels, int32_t* out) {
  const Grid go = to_grid(gout), gl = to_grid(glab);
  const int D = go.D;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < go.n[2]; ++z)
    for (int64_t y = 0; y < go.n[1]; ++y)
      for (int64_t x = 0; x < go.n[0]; ++x) {
        const int64_t i = x + go.n[0] * (y + go.n[1] * z);
        double p[3], c[3];
        grid_point(go, x, y, z, p);
        if (u)
          for (int d = 0; d < D; ++d) p[d] += static_cast<double>(u[i * D + d]);
        to_cidx(gl, p, c);
        int32_t v = 0;
        if (inside(gl, c)) {
          int64_t ix[3] = {0, 0, 0};
          for (int d = 0; d < D; ++d) ix[d] = static_cast<int64_t>(std::floor(c[d] + 0.5));
          v = labels[ix[0] + gl.n[0] * (ix[1] + gl.n[1] * ix[2])];
        }
        out[i] = v;
      }
}

int og_syn_register_f64(const og_params* P, const og_grid* gf, const double* fixed, const og_grid* gm, const double* moving,
                        double* out_fwd, double* out_inv, og_trace* trace, double* state4, double eps) {
  return syn_register<double>(P, to_grid(gf), fixed, to_grid(gm), moving, out_fwd, out_inv, trace, state4, eps);
}

// Same, plus the result of every level: lvl_fwd / lvl_inv hold n_levels
// fields on the last level grid, level l's being the output of the same
// schedule with no iterations after level l.
int og_syn_register_levels_f64(const og_params* P, const og_grid* gf, const double* fixed, const og_grid* gm,
                               const double* moving, double* out_fwd, double* out_inv, og_trace* trace,
                               double* lvl_fwd, double* lvl_inv, double eps) {
  return syn_register<double>(P, to_grid(gf), fixed, to_grid(gm), moving, out_fwd, out_inv, trace, nullptr, eps,
                              nullptr, 0, nullptr, 0, nullptr, nullptr, lvl_fwd, lvl_inv);
}

int og_syn_register_f32(const og_params* P, const og_grid* gf, const float* fixed, const og_grid* gm, const float* moving,
                        float* out_fwd, float* out_inv, og_trace* trace, float* state4, double eps) {
  return syn_register<float>(P, to_grid(gf), fixed, to_grid(gm), moving, out_fwd, out_inv, trace, state4, eps);
}

// The SyN stage with fixed / moving initial transforms (identity, affine or
// displacement field), each list in the order a point passes through it.
int og_syn_register_init_f64(const og_params* P, const og_grid* gf, const double* fixed, const og_grid* gm,
                             const double* moving, const og_xform* fx, int nfx, const og_xform* mx, int nmx,
                             double* out_fwd, double* out_inv, og_trace* trace, double* state4, double eps) {
  return syn_register<double>(P, to_grid(gf), fixed, to_grid(gm), moving, out_fwd, out_inv, trace, state4, eps, fx, nfx,
                              mx, nmx);
}

int og_syn_register_init_f32(const og_params* P, const og_grid* gf, const float* fixed, const og_grid* gm,
                             const float* moving, const og_xform* fx, int nfx, const og_xform* mx, int nmx,
                             float* out_fwd, float* out_inv, og_trace* trace, float* state4, double eps) {
  return syn_register<float>(P, to_grid(gf), fixed, to_grid(gm), moving, out_fwd, out_inv, trace, state4, eps, fx, nfx,
                             mx, nmx);
}

// The SyN stage restored from a saved state: state4 holds FixedToMiddle
// (fwd, inv) and MovingToMiddle (fwd, inv) on grid gs, one after another.
int og_syn_register_restore_f64(const og_params* P, const og_grid* gf, const double* fixed, const og_grid* gm,
                                const double* moving, const og_grid* gs, const double* state_in, const og_xform* mx,
                                int nmx, double* out_fwd, double* out_inv, og_trace* trace, double* state4,
                                double eps) {
  const Grid g = to_grid(gs);
  return syn_register<double>(P, to_grid(gf), fixed, to_grid(gm), moving, out_fwd, out_inv, trace, state4, eps, nullptr,
                              0, mx, nmx, &g, state_in);
}

void og_warp_image_xf_f64(const og_grid* gi, const double* img, const og_grid* go, const double* u,
                          const og_xform* xf, int nt, double* out) {
  warp_image<double>(to_grid(gi), img, to_grid(go), u, out, xf, nt);
}

int og_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
