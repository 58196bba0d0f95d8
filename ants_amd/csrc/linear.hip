// Host engine and C-ABI of the linear stages (antsRegistration
// -t Translation[step] / Rigid[step] / Affine[step] with MeanSquares, GC or
// Mattes), the stages antsRegistrationSyN.sh runs before SyN.
//
// RegistrationHelper builds each one as ImageRegistrationMethodv4 with a
// ConjugateGradientLineSearchOptimizerv4 (Examples/itkantsRegistrationHelper.hxx:
// 1290-1372, 1464-1732; itkantsRegistrationHelper.h:905-1097).  Per level:
// window both images to [0, 1], smooth them (DiscreteGaussian), shrink the
// virtual domain (the fixed grid), draw the metric's sample points, estimate
// the parameter scales from physical shifts of the virtual domain's corners,
// then iterate: metric value + derivative (one GPU pass; two for Mattes),
// window convergence test, Polak-Ribiere direction, golden-section line
// search over [0, 2 lr] (a GPU value pass per trial), parameter update.
// The GPU passes reduce to a few dozen doubles; the optimizer is host code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ants_syn.h"
#include "kernels.hpp"

namespace ants {
// Shared with the SyN engine (engine.hip).
int host_error(int code, const char* msg);
int host_geom(const ants_image_desc* d, GeomDev* g);
int host_level_desc(const ants_image_desc* fixed, int factor, ants_image_desc* out);
}  // namespace ants

using namespace ants;

namespace {

constexpr double kDblEps = std::numeric_limits<double>::epsilon();
constexpr double kPi = 3.14159265358979323846;

int err(int code, const std::string& msg) { return host_error(code, msg.c_str()); }

// Host threads for per-point set-up work: OMP_NUM_THREADS when set (the
// share of the machine the process was given), else the hardware count, at
// most 64.
int64_t host_threads() {
  int64_t t = 0;
  if (const char* e = std::getenv("OMP_NUM_THREADS")) t = std::atoll(e);
  if (t <= 0) t = static_cast<int64_t>(std::thread::hardware_concurrency());
  return std::max<int64_t>(1, std::min<int64_t>(t, 64));
}

#define LIN_TRY(expr)               \
  do {                              \
    int rc_ = (expr);               \
    if (rc_ != ANTS_OK) return rc_; \
  } while (0)
#define LIN_HIP(expr)                                                                               \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return err(ANTS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// MT19937 (Matsumoto & Nishimura 1998, init_genrand seeding) with the
// variates ITK's MersenneTwisterRandomVariateGenerator provides.
class Twister {
 public:
  explicit Twister(uint32_t s) {
    mt_[0] = s;
    for (int i = 1; i < 624; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + static_cast<uint32_t>(i);
    idx_ = 624;
  }
  uint32_t next() {
    if (idx_ >= 624) twist();
    uint32_t y = mt_[idx_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  // uniform in [0, n]: mask to the bits n uses, reject values above n
  uint32_t below_or_equal(uint32_t n) {
    uint32_t used = n;
    used |= used >> 1;
    used |= used >> 2;
    used |= used >> 4;
    used |= used >> 8;
    used |= used >> 16;
    uint32_t i;
    do i = next() & used;
    while (i > n);
    return i;
  }

 private:
  static uint32_t mix(uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return c ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }
  // the recurrence in three runs so that no index wraps: words 0..226 read
  // the old word 397 ahead, 227..622 the new word 227 behind, 623 word 0
  void twist() {
    int i = 0;
    for (; i < 227; ++i) mt_[i] = mix(mt_[i], mt_[i + 1], mt_[i + 397]);
    for (; i < 623; ++i) mt_[i] = mix(mt_[i], mt_[i + 1], mt_[i - 227]);
    mt_[623] = mix(mt_[623], mt_[0], mt_[396]);
    idx_ = 0;
  }
  uint32_t mt_[624];
  int idx_;
};

// Box-Muller normal from two consecutive draws: u1 on [0, 1], u2 on [0, 1)
inline double box_muller(uint32_t a, uint32_t b) {
  const double u1 = static_cast<double>(a) * (1.0 / 4294967295.0);
  const double u2 = static_cast<double>(b) * (1.0 / 4294967296.0);
  const double r = std::sqrt(-2.0 * std::log(1.0 - u1) * 1.0);
  return 0.0 + r * std::cos(2.0 * kPi * u2);
}

// ---------------------------------------------------------------------------
// 3x3 row-major helpers (D x D used, identity-padded).
struct M3 {
  double a[9];
};
M3 eye3() {
  M3 m{};
  m.a[0] = m.a[4] = m.a[8] = 1.0;
  return m;
}
M3 mul(const M3& x, const M3& y) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += x.a[i * 3 + k] * y.a[k * 3 + j];
      r.a[i * 3 + j] = s;
    }
  return r;
}
void mulv(const M3& m, const double v[3], double out[3]) {
  double t[3];
  for (int i = 0; i < 3; ++i) t[i] = m.a[i * 3 + 0] * v[0] + m.a[i * 3 + 1] * v[1] + m.a[i * 3 + 2] * v[2];
  for (int i = 0; i < 3; ++i) out[i] = t[i];
}

int n_params(int kind, int D) {
  if (kind == ANTS_LINEAR_TRANSLATION) return D;
  if (kind == ANTS_LINEAR_RIGID) return D == 2 ? 3 : 6;
  return D * D + D;
}

void rot3(double ax, double ay, double az, M3& rz, M3& rx, M3& ry) {
  const double cx = std::cos(ax), sx = std::sin(ax), cy = std::cos(ay), sy = std::sin(ay), cz = std::cos(az),
               sz = std::sin(az);
  rx = M3{{1, 0, 0, 0, cx, -sx, 0, sx, cx}};
  ry = M3{{cy, 0, sy, 0, 1, 0, -sy, 0, cy}};
  rz = M3{{cz, -sz, 0, sz, cz, 0, 0, 0, 1}};
}

// Stage transform y = A (x - c) + c + t: Translation (TranslationTransform),
// Rigid (Euler2DTransform; Euler3DTransform with R = Rz Rx Ry), Affine
// (AffineTransform, matrix row-major then translation).
void stage_matrix(int kind, const double* p, int D, M3& A, double t[3]) {
  A = eye3();
  t[0] = t[1] = t[2] = 0.0;
  if (kind == ANTS_LINEAR_TRANSLATION) {
    for (int d = 0; d < D; ++d) t[d] = p[d];
  } else if (kind == ANTS_LINEAR_AFFINE) {
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) A.a[i * 3 + j] = p[i * D + j];
    for (int d = 0; d < D; ++d) t[d] = p[D * D + d];
  } else if (D == 2) {
    const double c = std::cos(p[0]), s = std::sin(p[0]);
    A.a[0] = c;
    A.a[1] = -s;
    A.a[3] = s;
    A.a[4] = c;
    t[0] = p[1];
    t[1] = p[2];
  } else {
    M3 rz, rx, ry;
    rot3(p[0], p[1], p[2], rz, rx, ry);
    A = mul(mul(rz, rx), ry);
    for (int d = 0; d < 3; ++d) t[d] = p[3 + d];
  }
}

// dS/dp_p at x = L_p (x - c) + tau_p
void jacobian_forms(int kind, const double* p, int D, LinJac& J) {
  std::memset(&J, 0, sizeof(J));
  if (kind == ANTS_LINEAR_TRANSLATION) {
    for (int i = 0; i < D; ++i) J.tau[i][i] = 1.0;
  } else if (kind == ANTS_LINEAR_AFFINE) {
    for (int i = 0; i < D; ++i) {
      for (int j = 0; j < D; ++j) J.L[i * D + j][i * 3 + j] = 1.0;
      J.tau[D * D + i][i] = 1.0;
    }
  } else if (D == 2) {
    const double c = std::cos(p[0]), s = std::sin(p[0]);
    J.L[0][0] = -s;
    J.L[0][1] = -c;
    J.L[0][3] = c;
    J.L[0][4] = -s;
    J.tau[1][0] = J.tau[2][1] = 1.0;
  } else {
    M3 rz, rx, ry;
    rot3(p[0], p[1], p[2], rz, rx, ry);
    const double cx = std::cos(p[0]), sx = std::sin(p[0]), cy = std::cos(p[1]), sy = std::sin(p[1]),
                 cz = std::cos(p[2]), sz = std::sin(p[2]);
    const M3 drx{{0, 0, 0, 0, -sx, -cx, 0, cx, -sx}};
    const M3 dry{{-sy, 0, cy, 0, 0, 0, -cy, 0, -sy}};
    const M3 drz{{-sz, -cz, 0, cz, -sz, 0, 0, 0, 0}};
    const M3 l0 = mul(mul(rz, drx), ry), l1 = mul(mul(rz, rx), dry), l2 = mul(mul(drz, rx), ry);
    std::memcpy(J.L[0], l0.a, sizeof(l0.a));
    std::memcpy(J.L[1], l1.a, sizeof(l1.a));
    std::memcpy(J.L[2], l2.a, sizeof(l2.a));
    for (int i = 0; i < 3; ++i) J.tau[3 + i][i] = 1.0;
  }
}

// The affines of a list applied in order, folded: x -> A x + b.
int affine_chain(const ants_transform* ts, int n, int D, M3& A, double b[3]) {
  A = eye3();
  b[0] = b[1] = b[2] = 0.0;
  if (n < 0 || n > ANTS_MAX_TRANSFORMS || (n > 0 && !ts)) return err(ANTS_ERR_INVALID, "bad initial transform list");
  for (int k = 0; k < n; ++k) {
    const ants_transform& T = ts[k];
    if (T.kind == ANTS_TRANSFORM_IDENTITY) continue;
    if (T.kind != ANTS_TRANSFORM_AFFINE)
      return err(ANTS_ERR_UNSUPPORTED, "linear stages take affine initial transforms only");
    M3 M = eye3();
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) M.a[i * 3 + j] = T.matrix[i * D + j];
    double off[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < D; ++i) {
      double mc = 0.0;
      for (int j = 0; j < D; ++j) mc += M.a[i * 3 + j] * T.center[j];
      off[i] = T.translation[i] + T.center[i] - mc;
    }
    double nb[3];
    mulv(M, b, nb);
    for (int i = 0; i < 3; ++i) b[i] = nb[i] + off[i];
    A = mul(M, A);
  }
  return ANTS_OK;
}

int validate(const ants_linear_params* p) {
  if (!p) return err(ANTS_ERR_INVALID, "null parameters");
  if (p->dim != 2 && p->dim != 3) return err(ANTS_ERR_INVALID, "dim must be 2 or 3");
  if (p->transform < ANTS_LINEAR_TRANSLATION || p->transform > ANTS_LINEAR_AFFINE)
    return err(ANTS_ERR_UNSUPPORTED, "linear transform must be Translation, Rigid or Affine");
  if (!(p->gradient_step > 0.0)) return err(ANTS_ERR_INVALID, "gradient step must be > 0");
  if (p->metric < ANTS_METRIC_MEAN_SQUARES || p->metric > ANTS_METRIC_MATTES)
    return err(ANTS_ERR_UNSUPPORTED, "linear-stage metric must be MeanSquares, GC or Mattes");
  if (p->metric == ANTS_METRIC_MATTES && (p->bins < 5 || p->bins > 1024))
    return err(ANTS_ERR_INVALID, "Mattes bins must be 5..1024");
  if (p->sampling < ANTS_SAMPLING_NONE || p->sampling > ANTS_SAMPLING_RANDOM)
    return err(ANTS_ERR_INVALID, "unknown sampling strategy");
  if (p->sampling != ANTS_SAMPLING_NONE && !(p->sampling_percentage > 0.0 && p->sampling_percentage <= 1.0))
    return err(ANTS_ERR_INVALID, "sampling percentage must be in (0, 1]");
  if (p->n_levels < 1 || p->n_levels > ANTS_MAX_LEVELS) return err(ANTS_ERR_INVALID, "n_levels must be 1..8");
  for (int l = 0; l < p->n_levels; ++l) {
    if (p->shrink[l] < 1) return err(ANTS_ERR_INVALID, "shrink factors must be >= 1");
    if (!(p->sigma[l] >= 0.0)) return err(ANTS_ERR_INVALID, "smoothing sigmas must be >= 0");
    if (p->iterations[l] < 0) return err(ANTS_ERR_INVALID, "iterations must be >= 0");
  }
  if (p->convergence_window < 1) return err(ANTS_ERR_INVALID, "convergence window must be >= 1");
  return ANTS_OK;
}

// Sample points of a level: the draws come off the one MT19937 stream in
// order (per point: for random sampling the voxel draw(s), then two draws per
// axis for the jitter normals), which is sequential; turning them into
// points (Box-Muller, index to physical point) is independent per point and
// runs on host threads, so the result is the same bits on any thread count.
// The stream is drawn in chunks into two reused buffers: while the threads
// turn chunk c into points, this thread draws chunk c + 1.
int64_t level_point_count(const ants_linear_params* p, const ants_image_desc& lv) {
  const int64_t N = lv.size[0] * lv.size[1] * (lv.dim == 3 ? lv.size[2] : 1);
  if (p->sampling == ANTS_SAMPLING_NONE) return 0;
  if (p->sampling == ANTS_SAMPLING_REGULAR) {
    const int64_t step = static_cast<int64_t>(std::ceil(1.0 / p->sampling_percentage));
    return (N + step - 1) / step;
  }
  return static_cast<int64_t>(std::floor(static_cast<double>(N) * p->sampling_percentage));
}

// writes level_point_count() points, D doubles each, to out
int level_points(const ants_linear_params* p, const ants_image_desc& lv, double* out) {
  const int D = lv.dim;
  const int64_t nx = lv.size[0], ny = lv.size[1], nz = D == 3 ? lv.size[2] : 1;
  const int64_t N = nx * ny * nz;
  if (p->sampling == ANTS_SAMPLING_NONE) return ANTS_OK;
  GeomDev g;
  LIN_TRY(host_geom(&lv, &g));
  const int64_t n = level_point_count(p, lv);
  const int64_t step =
      p->sampling == ANTS_SAMPLING_REGULAR ? static_cast<int64_t>(std::ceil(1.0 / p->sampling_percentage)) : 0;
  if (step == 0 && N - 1 > static_cast<int64_t>(0xffffffffu))
    return err(ANTS_ERR_UNSUPPORTED, "grid too large for random sampling");
  double third[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < D; ++d) third[d] = lv.spacing[d] / 3.0;
  const int per = 2 * D;
  const int64_t kChunk = 262144;  // points per chunk
  struct Buf {
    std::vector<uint32_t> draws, vox;
  } buf[2];
  for (auto& b : buf) {
    b.draws.resize(static_cast<size_t>(std::min(n, kChunk)) * per);
    if (step == 0) b.vox.resize(static_cast<size_t>(std::min(n, kChunk)));
  }
  Twister tw(p->seed);
  auto draw = [&](Buf& b, int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; ++k) {
      if (step == 0) b.vox[static_cast<size_t>(k - k0)] = tw.below_or_equal(static_cast<uint32_t>(N - 1));
      uint32_t* d = &b.draws[static_cast<size_t>(k - k0) * per];
      for (int q = 0; q < per; ++q) d[q] = tw.next();
    }
  };
  auto points = [&](const Buf& b, int64_t c0, int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; ++k) {
      const int64_t lin = step ? k * step : static_cast<int64_t>(b.vox[static_cast<size_t>(k - c0)]);
      const double ix[3] = {static_cast<double>(lin % nx), static_cast<double>((lin / nx) % ny),
                            static_cast<double>(lin / (nx * ny))};
      const uint32_t* dr = &b.draws[static_cast<size_t>(k - c0) * per];
      for (int r = 0; r < D; ++r) {
        double x = g.org[r];
        for (int c = 0; c < D; ++c) x = x + g.m[r * 3 + c] * ix[c];
        out[static_cast<size_t>(k) * D + r] = x + box_muller(dr[2 * r], dr[2 * r + 1]) * third[r];
      }
    }
  };
  const int64_t T = host_threads();
  std::vector<std::thread> th;
  int64_t cur0 = 0, cur1 = std::min(n, kChunk);
  if (n > 0) draw(buf[0], cur0, cur1);
  for (int c = 0; cur0 < n; ++c) {
    Buf& b = buf[c & 1];
    const int64_t span = cur1 - cur0;
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(T, span / 4096));
    const int64_t part = (span + nt - 1) / nt;
    for (int64_t t = 0; t < nt; ++t) {
      const int64_t k0 = cur0 + t * part, k1 = std::min(cur1, k0 + part);
      if (k0 < k1) th.emplace_back(points, std::cref(b), cur0, k0, k1);
    }
    const int64_t nxt0 = cur1, nxt1 = std::min(n, cur1 + kChunk);
    if (nxt0 < n) draw(buf[(c + 1) & 1], nxt0, nxt1);
    for (auto& x : th) x.join();
    th.clear();
    cur0 = nxt0;
    cur1 = nxt1;
  }
  return ANTS_OK;
}

// Largest physical shift of the points under a parameter change: the
// composite (stage, then the moving initial chain) at p + step against p.
double max_shift(int kind, int D, const double* p, const double* step, const double c[3], const M3& MA,
                 const std::vector<double>& pts) {
  const int P = n_params(kind, D);
  double q[kLinMaxParams];
  for (int i = 0; i < P; ++i) q[i] = p[i] + step[i];
  M3 A0, A1;
  double t0[3], t1[3];
  stage_matrix(kind, p, D, A0, t0);
  stage_matrix(kind, q, D, A1, t1);
  double mx = 0.0;
  const size_t n = pts.size() / D;
  for (size_t k = 0; k < n; ++k) {
    double r[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < D; ++d) r[d] = pts[k * D + d] - c[d];
    double y0[3], y1[3];
    mulv(A0, r, y0);
    mulv(A1, r, y1);
    double dv[3];
    for (int d = 0; d < 3; ++d) dv[d] = (y1[d] + t1[d]) - (y0[d] + t0[d]);
    double dz[3];
    mulv(MA, dv, dz);
    double s = 0.0;
    for (int d = 0; d < D; ++d) s += dz[d] * dz[d];
    mx = std::max(mx, std::sqrt(s));
  }
  return mx;
}

std::vector<double> corner_points(const ants_image_desc& lv) {
  const int D = lv.dim;
  GeomDev g;
  std::vector<double> pts;
  if (host_geom(&lv, &g) != ANTS_OK) return pts;
  for (int i = 0; i < (1 << D); ++i) {
    double ix[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < D; ++d) ix[d] = ((i >> d) & 1) ? static_cast<double>(lv.size[d] - 1) : 0.0;
    for (int r = 0; r < D; ++r) {
      double s = g.org[r];
      for (int c = 0; c < D; ++c) s = s + g.m[r * 3 + c] * ix[c];
      pts.push_back(s);
    }
  }
  return pts;
}

// RegistrationParameterScalesFromPhysicalShift::EstimateScales with a 0.01
// parameter change.
void estimate_scales(int kind, int D, const double* p, const double c[3], const M3& MA,
                     const std::vector<double>& corners, double* scales) {
  const int P = n_params(kind, D);
  const double small = 0.01;
  double sh[kLinMaxParams];
  double mn = std::numeric_limits<double>::max();
  for (int i = 0; i < P; ++i) {
    double d[kLinMaxParams] = {0.0};
    d[i] = small;
    sh[i] = max_shift(kind, D, p, d, c, MA, corners);
    if (sh[i] > kDblEps && sh[i] < mn) mn = sh[i];
  }
  for (int i = 0; i < P; ++i) {
    if (mn == std::numeric_limits<double>::max()) {
      scales[i] = 1.0;
      continue;
    }
    scales[i] = (sh[i] <= kDblEps ? mn * mn : sh[i] * sh[i]) / (small * small);
  }
}

struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t b) {
    if (b <= bytes) return ANTS_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, b) != hipSuccess) return err(ANTS_ERR_OOM, "device allocation failed");
    bytes = b;
    return ANTS_OK;
  }
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

struct PinnedMem {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t b) {
    if (b <= bytes) return ANTS_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    if (hipHostMalloc(&p, b, hipHostMallocDefault) != hipSuccess) return err(ANTS_ERR_OOM, "pinned allocation failed");
    bytes = b;
    return ANTS_OK;
  }
  ~PinnedMem() {
    if (p) (void)hipHostFree(p);
  }
};

// Device copies of level sample points, shared between stages.  The points
// depend only on the seed, the sampling strategy and percentage and the
// level grid, so consecutive stages of one registration (the Rigid and
// Affine stages of antsRegistrationSyN.sh share all of them) reuse one
// upload.  At most ANTS_LINEAR_POINT_CACHE sets (default 8, 0 disables) are
// kept per process, least recently used first out; a stage holds its set by
// shared ownership, so an eviction never frees points in use.
struct PointKey {
  int device, dim, sampling;
  uint32_t seed;
  double percentage;
  int64_t size[3];
  double spacing[3], origin[3], direction[9];
};
struct PointSet {
  DevMem mem;
  int64_t n = 0;
};
struct PointCache {
  std::mutex mu;
  std::vector<std::pair<PointKey, std::shared_ptr<PointSet>>> sets;  // most recent first
};
PointCache& point_cache() {
  static PointCache* c = new PointCache();  // never destroyed: no device frees during process exit
  return *c;
}
size_t point_cache_capacity() {
  static const size_t cap = [] {
    const char* e = std::getenv("ANTS_LINEAR_POINT_CACHE");
    return e ? static_cast<size_t>(std::max(0, std::atoi(e))) : static_cast<size_t>(8);
  }();
  return cap;
}
PointKey point_key(const ants_linear_params* p, const ants_image_desc& lv) {
  PointKey k;
  std::memset(&k, 0, sizeof(k));  // padding too: keys compare with memcmp
  int dev = 0;
  (void)hipGetDevice(&dev);
  k.device = dev;
  k.dim = lv.dim;
  k.sampling = p->sampling;
  k.seed = p->seed;
  k.percentage = p->sampling_percentage;
  for (int d = 0; d < lv.dim; ++d) {
    k.size[d] = lv.size[d];
    k.spacing[d] = lv.spacing[d];
    k.origin[d] = lv.origin[d];
  }
  for (int i = 0; i < lv.dim * lv.dim; ++i) k.direction[i] = lv.direction[i];
  return k;
}

// One stage's device state and its metric evaluations.
struct LinearStage {
  ants_linear_params P{};
  int D = 3, np = 0;
  hipStream_t s = nullptr;
  ants_image_desc fdesc{}, mdesc{};
  GeomDev fgeom{}, mgeom{};
  M3 MA = eye3(), FA = eye3();
  double mb[3] = {0, 0, 0}, fbv[3] = {0, 0, 0};
  double c[3] = {0, 0, 0};
  DevMem fpre, mpre, flev, mlev, partials, res, wtab, minmax;
  // res holds one evaluation's results contiguously: the Qmax sums, then for
  // Mattes the B x B joint histogram (u64) and the B fixed-bin counts (u32);
  // one copy brings them to the pinned mirror hres
  PinnedMem hres, hw;
  size_t hist_off = 0, fcnt_off = 0, res_bytes = 0;
  std::shared_ptr<PointSet> pts;  // the level's sample points (shared, see PointCache)
  // level state
  ants_image_desc lv{};
  int64_t npts = 0;   // 0: the level grid
  double fmin = 0, fmax = 0, mmin = 0, mmax = 0;
  int64_t evaluations = 0;

  int init(const ants_linear_params* p, const ants_image_desc* f, const float* d_f, const ants_image_desc* m,
           const float* d_m, const ants_transform* mi, int nmi, const ants_transform* fi, int nfi,
           const double* center, void* stream) {
    LIN_TRY(validate(p));
    P = *p;
    D = p->dim;
    np = n_params(P.transform, D);
    s = static_cast<hipStream_t>(stream);
    if (!f || !m || !d_f || !d_m) return err(ANTS_ERR_INVALID, "null image");
    if (f->dim != D || m->dim != D) return err(ANTS_ERR_INVALID, "image dimension differs from params");
    fdesc = *f;
    mdesc = *m;
    LIN_TRY(host_geom(f, &fgeom));
    LIN_TRY(host_geom(m, &mgeom));
    LIN_TRY(affine_chain(mi, nmi, D, MA, mb));
    LIN_TRY(affine_chain(fi, nfi, D, FA, fbv));
    for (int d = 0; d < 3; ++d) c[d] = (center && d < D) ? center[d] : 0.0;
    const int64_t nf = static_cast<int64_t>(fgeom.n[0]) * fgeom.n[1] * fgeom.n[2];
    const int64_t nm = static_cast<int64_t>(mgeom.n[0]) * mgeom.n[1] * mgeom.n[2];
    LIN_TRY(fpre.ensure(nf * sizeof(float)));
    LIN_TRY(mpre.ensure(nm * sizeof(float)));
    LIN_TRY(flev.ensure(nf * sizeof(float)));
    LIN_TRY(mlev.ensure(nm * sizeof(float)));
    if (P.preprocess) {
      LIN_TRY(ants_op_preprocess(nf, d_f, static_cast<float*>(fpre.p), stream));
      LIN_TRY(ants_op_preprocess(nm, d_m, static_cast<float*>(mpre.p), stream));
    } else {
      LIN_HIP(hipMemcpyAsync(fpre.p, d_f, nf * sizeof(float), hipMemcpyDeviceToDevice, s));
      LIN_HIP(hipMemcpyAsync(mpre.p, d_m, nm * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    const int Qmax = 6 + 3 * kLinMaxParams;
    LIN_TRY(partials.ensure(static_cast<size_t>(kLinMaxBlocks) * Qmax * sizeof(double)));
    hist_off = Qmax * sizeof(double);
    fcnt_off = hist_off;
    res_bytes = hist_off;
    if (P.metric == ANTS_METRIC_MATTES) {
      const size_t B = static_cast<size_t>(P.bins);
      fcnt_off = hist_off + B * B * sizeof(unsigned long long);
      res_bytes = fcnt_off + B * sizeof(unsigned int);
      LIN_TRY(wtab.ensure(B * B * sizeof(double)));
      LIN_TRY(hw.ensure(B * B * sizeof(double)));
    }
    LIN_TRY(res.ensure(res_bytes));
    LIN_TRY(hres.ensure(res_bytes));
    return ANTS_OK;
  }

  // the level's sample points: from the cache, or made on the host and
  // uploaded (and then offered to the cache)
  int level_point_set() {
    pts.reset();
    npts = 0;
    if (P.sampling == ANTS_SAMPLING_NONE) return ANTS_OK;
    const PointKey key = point_key(&P, lv);
    const size_t cap = point_cache_capacity();
    PointCache& pc = point_cache();
    if (cap > 0) {
      std::lock_guard<std::mutex> lk(pc.mu);
      for (size_t i = 0; i < pc.sets.size(); ++i)
        if (std::memcmp(&pc.sets[i].first, &key, sizeof(key)) == 0) {
          auto hit = pc.sets[i];
          pc.sets.erase(pc.sets.begin() + static_cast<std::ptrdiff_t>(i));
          pc.sets.insert(pc.sets.begin(), hit);
          pts = hit.second;
          npts = pts->n;
          return ANTS_OK;
        }
    }
    const int64_t n = level_point_count(&P, lv);
    if (n == 0) return err(ANTS_ERR_INVALID, "no sample points on this level");
    std::unique_ptr<double[]> hp(new double[static_cast<size_t>(n) * D]);
    LIN_TRY(level_points(&P, lv, hp.get()));
    auto set = std::make_shared<PointSet>();
    LIN_TRY(set->mem.ensure(static_cast<size_t>(n) * D * sizeof(double)));
    LIN_HIP(hipMemcpyAsync(set->mem.p, hp.get(), static_cast<size_t>(n) * D * sizeof(double), hipMemcpyHostToDevice,
                           s));
    LIN_HIP(hipStreamSynchronize(s));  // complete before any other stage may read it
    set->n = n;
    pts = set;
    npts = n;
    if (cap > 0) {
      std::lock_guard<std::mutex> lk(pc.mu);
      pc.sets.insert(pc.sets.begin(), std::make_pair(key, set));
      while (pc.sets.size() > cap) pc.sets.pop_back();
    }
    return ANTS_OK;
  }

  int image_range(const float* d, int64_t n, double& lo, double& hi) {
    LIN_TRY(minmax.ensure(2 * sizeof(uint32_t)));
    auto* mm = static_cast<uint32_t*>(minmax.p);
    launch_fill_u32(mm, 0xffffffffu, 1, s);
    launch_fill_u32(mm + 1, 0u, 1, s);
    launch_minmax(d, n, mm, s);
    uint32_t pair[2];
    LIN_HIP(hipMemcpyAsync(pair, mm, sizeof(pair), hipMemcpyDeviceToHost, s));
    LIN_HIP(hipStreamSynchronize(s));
    auto decode = [](uint32_t o) {
      uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
      float f;
      std::memcpy(&f, &u, 4);
      return static_cast<double>(f);
    };
    lo = decode(pair[0]);
    hi = decode(pair[1]);
    return ANTS_OK;
  }

  int set_level(int level) {
    LIN_TRY(host_level_desc(&fdesc, P.shrink[level], &lv));
    const int64_t nf = static_cast<int64_t>(fgeom.n[0]) * fgeom.n[1] * fgeom.n[2];
    const int64_t nm = static_cast<int64_t>(mgeom.n[0]) * mgeom.n[1] * mgeom.n[2];
    const double sg = P.sigma[level];
    if (sg > 0.0) {
      LIN_TRY(ants_op_smooth_image(&fdesc, static_cast<const float*>(fpre.p), static_cast<float*>(flev.p), sg,
                                   P.sigma_in_mm, s));
      LIN_TRY(ants_op_smooth_image(&mdesc, static_cast<const float*>(mpre.p), static_cast<float*>(mlev.p), sg,
                                   P.sigma_in_mm, s));
    } else {
      LIN_HIP(hipMemcpyAsync(flev.p, fpre.p, nf * sizeof(float), hipMemcpyDeviceToDevice, s));
      LIN_HIP(hipMemcpyAsync(mlev.p, mpre.p, nm * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    LIN_TRY(level_point_set());
    if (P.metric == ANTS_METRIC_MATTES) {
      LIN_TRY(image_range(static_cast<const float*>(flev.p), nf, fmin, fmax));
      LIN_TRY(image_range(static_cast<const float*>(mlev.p), nm, mmin, mmax));
      if (!(fmax > fmin) || !(mmax > mmin)) return err(ANTS_ERR_INVALID, "constant image: Mattes bins undefined");
    }
    LIN_HIP(hipStreamSynchronize(s));
    return ANTS_OK;
  }

  // value and -dvalue/dparams at stage parameters prm on the current level
  int evaluate(const double* prm, bool deriv, double* value, double* der, int64_t* nvalid) {
    ++evaluations;
    LinDev L;
    std::memset(&L, 0, sizeof(L));
    LIN_TRY(host_geom(&lv, &L.vg));
    L.pts = P.sampling != ANTS_SAMPLING_NONE ? static_cast<const double*>(pts->mem.p) : nullptr;
    L.npts = npts;
    L.fg = fgeom;
    L.mg = mgeom;
    L.fimg = static_cast<const float*>(flev.p);
    L.mimg = static_cast<const float*>(mlev.p);
    M3 A;
    double t[3];
    stage_matrix(P.transform, prm, D, A, t);
    // y = MA (A (x - c) + c + t) + mb = (MA A) x + MA (c + t - A c) + mb
    const M3 Y = mul(MA, A);
    double ac[3], v[3], yb[3];
    mulv(A, c, ac);
    for (int d = 0; d < 3; ++d) v[d] = c[d] + t[d] - ac[d];
    mulv(MA, v, yb);
    for (int i = 0; i < 3; ++i) {
      L.yb[i] = i < D ? yb[i] + mb[i] : 0.0;
      L.fb[i] = i < D ? fbv[i] : 0.0;
      L.c[i] = c[i];
      L.msp[i] = i < D ? mdesc.spacing[i] : 1.0;
      for (int j = 0; j < 3; ++j) {
        L.ya[i * 3 + j] = Y.a[i * 3 + j];
        L.fa[i * 3 + j] = FA.a[i * 3 + j];
        L.G[i * 3 + j] = MA.a[i * 3 + j];
      }
    }
    LinJac J;
    jacobian_forms(P.transform, prm, D, J);
    double* dout = static_cast<double*>(res.p);
    const double* host = static_cast<const double*>(hres.p);
    if (P.metric != ANTS_METRIC_MATTES) {
      const int mode = P.metric == ANTS_METRIC_MEAN_SQUARES ? kLinMS : kLinGC;
      if (!launch_lin_metric(L, J, D, np, mode, deriv, static_cast<double*>(partials.p), dout, nullptr, nullptr, s))
        return err(ANTS_ERR_UNSUPPORTED, "parameter count not supported");
      LIN_HIP(hipGetLastError());
      const int Q = lin_quantities(mode, deriv, np);
      LIN_HIP(hipMemcpyAsync(hres.p, dout, Q * sizeof(double), hipMemcpyDeviceToHost, s));
      LIN_HIP(hipStreamSynchronize(s));
      const double N = host[0];
      *nvalid = static_cast<int64_t>(N);
      if (N < 1.0) return err(ANTS_ERR_INVALID, "no sample point maps inside both images");
      if (mode == kLinMS) {
        *value = host[1] / N;
        if (deriv)
          for (int q = 0; q < np; ++q) der[q] = host[2 + q] / N;
        return ANTS_OK;
      }
      const double fb = host[1] / N, mbar = host[2] / N;
      const double fm = host[5] - N * fb * mbar, ff = host[3] - N * fb * fb, mm = host[4] - N * mbar * mbar;
      if (!(ff > 0.0 && mm > 0.0)) return err(ANTS_ERR_INVALID, "constant image over the valid points (GC)");
      *value = -fm * fm / (ff * mm);
      if (deriv)
        for (int q = 0; q < np; ++q) {
          const double sdm = host[6 + q];
          const double fdm = host[6 + np + q] - fb * sdm, mdm = host[6 + 2 * np + q] - mbar * sdm;
          der[q] = 2.0 * fm / (ff * mm) * (fdm - fm / mm * mdm);
        }
      return ANTS_OK;
    }
    // Mattes: joint histogram pass, PDFs and value on the host, derivative pass
    const int B = P.bins;
    const int pad = 2;
    L.bins = B;
    L.fbs = (fmax - fmin) / static_cast<double>(B - 2 * pad);
    L.mbs = (mmax - mmin) / static_cast<double>(B - 2 * pad);
    L.fnm = fmin / L.fbs - pad;
    L.mnm = mmin / L.mbs - pad;
    char* rb = static_cast<char*>(res.p);
    LIN_HIP(hipMemsetAsync(rb + hist_off, 0, res_bytes - hist_off, s));
    launch_lin_metric(L, J, D, np, kLinHist, false, static_cast<double*>(partials.p), dout,
                      reinterpret_cast<unsigned long long*>(rb + hist_off), reinterpret_cast<unsigned int*>(rb + fcnt_off),
                      s);
    LIN_HIP(hipGetLastError());
    // the count (sum 0) and everything from the histogram on, in one copy
    LIN_HIP(hipMemcpyAsync(hres.p, res.p, res_bytes, hipMemcpyDeviceToHost, s));
    LIN_HIP(hipStreamSynchronize(s));
    const auto* hhost = reinterpret_cast<const unsigned long long*>(static_cast<const char*>(hres.p) + hist_off);
    const auto* chost = reinterpret_cast<const unsigned int*>(static_cast<const char*>(hres.p) + fcnt_off);
    double* whost = static_cast<double*>(hw.p);
    const double N = host[0];
    *nvalid = static_cast<int64_t>(N);
    if (N < 1.0) return err(ANTS_ERR_INVALID, "no sample point maps inside both images");
    std::vector<double> pm(B, 0.0), pf(B, 0.0);
    std::vector<double> pj(static_cast<size_t>(B) * B);
    for (int i = 0; i < B; ++i) {
      pf[i] = static_cast<double>(chost[i]) / N;
      for (int j = 0; j < B; ++j) pj[i * B + j] = static_cast<double>(hhost[i * B + j]) / kLinHistScale / N;
    }
    for (int j = 0; j < B; ++j) {
      double sm = 0.0;
      for (int i = 0; i < B; ++i) sm += pj[i * B + j];
      pm[j] = sm;
    }
    double val = 0.0;
    for (int i = 0; i < B; ++i)
      for (int j = 0; j < B; ++j) {
        double w = 0.0;
        const double pij = pj[i * B + j];
        if (pij > kDblEps && pm[j] > kDblEps) {
          const double ratio = std::log(pij / pm[j]);
          w = ratio;
          if (pf[i] > kDblEps) val += pij * (ratio - std::log(pf[i]));
        }
        whost[i * B + j] = w;
      }
    *value = -val;
    if (!deriv) return ANTS_OK;
    LIN_HIP(hipMemcpyAsync(wtab.p, whost, static_cast<size_t>(B) * B * sizeof(double), hipMemcpyHostToDevice, s));
    L.wtab = static_cast<const double*>(wtab.p);
    if (!launch_lin_metric(L, J, D, np, kLinHistDeriv, true, static_cast<double*>(partials.p), dout, nullptr, nullptr,
                           s))
      return err(ANTS_ERR_UNSUPPORTED, "parameter count not supported");
    LIN_HIP(hipGetLastError());
    LIN_HIP(hipMemcpyAsync(hres.p, dout, (1 + np) * sizeof(double), hipMemcpyDeviceToHost, s));
    LIN_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < np; ++q) der[q] = host[1 + q] / (L.mbs * N);
    return ANTS_OK;
  }
};

// Recursive golden-section search over [a, c] with interior b (published
// recursive form; epsilon 0.2, at most 20 steps, as the helper configures
// ConjugateGradientLineSearchOptimizerv4).
struct LineSearch {
  LinearStage* st;
  const double* base;
  const double* dir;
  int np;
  int it = 0;
  int rc = ANTS_OK;
  double value_at(double step) {
    double q[kLinMaxParams];
    for (int i = 0; i < np; ++i) q[i] = base[i] + step * dir[i];
    double v = 0.0;
    int64_t n = 0;
    const int r = st->evaluate(q, false, &v, nullptr, &n);
    if (r != ANTS_OK && rc == ANTS_OK) rc = r;
    return v;
  }
  double search(double a, double b, double c, double fb, bool have_fb) {
    const double resphi = 2.0 - (1.0 + std::sqrt(5.0)) / 2.0;
    if (it > 20) return (c + a) / 2.0;
    ++it;
    const double x = (c - b > b - a) ? b + resphi * (c - b) : b - resphi * (b - a);
    if (std::fabs(c - a) < 0.2 * (std::fabs(b) + std::fabs(x))) return (c + a) / 2.0;
    const double fx = value_at(x);
    if (rc != ANTS_OK) return b;
    if (!have_fb) fb = value_at(b);
    if (rc != ANTS_OK) return b;
    if (fx < fb) return (c - b > b - a) ? search(b, x, c, fx, true) : search(a, x, b, fx, true);
    return (c - b > b - a) ? search(a, b, x, fb, true) : search(x, b, c, fb, true);
  }
};

int run_stage(LinearStage& st, const double* p0, ants_linear_result* res, std::vector<ants_linear_record>& trace) {
  const ants_linear_params& P = st.P;
  const int D = st.D, np = st.np;
  std::vector<double> prm(np, 0.0);
  if (p0) {
    for (int i = 0; i < np; ++i) prm[i] = p0[i];
  } else if (P.transform == ANTS_LINEAR_AFFINE) {
    for (int i = 0; i < D; ++i) prm[i * D + i] = 1.0;
  }
  std::memset(res, 0, sizeof(*res));
  res->n_params = np;
  double lr = P.gradient_step;
  // ANTS_LINEAR_VERBOSE: per-level host timings on stderr (set-up: pyramid
  // image, sample points, Mattes ranges; then the iterations)
  const bool verbose = std::getenv("ANTS_LINEAR_VERBOSE") != nullptr;
  using clk = std::chrono::steady_clock;
  for (int level = 0; level < P.n_levels; ++level) {
    const auto t_set = clk::now();
    const int64_t ev0 = st.evaluations;
    LIN_TRY(st.set_level(level));
    const auto t_iter = clk::now();
    const std::vector<double> corners = corner_points(st.lv);
    double scales[kLinMaxParams];
    estimate_scales(P.transform, D, prm.data(), st.c, st.MA, corners, scales);
    std::vector<double> energies;
    std::vector<double> conj(np, 0.0), last(np, 0.0), g(np), der(np);
    int it = 0;
    while (it < P.iterations[level]) {
      double val = 0.0;
      int64_t n = 0;
      LIN_TRY(st.evaluate(prm.data(), true, &val, der.data(), &n));
      energies.push_back(val);
      double conv = 0.0;
      LIN_TRY(ants_syn_convergence_value(energies.data(), static_cast<int>(energies.size()), P.convergence_window,
                                         &conv));
      if (conv <= P.convergence_threshold) break;
      for (int i = 0; i < np; ++i) g[i] = der[i] / scales[i];
      if (it == 0) {
        std::fill(conj.begin(), conj.end(), 0.0);
        std::fill(last.begin(), last.end(), 0.0);
        const double ss = max_shift(P.transform, D, prm.data(), g.data(), st.c, st.MA, corners);
        lr = ss <= kDblEps ? 1.0 : P.gradient_step / ss;
      }
      double den = 0.0, num = 0.0;
      for (int i = 0; i < np; ++i) den += last[i] * last[i];
      double gamma = 0.0;
      if (den > kDblEps) {
        for (int i = 0; i < np; ++i) num += (g[i] - last[i]) * g[i];
        gamma = num / den;
      }
      if (gamma < 0.0 || gamma > 5.0) gamma = 0.0;
      for (int i = 0; i < np; ++i) {
        last[i] = g[i];
        conj[i] = g[i] + gamma * conj[i];
      }
      LineSearch ls{&st, prm.data(), conj.data(), np};
      const double step = ls.search(lr * 0.0, lr, lr * 2.0, 0.0, false);
      if (ls.rc != ANTS_OK) return ls.rc;
      lr = step;
      for (int i = 0; i < np; ++i) prm[i] += lr * conj[i];
      trace.push_back(ants_linear_record{level, it + 1, val, conv, lr});
      ++it;
    }
    res->iterations[level] = it;
    if (verbose) {
      const auto t_end = clk::now();
      const double ms_set = std::chrono::duration<double, std::milli>(t_iter - t_set).count();
      const double ms_it = std::chrono::duration<double, std::milli>(t_end - t_iter).count();
      const int64_t ne = st.evaluations - ev0;
      std::fprintf(stderr, "linear level %d: %lld points, set-up %.2f ms, %d iterations, %lld evaluations in %.2f ms "
                   "(%.3f ms each)\n", level, static_cast<long long>(st.npts), ms_set, it, static_cast<long long>(ne),
                   ms_it, ne > 0 ? ms_it / static_cast<double>(ne) : 0.0);
    }
  }
  for (int i = 0; i < np; ++i) res->params[i] = prm[i];
  M3 A;
  double t[3];
  stage_matrix(P.transform, prm.data(), D, A, t);
  ants_transform& T = res->transform;
  std::memset(&T, 0, sizeof(T));
  T.kind = ANTS_TRANSFORM_AFFINE;
  for (int i = 0; i < D; ++i) {
    for (int j = 0; j < D; ++j) T.matrix[i * D + j] = A.a[i * 3 + j];
    T.translation[i] = t[i];
    T.center[i] = st.c[i];
  }
  res->metric_evaluations = st.evaluations;
  return ANTS_OK;
}

}  // namespace

extern "C" {

int ants_linear_register(const ants_linear_params* p, const ants_image_desc* fixed, const float* d_fixed,
                         const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                         int n_moving, const ants_transform* fixed_initial, int n_fixed, const double* center,
                         const double* initial_params, ants_linear_result* out, ants_linear_record* trace,
                         int64_t trace_capacity, int64_t* trace_count, void* stream) {
  if (!out) return err(ANTS_ERR_INVALID, "null result");
  if (trace_capacity < 0 || (trace_capacity > 0 && !trace)) return err(ANTS_ERR_INVALID, "bad trace buffer");
  LinearStage st;
  LIN_TRY(st.init(p, fixed, d_fixed, moving, d_moving, moving_initial, n_moving, fixed_initial, n_fixed, center,
                  stream));
  std::vector<ants_linear_record> tr;
  const int rc = run_stage(st, initial_params, out, tr);
  const int64_t w = std::min<int64_t>(trace_capacity, static_cast<int64_t>(tr.size()));
  for (int64_t i = 0; i < w; ++i) trace[i] = tr[i];
  if (trace_count) *trace_count = static_cast<int64_t>(tr.size());
  return rc;
}

int ants_linear_metric(const ants_linear_params* p, int level, const ants_image_desc* fixed, const float* d_fixed,
                       const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                       int n_moving, const ants_transform* fixed_initial, int n_fixed, const double* center,
                       const double* params, int derivative, double* value, double* der, int64_t* n_valid,
                       void* stream) {
  if (!value || !n_valid || !params || (derivative && !der)) return err(ANTS_ERR_INVALID, "null output");
  LinearStage st;
  LIN_TRY(st.init(p, fixed, d_fixed, moving, d_moving, moving_initial, n_moving, fixed_initial, n_fixed, center,
                  stream));
  if (level < 0 || level >= p->n_levels) return err(ANTS_ERR_INVALID, "level out of range");
  LIN_TRY(st.set_level(level));
  return st.evaluate(params, derivative != 0, value, der, n_valid);
}

int ants_linear_sample_points(const ants_linear_params* p, int level, const ants_image_desc* fixed, double* out,
                              int64_t capacity, int64_t* count) {
  LIN_TRY(validate(p));
  if (!count || capacity < 0 || (capacity > 0 && !out)) return err(ANTS_ERR_INVALID, "bad output buffer");
  if (level < 0 || level >= p->n_levels) return err(ANTS_ERR_INVALID, "level out of range");
  if (!fixed || fixed->dim != p->dim) return err(ANTS_ERR_INVALID, "fixed image dimension differs from params");
  ants_image_desc lv;
  LIN_TRY(host_level_desc(fixed, p->shrink[level], &lv));
  const int64_t n = level_point_count(p, lv);
  *count = n;
  if (capacity == 0 || n == 0) return ANTS_OK;
  if (capacity >= n) return level_points(p, lv, out);
  std::unique_ptr<double[]> pts(new double[static_cast<size_t>(n) * p->dim]);
  LIN_TRY(level_points(p, lv, pts.get()));
  std::memcpy(out, pts.get(), static_cast<size_t>(capacity) * p->dim * sizeof(double));
  return ANTS_OK;
}

int ants_linear_estimate_scales(const ants_linear_params* p, int level, const ants_image_desc* fixed,
                                const ants_transform* moving_initial, int n_moving, const double* center,
                                const double* params, double* scales) {
  LIN_TRY(validate(p));
  if (!scales || !params) return err(ANTS_ERR_INVALID, "null argument");
  if (level < 0 || level >= p->n_levels) return err(ANTS_ERR_INVALID, "level out of range");
  if (!fixed || fixed->dim != p->dim) return err(ANTS_ERR_INVALID, "fixed image dimension differs from params");
  const int D = p->dim;
  M3 MA;
  double mb[3];
  LIN_TRY(affine_chain(moving_initial, n_moving, D, MA, mb));
  ants_image_desc lv;
  LIN_TRY(host_level_desc(fixed, p->shrink[level], &lv));
  double c[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < D; ++d) c[d] = center ? center[d] : 0.0;
  estimate_scales(p->transform, D, params, c, MA, corner_points(lv), scales);
  return ANTS_OK;
}

int ants_linear_release_points(void) {
  PointCache& pc = point_cache();
  std::lock_guard<std::mutex> lk(pc.mu);
  pc.sets.clear();
  return ANTS_OK;
}

}  // extern "C"
