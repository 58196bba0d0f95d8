// Host engine and C-ABI of the SyN + CC registration library.
//
// The engine replays SyNImageRegistrationMethod as RegistrationHelper
// configures it (Examples/itkantsRegistrationHelper.hxx:1983-2220): per
// level, smooth both full-resolution images, bring the four middle fields
// to the level grid, then iterate: warp, CC value + both derivatives,
// update-field smoothing + max norm, scaled composition onto the forward
// middle fields, total-field smoothing, four fixed-point inversions, energy
// and convergence.  All of it runs as HIP kernels on one stream; the host
// reads back one small statistics block per iteration (energy and inner
// step counts) to drive the convergence monitor and the observer.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ants_syn.h"
#include "kernels.hpp"

using namespace ants;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(ANTS_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define ANTS_TRY(expr)         \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != ANTS_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// Gaussian operator coefficients (ITK GaussianOperator::GenerateCoefficients):
// c_i = e^-t I_i(t) grown until 1 + 2 sum c_i >= 1 - maxError, normalised.
// Modified Bessel functions by the method of Numerical Recipes in C sec. 6.6
// (polynomial fits of Abramowitz & Stegun 9.8.1-9.8.4, Miller recurrence).
double bi0(double x) {
  double ax = std::fabs(x);
  if (ax < 3.75) {
    double y = (x / 3.75) * (x / 3.75);
    return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  double y = 3.75 / ax;
  return (std::exp(ax) / std::sqrt(ax)) *
         (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}
double bi1(double x) {
  double ax = std::fabs(x), r;
  if (ax < 3.75) {
    double y = (x / 3.75) * (x / 3.75);
    r = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    double y = 3.75 / ax;
    r = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    r = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y * (-0.1031555e-1 + y * r))));
    r *= std::exp(ax) / std::sqrt(ax);
  }
  return x < 0.0 ? -r : r;
}
double bin(int n, double x) {
  if (n == 0) return bi0(x);
  if (n == 1) return bi1(x);
  if (x == 0.0) return 0.0;
  const double tox = 2.0 / std::fabs(x);
  double bip = 0.0, ans = 0.0, b = 1.0;
  for (int j = 2 * (n + static_cast<int>(std::sqrt(40.0 * n))); j > 0; --j) {
    const double bim = bip + j * tox * b;
    bip = b;
    b = bim;
    if (std::fabs(b) > 1.0e10) {
      ans *= 1.0e-10;
      b *= 1.0e-10;
      bip *= 1.0e-10;
    }
    if (j == n) ans = bip;
  }
  ans *= bi0(x) / b;
  return (x < 0.0 && (n & 1)) ? -ans : ans;
}

int gaussian_taps(double variance, double max_error, int64_t max_width, Taps* t) {
  std::vector<double> c;
  const double et = std::exp(-variance), cap = 1.0 - max_error;
  double sum = 0.0;
  c.push_back(et * bi0(variance));
  sum += c[0];
  c.push_back(et * bi1(variance));
  sum += 2.0 * c[1];
  for (int i = 2; sum < cap; ++i) {
    c.push_back(et * bin(i, variance));
    sum += 2.0 * c[i];
    if (c[i] < sum * std::numeric_limits<double>::epsilon()) break;
    if (static_cast<int64_t>(c.size()) > max_width) break;
  }
  if (static_cast<int>(c.size()) > kMaxTaps)
    return fail(ANTS_ERR_UNSUPPORTED, "Gaussian kernel radius %zu exceeds %d", c.size() - 1, kMaxTaps - 1);
  std::memset(t, 0, sizeof(*t));
  t->r = static_cast<int>(c.size()) - 1;
  for (size_t i = 0; i < c.size(); ++i) t->c[i] = static_cast<float>(c[i] / sum);
  return ANTS_OK;
}

// ---------------------------------------------------------------------------
// Image geometry: point = org + dir * (index .* sp) (itk::ImageBase).
struct HostGrid {
  int D;
  int64_t n[3];
  double sp[3], org[3];
  double dir[9];  // 3x3 row-major, identity-padded
  bool oblique;
  int64_t count() const { return n[0] * n[1] * n[2]; }
};


int to_host_grid(const ants_image_desc* d, HostGrid* g) {
  if (!d) return fail(ANTS_ERR_INVALID, "null image descriptor");
  if (d->dim != 2 && d->dim != 3) return fail(ANTS_ERR_INVALID, "dim must be 2 or 3 (got %d)", d->dim);
  g->D = d->dim;
  // direction: row-major dim x dim, all zero meaning identity; orthonormal
  const int D = d->dim;
  bool all_zero = true;
  for (int k = 0; k < D * D; ++k)
    if (d->direction[k] != 0.0) all_zero = false;
  g->oblique = false;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const double id = a == b ? 1.0 : 0.0;
      const double v = (a < D && b < D && !all_zero) ? d->direction[a * D + b] : id;
      g->dir[a * 3 + b] = v;
      if (v != id) g->oblique = true;
    }
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      double dot = 0.0;
      for (int k = 0; k < D; ++k) dot += g->dir[k * 3 + a] * g->dir[k * 3 + b];
      if (std::fabs(dot - (a == b ? 1.0 : 0.0)) > 1e-6)
        return fail(ANTS_ERR_UNSUPPORTED, "image direction must be orthonormal");
    }
  for (int k = 0; k < 3; ++k) {
    g->n[k] = k < d->dim ? d->size[k] : 1;
    g->sp[k] = k < d->dim ? d->spacing[k] : 1.0;
    g->org[k] = k < d->dim ? d->origin[k] : 0.0;
    if (g->n[k] < 1) return fail(ANTS_ERR_INVALID, "image size must be positive");
    if (!(g->sp[k] > 0.0)) return fail(ANTS_ERR_INVALID, "image spacing must be positive");
  }
  if (g->count() > (static_cast<int64_t>(1) << 31) - 1) return fail(ANTS_ERR_UNSUPPORTED, "image too large");
  return ANTS_OK;
}

GridDev to_dev(const HostGrid& h) {
  GridDev g{};
  g.D = h.D;
  for (int k = 0; k < 3; ++k) {
    g.n[k] = static_cast<int>(h.n[k]);
    g.sp[k] = static_cast<float>(h.sp[k]);
    g.isp[k] = h.sp[k] > 0.0 ? static_cast<float>(1.0 / h.sp[k]) : 0.f;
    g.org[k] = static_cast<float>(h.org[k]);
    g.dsp[k] = h.sp[k];
    g.dorg[k] = h.org[k];
  }
  g.oblique = h.oblique ? 1 : 0;
  for (int k = 0; k < 9; ++k) {
    g.dir[k] = static_cast<float>(h.dir[k]);
    g.ddir[k] = h.dir[k];
  }
  set_divisors(g);
  return g;
}

void to_desc(const HostGrid& h, ants_image_desc* d) {
  std::memset(d, 0, sizeof(*d));
  d->dim = h.D;
  for (int k = 0; k < 3; ++k) {
    d->size[k] = h.n[k];
    d->spacing[k] = h.sp[k];
    d->origin[k] = h.org[k];
  }
  for (int r = 0; r < h.D; ++r)
    for (int c = 0; c < h.D; ++c) d->direction[r * h.D + c] = h.dir[r * 3 + c];
}

// RegistrationHelper::CalculateShrinkFactorsPerDimension + ShrinkImageFilter
// output information (Examples/itkantsRegistrationHelper.hxx:16-39, 593-641).
HostGrid level_grid(const HostGrid& g, int factor) {
  int mi = 0;
  for (int d = 1; d < g.D; ++d)
    if (g.sp[mi] > g.sp[d]) mi = d;
  int f[3] = {1, 1, 1};
  f[mi] = factor;
  const double target = g.sp[mi] * factor;
  for (int d = 0; d < g.D; ++d) {
    if (d == mi) continue;
    double best = std::fabs(g.sp[d] * factor - target);
    int bf = factor;
    for (int k = factor - 1; k > 0; --k) {
      const double diff = std::fabs(g.sp[d] * k - target);
      if (diff <= best) {  // the smaller factor wins ties
        best = diff;
        bf = k;
      }
    }
    f[d] = bf;
  }
  HostGrid o = g;
  double delta[3] = {0.0, 0.0, 0.0};  // index-aligned offset of the output origin
  for (int d = 0; d < g.D; ++d) {
    o.sp[d] = g.sp[d] * f[d];
    o.n[d] = std::max<int64_t>(1, static_cast<int64_t>(std::floor(static_cast<double>(g.n[d]) / f[d])));
    const double in_centre = g.org[d] + (g.n[d] - 1) / 2.0 * g.sp[d];
    const double out_centre = g.org[d] + (o.n[d] - 1) / 2.0 * o.sp[d];
    o.org[d] = g.org[d] + (in_centre - out_centre);
    delta[d] = (g.n[d] - 1) / 2.0 * g.sp[d] - (o.n[d] - 1) / 2.0 * o.sp[d];
  }
  // with a direction the centres coincide in physical space: origin + dir * delta
  if (g.oblique)
    for (int r = 0; r < g.D; ++r) {
      double sum = g.org[r];
      for (int k = 0; k < g.D; ++k) sum += g.dir[r * 3 + k] * delta[k];
      o.org[r] = sum;
    }
  return o;
}

int validate_params(const ants_syn_params* p) {
  if (!p) return fail(ANTS_ERR_INVALID, "null parameters");
  if (p->dim != 2 && p->dim != 3) return fail(ANTS_ERR_INVALID, "dim must be 2 or 3");
  if (p->n_levels < 1 || p->n_levels > ANTS_MAX_LEVELS)
    return fail(ANTS_ERR_INVALID, "number of levels must be in 1..%d", ANTS_MAX_LEVELS);
  for (int l = 0; l < p->n_levels; ++l) {
    if (p->shrink[l] < 1) return fail(ANTS_ERR_INVALID, "shrink factors must be >= 1");
    if (p->iterations[l] < 0) return fail(ANTS_ERR_INVALID, "iterations must be >= 0");
    if (!(p->sigma[l] >= 0.0)) return fail(ANTS_ERR_INVALID, "smoothing sigmas must be >= 0");
  }
  if (p->cc_radius < 1 || p->cc_radius > 16) return fail(ANTS_ERR_INVALID, "CC radius must be in 1..16");
  if (!(p->learning_rate > 0.0)) return fail(ANTS_ERR_INVALID, "SyN gradient step must be > 0");
  if (p->convergence_window < 1) return fail(ANTS_ERR_INVALID, "convergence window must be >= 1");
  if (p->precision != ANTS_PRECISION_FP32 && p->precision != ANTS_PRECISION_BF16_PYRAMID)
    return fail(ANTS_ERR_INVALID, "unknown precision mode");
  if (p->model != ANTS_MODEL_SYN && p->model != ANTS_MODEL_BSPLINE_SYN)
    return fail(ANTS_ERR_INVALID, "unknown field model %d", p->model);
  if (p->model == ANTS_MODEL_BSPLINE_SYN) {
    if (p->spline_order < 1 || p->spline_order > 3) return fail(ANTS_ERR_UNSUPPORTED, "B-spline order must be 1..3");
    int zeros = 0;
    for (int d = 0; d < p->dim; ++d) {
      if (p->update_mesh[d] < 1) return fail(ANTS_ERR_INVALID, "BSplineSyN update mesh sizes must be >= 1");
      if (p->total_mesh[d] < 0) return fail(ANTS_ERR_INVALID, "BSplineSyN total mesh sizes must be >= 0");
      zeros += p->total_mesh[d] == 0;
      // the helper doubles the mesh at every level
      const int64_t lim = int64_t{1} << 16;
      if ((static_cast<int64_t>(p->update_mesh[d]) << (p->n_levels - 1)) > lim ||
          (static_cast<int64_t>(p->total_mesh[d]) << (p->n_levels - 1)) > lim)
        return fail(ANTS_ERR_INVALID, "BSplineSyN mesh sizes too large");
    }
    if (zeros != 0 && zeros != p->dim)
      return fail(ANTS_ERR_INVALID, "BSplineSyN total mesh must be zero on every axis or on none");
  }
  return ANTS_OK;
}

// WindowConvergenceMonitoringFunction (see oracle): window of the last W
// energies, running |E| total over the level, order-1 two-control-point
// B-spline by multilevel B-spline approximation, value = -slope at t = 1.
struct ConvergenceMonitor {
  int window;
  std::vector<double> e;
  double total = 0.0;
  explicit ConvergenceMonitor(int w) : window(w) {}
  void add(double v) {
    e.push_back(v);
    if (static_cast<int>(e.size()) > window) e.erase(e.begin());
    total += std::fabs(v);
  }
  double value() const {
    if (static_cast<int>(e.size()) < window || window < 2) return std::numeric_limits<double>::max();
    double num0 = 0, num1 = 0, den0 = 0, den1 = 0;
    for (int n = 0; n < window; ++n) {
      const double t = static_cast<double>(n) / (window - 1);
      const double v = e[n] / total;
      const double w0 = 1.0 - t, w1 = t, w2 = w0 * w0 + w1 * w1;
      num0 += w0 * w0 * (w0 * v / w2);
      num1 += w1 * w1 * (w1 * v / w2);
      den0 += w0 * w0;
      den1 += w1 * w1;
    }
    const double p0 = den0 > 0 ? num0 / den0 : 0.0, p1 = den1 > 0 ? num1 / den1 : 0.0;
    return -(p1 - p0);
  }
};

// Device buffer owned by the engine.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t b) {
    if (b <= bytes) return ANTS_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(ANTS_ERR_OOM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
    }
    bytes = b;
    return ANTS_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  float* f() const { return static_cast<float*>(p); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
};

// HIP-event timing of kernel classes on one stream.  Events are recorded
// around each launch of an enabled class and read once the iteration has
// published its statistics; the bytes of inversion launches are attributed afterwards
// from the per-step `computed` flags.  Launches of steps that found their
// inversion already converged count with their time and zero bytes, so the
// class's mean launch time is the mean over every dispatch of the kernel.
struct Profiler {
  unsigned int mask = 0;    // classes timed (a nonzero mask makes iterations launch eagerly)
  unsigned int active = 0;  // classes timed in the current iteration
  int every = 1;            // time iterations 1, 1 + every, 1 + 2 every, ... of each level
  hipStream_t s = nullptr;
  ants_syn_kernel_stats acc{};
  ants_syn_kernel_stats acc_level[ANTS_MAX_LEVELS]{};  // the same per pyramid level
  int level = 0;                                       // level of the spans being recorded
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Rec {
    int cls;
    hipEvent_t a, b;
    double bytes;
    int launches;
    int batches, j0, j1;  // inversion spans: batch bit mask, steps [j0, j1)
    int level;
  };
  std::vector<Rec> recs;
  hipEvent_t a_ = nullptr;

  bool on(int cls) const { return (active >> cls) & 1u; }
  hipEvent_t next() {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  void begin(int cls) {
    a_ = nullptr;
    if (!on(cls)) return;
    a_ = next();
    if (a_) (void)hipEventRecord(a_, s);
  }
  void end(int cls, double bytes, int launches = 1, int batches = 0, int j0 = 0, int j1 = 0) {
    if (!on(cls) || !a_) return;
    hipEvent_t b = next();
    if (!b) return;
    (void)hipEventRecord(b, s);
    recs.push_back(Rec{cls, a_, b, bytes, launches, batches, j0, j1, level});
    a_ = nullptr;
  }
  // stats: the iteration's statistics block (host copy); N voxels, D dims
  void collect(const IterStats* st, int64_t N, int D) {
    for (const Rec& r : recs) {
      double bytes = r.bytes;
      if (r.cls == ANTS_KC_INV) {
        bytes = 0.0;
        const int qs[2][2] = {{0, 2}, {1, 3}};
        for (int bt = 0; bt < 2; ++bt) {
          if (!((r.batches >> bt) & 1)) continue;
          for (int j = r.j0; j < r.j1; ++j)
            for (int k = 0; k < 2; ++k) {
            const RedSlot* S = st->inv[qs[bt][k]];
            const bool upd = j >= 1 && S[j - 1].computed;
            const bool nrm = S[j].computed != 0;
            // algorithmic bytes of an inversion with K norm passes:
            // (3K + 1) field arrays (read inv, gather field, write inv per
            // step, plus the first norm pass); the stored residual this
            // kernel also moves is overhead, not algorithmic
            if (upd)
              bytes += 3.0 * D * 4.0 * N;
            else if (nrm)
              bytes += 1.0 * D * 4.0 * N;
          }
        }
      }
      float ms = 0.f;
      // the span ended before the statistics were published: complete on
      // the device; the runtime may not have noted it yet
      (void)hipEventSynchronize(r.b);
      const hipError_t e = hipEventElapsedTime(&ms, r.a, r.b);
      if (e != hipSuccess) {
        (void)hipGetLastError();  // not sticky: a later launch check must not see it
        if (std::getenv("ANTS_SYN_VERBOSE"))
          std::fprintf(stderr, "ants_syn: no time for class %d span (%s)\n", r.cls, hipGetErrorString(e));
        continue;
      }
      acc.ms[r.cls] += ms;
      acc.launches[r.cls] += r.launches;
      acc.bytes[r.cls] += bytes;
      ants_syn_kernel_stats& al = acc_level[r.level];
      al.ms[r.cls] += ms;
      al.launches[r.cls] += r.launches;
      al.bytes[r.cls] += bytes;
    }
    recs.clear();
    used = 0;
  }
  ~Profiler() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// A captured iteration: the graph plus the buffer rotation it leaves behind
// and the profiler records (events) it contains.
struct GraphKey {
  uint64_t setup;  // hash of the image geometry and stage parameters
  int level;
  const void* p[8];
  unsigned int prof_mask;
  int inv_launches;
  bool operator<(const GraphKey& o) const {
    if (setup != o.setup) return setup < o.setup;
    if (level != o.level) return level < o.level;
    if (inv_launches != o.inv_launches) return inv_launches < o.inv_launches;
    if (prof_mask != o.prof_mask) return prof_mask < o.prof_mask;
    return std::lexicographical_compare(p, p + 8, o.p, o.p + 8);
  }
};

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
  const unsigned char* b = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  float* nst[4];
  float* nwk[4];
};

// Waits until the iteration enqueued on `s` has published its statistics
// (k_publish_stats sets *flag in pinned host memory).  Spinning on the flag
// returns a few microseconds sooner than a stream synchronisation; the
// stream is polled now and then so that a failed launch ends the wait.
static int wait_published(const uint32_t* flag, hipStream_t s) {
  const volatile uint32_t* f = flag;
  for (uint64_t spin = 1;; ++spin) {
    if (*f) break;
    if ((spin & 255) == 0) {
      const hipError_t q = hipStreamQuery(s);
      if (q == hipSuccess) {
        if (*f) break;
        return fail(ANTS_ERR_HIP, "iteration finished without publishing its statistics");
      }
      if (q != hipErrorNotReady) return fail(ANTS_ERR_HIP, "iteration failed: %s", hipGetErrorString(q));
      if (spin > 4096) std::this_thread::yield();
    }
    __builtin_ia32_pause();
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return ANTS_OK;
}

// Initial transforms (fixed side, moving side) composed into y = A p + o;
// on[k] false: identity.
struct InitialXforms {
  bool on[2] = {false, false};
  double A[2][9] = {};
  double o[2][3] = {};
  // a list holding a displacement field is kept as it is (the warp then
  // applies it transform by transform); n[k] entries in list[k]
  bool general[2] = {false, false};
  int n[2] = {0, 0};
  ants_transform list[2][ANTS_MAX_TRANSFORMS - 1] = {};
};

namespace {
int geom_from_desc(const ants_image_desc* d, GeomDev* g);
int xform_to_dev(const ants_transform& T, int D, XformDev* X);
}  // namespace

// Composes each list (the point passes through list[0] first) with
// MatrixOffsetTransformBase's y = B (x - c) + t + c, in double.
int compose_initial(int D, const ants_transform* fixed, int n_fixed, const ants_transform* moving, int n_moving,
                    InitialXforms* out) {
  const ants_transform* lists[2] = {fixed, moving};
  const int counts[2] = {n_fixed, n_moving};
  InitialXforms x;
  for (int k = 0; k < 2; ++k) {
    const int n = counts[k];
    if (n < 0 || n > ANTS_MAX_TRANSFORMS || (n > 0 && !lists[k]))
      return fail(ANTS_ERR_INVALID, "initial transform count must be 0..%d", ANTS_MAX_TRANSFORMS);
    double Ac[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, oc[3] = {0, 0, 0};
    for (int t = 0; t < n; ++t) {
      const ants_transform& T = lists[k][t];
      if (T.kind == ANTS_TRANSFORM_IDENTITY) continue;
      if (T.kind == ANTS_TRANSFORM_DISPLACEMENT) {
        // checked here (geometry, buffer); the list is kept whole below
        XformDev X;
        ANTS_TRY(xform_to_dev(T, D, &X));
        x.general[k] = true;
        continue;
      }
      if (T.kind != ANTS_TRANSFORM_AFFINE) return fail(ANTS_ERR_INVALID, "unknown transform kind %d", T.kind);
      double B[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, ob[3] = {0, 0, 0};
      for (int r = 0; r < D; ++r) {
        double bc = 0.0;
        for (int c = 0; c < D; ++c) {
          B[r * 3 + c] = T.matrix[r * D + c];
          bc += T.matrix[r * D + c] * T.center[c];
        }
        ob[r] = T.translation[r] + T.center[r] - bc;
      }
      // running composite: y = B (Ac p + oc) + ob
      double An[9], on[3];
      for (int r = 0; r < 3; ++r) {
        double v = ob[r];
        for (int c = 0; c < 3; ++c) {
          double a = 0.0;
          for (int m = 0; m < 3; ++m) a += B[r * 3 + m] * Ac[m * 3 + c];
          An[r * 3 + c] = a;
          v += B[r * 3 + c] * oc[c];
        }
        on[r] = v;
      }
      std::memcpy(Ac, An, sizeof(Ac));
      std::memcpy(oc, on, sizeof(oc));
      x.on[k] = true;
    }
    std::memcpy(x.A[k], Ac, sizeof(Ac));
    std::memcpy(x.o[k], oc, sizeof(oc));
    if (x.general[k]) {
      // the displaced point goes through the list after the SyN field: one
      // slot of the warp's transform list is taken by that field
      if (n > ANTS_MAX_TRANSFORMS - 1)
        return fail(ANTS_ERR_UNSUPPORTED, "at most %d initial transforms with a displacement field",
                    ANTS_MAX_TRANSFORMS - 1);
      x.n[k] = n;
      for (int t = 0; t < n; ++t) x.list[k][t] = lists[k][t];
    }
  }
  *out = x;
  return ANTS_OK;
}

struct ants_syn_handle {
  ants_syn_params P{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  DevBuf F, M, Fs, Ms, tmp, Fw, Mw, Fb, Mb;  // Fb, Mb: BF16 pyramid images
  DevBuf fld[8];  // 4 middle fields, 4 work fields
  DevBuf blend;   // 2 fields: the separable smoothing's middle pass when a 3-D variance is below 0.5
  // BSplineSyN: fit tables of every level (update and total fields; host
  // copy kept for the asynchronous upload) and the lattices of the update
  // fits (front end) and total fits (back end), two fields each
  DevBuf bsp_tab, bsp_upd, bsp_tot, bsp_part;  // bsp_part: max-norm partials (front end, back end)
  // the observer's full-scale CC (--print-similarity-measure-interval)
  int fscc_interval = 0;
  double fscc_last = 0.0;  // of the latest iteration; 0 when not computed
  DevBuf fscc_buf;
  std::vector<unsigned char> bsp_host;
  DevBuf stats, partials_cc, partials_inv, masks_inv, minmax;
  IterStats* host_stats = nullptr;  // pinned
  uint32_t* host_flag = nullptr;    // pinned, coherent: set when an iteration's statistics are published
  ants_syn_stats run_stats{};
  std::vector<ants_syn_iteration_record> trace;  // the last run's iteration records
  HostGrid last_grid{};
  bool have_state = false;
  float* state[4] = {nullptr, nullptr, nullptr, nullptr};  // f2m fwd, f2m inv, m2m fwd, m2m inv
  Profiler prof;
  std::map<GraphKey, GraphEntry> graphs;
  bool use_graphs = true;
  // host-buffer entry: device copies of the inputs and outputs (grow-only)
  // and two pinned staging chunks for the transfers
  DevBuf io_f, io_m, io_fwd, io_inv;
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  // second stream and events of the pipelined iteration loop (created on first use)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_main = nullptr, ev_front = nullptr, ev_inv0 = nullptr;
  int inv_launches = kInvLaunchesDefault;  // starting value per level
  bool inv_launches_fixed = false;          // set by ANTS_SYN_INV_LAUNCHES: no adaptation
  std::vector<const void*> graph_signature;  // buffers + geometry the cached graphs were built for
  // initial transforms (fixed side, moving side) composed into y = A p + o;
  // on[k] false: identity
  bool init_on[2] = {false, false};
  double init_A[2][9] = {};
  double init_o[2][3] = {};
  bool init_general[2] = {false, false};
  int init_n[2] = {0, 0};
  ants_transform init_list[2][ANTS_MAX_TRANSFORMS - 1] = {};

  // restored state (F2M fwd, F2M inv, M2M fwd, M2M inv on restore_grid),
  // resampled onto the level-0 grid at the start of a run
  DevBuf restore_buf;
  HostGrid restore_grid{};
  bool restore_on = false;

  void set_initial(const InitialXforms& x) {
    for (int k = 0; k < 2; ++k) {
      init_on[k] = x.on[k];
      std::memcpy(init_A[k], x.A[k], sizeof(init_A[k]));
      std::memcpy(init_o[k], x.o[k], sizeof(init_o[k]));
      init_general[k] = x.general[k];
      init_n[k] = x.n[k];
      for (int t = 0; t < x.n[k]; ++t) init_list[k][t] = x.list[k][t];
    }
  }

  void clear_graphs() {
    for (auto& kv : graphs)
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    graphs.clear();
  }

  // device bytes the handle's grow-only workspace holds
  int64_t device_bytes() const {
    const DevBuf* all[] = {&F,      &M,           &Fs,          &Ms,        &tmp,    &Fw,          &Mw,
                           &Fb,     &Mb,          &stats,       &partials_cc, &partials_inv, &masks_inv,
                           &minmax, &restore_buf, &io_f,        &io_m,      &io_fwd, &io_inv, &blend,
                           &bsp_tab, &bsp_upd,    &bsp_tot,     &bsp_part,  &fscc_buf};
    int64_t b = 0;
    for (const DevBuf* x : all) b += static_cast<int64_t>(x->bytes);
    for (const auto& x : fld) b += static_cast<int64_t>(x.bytes);
    return b;
  }
  // frees the workspace (and the graphs that refer to it); the restored
  // state, set by the caller, is kept
  void release_workspace() {
    clear_graphs();
    graph_signature.clear();
    DevBuf* all[] = {&F,     &M,           &Fs,          &Ms,        &tmp,    &Fw,   &Mw,   &Fb,     &Mb,
                     &stats, &partials_cc, &partials_inv, &masks_inv, &minmax, &io_f, &io_m, &io_fwd, &io_inv,
                     &blend, &bsp_tab, &bsp_upd, &bsp_tot, &bsp_part, &fscc_buf};
    for (DevBuf* b : all) b->release();
    for (auto& b : fld) b.release();
    have_state = false;
  }

  ~ants_syn_handle() {
    release_workspace();
    restore_buf.release();
    if (host_stats) (void)hipHostFree(host_stats);
    if (host_flag) (void)hipHostFree(host_flag);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    for (hipEvent_t e : {ev_main, ev_front, ev_inv0, pin_ev[0], pin_ev[1]})
      if (e) (void)hipEventDestroy(e);
    for (void* q : pin)
      if (q) (void)hipHostFree(q);
    if (stream2) (void)hipStreamDestroy(stream2);
  }

  int preprocess(const float* in, float* out, int64_t n);
  int smooth_image(const HostGrid& g, const float* in, float* out, double sigma);
  int smooth_fields(const HostGrid& g, int np, float* const* in, float* const* work, float* const* out,
                    double variance, unsigned int* maxbits);
  int run(const ants_image_desc* fd, const float* d_fixed, const ants_image_desc* md, const float* d_moving,
          float* d_out_fwd, float* d_out_inv, ants_syn_iteration_cb cb, void* user);
  int staged_upload(void* d_dst, const void* host_src, size_t bytes);
  int staged_download(void* host_dst, const void* d_src, size_t bytes);
};

namespace {

// WarpMap of (level grid gl) -> [x + u] -> [y = A p + o] -> continuous index
// in image grid gi, formed in double: with K = diag(1/sp_i) dir_i^T,
// G = K A dir_l diag(sp_l), Q = K A, h = K (A org_l + o - org_i).
WarpMap make_warp_map(const HostGrid& gl, const double* A, const double* o, const HostGrid& gi) {
  double K[9], KA[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) K[r * 3 + c] = gi.dir[c * 3 + r] / gi.sp[r];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) t += K[r * 3 + k] * A[k * 3 + c];
      KA[r * 3 + c] = t;
    }
  WarpMap m{};
  double p0[3];
  for (int r = 0; r < 3; ++r) {
    double t = o[r] - gi.org[r];
    for (int k = 0; k < 3; ++k) t += A[r * 3 + k] * gl.org[k];
    p0[r] = t;
  }
  for (int r = 0; r < 3; ++r) {
    double h = 0.0;
    for (int k = 0; k < 3; ++k) h += K[r * 3 + k] * p0[k];
    m.h[r] = static_cast<float>(h);
    for (int c = 0; c < 3; ++c) {
      m.Q[r * 3 + c] = static_cast<float>(KA[r * 3 + c]);
      double g = 0.0;
      for (int k = 0; k < 3; ++k) g += KA[r * 3 + k] * gl.dir[k * 3 + c];
      m.G[r * 3 + c] = static_cast<float>(g * gl.sp[c]);
    }
  }
  return m;
}

// Device-side PreprocessImage: min/max reduction, host computes the window
// exactly as the CPU path does, then one elementwise pass.
int preprocess_on(hipStream_t s, DevBuf& mm, const float* in, float* out, int64_t n, uint32_t* host_pair) {
  ANTS_TRY(mm.ensure(2 * sizeof(uint32_t)));
  uint32_t* d = static_cast<uint32_t*>(mm.p);
  launch_fill_u32(d, 0xffffffffu, 1, s);
  launch_fill_u32(d + 1, 0u, 1, s);
  launch_minmax(in, n, d, s);
  HIP_TRY(hipMemcpyAsync(host_pair, d, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  auto decode = [](uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  const double mn = decode(host_pair[0]), mx = decode(host_pair[1]);
  const float lo = static_cast<float>(mn);
  const float hi = static_cast<float>(mx + (mx - mn) / 2560.0);
  const double scale = hi > lo ? 1.0 / (static_cast<double>(hi) - static_cast<double>(lo)) : 0.0;
  const double shift = -static_cast<double>(lo) * scale;
  launch_window(in, out, n, lo, hi, scale, shift, s);
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}

int smooth_image_on(hipStream_t s, const HostGrid& g, const float* in, float* out, float* tmp, double sigma,
                    bool in_mm) {
  const GridDev gd = to_dev(g);
  const int64_t n = g.count();
  if (!(sigma > 0.0)) {
    if (out != in) HIP_TRY(hipMemcpyAsync(out, in, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    return ANTS_OK;
  }
  // one fused 3-D pass when every axis has the same taps (voxel sigmas or
  // isotropic spacing)
  if (g.D == 3 && (!in_mm || (g.sp[0] == g.sp[1] && g.sp[1] == g.sp[2]))) {
    double var = sigma * sigma;
    if (in_mm) var /= g.sp[0] * g.sp[0];
    Taps t;
    ANTS_TRY(gaussian_taps(var, 0.01, 32, &t));
    if (launch_smooth3d_image(gd, in, out, t, s)) {
      HIP_TRY(hipGetLastError());
      return ANTS_OK;
    }
  }
  // ping-pong so that the last pass lands in `out`
  const float* src = in;
  for (int d = 0; d < g.D; ++d) {
    double var = sigma * sigma;
    if (in_mm) var /= g.sp[d] * g.sp[d];
    Taps t;
    ANTS_TRY(gaussian_taps(var, 0.01, 32, &t));
    float* dst = ((g.D - 1 - d) % 2 == 0) ? out : tmp;
    launch_gauss_axis(gd, 1, src, dst, t, d, s);
    src = dst;
  }
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}

// SyN GaussianSmoothDisplacementField for np fields; maxbits (optional)
// receives each result's max |u/spacing|.  work/out may alias in.
// Grids below this size take the separable passes instead of the fused one.
#ifndef ANTS_SMOOTH_FUSED_MIN_VOXELS
#define ANTS_SMOOTH_FUSED_MIN_VOXELS 0
#endif
constexpr int64_t kSmoothFusedMinVoxels = ANTS_SMOOTH_FUSED_MIN_VOXELS;

// `extra` (np fields; needed only for 3-D fields with 0 < variance < 0.5 off
// the fused path) keeps the middle pass out of `in`, which the blend reads.
int smooth_fields_on(hipStream_t s, const HostGrid& g, int np, float* const* in, float* const* work,
                     float* const* out, double variance, unsigned int* maxbits, Profiler* prof = nullptr,
                     float* const* extra = nullptr) {
  const GridDev gd = to_dev(g);
  const int D = g.D;
  const double pass_bytes = 2.0 * D * 4.0 * static_cast<double>(g.count());
  if (!(variance > 0.0)) {
    // plain copy (no boundary handling); the max norm still comes from a pass
    Taps t{};
    t.r = 0;
    t.c[0] = 1.f;
    if (prof) prof->begin(ANTS_KC_SMOOTH);
    launch_gauss_field_final(gd, np, in, in, out, t, 0, 1.f, 0.f, false, maxbits, s);
    if (prof) prof->end(ANTS_KC_SMOOTH, np * pass_bytes);
    HIP_TRY(hipGetLastError());
    return ANTS_OK;
  }
  // one fused pass when the kernel radius and dimension allow (taps per axis
  // differ only through MaxWidth = size[d], checked below)
  if (D == 3 && g.count() >= kSmoothFusedMinVoxels) {
    Taps tx, ty, tz;
    ANTS_TRY(gaussian_taps(variance, 0.001, g.n[0], &tx));
    ANTS_TRY(gaussian_taps(variance, 0.001, g.n[1], &ty));
    ANTS_TRY(gaussian_taps(variance, 0.001, g.n[2], &tz));
    const bool same = tx.r == ty.r && ty.r == tz.r &&
                      std::memcmp(tx.c, ty.c, sizeof(float) * (tx.r + 1)) == 0 &&
                      std::memcmp(ty.c, tz.c, sizeof(float) * (ty.r + 1)) == 0;
    const float w1f = variance < 0.5 ? static_cast<float>(1.0 - variance / 0.5) : 1.f;
    if (same && out[0] != in[0]) {
      const float* cin[4];
      for (int k = 0; k < np; ++k) cin[k] = in[k];
      if (prof) prof->begin(ANTS_KC_SMOOTH);
      const bool ok = launch_smooth3d(gd, np, cin, cin, out, tx, w1f, 1.f - w1f, true, maxbits, s);
      if (prof) prof->end(ANTS_KC_SMOOTH, np * pass_bytes);
      if (ok) {
        HIP_TRY(hipGetLastError());
        return ANTS_OK;
      }
    }
  }
  // axes 0..D-2 ping-pong between work and in (or, when the blend needs the
  // input of a 3-D field, between work and extra); the last axis lands in out
  const float w1 = variance < 0.5 ? static_cast<float>(1.0 - variance / 0.5) : 1.f;
  const float w2 = 1.f - w1;
  const bool keep_in = w2 != 0.f && D % 2 == 1;
  if (keep_in && !extra)
    return fail(ANTS_ERR_UNSUPPORTED, "update/total variance below 0.5 needs a blend buffer here");
  const float* cur[4];
  for (int k = 0; k < np; ++k) cur[k] = in[k];
  for (int d = 0; d < D - 1; ++d) {
    Taps t;
    ANTS_TRY(gaussian_taps(variance, 0.001, g.n[d], &t));
    for (int k = 0; k < np; ++k) {
      float* dst = (d % 2 == 0) ? work[k] : (keep_in ? extra[k] : in[k]);
      if (prof) prof->begin(ANTS_KC_SMOOTH);
      launch_gauss_axis(gd, D, cur[k], dst, t, d, s);
      if (prof) prof->end(ANTS_KC_SMOOTH, pass_bytes);
      cur[k] = dst;
    }
  }
  Taps t;
  ANTS_TRY(gaussian_taps(variance, 0.001, g.n[D - 1], &t));
  const float* orig[4];
  for (int k = 0; k < np; ++k) orig[k] = in[k];
  if (prof) prof->begin(ANTS_KC_SMOOTH);
  launch_gauss_field_final(gd, np, cur, orig, out, t, D - 1, w1, w2, true, maxbits, s);
  if (prof) prof->end(ANTS_KC_SMOOTH, np * pass_bytes);
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}

// ---- BSplineSyN: tables and passes of the B-spline field fit ----
// BSplineSmoothDisplacementField of BSplineSyN (ITK's DisplacementField-
// ToBSplineImageFilter with one level of BSplineScatteredDataPointSetTo-
// ImageFilter, stationary boundary, no confidence image; configured at
// Examples/itkantsRegistrationHelper.hxx:2222-2340).  Voxel i of an axis of n
// voxels and S spans sits at p = i S / (n - 1) (p = S moved to S - eps); its
// span is floor(p) and its basis values are ITK's centred BSplineKernel-
// Function of order q at t - j + (q - 1) / 2, t = p - floor(p), j = 0..q.

double bsp_basis(int q, double u) {
  u = std::fabs(u);
  if (q == 1) return u < 1.0 ? 1.0 - u : 0.0;
  if (q == 2) return u < 0.5 ? 0.75 - u * u : (u < 1.5 ? 0.5 * (1.5 - u) * (1.5 - u) : 0.0);
  if (u < 1.0) return (4.0 + u * u * (3.0 * u - 6.0)) / 6.0;
  return u < 2.0 ? (2.0 - u) * (2.0 - u) * (2.0 - u) / 6.0 : 0.0;
}

// The filter's epsilon for points at the domain end: 100 FLT_EPSILON, times
// ten while float(S) - eps still rounds to float(S) (S: largest span count).
double bsp_end_eps(int max_spans) {
  volatile float eps = 100.0f * std::numeric_limits<float>::epsilon();
  const volatile float S = static_cast<float>(max_spans);
  while (static_cast<float>(S - eps) == static_cast<float>(S)) eps = eps * 10.0f;
  return static_cast<double>(eps);
}

// Layout of one fit's tables inside a table blob (byte offsets per axis:
// span, bw, lo, goff, gw, A, Ai) and the sizes of its passes.
struct BspLayout {
  int D = 0, q = 3;
  int n[3] = {1, 1, 1}, m[3] = {1, 1, 1};
  int gtot[3] = {0, 0, 0};
  size_t off[3][7] = {};
  // floats per field of the intermediate lattices (x-fitted, xy-fitted, full)
  int64_t field_floats() const {
    return (static_cast<int64_t>(m[0]) * n[1] * n[2] + static_cast<int64_t>(m[0]) * m[1] * n[2] +
            static_cast<int64_t>(m[0]) * m[1] * m[2]) *
           D;
  }
};

template <class T>
size_t blob_put(std::vector<unsigned char>& blob, const std::vector<T>& v) {
  size_t o = (blob.size() + 15) & ~size_t(15);
  blob.resize(o + v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(blob.data() + o, v.data(), v.size() * sizeof(T));
  return o;
}

// Appends the tables of a fit of grid g (mesh spans per axis, order q) to blob.
int bsp_tables(const HostGrid& g, const int mesh[3], int q, std::vector<unsigned char>& blob, BspLayout* out) {
  BspLayout L;
  L.D = g.D;
  L.q = q;
  int max_spans = 0;
  for (int a = 0; a < g.D; ++a) max_spans = std::max(max_spans, mesh[a]);
  const double eps = bsp_end_eps(max_spans);
  for (int a = 0; a < g.D; ++a) {
    const int64_t n = g.n[a];
    const int S = mesh[a];
    if (n < 2) return fail(ANTS_ERR_UNSUPPORTED, "B-spline field fits need at least 2 voxels per axis");
    if (S < 1) return fail(ANTS_ERR_INVALID, "B-spline mesh sizes must be >= 1");
    const int m = S + q;
    L.n[a] = static_cast<int>(n);
    L.m[a] = m;
    std::vector<int> span(n), lo(m), goff(m + 1, 0);
    std::vector<float> bw(4 * n, 0.f), gw;
    std::vector<double> bd(4 * n, 0.0), fd(4 * n, 0.0), A(m, 0.0), Ai(m, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      double p = static_cast<double>(i) * S / static_cast<double>(n - 1);
      if (std::fabs(p - S) <= eps) p = S - eps;
      const int si = static_cast<int>(std::floor(p));
      const double t = p - si;
      span[i] = si;
      double sd = 0.0;
      for (int j = 0; j <= q; ++j) {
        bd[4 * i + j] = bsp_basis(q, t - j + 0.5 * (q - 1));
        sd += bd[4 * i + j] * bd[4 * i + j];
      }
      const bool interior = i > 0 && i < n - 1;
      for (int j = 0; j <= q; ++j) {
        const double b = bd[4 * i + j];
        bw[4 * i + j] = static_cast<float>(b);
        fd[4 * i + j] = interior ? b * b * b / sd : 0.0;
      }
    }
    // control point k: voxels whose span lies in [k - q, k] (contiguous:
    // spans do not decrease along the axis)
    for (int k = 0; k < m; ++k) {
      int l0 = 0;
      while (l0 < n && span[l0] < k - q) ++l0;
      int h0 = static_cast<int>(n) - 1;
      while (h0 >= 0 && span[h0] > k) --h0;
      lo[k] = l0;
      goff[k] = static_cast<int>(gw.size());
      for (int i = l0; i <= h0; ++i) {
        const double b = bd[4 * i + (k - span[i])];
        gw.push_back(static_cast<float>(fd[4 * i + (k - span[i])]));
        A[k] += b * b;
        if (i > 0 && i < n - 1) Ai[k] += b * b;
      }
    }
    goff[m] = static_cast<int>(gw.size());
    L.gtot[a] = goff[m];
    L.off[a][0] = blob_put(blob, span);
    L.off[a][1] = blob_put(blob, bw);
    L.off[a][2] = blob_put(blob, lo);
    L.off[a][3] = blob_put(blob, goff);
    L.off[a][4] = blob_put(blob, gw);
    L.off[a][5] = blob_put(blob, A);
    L.off[a][6] = blob_put(blob, Ai);
  }
  // the passes index lattice planes and voxels with 32-bit integers
  if (static_cast<int64_t>(L.m[0]) * L.m[1] * L.D >= (int64_t{1} << 31) ||
      static_cast<int64_t>(L.n[0]) * L.n[1] * L.n[2] >= (int64_t{1} << 31))
    return fail(ANTS_ERR_UNSUPPORTED, "B-spline lattice too large");
  *out = L;
  return ANTS_OK;
}

BspDev bsp_dev(const BspLayout& L, const void* dbase) {
  const unsigned char* b = static_cast<const unsigned char*>(dbase);
  BspDev T{};
  T.D = L.D;
  T.q = L.q;
  for (int a = 0; a < L.D; ++a) {
    BspAxisDev& X = T.ax[a];
    X.span = reinterpret_cast<const int*>(b + L.off[a][0]);
    X.bw = reinterpret_cast<const float*>(b + L.off[a][1]);
    X.lo = reinterpret_cast<const int*>(b + L.off[a][2]);
    X.goff = reinterpret_cast<const int*>(b + L.off[a][3]);
    X.gw = reinterpret_cast<const float*>(b + L.off[a][4]);
    X.A = reinterpret_cast<const double*>(b + L.off[a][5]);
    X.Ai = reinterpret_cast<const double*>(b + L.off[a][6]);
    X.n = L.n[a];
    X.m = L.m[a];
    X.gtot = L.gtot[a];
  }
  return T;
}

// The B-spline fit of np fields: fit along x, y (, z; the last fit pass
// divides by the weights), then evaluate along z (3-D) and along y and x in
// one fused pass (the separate y and x passes where a row's tables exceed
// LDS); the x evaluation writes `out` and each field's max |u / spacing|
// into maxbits[k] through the per-block partials in `part`.  buf[k] holds
// field_floats() floats of lattices for field k.  in and out may not alias.
int bsp_fields_on(hipStream_t s, const HostGrid& g, const BspDev& T, const BspLayout& L, int np,
                  const float* const* in, float* const* out, float* const* buf, unsigned int* maxbits, float* part,
                  Profiler* prof = nullptr) {
  const int D = g.D;
  const int n0 = L.n[0], n1 = L.n[1], n2 = D == 3 ? L.n[2] : 1;
  const int m0 = L.m[0], m1 = L.m[1], m2 = D == 3 ? L.m[2] : 1;
  float *t1[4], *t2[4], *lat[4];
  for (int k = 0; k < np; ++k) {
    t1[k] = buf[k];
    t2[k] = t1[k] + static_cast<int64_t>(m0) * n1 * n2 * D;
    lat[k] = t2[k] + static_cast<int64_t>(m0) * m1 * n2 * D;
  }
  const float sp[3] = {static_cast<float>(g.sp[0]), static_cast<float>(g.sp[1]), static_cast<float>(g.sp[2])};
  if (prof) prof->begin(ANTS_KC_SMOOTH);
  int launches = 0;
  const float* c1[4];
  const float* c2[4];
  for (int k = 0; k < np; ++k) {
    c1[k] = t1[k];
    c2[k] = t2[k];
  }
  const int d_in[3] = {n0, n1, n2};
  const int d_x[3] = {m0, n1, n2};
  launch_bsp_fit_axis(T, np, in, t1, d_in, 0, false, s);
  if (D == 3) {
    launch_bsp_fit_axis(T, np, c1, t2, d_x, 1, false, s);
    const int d_xy[3] = {m0, m1, n2};
    launch_bsp_fit_axis(T, np, c2, lat, d_xy, 2, true, s);
    launches += 3;
  } else {
    launch_bsp_fit_axis(T, np, c1, lat, d_x, 1, true, s);
    launches += 2;
  }
  const float* cl[4];
  for (int k = 0; k < np; ++k) cl[k] = lat[k];
  // evaluation: z (3-D) into t2, then y and x
  const float* ylat[4];
  if (D == 3) {
    const int d_l[3] = {m0, m1, m2};
    launch_bsp_eval_axis(T, np, cl, t2, d_l, 2, sp, nullptr, nullptr, s);  // (m0, m1, n2)
    launches += 1;
    for (int k = 0; k < np; ++k) ylat[k] = t2[k];
  } else {
    for (int k = 0; k < np; ++k) ylat[k] = lat[k];
  }
  if (launch_bsp_eval_yx(T, np, ylat, out, n2, sp, maxbits, part, s)) {
    launches += maxbits ? 2 : 1;
  } else {
    const int d_y[3] = {m0, m1, n2};
    launch_bsp_eval_axis(T, np, ylat, t1, d_y, 1, sp, nullptr, nullptr, s);  // (m0, n1, n2)
    launch_bsp_eval_axis(T, np, c1, out, d_x, 0, sp, maxbits, part, s);
    launches += maxbits ? 3 : 2;
  }
  if (prof) prof->end(ANTS_KC_SMOOTH, np * 2.0 * D * 4.0 * static_cast<double>(g.count()), launches);
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}

}  // namespace

int ants_syn_handle::run(const ants_image_desc* fd, const float* d_fixed, const ants_image_desc* md,
                         const float* d_moving, float* d_out_fwd, float* d_out_inv, ants_syn_iteration_cb cb,
                         void* user) {
  HostGrid gF, gM;
  ANTS_TRY(to_host_grid(fd, &gF));
  ANTS_TRY(to_host_grid(md, &gM));
  if (gF.D != P.dim || gM.D != P.dim) return fail(ANTS_ERR_INVALID, "image dimension does not match parameters");
  if (!d_fixed || !d_moving) return fail(ANTS_ERR_INVALID, "null image buffer");
  const int D = P.dim;
  const int64_t NF = gF.count(), NM = gM.count();
  std::vector<HostGrid> lg(P.n_levels);
  int64_t maxN = 0;
  for (int l = 0; l < P.n_levels; ++l) {
    lg[l] = level_grid(gF, P.shrink[l]);
    maxN = std::max(maxN, lg[l].count());
  }
  const size_t fbytes = static_cast<size_t>(maxN) * D * sizeof(float);
  ANTS_TRY(F.ensure(NF * sizeof(float)));
  ANTS_TRY(M.ensure(NM * sizeof(float)));
  ANTS_TRY(Fs.ensure(NF * sizeof(float)));
  ANTS_TRY(Ms.ensure(NM * sizeof(float)));
  ANTS_TRY(tmp.ensure(std::max(NF, NM) * sizeof(float)));
  const bool bf16 = P.precision == ANTS_PRECISION_BF16_PYRAMID;
  if (bf16) {
    ANTS_TRY(Fb.ensure(NF * sizeof(uint16_t)));
    ANTS_TRY(Mb.ensure(NM * sizeof(uint16_t)));
  }
  ANTS_TRY(Fw.ensure(maxN * sizeof(float)));
  ANTS_TRY(Mw.ensure(maxN * sizeof(float)));
  for (auto& b : fld) ANTS_TRY(b.ensure(fbytes));
  auto small_var = [](double v) { return v > 0.0 && v < 0.5; };
  const bool need_blend = D == 3 && (small_var(P.update_variance) || small_var(P.total_variance));
  if (need_blend) ANTS_TRY(blend.ensure(2 * fbytes));
  float* const blend_f[2] = {need_blend ? blend.f() : nullptr,
                             need_blend ? blend.f() + static_cast<size_t>(maxN) * D : nullptr};
  // BSplineSyN: the fit tables of every level (mesh << level, as the
  // helper's adaptors set it), uploaded once per run, and the lattices
  const bool bsp = P.model == ANTS_MODEL_BSPLINE_SYN;
  const bool bsp_total = bsp && P.total_mesh[0] > 0;
  std::vector<BspLayout> bsp_lu(P.n_levels), bsp_lt(P.n_levels);
  std::vector<size_t> bsp_bu(P.n_levels, 0), bsp_bt(P.n_levels, 0);
  int64_t bsp_fu = 0, bsp_ft = 0;  // floats per field of the update / total lattices
  if (bsp) {
    bsp_host.clear();
    for (int l = 0; l < P.n_levels; ++l) {
      int um[3] = {1, 1, 1}, tm[3] = {1, 1, 1};
      for (int d = 0; d < D; ++d) {
        um[d] = P.update_mesh[d] << l;
        tm[d] = P.total_mesh[d] << l;
      }
      ANTS_TRY(bsp_tables(lg[l], um, P.spline_order, bsp_host, &bsp_lu[l]));
      bsp_fu = std::max(bsp_fu, bsp_lu[l].field_floats());
      if (bsp_total) {
        ANTS_TRY(bsp_tables(lg[l], tm, P.spline_order, bsp_host, &bsp_lt[l]));
        bsp_ft = std::max(bsp_ft, bsp_lt[l].field_floats());
      }
    }
    ANTS_TRY(bsp_tab.ensure(bsp_host.size()));
    HIP_TRY(hipMemcpyAsync(bsp_tab.p, bsp_host.data(), bsp_host.size(), hipMemcpyHostToDevice, stream));
    ANTS_TRY(bsp_upd.ensure(2 * bsp_fu * sizeof(float)));
    if (bsp_total) ANTS_TRY(bsp_tot.ensure(2 * bsp_ft * sizeof(float)));
    ANTS_TRY(bsp_part.ensure(4 * bsp_partials_needed() * sizeof(float)));
  }
  if (fscc_interval > 0) {
    // two resampled images, two derivative fields (CC writes them), a slot and the CC partials
    const size_t pb = static_cast<size_t>(cc_partials_needed(to_dev(gF), 4, false)) * sizeof(Partial);
    ANTS_TRY(fscc_buf.ensure((2 * NF + 2 * NF * D) * sizeof(float) + sizeof(RedSlot) + pb));
  }
  ANTS_TRY(stats.ensure(2 * sizeof(IterStats)));  // two iteration parities
  int64_t ccp = 0, invp = 0, invw = 0;
  for (int l = 0; l < P.n_levels; ++l) {
    ccp = std::max(ccp, cc_partials_needed(to_dev(lg[l]), P.cc_radius, bf16));
    invp = std::max(invp, inv_partials_needed(to_dev(lg[l])));
    invw = std::max(invw, inv_mask_words(to_dev(lg[l])));
  }
  ANTS_TRY(partials_cc.ensure(2 * ccp * sizeof(Partial)));  // two iteration parities
  ANTS_TRY(partials_inv.ensure(4 * invp * sizeof(Partial)));  // 2 step parities x 2 problems
  ANTS_TRY(masks_inv.ensure(4 * invw * sizeof(uint64_t)));    // 2 step parities x 2 problems
  if (!host_stats)
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&host_stats), 2 * sizeof(IterStats), hipHostMallocDefault));
  if (!host_flag) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&host_flag), 256, hipHostMallocCoherent));

  hipStream_t s = stream;
  IterStats* dstats = static_cast<IterStats*>(stats.p);
  // cached iteration graphs bake in buffer addresses (cleared when any
  // buffer moves) and are keyed by the geometry and stage parameters
  uint64_t setup_hash = 0;
  {
    std::vector<const void*> sig = {F.p, M.p, Fs.p, Ms.p, Fw.p, Mw.p, Fb.p, Mb.p, stats.p, partials_cc.p,
                                    partials_inv.p, host_stats, host_flag, static_cast<const void*>(s)};
    for (auto& b : fld) sig.push_back(b.p);
    for (const DevBuf* b : {&bsp_tab, &bsp_upd, &bsp_tot, &bsp_part}) sig.push_back(b->p);
    if (sig != graph_signature || graphs.size() > 512) {
      clear_graphs();
      graph_signature = sig;
    }
    // field by field: struct padding is not hashed
    auto hg = [&](const HostGrid& g) {
      setup_hash = fnv1a(&g.D, sizeof(g.D), setup_hash);
      setup_hash = fnv1a(g.n, sizeof(g.n), setup_hash);
      setup_hash = fnv1a(g.sp, sizeof(g.sp), setup_hash);
      setup_hash = fnv1a(g.org, sizeof(g.org), setup_hash);
    };
    setup_hash = 1469598103934665603ull;
    hg(gF);
    hg(gM);
    const int ints[] = {P.dim, P.n_levels, P.sigma_in_mm, P.cc_radius, P.precision};
    setup_hash = fnv1a(ints, sizeof(ints), setup_hash);
    const uint64_t tg = tune_generation();  // launch shapes baked into captured graphs
    setup_hash = fnv1a(&tg, sizeof(tg), setup_hash);
    setup_hash = fnv1a(P.shrink, sizeof(P.shrink), setup_hash);
    setup_hash = fnv1a(P.sigma, sizeof(P.sigma), setup_hash);
    const double dbl[] = {P.learning_rate, P.update_variance, P.total_variance};
    setup_hash = fnv1a(dbl, sizeof(dbl), setup_hash);
    const int bints[] = {P.model,         P.update_mesh[0], P.update_mesh[1], P.update_mesh[2],
                         P.total_mesh[0], P.total_mesh[1],  P.total_mesh[2],  P.spline_order};
    setup_hash = fnv1a(bints, sizeof(bints), setup_hash);
    for (int k = 0; k < 2; ++k) {
      const int on = init_on[k] ? 1 : 0;
      setup_hash = fnv1a(&on, sizeof(on), setup_hash);
      if (on) {
        setup_hash = fnv1a(init_A[k], sizeof(init_A[k]), setup_hash);
        setup_hash = fnv1a(init_o[k], sizeof(init_o[k]), setup_hash);
      }
      const int gen = init_general[k] ? init_n[k] : -1;
      setup_hash = fnv1a(&gen, sizeof(gen), setup_hash);
      for (int t = 0; t < (init_general[k] ? init_n[k] : 0); ++t) {
        const ants_transform& T = init_list[k][t];
        setup_hash = fnv1a(&T.kind, sizeof(T.kind), setup_hash);
        setup_hash = fnv1a(T.matrix, sizeof(T.matrix), setup_hash);
        setup_hash = fnv1a(T.translation, sizeof(T.translation), setup_hash);
        setup_hash = fnv1a(T.center, sizeof(T.center), setup_hash);
        setup_hash = fnv1a(&T.d_field, sizeof(T.d_field), setup_hash);
        setup_hash = fnv1a(T.field.size, sizeof(T.field.size), setup_hash);
        setup_hash = fnv1a(T.field.spacing, sizeof(T.field.spacing), setup_hash);
        setup_hash = fnv1a(T.field.origin, sizeof(T.field.origin), setup_hash);
        setup_hash = fnv1a(T.field.direction, sizeof(T.field.direction), setup_hash);
      }
    }
    // Iterations launch kernel by kernel unless ANTS_SYN_GRAPHS=1: replaying
    // captured iteration graphs measured slower on this runtime (41.7 against
    // 41.1 ms per 192x224x192 pair; 69.9 against 73.7 pairs/s with eight
    // registrations in flight).  Profiled runs always launch eagerly: events
    // recorded by the event nodes of a replayed graph yield no elapsed times.
    const char* env = std::getenv("ANTS_SYN_GRAPHS");
    use_graphs = (env && env[0] == '1') && prof.mask == 0;
    // inversion steps launched per batch inside an iteration's sequence; a
    // batch that still has work after them is finished from the host
    const char* ks = std::getenv("ANTS_SYN_INV_LAUNCHES");
    inv_launches = ks ? std::atoi(ks) : kInvLaunchesDefault;
    inv_launches = std::min(std::max(inv_launches, 1), kInvMaxIter + 1);
    inv_launches_fixed = ks != nullptr;
  }
  // ANTS_SYN_PIPELINE=1 pipelines iterations over two streams (the next
  // iteration's front end beside this one's inversions; not with graph
  // replay, nor while the profiler times front-end kernel classes, whose
  // spans live on the main stream).  Same bits, but measured no faster:
  // 41.9 against 41.0 ms per 192x224x192 pair.  The overlapped kernels slow
  // each other down (at 96x112x96 CC 31 -> 63 us, smoothing 45 -> 70 us,
  // inversion steps 18 -> 35-49 us), so the small levels are bound by
  // per-CU issue, not by launch latency.  Off by default.
  bool pipelined = false;
  {
    const char* pe = std::getenv("ANTS_SYN_PIPELINE");
    const unsigned int front_classes = (1u << ANTS_KC_WARP) | (1u << ANTS_KC_CC) | (1u << ANTS_KC_SMOOTH);
    pipelined = !use_graphs && (prof.mask & front_classes) == 0 && (pe && pe[0] == '1');
    if (pipelined && !stream2) HIP_TRY(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    for (hipEvent_t* e : {&ev_main, &ev_front, &ev_inv0})
      if (pipelined && !*e) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  hipStream_t s2 = stream2;
  // whatever path leaves run(), nothing of it is still running on the second stream
  struct SecondStreamDrain {
    hipStream_t s2;
    ~SecondStreamDrain() {
      if (s2) (void)hipStreamSynchronize(s2);
    }
  } drain{pipelined ? s2 : nullptr};
  std::memset(&run_stats, 0, sizeof(run_stats));
  trace.clear();
  std::memset(&prof.acc, 0, sizeof(prof.acc));
  std::memset(prof.acc_level, 0, sizeof(prof.acc_level));
  prof.s = s;
  prof.recs.clear();
  prof.used = 0;
  have_state = false;

  // preprocessing (PreprocessImage) or plain copies
  uint32_t* pair = reinterpret_cast<uint32_t*>(host_stats);
  if (P.preprocess) {
    ANTS_TRY(preprocess_on(s, minmax, d_fixed, F.f(), NF, pair));
    ANTS_TRY(preprocess_on(s, minmax, d_moving, M.f(), NM, pair));
  } else {
    HIP_TRY(hipMemcpyAsync(F.p, d_fixed, NF * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(M.p, d_moving, NM * sizeof(float), hipMemcpyDeviceToDevice, s));
  }

  // field slots: state (f2m_f, f2m_i, m2m_f, m2m_i) and work (4)
  float* st[4] = {fld[0].f(), fld[1].f(), fld[2].f(), fld[3].f()};
  float* wk[4] = {fld[4].f(), fld[5].f(), fld[6].f(), fld[7].f()};
  const GridDev gFd = to_dev(gF), gMd = to_dev(gM);
  const float eps = std::numeric_limits<float>::epsilon();
  HostGrid prev{};
  const bool general = init_general[0] || init_general[1];
  const bool use_maps = !general && (init_on[0] || init_on[1]);
  for (int l = 0; l < P.n_levels; ++l) {
    const HostGrid& g = lg[l];
    const GridDev gd = to_dev(g);
    const int64_t N = g.count();
    const BspDev bsp_u_dev = bsp ? bsp_dev(bsp_lu[l], bsp_tab.p) : BspDev{};
    const BspDev bsp_t_dev = bsp_total ? bsp_dev(bsp_lt[l], bsp_tab.p) : BspDev{};
    WarpMap maps[2];
    // with a displacement field among the initial transforms each image is
    // warped by the antsApplyTransforms kernel through [the SyN inverse
    // middle field on the level grid, then the side's initial transforms]
    ApplyDev gw[2];
    GridDev gw_out{};
    // side k's image (0 fixed, 1 moving) sampled on grid `out` through [the
    // SyN inverse middle field on this level's grid, then the side's initial
    // transforms], as ResampleImageFilter with that composite does; the
    // field's buffer is set where the apply kernel is launched
    auto side_apply = [&](int k, const HostGrid& out, ApplyDev* A) -> int {
      ants_image_desc od, ld, id;
      to_desc(out, &od);
      to_desc(g, &ld);
      to_desc(k == 0 ? gF : gM, &id);
      std::memset(A, 0, sizeof(*A));
      ANTS_TRY(geom_from_desc(&od, &A->ref));
      ANTS_TRY(geom_from_desc(&id, &A->in));
      A->defv = 0.f;
      A->t[0].kind = ANTS_TRANSFORM_DISPLACEMENT;
      ANTS_TRY(geom_from_desc(&ld, &A->t[0].fg));
      int nt = 1;
      if (init_general[k]) {
        for (int t = 0; t < init_n[k]; ++t) ANTS_TRY(xform_to_dev(init_list[k][t], D, &A->t[nt++]));
      } else if (init_on[k]) {
        XformDev& X = A->t[nt++];
        X.kind = ANTS_TRANSFORM_AFFINE;
        for (int r = 0; r < 3; ++r) {
          X.off[r] = init_o[k][r];
          for (int c = 0; c < 3; ++c) X.a[r * 3 + c] = init_A[k][r * 3 + c];
        }
      }
      A->nt = nt;
      return ANTS_OK;
    };
    auto grid_of = [&](const HostGrid& h) {
      GridDev o{};
      o.D = D;
      for (int k = 0; k < 3; ++k) o.n[k] = static_cast<int>(h.n[k]);
      set_divisors(o);
      return o;
    };
    if (general) {
      for (int k = 0; k < 2; ++k) ANTS_TRY(side_apply(k, g, &gw[k]));
      gw_out = grid_of(g);
    }
    // the observer's full-scale CC: both original images on the fixed grid
    ApplyDev fs_apply[2];
    const GridDev fs_out = grid_of(gF);
    if (fscc_interval > 0)
      for (int k = 0; k < 2; ++k) ANTS_TRY(side_apply(k, gF, &fs_apply[k]));
    if (use_maps) {
      static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Z[3] = {0, 0, 0};
      maps[0] = make_warp_map(g, init_on[0] ? init_A[0] : I, init_on[0] ? init_o[0] : Z, gF);
      maps[1] = make_warp_map(g, init_on[1] ? init_A[1] : I, init_on[1] ? init_o[1] : Z, gM);
    }
    if (l == 0 && restore_on) {
      // SyN restored from a saved state: the level-0 parameter adaptor
      // resamples the four fields onto the level grid
      if (restore_grid.D != D) return fail(ANTS_ERR_INVALID, "restored state dimension does not match");
      const float* rb = restore_buf.f();
      const int64_t nr = restore_grid.count() * D;
      const float* in[4] = {rb, rb + nr, rb + 2 * nr, rb + 3 * nr};
      launch_resample_fields(to_dev(restore_grid), gd, 4, in, st, s);
      HIP_TRY(hipGetLastError());
    } else if (l == 0) {
      for (int k = 0; k < 4; ++k) HIP_TRY(hipMemsetAsync(st[k], 0, N * D * sizeof(float), s));
    } else {
      const float* in[4] = {st[0], st[1], st[2], st[3]};
      launch_resample_fields(to_dev(prev), gd, 4, in, wk, s);
      HIP_TRY(hipGetLastError());
      for (int k = 0; k < 4; ++k) std::swap(st[k], wk[k]);
    }
    prev = g;
    prof.level = l;
    run_stats.levels_run = l + 1;
    run_stats.level_voxels[l] = N;
    if (P.iterations[l] <= 0) continue;
    ANTS_TRY(smooth_image_on(s, gF, F.f(), Fs.f(), tmp.f(), P.sigma[l], P.sigma_in_mm != 0));
    ANTS_TRY(smooth_image_on(s, gM, M.f(), Ms.f(), tmp.f(), P.sigma[l], P.sigma_in_mm != 0));
    if (bf16) {
      launch_to_bf16(Fs.f(), static_cast<uint16_t*>(Fb.p), NF, s);
      launch_to_bf16(Ms.f(), static_cast<uint16_t*>(Mb.p), NM, s);
      HIP_TRY(hipGetLastError());
    }

    ConvergenceMonitor mon(P.convergence_window);
    // inversion steps launched per batch with each iteration at this level;
    // raised by one after an iteration that needed a host tail
    int kl = inv_launches;
    bool converged = false;
    int iter = 0;
    // algorithmic bytes per voxel-iteration: image gathers + field passes
    int s_max = 1;
    for (int d = 0; d < D; ++d) s_max = std::max(s_max, static_cast<int>(std::lround(g.sp[d] / gF.sp[d])));
    double gsz = 1.0;
    for (int d = 0; d < D; ++d) gsz *= std::min(2, s_max);
    const double e_i = bf16 ? 2.0 : 4.0, e_f = 4.0;
    // One SyN iteration as a fixed launch sequence over buffer roles derived
    // from the rotation (cst, cwk); it reports the rotation after the
    // iteration in (nst, nwk).  Host work here is only pointer bookkeeping,
    // so the sequence can be captured once into a hipGraph and replayed.
    // inversion steps j0..j1-1 of one batch (two problems, one launch per step)
    // Statistics are double-buffered by iteration parity `par`: while
    // iteration i's inversions run, iteration i+1's front end (warp, CC,
    // smoothing) may already fill the other block.
    constexpr int kStatWords = static_cast<int>(sizeof(IterStats) / 4);
    auto stats_of = [&](int par) { return dstats + par; };
    auto hstats_of = [&](int par) { return host_stats + par; };
    auto flag_of = [&](int par) { return host_flag + 32 * par; };
    auto parts_cc_of = [&](int par) { return static_cast<Partial*>(partials_cc.p) + par * ccp; };
    // inversion steps j0..j1-1 of one batch (two problems, one launch per step)
    auto inv_steps_on = [&](int batch, const float* const* field, float* const* inv, int j0, int j1,
                            const InvGate& gate, int par) -> int {
      RedSlot* slots[2] = {stats_of(par)->inv[batch], stats_of(par)->inv[batch + 2]};
      Partial* parts = static_cast<Partial*>(partials_inv.p);
      uint64_t* masks = static_cast<uint64_t*>(masks_inv.p);
      // step 0 is preceded by the sample launch that bounds its max norm
      if (j0 == 0) launch_inv_step(gd, 2, field, inv, slots, parts, masks, -1, s, gate);
      for (int j = j0; j < j1; ++j) launch_inv_step(gd, 2, field, inv, slots, parts, masks, j, s, gate);
      HIP_TRY(hipGetLastError());
      return ANTS_OK;
    };
    // host-completed steps: one timed span per call
    auto inv_tail_on = [&](int batch, const float* const* field, float* const* inv, int j0, int par) -> int {
      prof.begin(ANTS_KC_INV);
      ANTS_TRY(inv_steps_on(batch, field, inv, j0, kInvMaxIter + 1, kNoGate, par));
      prof.end(ANTS_KC_INV, 0.0, kInvMaxIter + 1 - j0 + (j0 == 0 && inv_speculates(gd) ? 1 : 0), 1 << batch, j0,
               kInvMaxIter + 1);
      return ANTS_OK;
    };
    // After the iteration's sequence: a batch whose problems computed a norm
    // at the last launched step has further steps to run.  If the first
    // batch is unfinished the gated second batch did nothing, and both are
    // run to the end here (first0 = true: the inverse middle fields changed
    // after the launches that were enqueued with the iteration); otherwise
    // only the second batch's tail.  In terms of the rotated buffers: batch
    // 0 inverts st[0]/st[2] into st[1]/st[3], batch 1 inverts st[1]/st[3]
    // into st[0]/st[2].
    auto inv_finish = [&](bool& extra, bool& first0, int par) -> int {
      extra = false;
      first0 = false;
      const int K = kl;
      if (K > kInvMaxIter) return ANTS_OK;
      const IterStats* hs = hstats_of(par);
      auto pending = [&](int batch) {
        return hs->inv[batch][K - 1].computed || hs->inv[batch + 2][K - 1].computed;
      };
      const float* f0[2] = {st[0], st[2]};
      float* i0[2] = {st[1], st[3]};
      const float* f1[2] = {st[1], st[3]};
      float* i1[2] = {st[0], st[2]};
      if (pending(0)) {
        // the gated launches of batch 1 did nothing: they carry no bytes
        for (auto& r : prof.recs)
          if (r.cls == ANTS_KC_INV) r.batches &= ~2;
        ANTS_TRY(inv_tail_on(0, f0, i0, K, par));
        ANTS_TRY(inv_tail_on(1, f1, i1, 0, par));
        extra = true;
        first0 = true;
      } else if (pending(1)) {
        ANTS_TRY(inv_tail_on(1, f1, i1, K, par));
        extra = true;
      }
      if (extra) {
        HIP_TRY(hipMemcpyAsync(hstats_of(par), stats_of(par), sizeof(IterStats), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        run_stats.inv_tails += 1;
      }
      return ANTS_OK;
    };

    const double Nd = static_cast<double>(N);
    // Front end of an iteration on stream fs: warp both smoothed images
    // through the inverse middle fields cst[1], cst[3] (the warp's first
    // block also zeroes the iteration's statistics), CC value and both
    // derivatives into wk[0], wk[1], update-field smoothing + max norms
    // (result in wk[0], wk[1]; scratch wk[2], wk[3]).  It writes only Fw, Mw,
    // the four work buffers and its own statistics block; wk is updated to
    // the rotation it leaves.
    auto front = [&](int par, float* const* cst, float* wk_[4], hipStream_t fs, int& n_cc) -> int {
      uint32_t* stat_words = reinterpret_cast<uint32_t*>(stats_of(par));
      prof.begin(ANTS_KC_WARP);
      if (general) {
        launch_fill_u32(stat_words, 0u, kStatWords, fs);
        gw[0].t[0].f = cst[1];
        gw[1].t[0].f = cst[3];
        launch_apply_transforms(gw_out, gw[0], ANTS_INTERP_LINEAR, Fs.f(), Fw.f(), fs);
        launch_apply_transforms(gw_out, gw[1], ANTS_INTERP_LINEAR, Ms.f(), Mw.f(), fs);
      } else if (bf16)
        launch_warp_pair_bf16(gd, gFd, static_cast<const uint16_t*>(Fb.p), cst[1], Fw.f(), gMd,
                              static_cast<const uint16_t*>(Mb.p), cst[3], Mw.f(), fs, stat_words, kStatWords,
                              use_maps ? maps : nullptr);
      else
        launch_warp_pair(gd, gFd, Fs.f(), cst[1], Fw.f(), gMd, Ms.f(), cst[3], Mw.f(), fs, stat_words, kStatWords,
                         use_maps ? maps : nullptr);
      prof.end(ANTS_KC_WARP, Nd * (2.0 * D * e_f + 2.0 * gsz * e_i + 2.0 * e_i));
      prof.begin(ANTS_KC_CC);
      // the energy slot is reduced when the iteration's statistics are published
      n_cc = launch_cc(gd, Fw.f(), Mw.f(), P.cc_radius, eps, wk_[0], wk_[1], parts_cc_of(par), &stats_of(par)->energy,
                       fs, false, bf16);
      prof.end(ANTS_KC_CC, Nd * (2.0 * e_i + 2.0 * D * e_f));
      HIP_TRY(hipGetLastError());
      float* in[2] = {wk_[0], wk_[1]};
      float* work[2] = {wk_[2], wk_[3]};
      if (bsp) {
        // BSplineSyN: the update fields are B-spline fits (into wk[2], wk[3])
        float* lat[2] = {bsp_upd.f(), bsp_upd.f() + bsp_fu};
        ANTS_TRY(bsp_fields_on(fs, g, bsp_u_dev, bsp_lu[l], 2, in, work, lat, stats_of(par)->maxnorm, bsp_part.f(),
                               &prof));
        std::swap(wk_[0], wk_[2]);
        std::swap(wk_[1], wk_[3]);
        return ANTS_OK;
      }
      float* out[2] = {D % 2 == 1 ? wk_[2] : wk_[0], D % 2 == 1 ? wk_[3] : wk_[1]};
      ANTS_TRY(smooth_fields_on(fs, g, 2, in, work, out, P.update_variance, stats_of(par)->maxnorm, &prof,
                                need_blend ? blend_f : nullptr));
      if (out[0] != wk_[0]) {
        std::swap(wk_[0], wk_[2]);
        std::swap(wk_[1], wk_[3]);
      }
      return ANTS_OK;
    };
    // Back end on the main stream: scaled composition onto the forward
    // middle fields (t -> wk[2], wk[3]), total-field smoothing, the first kl
    // steps of both inversion batches (inverse middle fields from the
    // previous inverses in place, then forward fields as inverses of the new
    // inverses, started from t, in place; the second batch gated on the
    // first having finished, the host completing whatever is left), and the
    // publication of the statistics.  `ev_inv` (optional) is recorded once
    // the first batch's launches are enqueued: from there on the inverse
    // middle fields are final unless that batch needs a host tail.  st_/wk_
    // are rotated in place.
    auto back = [&](int par, float* st_[4], float* wk_[4], int n_cc, hipEvent_t ev_inv) -> int {
      {
        const float* disp[2] = {wk_[0], wk_[1]};
        const float* warp[2] = {st_[0], st_[2]};
        float* out[2] = {wk_[2], wk_[3]};
        prof.begin(ANTS_KC_COMPOSE);
        launch_compose(gd, 2, disp, warp, out, stats_of(par)->maxnorm, static_cast<float>(P.learning_rate), s);
        prof.end(ANTS_KC_COMPOSE, 2.0 * Nd * 3.0 * D * e_f);
        HIP_TRY(hipGetLastError());
      }
      if (bsp_total) {
        // BSplineSyN with a total-field mesh: fit the composed fields
        const float* in[2] = {wk_[2], wk_[3]};
        float* out[2] = {wk_[0], wk_[1]};
        float* lat[2] = {bsp_tot.f(), bsp_tot.f() + bsp_ft};
        ANTS_TRY(bsp_fields_on(s, g, bsp_t_dev, bsp_lt[l], 2, in, out, lat, stats_of(par)->maxnorm + 2,
                               bsp_part.f() + 2 * bsp_partials_needed(), &prof));
        std::swap(wk_[0], wk_[2]);
        std::swap(wk_[1], wk_[3]);
      } else if (!bsp && P.total_variance > 0.0) {
        float* in[2] = {wk_[2], wk_[3]};
        float* work[2] = {wk_[0], wk_[1]};
        float* out[2] = {D % 2 == 1 ? wk_[0] : wk_[2], D % 2 == 1 ? wk_[1] : wk_[3]};
        ANTS_TRY(smooth_fields_on(s, g, 2, in, work, out, P.total_variance, stats_of(par)->maxnorm + 2, &prof,
                                  need_blend ? blend_f : nullptr));
        if (out[0] != wk_[2]) {
          std::swap(wk_[0], wk_[2]);
          std::swap(wk_[1], wk_[3]);
        }
      }
      prof.begin(ANTS_KC_INV);
      for (int batch = 0; batch < 2; ++batch) {
        const float* field[2] = {batch == 0 ? wk_[2] : st_[1], batch == 0 ? wk_[3] : st_[3]};
        float* inv[2] = {batch == 0 ? st_[1] : wk_[2], batch == 0 ? st_[3] : wk_[3]};
        InvGate gate = kNoGate;
        if (batch == 1 && kl <= kInvMaxIter)
          gate = InvGate{{stats_of(par)->inv[0], stats_of(par)->inv[2]}, 2, kl - 1};
        ANTS_TRY(inv_steps_on(batch, field, inv, 0, kl, gate, par));
        if (batch == 0 && ev_inv) HIP_TRY(hipEventRecord(ev_inv, s));
      }
      prof.end(ANTS_KC_INV, 0.0, 2 * (kl + (inv_speculates(gd) ? 1 : 0)), 3, 0, kl);  // per batch: [sample] + kl steps
      std::swap(st_[0], wk_[2]);
      std::swap(st_[2], wk_[3]);
      launch_publish_stats(parts_cc_of(par), n_cc, &stats_of(par)->energy,
                           reinterpret_cast<uint32_t*>(stats_of(par)), reinterpret_cast<uint32_t*>(hstats_of(par)),
                           kStatWords, flag_of(par), s);
      HIP_TRY(hipGetLastError());
      return ANTS_OK;
    };
    // One SyN iteration as a fixed launch sequence on the main stream over
    // buffer roles derived from the rotation (cst, cwk); it reports the
    // rotation after the iteration in (nst, nwk).  Host work here is only
    // pointer bookkeeping, so the sequence can be captured once into a
    // hipGraph and replayed.
    auto enqueue = [&](float* const* cst, float* const* cwk, float** nst, float** nwk) -> int {
      float* st_[4] = {cst[0], cst[1], cst[2], cst[3]};
      float* wk_[4] = {cwk[0], cwk[1], cwk[2], cwk[3]};
      int n_cc = 0;
      ANTS_TRY(front(0, st_, wk_, s, n_cc));
      ANTS_TRY(back(0, st_, wk_, n_cc, nullptr));
      for (int k = 0; k < 4; ++k) {
        nst[k] = st_[k];
        nwk[k] = wk_[k];
      }
      return ANTS_OK;
    };

    // 7. energy, convergence, observer, work counters of a finished iteration
    auto account = [&](int par, bool& stop) -> int {
      const IterStats* hs = hstats_of(par);
      prof.collect(hs, N, D);
      const double value = -hs->energy.sum / static_cast<double>(N);
      const double E = 0.5 * (value + value);
      int kin[4];
      for (int q = 0; q < 4; ++q) {
        int c = 0;
        for (int j = 0; j < kInvSlots; ++j) c += hs->inv[q][j].computed ? 1 : 0;
        kin[q] = c;
      }
      mon.add(E);
      const double conv = mon.value();
      if (!P.force_iterations && conv < P.convergence_threshold) stop = true;
      run_stats.iterations[l] = iter;
      const int ksum = kin[0] + kin[1] + kin[2] + kin[3];
      run_stats.inv_steps[l] += ksum;
      run_stats.voxel_iterations += static_cast<double>(N);
      run_stats.algorithmic_bytes +=
          static_cast<double>(N) * ((2.0 * gsz * e_i + 14.0 * D * e_f) + (3.0 * ksum + 4.0) * D * e_f);
      // antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate
      // (Examples/...CommandIterationUpdate.h:151-159, 252-397): at iteration
      // 1, every interval-th and the level's last, CC r=4 between the original
      // images, each resampled onto the fixed grid through its side's
      // composite [initial transforms, middle transform inverse]
      fscc_last = 0.0;
      if (fscc_interval > 0 && (iter == 1 || iter % fscc_interval == 0 || iter == P.iterations[l])) {
        const int64_t NFF = gF.count();
        float* img = fscc_buf.f();
        float* grad = img + 2 * NFF;
        fs_apply[0].t[0].f = st[1];
        fs_apply[1].t[0].f = st[3];
        launch_apply_transforms(fs_out, fs_apply[0], ANTS_INTERP_LINEAR, d_fixed, img, s);
        launch_apply_transforms(fs_out, fs_apply[1], ANTS_INTERP_LINEAR, d_moving, img + NFF, s);
        RedSlot* slot = reinterpret_cast<RedSlot*>(fscc_buf.f() + 2 * NFF + 2 * NFF * D);
        Partial* parts = reinterpret_cast<Partial*>(slot + 1);
        launch_cc(gFd, img, img + NFF, 4, eps, grad, grad + NFF * D, parts, slot, s, true, false);
        HIP_TRY(hipGetLastError());
        RedSlot hslot{};
        HIP_TRY(hipMemcpyAsync(&hslot, slot, sizeof(RedSlot), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        fscc_last = -hslot.sum / static_cast<double>(NFF);
      }
      trace.push_back(ants_syn_iteration_record{l, iter, E, conv, {kin[0], kin[1], kin[2], kin[3]}, fscc_last});
      if (cb && cb(user, l, iter, E, conv, kin) != 0) return fail(ANTS_ERR_CANCELLED, "cancelled by the observer");
      return ANTS_OK;
    };

    if (pipelined) {
      // Iteration i+1's front end runs on the second stream while iteration
      // i's inversions finish: it needs only the new inverse middle fields
      // (final once the first inversion batch is done) and writes only
      // buffers that the second batch does not touch.  It is enqueued before
      // iteration i's convergence is known; past the level's last iteration
      // its results are simply not used.  When the first batch needs a host
      // tail, the front end is run again after it.
      HIP_TRY(hipEventRecord(ev_main, s));
      HIP_TRY(hipStreamWaitEvent(s2, ev_main, 0));
      int par = 0, n_cc = 0;
      if (iter < P.iterations[l]) {
        ANTS_TRY(front(par, st, wk, s2, n_cc));
        HIP_TRY(hipEventRecord(ev_front, s2));
      }
      while (iter++ < P.iterations[l] && !converged) {
        prof.active = (iter - 1) % prof.every == 0 ? prof.mask : 0u;
        HIP_TRY(hipStreamWaitEvent(s, ev_front, 0));
        *reinterpret_cast<volatile uint32_t*>(flag_of(par)) = 0u;
        ANTS_TRY(back(par, st, wk, n_cc, ev_inv0));
        run_stats.eager_iterations += 1;
        const bool next = iter < P.iterations[l];
        float* wk_next[4] = {wk[0], wk[1], wk[2], wk[3]};
        int n_cc_next = 0;
        if (next) {
          HIP_TRY(hipStreamWaitEvent(s2, ev_inv0, 0));
          ANTS_TRY(front(par ^ 1, st, wk_next, s2, n_cc_next));
          HIP_TRY(hipEventRecord(ev_front, s2));
        }
        ANTS_TRY(wait_published(flag_of(par), s));
        bool extra = false, first0 = false;
        ANTS_TRY(inv_finish(extra, first0, par));
        if (extra && !inv_launches_fixed) kl = std::min(kl + 1, kInvMaxIter + 1);
        if (first0 && next) {
          // the inverse middle fields changed after the front end started
          for (int k = 0; k < 4; ++k) wk_next[k] = wk[k];
          HIP_TRY(hipEventRecord(ev_main, s));
          HIP_TRY(hipStreamWaitEvent(s2, ev_main, 0));
          ANTS_TRY(front(par ^ 1, st, wk_next, s2, n_cc_next));
          HIP_TRY(hipEventRecord(ev_front, s2));
        }
        for (int k = 0; k < 4; ++k) wk[k] = wk_next[k];
        n_cc = n_cc_next;
        bool stop = false;
        const int rc = account(par, stop);
        if (rc != ANTS_OK) {
          (void)hipStreamSynchronize(s2);
          return rc;
        }
        if (stop) converged = true;
        par ^= 1;
      }
      // a front end enqueued past the level's last iteration finishes before
      // the level's buffers are reused
      HIP_TRY(hipEventRecord(ev_main, s2));
      HIP_TRY(hipStreamWaitEvent(s, ev_main, 0));
    } else {
      while (iter++ < P.iterations[l] && !converged) {
        float* nst[4];
        float* nwk[4];
        prof.active = (iter - 1) % prof.every == 0 ? prof.mask : 0u;
        GraphKey key{setup_hash, l, {st[0], st[1], st[2], st[3], wk[0], wk[1], wk[2], wk[3]}, prof.mask, kl};
        auto it = use_graphs ? graphs.find(key) : graphs.end();
        if (use_graphs && it == graphs.end()) {
          // capture this rotation's sequence once
          GraphEntry ge{};
          int rc = ANTS_OK;
          if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            rc = enqueue(st, wk, ge.nst, ge.nwk);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &graph);
            hipError_t ei = hipErrorUnknown;
            if (rc == ANTS_OK && e == hipSuccess && graph)
              ei = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
            if (ei != hipSuccess && std::getenv("ANTS_SYN_VERBOSE"))
              std::fprintf(stderr, "ants_syn: iteration capture at level %d failed (enqueue %d: %s; end %s; instantiate %s)\n",
                           l, rc, rc == ANTS_OK ? "" : g_last_error.c_str(), hipGetErrorString(e), hipGetErrorString(ei));
            if (ei == hipSuccess) it = graphs.emplace(key, std::move(ge)).first;
            if (graph) (void)hipGraphDestroy(graph);
          }
          (void)hipGetLastError();
          if (it == graphs.end()) use_graphs = false;  // capture not possible here: launch eagerly
        }
        *reinterpret_cast<volatile uint32_t*>(flag_of(0)) = 0u;
        if (use_graphs && it != graphs.end()) {
          HIP_TRY(hipGraphLaunch(it->second.exec, s));
          run_stats.graph_launches += 1;
          for (int k = 0; k < 4; ++k) {
            nst[k] = it->second.nst[k];
            nwk[k] = it->second.nwk[k];
          }
        } else {
          ANTS_TRY(enqueue(st, wk, nst, nwk));
          run_stats.eager_iterations += 1;
        }
        for (int k = 0; k < 4; ++k) {
          st[k] = nst[k];
          wk[k] = nwk[k];
        }
        ANTS_TRY(wait_published(flag_of(0), s));
        bool extra = false, first0 = false;
        ANTS_TRY(inv_finish(extra, first0, 0));
        if (extra && !inv_launches_fixed) kl = std::min(kl + 1, kInvMaxIter + 1);
        bool stop = false;
        ANTS_TRY(account(0, stop));
        if (stop) converged = true;
      }
    }
  }
  // final composition: fwd = M2M.inv o F2M.fwd ; inv = F2M.inv o M2M.fwd
  const HostGrid& gl = prev;
  const GridDev gld = to_dev(gl);
  {
    const float* disp[2] = {st[3], st[1]};
    const float* warp[2] = {st[0], st[2]};
    float* out[2] = {d_out_fwd ? d_out_fwd : wk[0], d_out_inv ? d_out_inv : wk[1]};
    launch_compose(gld, 2, disp, warp, out, nullptr, 1.f, s);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(s));
  for (int k = 0; k < 4; ++k) state[k] = st[k];
  last_grid = gl;
  have_state = true;
  return ANTS_OK;
}

// ---------------------------------------------------------------------------
// Host-buffer transfers.  Copies from and to the caller's pageable memory go
// through two pinned chunks: threads copy one chunk between the caller's
// buffer and pinned memory while the DMA engine moves the other (a pageable
// hipMemcpy stages through one host thread).
namespace {

constexpr size_t kPinChunk = size_t{16} << 20;

void parallel_copy(void* dst, const void* src, size_t bytes) {
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nt = std::min<size_t>(std::max<size_t>(hw / 2, 1), std::min<size_t>(8, bytes / (size_t{1} << 20) + 1));
  if (nt <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = (bytes / nt + 63) / 64 * 64;
  std::vector<std::thread> th;
  for (size_t k = 1; k < nt; ++k) {
    const size_t o = k * part;
    if (o >= bytes) break;
    th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o,
                                      std::min(part, bytes - o)); });
  }
  std::memcpy(dst, src, std::min(part, bytes));
  for (auto& t : th) t.join();
}

}  // namespace

static int ensure_pinned(ants_syn_handle* h) {
  for (int k = 0; k < 2; ++k) {
    if (!h->pin[k]) HIP_TRY(hipHostMalloc(&h->pin[k], kPinChunk, hipHostMallocDefault));
    if (!h->pin_ev[k]) HIP_TRY(hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming));
  }
  return ANTS_OK;
}

int ants_syn_handle::staged_upload(void* d_dst, const void* host_src, size_t bytes) {
  ANTS_TRY(ensure_pinned(this));
  const char* src = static_cast<const char*>(host_src);
  char* dst = static_cast<char*>(d_dst);
  for (size_t o = 0, k = 0; o < bytes; o += kPinChunk, ++k) {
    const int b = static_cast<int>(k & 1);
    // the last DMA out of this chunk (this call's or an earlier one's) has finished
    HIP_TRY(hipEventSynchronize(pin_ev[b]));
    const size_t n = std::min(kPinChunk, bytes - o);
    parallel_copy(pin[b], src + o, n);
    HIP_TRY(hipMemcpyAsync(dst + o, pin[b], n, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipEventRecord(pin_ev[b], stream));
  }
  return ANTS_OK;
}

int ants_syn_handle::staged_download(void* host_dst, const void* d_src, size_t bytes) {
  ANTS_TRY(ensure_pinned(this));
  const char* src = static_cast<const char*>(d_src);
  char* dst = static_cast<char*>(host_dst);
  const size_t nchunks = (bytes + kPinChunk - 1) / kPinChunk;
  auto issue = [&](size_t k) -> int {
    const int b = static_cast<int>(k & 1);
    const size_t o = k * kPinChunk, n = std::min(kPinChunk, bytes - o);
    HIP_TRY(hipMemcpyAsync(pin[b], src + o, n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipEventRecord(pin_ev[b], stream));
    return ANTS_OK;
  };
  if (nchunks > 0) ANTS_TRY(issue(0));
  for (size_t k = 0; k < nchunks; ++k) {
    const int b = static_cast<int>(k & 1);
    HIP_TRY(hipEventSynchronize(pin_ev[b]));
    if (k + 1 < nchunks) ANTS_TRY(issue(k + 1));  // the other chunk: its previous copy-out is done
    const size_t o = k * kPinChunk;
    parallel_copy(dst + o, pin[b], std::min(kPinChunk, bytes - o));
  }
  return ANTS_OK;
}

// ---------------------------------------------------------------------------
// C-ABI
extern "C" {

const char* ants_last_error(void) { return g_last_error.c_str(); }
int ants_version(void) { return 1; }

int ants_device_malloc(void** d_ptr, size_t bytes) {
  if (!d_ptr) return fail(ANTS_ERR_INVALID, "null output pointer");
  hipError_t e = hipMalloc(d_ptr, bytes);
  if (e != hipSuccess) return fail(ANTS_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return ANTS_OK;
}
int ants_device_free(void* d_ptr) {
  HIP_TRY(hipFree(d_ptr));
  return ANTS_OK;
}
int ants_memcpy_h2d(void* d_dst, const void* src, size_t bytes, void* stream) {
  HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
  HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return ANTS_OK;
}
int ants_memcpy_d2h(void* dst, const void* d_src, size_t bytes, void* stream) {
  HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
  HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return ANTS_OK;
}
int ants_memset_device(void* d_dst, int value, size_t bytes, void* stream) {
  HIP_TRY(hipMemsetAsync(d_dst, value, bytes, static_cast<hipStream_t>(stream)));
  return ANTS_OK;
}
int ants_copy_calibration(void* d_dst, const void* d_src, size_t bytes, void* stream) {
  if (!d_dst || !d_src || bytes % 16 != 0)
    return fail(ANTS_ERR_INVALID, "copy_calibration: null pointer or size not a multiple of 16");
  if (bytes == 0) return ANTS_OK;
  launch_copy_calibration(d_src, d_dst, static_cast<int64_t>(bytes / 16), static_cast<hipStream_t>(stream));
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}
int ants_stream_create(void** stream) {
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return ANTS_OK;
}
int ants_stream_destroy(void* stream) {
  HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return ANTS_OK;
}
int ants_stream_synchronize(void* stream) {
  HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return ANTS_OK;
}
int ants_device_memory_info(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return ANTS_OK;
}
int ants_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return ANTS_OK;
}
int ants_event_create(void** event) {
  hipEvent_t e;
  HIP_TRY(hipEventCreate(&e));
  *event = e;
  return ANTS_OK;
}
int ants_event_destroy(void* event) {
  HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return ANTS_OK;
}
int ants_event_record(void* event, void* stream) {
  HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return ANTS_OK;
}
int ants_event_elapsed_ms(float* ms, void* start, void* stop) {
  HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return ANTS_OK;
}

int ants_syn_level_geometry(const ants_syn_params* p, const ants_image_desc* fixed, int level, ants_image_desc* out) {
  ANTS_TRY(validate_params(p));
  if (level < 0 || level >= p->n_levels) return fail(ANTS_ERR_INVALID, "level out of range");
  HostGrid g;
  ANTS_TRY(to_host_grid(fixed, &g));
  if (!out) return fail(ANTS_ERR_INVALID, "null output descriptor");
  to_desc(level_grid(g, p->shrink[level]), out);
  return ANTS_OK;
}

int ants_syn_create(const ants_syn_params* p, ants_syn_handle** out) {
  if (!out) return fail(ANTS_ERR_INVALID, "null handle pointer");
  *out = nullptr;
  ANTS_TRY(validate_params(p));
  auto* h = new (std::nothrow) ants_syn_handle();
  if (!h) return fail(ANTS_ERR_OOM, "out of host memory");
  h->P = *p;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(ANTS_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  h->own_stream = true;
  *out = h;
  return ANTS_OK;
}

void ants_syn_destroy(ants_syn_handle* h) { delete h; }

int ants_syn_set_stream(ants_syn_handle* h, void* stream) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  if (h->own_stream && h->stream) HIP_TRY(hipStreamDestroy(h->stream));
  h->stream = static_cast<hipStream_t>(stream);
  h->own_stream = false;
  return ANTS_OK;
}

int ants_syn_set_initial_transforms(ants_syn_handle* h, const ants_transform* moving, int n_moving,
                                    const ants_transform* fixed, int n_fixed) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  InitialXforms x;
  ANTS_TRY(compose_initial(h->P.dim, fixed, n_fixed, moving, n_moving, &x));
  h->set_initial(x);
  return ANTS_OK;
}

int ants_syn_register_device(ants_syn_handle* h, const ants_image_desc* fixed, const float* d_fixed,
                             const ants_image_desc* moving, const float* d_moving, float* d_out_fwd,
                             float* d_out_inv, ants_syn_iteration_cb cb, void* user) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  return h->run(fixed, d_fixed, moving, d_moving, d_out_fwd, d_out_inv, cb, user);
}

int ants_syn_register(ants_syn_handle* h, const ants_image_desc* fixed, const float* host_fixed,
                      const ants_image_desc* moving, const float* host_moving, float* host_out_fwd,
                      float* host_out_inv, ants_syn_iteration_cb cb, void* user) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  HostGrid gF, gM;
  ANTS_TRY(to_host_grid(fixed, &gF));
  ANTS_TRY(to_host_grid(moving, &gM));
  if (!host_fixed || !host_moving) return fail(ANTS_ERR_INVALID, "null image buffer");
  const HostGrid gl = level_grid(gF, h->P.shrink[h->P.n_levels - 1]);
  const size_t fb = static_cast<size_t>(gl.count()) * gF.D * sizeof(float);
  // the handle keeps the device copies between calls (grow-only)
  ANTS_TRY(h->io_f.ensure(gF.count() * sizeof(float)));
  ANTS_TRY(h->io_m.ensure(gM.count() * sizeof(float)));
  ANTS_TRY(h->io_fwd.ensure(fb));
  ANTS_TRY(h->io_inv.ensure(fb));
  ANTS_TRY(h->staged_upload(h->io_f.p, host_fixed, gF.count() * sizeof(float)));
  ANTS_TRY(h->staged_upload(h->io_m.p, host_moving, gM.count() * sizeof(float)));
  int rc = h->run(fixed, h->io_f.f(), moving, h->io_m.f(), h->io_fwd.f(), h->io_inv.f(), cb, user);
  if (rc == ANTS_OK) {
    if (host_out_fwd) ANTS_TRY(h->staged_download(host_out_fwd, h->io_fwd.p, fb));
    if (host_out_inv) ANTS_TRY(h->staged_download(host_out_inv, h->io_inv.p, fb));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  return rc;
}

int ants_syn_get_state_device(ants_syn_handle* h, float* d_f2m_fwd, float* d_f2m_inv, float* d_m2m_fwd,
                              float* d_m2m_inv) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  if (!h->have_state) return fail(ANTS_ERR_INVALID, "no registration has completed on this handle");
  float* dst[4] = {d_f2m_fwd, d_f2m_inv, d_m2m_fwd, d_m2m_inv};
  const size_t b = static_cast<size_t>(h->last_grid.count()) * h->last_grid.D * sizeof(float);
  for (int k = 0; k < 4; ++k)
    if (dst[k]) HIP_TRY(hipMemcpyAsync(dst[k], h->state[k], b, hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return ANTS_OK;
}

int ants_syn_set_initial_state_device(ants_syn_handle* h, const ants_image_desc* grid, const float* d_f2m_fwd,
                                      const float* d_f2m_inv, const float* d_m2m_fwd, const float* d_m2m_inv) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  const float* src[4] = {d_f2m_fwd, d_f2m_inv, d_m2m_fwd, d_m2m_inv};
  int given = 0;
  for (const float* p : src) given += p ? 1 : 0;
  if (given == 0) {
    h->restore_on = false;
    return ANTS_OK;
  }
  // the reference requires both fields of both transforms ("Invalid state restoration" otherwise)
  if (given != 4) return fail(ANTS_ERR_INVALID, "a restored state needs all four fields");
  HostGrid g;
  ANTS_TRY(to_host_grid(grid, &g));
  if (g.D != h->P.dim) return fail(ANTS_ERR_INVALID, "state dimension does not match the stage");
  const size_t b = static_cast<size_t>(g.count()) * g.D * sizeof(float);
  ANTS_TRY(h->restore_buf.ensure(4 * b));
  char* dst = static_cast<char*>(h->restore_buf.p);
  for (int k = 0; k < 4; ++k) HIP_TRY(hipMemcpyAsync(dst + k * b, src[k], b, hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->restore_grid = g;
  h->restore_on = true;
  return ANTS_OK;
}

int ants_syn_convergence_value(const double* energies, int n, int window, double* out) {
  if (!out || n < 0 || (n > 0 && !energies)) return fail(ANTS_ERR_INVALID, "bad energy sequence");
  if (window < 1) return fail(ANTS_ERR_INVALID, "convergence window must be >= 1");
  ConvergenceMonitor mon(window);
  for (int i = 0; i < n; ++i) mon.add(energies[i]);
  *out = mon.value();
  return ANTS_OK;
}

int ants_syn_get_stats(const ants_syn_handle* h, ants_syn_stats* out) {
  if (!h || !out) return fail(ANTS_ERR_INVALID, "null argument");
  *out = h->run_stats;
  return ANTS_OK;
}

int ants_syn_get_trace(const ants_syn_handle* h, ants_syn_iteration_record* out, int64_t capacity, int64_t* count) {
  if (!h || !count) return fail(ANTS_ERR_INVALID, "null argument");
  if (capacity < 0 || (capacity > 0 && !out)) return fail(ANTS_ERR_INVALID, "bad output buffer");
  const int64_t n = static_cast<int64_t>(h->trace.size());
  const int64_t w = std::min(n, capacity);
  for (int64_t i = 0; i < w; ++i) out[i] = h->trace[i];
  *count = n;
  return ANTS_OK;
}

int ants_syn_set_profiling(ants_syn_handle* h, unsigned int class_mask) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  h->prof.mask = class_mask;
  return ANTS_OK;
}

int ants_syn_set_full_scale_cc_interval(ants_syn_handle* h, int interval) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  if (interval < 0) return fail(ANTS_ERR_INVALID, "similarity measure interval must be >= 0");
  h->fscc_interval = interval;
  return ANTS_OK;
}

int ants_syn_last_full_scale_cc(const ants_syn_handle* h, double* value) {
  if (!h || !value) return fail(ANTS_ERR_INVALID, "null argument");
  *value = h->fscc_last;
  return ANTS_OK;
}

int ants_syn_set_profiling_stride(ants_syn_handle* h, int every) {
  if (!h) return fail(ANTS_ERR_INVALID, "null handle");
  if (every < 1) return fail(ANTS_ERR_INVALID, "profiling stride must be >= 1");
  h->prof.every = every;
  return ANTS_OK;
}

int ants_syn_get_kernel_stats(const ants_syn_handle* h, ants_syn_kernel_stats* out) {
  if (!h || !out) return fail(ANTS_ERR_INVALID, "null argument");
  *out = h->prof.acc;
  return ANTS_OK;
}

int ants_syn_get_kernel_stats_level(const ants_syn_handle* h, int level, ants_syn_kernel_stats* out) {
  if (!h || !out) return fail(ANTS_ERR_INVALID, "null argument");
  if (level < 0 || level >= ANTS_MAX_LEVELS) return fail(ANTS_ERR_INVALID, "level out of range");
  *out = h->prof.acc_level[level];
  return ANTS_OK;
}

// ---- building blocks ----
namespace {
struct OpScratch {
  DevBuf stats, partials;
  ~OpScratch() {
    stats.release();
    partials.release();
  }
};
}  // namespace

int ants_op_smooth_image(const ants_image_desc* gdsc, const float* d_in, float* d_out, double sigma, int sigma_in_mm,
                         void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  if (!(sigma >= 0.0)) return fail(ANTS_ERR_INVALID, "sigma must be >= 0");
  DevBuf tmp;
  ANTS_TRY(tmp.ensure(g.count() * sizeof(float)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = smooth_image_on(s, g, d_in, d_out, tmp.f(), sigma, sigma_in_mm != 0);
  if (rc == ANTS_OK) HIP_TRY(hipStreamSynchronize(s));
  tmp.release();
  return rc;
}

int ants_op_smooth_field(const ants_image_desc* gdsc, const float* d_in, float* d_out, double variance,
                         void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  const size_t b = static_cast<size_t>(g.count()) * g.D * sizeof(float);
  DevBuf a, w, m;
  ANTS_TRY(a.ensure(b));
  ANTS_TRY(w.ensure(b));
  ANTS_TRY(m.ensure(16));
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemcpyAsync(a.p, d_in, b, hipMemcpyDeviceToDevice, s));
  float* in[1] = {a.f()};
  float* work[1] = {w.f()};
  float* out[1] = {d_out};
  DevBuf x;
  if (g.D == 3 && variance > 0.0 && variance < 0.5) ANTS_TRY(x.ensure(b));
  float* extra[1] = {x.f()};
  HIP_TRY(hipMemsetAsync(m.p, 0, 16, s));
  int rc = smooth_fields_on(s, g, 1, in, work, out, variance, static_cast<unsigned int*>(m.p), nullptr,
                            x.p ? extra : nullptr);
  if (rc == ANTS_OK) HIP_TRY(hipStreamSynchronize(s));
  return rc;
}

int ants_image_center_of_mass(const ants_image_desc* gdsc, const float* d_img, double* center, double* mass,
                              void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  if (!d_img || !center) return fail(ANTS_ERR_INVALID, "null buffer");
  DevBuf part;
  ANTS_TRY(part.ensure(static_cast<size_t>(kRedBlocks) * 4 * sizeof(double)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = launch_moments(to_dev(g), d_img, static_cast<double*>(part.p), s);
  HIP_TRY(hipGetLastError());
  std::vector<double> h(static_cast<size_t>(nb) * 4);
  HIP_TRY(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 4; ++k) m[k] += h[static_cast<size_t>(b) * 4 + k];
  if (mass) *mass = m[0];
  if (m[0] == 0.0) return fail(ANTS_ERR_INVALID, "the image's total intensity is zero: no centre of mass");
  // physical point of the mean index: origin + direction diag(spacing) index
  const double idx[3] = {m[1] / m[0], m[2] / m[0], m[3] / m[0]};
  for (int r = 0; r < g.D; ++r) {
    double v = g.org[r];
    for (int c = 0; c < g.D; ++c) v += g.dir[r * 3 + c] * g.sp[c] * idx[c];
    center[r] = v;
  }
  return ANTS_OK;
}

int ants_op_bspline_field(const ants_image_desc* gdsc, const float* d_in, float* d_out, const int* mesh, int order,
                          double* host_max, void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  if (!d_in || !d_out || !mesh) return fail(ANTS_ERR_INVALID, "null buffer or mesh");
  if (d_in == d_out) return fail(ANTS_ERR_INVALID, "the B-spline fit cannot run in place");
  if (order < 1 || order > 3) return fail(ANTS_ERR_UNSUPPORTED, "B-spline order must be 1..3");
  int m3[3] = {1, 1, 1};
  for (int d = 0; d < g.D; ++d) {
    if (mesh[d] < 1 || mesh[d] > (1 << 16)) return fail(ANTS_ERR_INVALID, "B-spline mesh sizes must be in 1..65536");
    m3[d] = mesh[d];
  }
  std::vector<unsigned char> blob;
  BspLayout L;
  ANTS_TRY(bsp_tables(g, m3, order, blob, &L));
  DevBuf tab, lat, m, part;
  ANTS_TRY(tab.ensure(blob.size()));
  ANTS_TRY(lat.ensure(static_cast<size_t>(L.field_floats()) * sizeof(float)));
  ANTS_TRY(m.ensure(16));
  ANTS_TRY(part.ensure(bsp_partials_needed() * sizeof(float)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemcpyAsync(tab.p, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(m.p, 0, 16, s));
  const BspDev T = bsp_dev(L, tab.p);
  const float* in[1] = {d_in};
  float* out[1] = {d_out};
  float* buf[1] = {lat.f()};
  ANTS_TRY(bsp_fields_on(s, g, T, L, 1, in, out, buf, static_cast<unsigned int*>(m.p), part.f()));
  unsigned int bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, m.p, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float mx;
  std::memcpy(&mx, &bits, 4);
  if (host_max) *host_max = mx;
  return ANTS_OK;
}

int ants_op_warp_image(const ants_image_desc* gi, const float* d_img, const ants_image_desc* go, const float* d_field,
                       float* d_out, void* stream) {
  HostGrid a, b;
  ANTS_TRY(to_host_grid(gi, &a));
  ANTS_TRY(to_host_grid(go, &b));
  if (a.D != b.D) return fail(ANTS_ERR_INVALID, "dimension mismatch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_warp_image(to_dev(a), d_img, to_dev(b), d_field, d_out, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return ANTS_OK;
}

int ants_op_warp_pair(const ants_image_desc* g_level, const ants_image_desc* g_fixed, const float* d_fixed,
                      const float* d_u_fixed, float* d_fixed_out, const ants_image_desc* g_moving,
                      const float* d_moving, const float* d_u_moving, float* d_moving_out, int bf16_images,
                      void* stream) {
  HostGrid gl, gF, gM;
  ANTS_TRY(to_host_grid(g_level, &gl));
  ANTS_TRY(to_host_grid(g_fixed, &gF));
  ANTS_TRY(to_host_grid(g_moving, &gM));
  if (gl.D != gF.D || gl.D != gM.D) return fail(ANTS_ERR_INVALID, "dimension mismatch");
  if (!d_fixed || !d_moving || !d_u_fixed || !d_u_moving || !d_fixed_out || !d_moving_out)
    return fail(ANTS_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (bf16_images) {
    DevBuf bF, bM;
    ANTS_TRY(bF.ensure(gF.count() * sizeof(uint16_t)));
    ANTS_TRY(bM.ensure(gM.count() * sizeof(uint16_t)));
    launch_to_bf16(d_fixed, static_cast<uint16_t*>(bF.p), gF.count(), s);
    launch_to_bf16(d_moving, static_cast<uint16_t*>(bM.p), gM.count(), s);
    launch_warp_pair_bf16(to_dev(gl), to_dev(gF), static_cast<const uint16_t*>(bF.p), d_u_fixed, d_fixed_out,
                          to_dev(gM), static_cast<const uint16_t*>(bM.p), d_u_moving, d_moving_out, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return ANTS_OK;
  }
  launch_warp_pair(to_dev(gl), to_dev(gF), d_fixed, d_u_fixed, d_fixed_out, to_dev(gM), d_moving, d_u_moving,
                   d_moving_out, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return ANTS_OK;
}

int ants_op_cc(const ants_image_desc* gdsc, const float* d_fixed, const float* d_moving, int radius, float* d_g_f2m,
               float* d_g_m2m, double* host_value, void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  if (radius < 1 || radius > 16) return fail(ANTS_ERR_INVALID, "radius must be in 1..16");
  const GridDev gd = to_dev(g);
  OpScratch sc;
  ANTS_TRY(sc.stats.ensure(sizeof(RedSlot)));
  ANTS_TRY(sc.partials.ensure(cc_partials_needed(gd, radius) * sizeof(Partial)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemsetAsync(sc.stats.p, 0, sizeof(RedSlot), s));
  launch_cc(gd, d_fixed, d_moving, radius, std::numeric_limits<float>::epsilon(), d_g_f2m, d_g_m2m,
            static_cast<Partial*>(sc.partials.p), static_cast<RedSlot*>(sc.stats.p), s);
  HIP_TRY(hipGetLastError());
  RedSlot r;
  HIP_TRY(hipMemcpyAsync(&r, sc.stats.p, sizeof(r), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (host_value) *host_value = -r.sum / static_cast<double>(g.count());
  return ANTS_OK;
}

int ants_op_inverse_consistency(const ants_image_desc* gdsc, const float* d_fwd, const float* d_inv, int margin,
                                double* host_max, double* host_mean, int64_t* host_count, void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  if (!d_fwd || !d_inv) return fail(ANTS_ERR_INVALID, "null field buffer");
  if (margin < 0) return fail(ANTS_ERR_INVALID, "margin must be >= 0");
  int64_t count = 1;
  for (int d = 0; d < g.D; ++d) count *= std::max<int64_t>(0, g.n[d] - 2 * static_cast<int64_t>(margin));
  const GridDev gd = to_dev(g);
  OpScratch sc;
  ANTS_TRY(sc.stats.ensure(sizeof(RedSlot)));
  ANTS_TRY(sc.partials.ensure(ic_partials_needed(gd) * sizeof(Partial)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_ic_residual(gd, d_fwd, d_inv, margin, static_cast<Partial*>(sc.partials.p), static_cast<RedSlot*>(sc.stats.p),
                     s);
  HIP_TRY(hipGetLastError());
  RedSlot r;
  HIP_TRY(hipMemcpyAsync(&r, sc.stats.p, sizeof(r), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float mx;
  std::memcpy(&mx, &r.maxbits, 4);
  if (host_max) *host_max = mx;
  if (host_mean) *host_mean = count > 0 ? r.sum / static_cast<double>(count) : 0.0;
  if (host_count) *host_count = count;
  return ANTS_OK;
}

int ants_op_compose(const ants_image_desc* gdsc, const float* d_disp, const float* d_warp, float* d_out,
                    void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float* disp[1] = {d_disp};
  const float* warp[1] = {d_warp};
  float* out[1] = {d_out};
  launch_compose(to_dev(g), 1, disp, warp, out, nullptr, 1.f, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return ANTS_OK;
}

int ants_op_scale_update(const ants_image_desc* gdsc, float* d_u, double rate, double* host_max, void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  const size_t b = static_cast<size_t>(g.count()) * g.D * sizeof(float);
  DevBuf a, zero, m;
  ANTS_TRY(a.ensure(b));
  ANTS_TRY(zero.ensure(b));
  ANTS_TRY(m.ensure(16));
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemsetAsync(m.p, 0, 16, s));
  HIP_TRY(hipMemsetAsync(zero.p, 0, b, s));
  Taps t{};
  t.r = 0;
  t.c[0] = 1.f;
  const GridDev gd = to_dev(g);
  // copy + max norm, then u = 0 + (rate/max) * copy(x + 0): the same scaling the engine fuses into composition
  const float* in[1] = {d_u};
  float* out[1] = {a.f()};
  launch_gauss_field_final(gd, 1, in, in, out, t, 0, 1.f, 0.f, false, static_cast<unsigned int*>(m.p), s);
  const float* disp[1] = {a.f()};
  const float* warp[1] = {zero.f()};
  float* res[1] = {d_u};
  launch_compose(gd, 1, disp, warp, res, static_cast<unsigned int*>(m.p), static_cast<float>(rate), s);
  HIP_TRY(hipGetLastError());
  unsigned int bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, m.p, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float mx;
  std::memcpy(&mx, &bits, 4);
  if (host_max) *host_max = mx;
  return ANTS_OK;
}

}  // extern "C"

namespace {

// The whole InvertDisplacementFieldImageFilter run on caller-provided
// scratch: step slots (kInvSlots RedSlots, zeroed here), per-block partials
// (2 step parities) and candidate masks (2 step parities).
int invert_on(hipStream_t s, const HostGrid& g, const float* d_field, const float* d_init, float* d_inv,
              RedSlot* d_slots, Partial* d_parts, uint64_t* d_masks, int* host_iters, double* host_max,
              double* host_mean) {
  if (!d_field || !d_inv) return fail(ANTS_ERR_INVALID, "null field buffer");
  const GridDev gd = to_dev(g);
  const size_t b = static_cast<size_t>(g.count()) * g.D * sizeof(float);
  HIP_TRY(hipMemsetAsync(d_slots, 0, sizeof(RedSlot) * kInvSlots, s));
  if (d_init)
    HIP_TRY(hipMemcpyAsync(d_inv, d_init, b, hipMemcpyDeviceToDevice, s));
  else
    HIP_TRY(hipMemsetAsync(d_inv, 0, b, s));
  const float* field[1] = {d_field};
  float* inv[1] = {d_inv};
  RedSlot* slots[1] = {d_slots};
  launch_inv_step(gd, 1, field, inv, slots, d_parts, d_masks, -1, s);
  for (int j = 0; j <= kInvMaxIter; ++j) launch_inv_step(gd, 1, field, inv, slots, d_parts, d_masks, j, s);
  HIP_TRY(hipGetLastError());
  RedSlot r[kInvSlots];
  HIP_TRY(hipMemcpyAsync(r, d_slots, sizeof(r), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  int k = 0;
  for (int j = 0; j < kInvSlots; ++j)
    if (r[j].computed) k = j + 1;
  if (host_iters) *host_iters = k;
  if (k > 0) {
    float mx;
    std::memcpy(&mx, &r[k - 1].maxbits, 4);
    if (host_max) *host_max = mx;
    if (host_mean) *host_mean = r[k - 1].sum / static_cast<double>(g.count());
  }
  return ANTS_OK;
}

// Scratch of ants_op_invert_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes the array's name
// to a kernel or copy (on the host it is only a stand-in for the symbol);
// it uses the per-device address that hipGetSymbolAddress returns.
constexpr size_t kNoAllocScratchBytes = size_t{8} << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch, looked up once per device.
int scratch_of_current_device(void** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ANTS_ERR_UNSUPPORTED, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ANTS_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ANTS_OK;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

extern "C" {

int ants_op_invert(const ants_image_desc* gdsc, const float* d_field, const float* d_init, float* d_inv,
                   int* host_iters, double* host_max, double* host_mean, void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  const GridDev gd = to_dev(g);
  OpScratch sc;
  ANTS_TRY(sc.stats.ensure(sizeof(RedSlot) * kInvSlots));
  ANTS_TRY(sc.partials.ensure(2 * inv_partials_needed(gd) * sizeof(Partial)));  // 2 step parities
  DevBuf masks;
  ANTS_TRY(masks.ensure(2 * inv_mask_words(gd) * sizeof(uint64_t)));
  return invert_on(static_cast<hipStream_t>(stream), g, d_field, d_init, d_inv, static_cast<RedSlot*>(sc.stats.p),
                   static_cast<Partial*>(sc.partials.p), static_cast<uint64_t*>(masks.p), host_iters, host_max,
                   host_mean);
}

int ants_op_invert_noalloc(const ants_image_desc* gdsc, const float* d_field, const float* d_init, float* d_inv,
                           int* host_iters, double* host_max, double* host_mean, void* stream) {
  HostGrid g;
  ANTS_TRY(to_host_grid(gdsc, &g));
  const GridDev gd = to_dev(g);
  // layout: step slots | partials (2 parities) | masks (2 parities)
  const size_t off_parts = align256(sizeof(RedSlot) * kInvSlots);
  const size_t off_masks = align256(off_parts + 2 * inv_partials_needed(gd) * sizeof(Partial));
  const size_t need = off_masks + 2 * inv_mask_words(gd) * sizeof(uint64_t);
  if (need > kNoAllocScratchBytes)
    return fail(ANTS_ERR_INVALID,
                "ants_op_invert_noalloc: a grid of %lld voxels needs %zu scratch bytes, the fixed scratch holds %zu "
                "(about 33 million voxels at most)",
                static_cast<long long>(g.count()), need, kNoAllocScratchBytes);
  void* base = nullptr;
  ANTS_TRY(scratch_of_current_device(&base));
  char* p = static_cast<char*>(base);
  return invert_on(static_cast<hipStream_t>(stream), g, d_field, d_init, d_inv, reinterpret_cast<RedSlot*>(p),
                   reinterpret_cast<Partial*>(p + off_parts), reinterpret_cast<uint64_t*>(p + off_masks), host_iters,
                   host_max, host_mean);
}

int ants_op_resample_field(const ants_image_desc* gi, const float* d_in, const ants_image_desc* go, float* d_out,
                           void* stream) {
  HostGrid a, b;
  ANTS_TRY(to_host_grid(gi, &a));
  ANTS_TRY(to_host_grid(go, &b));
  if (a.D != b.D) return fail(ANTS_ERR_INVALID, "dimension mismatch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float* in[1] = {d_in};
  float* out[1] = {d_out};
  launch_resample_fields(to_dev(a), to_dev(b), 1, in, out, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return ANTS_OK;
}

int ants_op_preprocess(int64_t n, const float* d_in, float* d_out, void* stream) {
  if (n < 1) return fail(ANTS_ERR_INVALID, "empty image");
  DevBuf mm;
  uint32_t pair[2];
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = preprocess_on(s, mm, d_in, d_out, n, pair);
  if (rc == ANTS_OK) HIP_TRY(hipStreamSynchronize(s));
  mm.release();
  return rc;
}

int ants_apply_transform_labels_nn(const ants_image_desc* go, const float* d_field, const ants_image_desc* gl,
                                   const int32_t* d_labels, int32_t* d_out, void* stream) {
  HostGrid a, b;
  ANTS_TRY(to_host_grid(go, &a));
  ANTS_TRY(to_host_grid(gl, &b));
  if (a.D != b.D) return fail(ANTS_ERR_INVALID, "dimension mismatch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_labels_nn(to_dev(a), d_field, to_dev(b), d_labels, d_out, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return ANTS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Launch-shape tuning table (tune_* in the kernel library).
namespace {

const char* const kKnobNames[kTuneKnobCount] = {"inv_blocks", "cc_zchunk", "smooth_zchunk", "cc_tile", "smooth_tile"};

int knob_of(const char* name) {
  for (int k = 0; k < kTuneKnobCount; ++k)
    if (std::strcmp(name, kKnobNames[k]) == 0) return k;
  return -1;
}

// Reads "knob bucket value" lines ('#' starts a comment).
int load_tuning_file(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(ANTS_ERR_INVALID, "cannot open tuning file %s", path);
  char line[256];
  int lineno = 0;
  while (std::fgets(line, sizeof(line), f)) {
    ++lineno;
    char* hash = std::strchr(line, '#');
    if (hash) *hash = 0;
    char name[64];
    int bucket = 0, value = 0;
    const int got = std::sscanf(line, "%63s %d %d", name, &bucket, &value);
    if (got <= 0) continue;
    const int k = got == 3 ? knob_of(name) : -1;
    if (k < 0 || tune_set(k, bucket, value) != 0) {
      std::fclose(f);
      return fail(ANTS_ERR_INVALID, "%s:%d: bad tuning entry", path, lineno);
    }
  }
  std::fclose(f);
  return ANTS_OK;
}

// At library load: the table committed beside the library (tuning.txt in
// the directory of libants_syn.so), or ANTS_SYN_TUNING=<file>; "none"
// leaves the table empty (the built-in launch models decide).
struct DefaultTuning {
  DefaultTuning() {
    const char* env = std::getenv("ANTS_SYN_TUNING");
    if (env && std::strcmp(env, "none") == 0) return;
    std::string path;
    if (env && env[0]) {
      path = env;
    } else {
      Dl_info info;
      if (!dladdr(reinterpret_cast<void*>(&ants_version), &info) || !info.dli_fname) return;
      path = info.dli_fname;
      const size_t slash = path.find_last_of('/');
      path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning.txt";
      FILE* f = std::fopen(path.c_str(), "r");
      if (!f) return;
      std::fclose(f);
    }
    if (load_tuning_file(path.c_str()) != ANTS_OK)
      std::fprintf(stderr, "ants_syn: tuning table not loaded: %s\n", g_last_error.c_str());
  }
};
DefaultTuning g_default_tuning;

}  // namespace

extern "C" {

int ants_tuning_set(int knob, int log2_voxels, int value) {
  if (tune_set(knob, log2_voxels, value) != 0)
    return fail(ANTS_ERR_INVALID, "bad tuning entry (knob %d, bucket %d, value %d)", knob, log2_voxels, value);
  return ANTS_OK;
}

int ants_tuning_get(int knob, int log2_voxels, int* value) {
  if (!value || knob < 0 || knob >= kTuneKnobCount || log2_voxels < 0 || log2_voxels >= kTuneBuckets)
    return fail(ANTS_ERR_INVALID, "bad tuning query");
  *value = tune_get(knob, int64_t{1} << log2_voxels);
  return ANTS_OK;
}

int ants_tuning_clear(void) {
  tune_clear();
  return ANTS_OK;
}

int ants_tuning_load(const char* path) {
  if (!path) return fail(ANTS_ERR_INVALID, "null path");
  return load_tuning_file(path);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// antsApplyTransforms subset.
namespace {

// Geometry with direction cosines: m = direction * diag(spacing) and its
// inverse diag(1/spacing) * direction^T (orthonormal directions only).
// One transform of a composite as the apply kernel takes it.
int xform_to_dev(const ants_transform& T, int D, XformDev* Xp) {
  XformDev& X = *Xp;
  std::memset(&X, 0, sizeof(X));
  X.kind = T.kind;
  if (T.kind == ANTS_TRANSFORM_IDENTITY) return ANTS_OK;
  if (T.kind == ANTS_TRANSFORM_AFFINE) {
    // MatrixOffsetTransformBase: offset = t + c - A c
    for (int r = 0; r < D; ++r) {
      double ac = 0.0;
      for (int k = 0; k < D; ++k) {
        X.a[r * 3 + k] = T.matrix[r * D + k];
        ac = ac + T.matrix[r * D + k] * T.center[k];
      }
      X.off[r] = T.translation[r] + T.center[r] - ac;
    }
    return ANTS_OK;
  }
  if (T.kind == ANTS_TRANSFORM_DISPLACEMENT) {
    ANTS_TRY(geom_from_desc(&T.field, &X.fg));
    if (X.fg.D != D) return fail(ANTS_ERR_INVALID, "displacement field dimension differs");
    if (!T.d_field) return fail(ANTS_ERR_INVALID, "null displacement field");
    X.f = T.d_field;
    return ANTS_OK;
  }
  return fail(ANTS_ERR_INVALID, "unknown transform kind %d", T.kind);
}

int geom_from_desc(const ants_image_desc* d, GeomDev* g) {
  if (!d) return fail(ANTS_ERR_INVALID, "null image descriptor");
  if (d->dim != 2 && d->dim != 3) return fail(ANTS_ERR_INVALID, "dim must be 2 or 3 (got %d)", d->dim);
  const int D = d->dim;
  std::memset(g, 0, sizeof(*g));
  g->D = D;
  for (int k = 0; k < 3; ++k) {
    const int64_t n = k < D ? d->size[k] : 1;
    if (n < 1 || n > (int64_t{1} << 30)) return fail(ANTS_ERR_INVALID, "image size out of range");
    g->n[k] = static_cast<int>(n);
    if (k < D && !(d->spacing[k] > 0.0)) return fail(ANTS_ERR_INVALID, "image spacing must be positive");
    g->org[k] = k < D ? d->origin[k] : 0.0;
  }
  // an all-zero direction means identity (header convention)
  double dir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool all_zero = true;
  for (int i = 0; i < D * D; ++i)
    if (d->direction[i] != 0.0) all_zero = false;
  for (int i = 0; i < D * D; ++i) dir[i] = all_zero ? (i % (D + 1) == 0 ? 1.0 : 0.0) : d->direction[i];
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) {
      double dot = 0.0;
      for (int k = 0; k < D; ++k) dot += dir[k * D + r] * dir[k * D + c];
      if (std::fabs(dot - (r == c ? 1.0 : 0.0)) > 1e-6)
        return fail(ANTS_ERR_UNSUPPORTED, "image direction must be orthonormal");
    }
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) {
      g->m[r * 3 + c] = dir[r * D + c] * d->spacing[c];
      g->minv[r * 3 + c] = dir[c * D + r] / d->spacing[r];
    }
  return ANTS_OK;
}

}  // namespace

// Host helpers shared with the linear-stage engine (linear.hip).
namespace ants {
int host_error(int code, const char* msg) { return fail(code, "%s", msg); }
int host_geom(const ants_image_desc* d, GeomDev* g) { return geom_from_desc(d, g); }
int host_level_desc(const ants_image_desc* fixed, int factor, ants_image_desc* out) {
  HostGrid g;
  ANTS_TRY(to_host_grid(fixed, &g));
  if (!out) return fail(ANTS_ERR_INVALID, "null output descriptor");
  to_desc(level_grid(g, factor), out);
  return ANTS_OK;
}
}  // namespace ants

extern "C" int ants_apply_transforms(const ants_image_desc* ref, const ants_image_desc* in, const float* d_in,
                                     const ants_transform* transforms, int n_transforms, int interpolation,
                                     double default_value, float* d_out, void* stream) {
  ApplyDev P;
  std::memset(&P, 0, sizeof(P));
  ANTS_TRY(geom_from_desc(ref, &P.ref));
  ANTS_TRY(geom_from_desc(in, &P.in));
  const int D = P.ref.D;
  if (P.in.D != D) return fail(ANTS_ERR_INVALID, "reference and input dimensions differ");
  if (!d_in || !d_out) return fail(ANTS_ERR_INVALID, "null image buffer");
  if (n_transforms < 0 || n_transforms > kMaxTransforms || (n_transforms > 0 && !transforms))
    return fail(ANTS_ERR_INVALID, "transform count must be 0..%d", kMaxTransforms);
  const bool bspline =
      interpolation >= ANTS_INTERP_BSPLINE_ORDER(0) && interpolation <= ANTS_INTERP_BSPLINE_ORDER(5);
  const bool sinc =
      interpolation >= ANTS_INTERP_COSINE_WINDOWED_SINC && interpolation <= ANTS_INTERP_WELCH_WINDOWED_SINC;
  if (!bspline && !sinc && (interpolation < ANTS_INTERP_LINEAR || interpolation > ANTS_INTERP_GENERIC_LABEL))
    return fail(ANTS_ERR_INVALID, "unknown interpolation %d", interpolation);
  P.nt = n_transforms;
  P.defv = static_cast<float>(default_value);
  for (int t = 0; t < n_transforms; ++t) ANTS_TRY(xform_to_dev(transforms[t], D, &P.t[t]));
  GridDev go{};
  go.D = D;
  for (int k = 0; k < 3; ++k) go.n[k] = P.ref.n[k];
  set_divisors(go);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!bspline) {
    launch_apply_transforms(go, P, interpolation, d_in, d_out, s);
    HIP_TRY(hipGetLastError());
    return ANTS_OK;
  }
  const int order = interpolation - ANTS_INTERP_BSPLINE_ORDER(0);
  const int64_t NI = static_cast<int64_t>(P.in.n[0]) * P.in.n[1] * (D == 3 ? P.in.n[2] : 1);
  void* coef = nullptr;
  HIP_TRY(hipMallocAsync(&coef, static_cast<size_t>(NI) * sizeof(double), s));
  P.coef = static_cast<const double*>(coef);
  launch_bspline_coefficients(P.in.n, D, d_in, static_cast<double*>(coef), order, s);
  launch_apply_transforms(go, P, 16 + order, d_in, d_out, s);
  const hipError_t le = hipGetLastError();
  const hipError_t fe = hipFreeAsync(coef, s);
  HIP_TRY(le);
  HIP_TRY(fe);
  return ANTS_OK;
}

extern "C" int ants_jacobian_determinant(const ants_image_desc* grid, const float* d_field, int do_log,
                                        int geometric, float* d_out, float* d_gradient, void* stream) {
  GeomDev g;
  ANTS_TRY(geom_from_desc(grid, &g));
  if (!d_field || !d_out) return fail(ANTS_ERR_INVALID, "null field or output buffer");
  if (geometric && d_gradient) return fail(ANTS_ERR_INVALID, "the geometric form has no gradient output");
  const int D = g.D;
  JacobianDev J;
  std::memset(&J, 0, sizeof(J));
  J.D = D;
  for (int k = 0; k < 3; ++k) {
    J.n[k] = g.n[k];
    J.isp[k] = k < D ? 1.0 / grid->spacing[k] : 0.0;
  }
  bool all_zero = true;
  for (int i = 0; i < D * D; ++i)
    if (grid->direction[i] != 0.0) all_zero = false;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      J.R[r * 3 + c] = (r < D && c < D && !all_zero) ? grid->direction[r * D + c] : (r == c ? 1.0 : 0.0);
  GridDev go{};
  go.D = D;
  for (int k = 0; k < 3; ++k) go.n[k] = g.n[k];
  set_divisors(go);
  if (geometric) {
    // the unmoved simplex: index offsets through direction and spacing
    const double s3 = 0.25 * std::sqrt(3.0), t2 = 0.5 / std::sqrt(2.0);
    const double o2[3][2] = {{0.0, s3}, {-0.5, -s3}, {0.5, -s3}};
    const double o3[4][3] = {{0.5, 0.0, -t2}, {-0.5, 0.0, -t2}, {0.0, 0.5, t2}, {0.0, -0.5, t2}};
    double q[4][3] = {};
    for (int v = 0; v <= D; ++v)
      for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) q[v][r] += J.R[r * 3 + c] * grid->spacing[c] * (D == 2 ? o2[v][c] : o3[v][c]);
    double vol;
    if (D == 2) {
      vol = 0.5 * std::fabs((q[0][0] - q[1][0]) * (q[0][1] - q[2][1]) - (q[0][0] - q[2][0]) * (q[0][1] - q[1][1]));
    } else {
      double a[3], b[3], c[3];
      for (int r = 0; r < 3; ++r) {
        a[r] = q[0][r] - q[3][r];
        b[r] = q[1][r] - q[3][r];
        c[r] = q[2][r] - q[3][r];
      }
      vol = std::fabs(a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) +
                      a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
    }
    launch_jacobian_geometric(go, J, d_field, do_log ? 1 : 0, vol, d_out, static_cast<hipStream_t>(stream));
  } else {
    launch_jacobian(go, J, d_field, do_log ? 1 : 0, d_out, d_gradient, static_cast<hipStream_t>(stream));
  }
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}

extern "C" int ants_apply_transforms_to_points(int dim, const double* d_in, int64_t n_points,
                                               const ants_transform* transforms, int n_transforms, double* d_out,
                                               void* stream) {
  if (dim != 2 && dim != 3) return fail(ANTS_ERR_INVALID, "dimension must be 2 or 3");
  if (n_points < 0) return fail(ANTS_ERR_INVALID, "negative point count");
  if (n_points > 0 && (!d_in || !d_out)) return fail(ANTS_ERR_INVALID, "null point buffer");
  if (n_transforms < 0 || n_transforms > kMaxTransforms || (n_transforms > 0 && !transforms))
    return fail(ANTS_ERR_INVALID, "transform count must be 0..%d", kMaxTransforms);
  ApplyDev P;
  std::memset(&P, 0, sizeof(P));
  P.nt = n_transforms;
  for (int t = 0; t < n_transforms; ++t) ANTS_TRY(xform_to_dev(transforms[t], dim, &P.t[t]));
  launch_transform_points(dim, P, d_in, d_out, n_points, static_cast<hipStream_t>(stream));
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}

extern "C" int ants_compose_transforms_to_field(const ants_image_desc* ref, const ants_transform* transforms,
                                                int n_transforms, float* d_out, void* stream) {
  ApplyDev P;
  std::memset(&P, 0, sizeof(P));
  ANTS_TRY(geom_from_desc(ref, &P.ref));
  const int D = P.ref.D;
  if (!d_out) return fail(ANTS_ERR_INVALID, "null output buffer");
  if (n_transforms < 0 || n_transforms > kMaxTransforms || (n_transforms > 0 && !transforms))
    return fail(ANTS_ERR_INVALID, "transform count must be 0..%d", kMaxTransforms);
  P.nt = n_transforms;
  for (int t = 0; t < n_transforms; ++t) ANTS_TRY(xform_to_dev(transforms[t], D, &P.t[t]));
  GridDev go{};
  go.D = D;
  for (int k = 0; k < 3; ++k) go.n[k] = P.ref.n[k];
  set_divisors(go);
  launch_compose_field(go, P, d_out, static_cast<hipStream_t>(stream));
  HIP_TRY(hipGetLastError());
  return ANTS_OK;
}

// ---------------------------------------------------------------------------
// Batched / served registration.
namespace {

double now_s() {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

int64_t footprint_bytes(const ants_syn_params& P, const HostGrid& gF, const HostGrid& gM) {
  int64_t maxN = 0;
  for (int l = 0; l < P.n_levels; ++l) maxN = std::max(maxN, level_grid(gF, P.shrink[l]).count());
  const int64_t NF = gF.count(), NM = gM.count(), D = P.dim;
  int64_t b = 0;
  b += (NF + NM) * 4 * 2;                // preprocessed + smoothed images
  b += std::max(NF, NM) * 4;             // smoothing scratch
  if (P.precision == ANTS_PRECISION_BF16_PYRAMID) b += (NF + NM) * 2;
  b += 2 * maxN * 4;                     // warped images
  b += 8 * maxN * D * 4;                 // middle and work fields
  auto small_var = [](double v) { return v > 0.0 && v < 0.5; };
  if (D == 3 && (small_var(P.update_variance) || small_var(P.total_variance)))
    b += 2 * maxN * D * 4;               // the blend buffers of a small 3-D variance
  if (P.model == ANTS_MODEL_BSPLINE_SYN) {
    // the B-spline lattices of the update and total fits (2 fields each;
    // the largest level of each) and the tables
    int64_t fu = 0, ft = 0;
    for (int l = 0; l < P.n_levels; ++l) {
      const HostGrid g = level_grid(gF, P.shrink[l]);
      int64_t nn[3] = {1, 1, 1}, mu[3] = {1, 1, 1}, mt[3] = {1, 1, 1};
      for (int d = 0; d < D; ++d) {
        nn[d] = g.n[d];
        mu[d] = (static_cast<int64_t>(P.update_mesh[d]) << l) + P.spline_order;
        mt[d] = (static_cast<int64_t>(P.total_mesh[d]) << l) + P.spline_order;
      }
      fu = std::max(fu, (mu[0] * nn[1] * nn[2] + mu[0] * mu[1] * nn[2] + mu[0] * mu[1] * mu[2]) * D);
      if (P.total_mesh[0] > 0)
        ft = std::max(ft, (mt[0] * nn[1] * nn[2] + mt[0] * mt[1] * nn[2] + mt[0] * mt[1] * mt[2]) * D);
      b += (nn[0] + nn[1] + nn[2]) * 2 * 40 + (mu[0] + mu[1] + mu[2] + mt[0] + mt[1] + mt[2]) * 24;
    }
    b += 2 * (fu + ft) * 4 + 4 * bsp_partials_needed() * 4;  // lattices, max-norm partials
  }
  b += 4 * ((maxN + 63) / 64) * 8;       // inversion candidate masks
  b += (NF + NM) * 4 + NM * 4;           // request inputs (images, labels)
  b += 2 * maxN * D * 4 + maxN * 4;      // outputs (fields, labels)
  return b;
}

// An immutable version of a resident fixed template.  ants_batch_set_template
// publishes a new version; queued and running requests keep the version that
// was current when they were submitted (shared ownership), so replacing a
// template never changes or frees data a request is using.
struct TemplateVer {
  ants_image_desc desc{};
  DevBuf img;
  std::atomic<int64_t>* held = nullptr;  // the batch's template byte count
  ~TemplateVer() {
    if (held) held->fetch_sub(static_cast<int64_t>(img.bytes));
  }
};

struct BatchItem {
  int64_t id;
  ants_syn_request req;
  int64_t footprint;
  double t_submit;
  InitialXforms init;  // composed at submit: the request's transform arrays need not outlive it
  std::shared_ptr<TemplateVer> templ;
};

// Per-slot staging buffers of a worker (request inputs and outputs).
struct SlotBufs {
  DevBuf fixed, moving, lab, fwd, inv, lab_out;
  int64_t bytes() const {
    return static_cast<int64_t>(fixed.bytes + moving.bytes + lab.bytes + fwd.bytes + inv.bytes + lab_out.bytes);
  }
  void release() {
    fixed.release();
    moving.release();
    lab.release();
    fwd.release();
    inv.release();
    lab_out.release();
  }
};

// Observer of a batched registration: a non-zero cancel flag stops it at
// the end of the current iteration.
int batch_cancel_cb(void* user, int, int, double, double, const int*) {
  return static_cast<std::atomic<int>*>(user)->load(std::memory_order_relaxed);
}

}  // namespace

// Admission accounting.  Every slot keeps its engine workspace and staging
// buffers between requests (they only grow), so the device memory the batch
// holds is: templates + for each slot, what it holds now, plus for a running
// slot the footprint of its request (the most its buffers can grow by).  A
// request is admitted when that total stays within the budget; an idle slot
// whose held buffers keep the queue head out frees them.
struct ants_batch {
  int slots = 1;
  int64_t budget = 0;
  int64_t next_id = 1;
  int64_t pending = 0;  // submitted, not yet returned to the caller
  bool stop = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<BatchItem> queue;
  std::deque<ants_request_result> done;
  std::vector<std::thread> workers;
  std::vector<hipStream_t> streams;
  std::vector<int64_t> held;     // per slot: device bytes of its engine + staging buffers
  std::vector<int64_t> running;  // per slot: footprint of the running request (0 when idle)
  std::vector<int64_t> run_id;   // per slot: id of the running request (0 when idle)
  std::unique_ptr<std::atomic<int>[]> cancel;  // per slot: cancel flag of the running request
  std::atomic<int64_t> template_bytes{0};
  std::shared_ptr<TemplateVer> templ[8];

  int64_t committed() const {  // under mu
    int64_t b = template_bytes.load();
    for (int k = 0; k < slots; ++k) b += held[k] + running[k];
    return b;
  }
  void worker(int slot);
};

void ants_batch::worker(int slot) {
  ants_syn_handle h;
  h.stream = streams[slot];
  h.own_stream = false;
  SlotBufs sb;
  auto slot_bytes = [&] { return h.device_bytes() + sb.bytes(); };
  for (;;) {
    BatchItem it;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (stop && queue.empty()) return;
        if (!queue.empty()) {
          const int64_t fp = queue.front().footprint;
          const int64_t others = committed() - held[slot];
          if (others + held[slot] + fp <= budget) break;  // fits with this slot's buffers kept
          if (held[slot] > 0) {
            // this slot's idle buffers keep the head out: free them (the
            // head then fits now, or once other slots finish or free theirs)
            h.release_workspace();
            sb.release();
            held[slot] = 0;
            cv_work.notify_all();
            continue;
          }
        }
        cv_work.wait(lk);
      }
      it = std::move(queue.front());
      queue.pop_front();
      running[slot] = it.footprint;
      run_id[slot] = it.id;
      cancel[slot].store(0);
    }
    ants_request_result res{};
    res.id = it.id;
    res.slot = slot;
    res.user = it.req.user;
    res.t_submit = it.t_submit;
    res.t_start = now_s();
    res.footprint_bytes = it.footprint;
    int rc = ANTS_OK;
    const ants_syn_request& r = it.req;
    const ants_image_desc* fd = it.templ ? &it.templ->desc : &r.fixed;
    HostGrid gF, gM;
    rc = validate_params(&r.params);
    if (rc == ANTS_OK) rc = to_host_grid(fd, &gF);
    if (rc == ANTS_OK) rc = to_host_grid(&r.moving, &gM);
    HostGrid gl{};
    if (rc == ANTS_OK) {
      h.P = r.params;
      h.set_initial(it.init);
      gl = level_grid(gF, r.params.shrink[r.params.n_levels - 1]);
      const size_t fb = static_cast<size_t>(gl.count()) * gF.D * sizeof(float);
      if (rc == ANTS_OK) rc = sb.moving.ensure(gM.count() * sizeof(float));
      if (rc == ANTS_OK && !it.templ) rc = sb.fixed.ensure(gF.count() * sizeof(float));
      if (rc == ANTS_OK) rc = sb.fwd.ensure(fb);
      if (rc == ANTS_OK) rc = sb.inv.ensure(fb);
      if (rc == ANTS_OK && r.host_moving_labels) rc = sb.lab.ensure(gM.count() * sizeof(int32_t));
      if (rc == ANTS_OK && r.host_out_labels) rc = sb.lab_out.ensure(gl.count() * sizeof(int32_t));
      hipStream_t s = h.stream;
      auto hip_ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == ANTS_OK) rc = fail(ANTS_ERR_HIP, "%s", hipGetErrorString(e));
      };
      if (rc == ANTS_OK) rc = h.staged_upload(sb.moving.p, r.host_moving, gM.count() * sizeof(float));
      if (rc == ANTS_OK && !it.templ) rc = h.staged_upload(sb.fixed.p, r.host_fixed, gF.count() * sizeof(float));
      if (rc == ANTS_OK && r.host_moving_labels)
        rc = h.staged_upload(sb.lab.p, r.host_moving_labels, gM.count() * sizeof(int32_t));
      const float* dfix = it.templ ? it.templ->img.f() : sb.fixed.f();
      if (rc == ANTS_OK)
        rc = h.run(fd, dfix, &r.moving, sb.moving.f(), sb.fwd.f(), sb.inv.f(), batch_cancel_cb, &cancel[slot]);
      if (rc == ANTS_OK && r.host_moving_labels && r.host_out_labels) {
        launch_labels_nn(to_dev(gl), sb.fwd.f(), to_dev(gM), static_cast<const int32_t*>(sb.lab.p),
                         static_cast<int32_t*>(sb.lab_out.p), s);
        hip_ok(hipGetLastError());
        if (rc == ANTS_OK) rc = h.staged_download(r.host_out_labels, sb.lab_out.p, gl.count() * sizeof(int32_t));
      }
      if (rc == ANTS_OK && r.host_out_fwd) rc = h.staged_download(r.host_out_fwd, sb.fwd.p, fb);
      if (rc == ANTS_OK && r.host_out_inv) rc = h.staged_download(r.host_out_inv, sb.inv.p, fb);
      hip_ok(hipStreamSynchronize(s));
    }
    res.status = rc;
    if (rc != ANTS_OK) std::snprintf(res.error, sizeof(res.error), "%s", g_last_error.c_str());
    res.voxel_iterations = h.run_stats.voxel_iterations;
    res.algorithmic_bytes = h.run_stats.algorithmic_bytes;
    res.t_end = now_s();
    it.templ.reset();  // a replaced template version is freed once no request holds it
    {
      std::lock_guard<std::mutex> lk(mu);
      held[slot] = slot_bytes();
      running[slot] = 0;
      run_id[slot] = 0;
      done.push_back(res);
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

double ants_batch_clock(void) { return now_s(); }

int64_t ants_syn_footprint(const ants_syn_params* p, const ants_image_desc* fixed, const ants_image_desc* moving) {
  if (validate_params(p) != ANTS_OK) return -1;
  HostGrid gF, gM;
  if (to_host_grid(fixed, &gF) != ANTS_OK || to_host_grid(moving, &gM) != ANTS_OK) return -1;
  return footprint_bytes(*p, gF, gM);
}

int ants_batch_create(const ants_batch_config* cfg, ants_batch** out) {
  if (!cfg || !out) return fail(ANTS_ERR_INVALID, "null argument");
  if (cfg->slots < 1 || cfg->slots > 64) return fail(ANTS_ERR_INVALID, "slots must be in 1..64");
  *out = nullptr;
  auto* b = new (std::nothrow) ants_batch();
  if (!b) return fail(ANTS_ERR_OOM, "out of host memory");
  b->slots = cfg->slots;
  b->budget = cfg->memory_budget;
  if (b->budget <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete b;
      return fail(ANTS_ERR_HIP, "hipMemGetInfo failed");
    }
    b->budget = static_cast<int64_t>(0.9 * static_cast<double>(fr));
  }
  // every slot's stream exists before any worker starts: a failure is the
  // caller's error, not a worker that silently never serves
  for (int i = 0; i < b->slots; ++i) {
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
      for (hipStream_t t : b->streams) (void)hipStreamDestroy(t);
      delete b;
      return fail(ANTS_ERR_HIP, "hipStreamCreateWithFlags for slot %d: %s", i, hipGetErrorString(e));
    }
    b->streams.push_back(s);
  }
  b->held.assign(b->slots, 0);
  b->running.assign(b->slots, 0);
  b->run_id.assign(b->slots, 0);
  b->cancel.reset(new std::atomic<int>[b->slots]);
  for (int i = 0; i < b->slots; ++i) b->cancel[i].store(0);
  for (int i = 0; i < b->slots; ++i) b->workers.emplace_back(&ants_batch::worker, b, i);
  *out = b;
  return ANTS_OK;
}

void ants_batch_destroy(ants_batch* b) {
  if (!b) return;
  {
    std::lock_guard<std::mutex> lk(b->mu);
    b->stop = true;
  }
  b->cv_work.notify_all();
  for (auto& t : b->workers) t.join();
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  delete b;
}

int ants_batch_set_template(ants_batch* b, int index, const ants_image_desc* g, const float* host_image) {
  if (!b || !g || !host_image) return fail(ANTS_ERR_INVALID, "null argument");
  if (index < 0 || index >= 8) return fail(ANTS_ERR_INVALID, "template index must be in 0..7");
  HostGrid hg;
  ANTS_TRY(to_host_grid(g, &hg));
  const int64_t bytes = hg.count() * static_cast<int64_t>(sizeof(float));
  {
    std::lock_guard<std::mutex> lk(b->mu);
    const int64_t replaced = b->templ[index] ? static_cast<int64_t>(b->templ[index]->img.bytes) : 0;
    if (b->committed() + bytes > b->budget + replaced)
      return fail(ANTS_ERR_OOM, "template of %lld bytes does not fit the budget of %lld", (long long)bytes,
                  (long long)b->budget);
  }
  // a new immutable version, complete on the device before it is published
  auto v = std::make_shared<TemplateVer>();
  ANTS_TRY(v->img.ensure(static_cast<size_t>(bytes)));
  HIP_TRY(hipMemcpy(v->img.p, host_image, static_cast<size_t>(bytes), hipMemcpyHostToDevice));
  v->desc = *g;
  v->held = &b->template_bytes;
  b->template_bytes.fetch_add(static_cast<int64_t>(v->img.bytes));
  std::shared_ptr<TemplateVer> old;
  {
    std::lock_guard<std::mutex> lk(b->mu);
    old = std::move(b->templ[index]);
    b->templ[index] = std::move(v);
  }
  old.reset();  // freed here unless queued or running requests still hold it
  b->cv_work.notify_all();
  return ANTS_OK;
}

int ants_batch_submit(ants_batch* b, const ants_syn_request* r, int64_t* id) {
  if (!b || !r) return fail(ANTS_ERR_INVALID, "null argument");
  ANTS_TRY(validate_params(&r->params));
  std::shared_ptr<TemplateVer> tv;
  if (r->use_template) {
    if (r->template_index < 0 || r->template_index >= 8) return fail(ANTS_ERR_INVALID, "template index out of range");
    std::lock_guard<std::mutex> lk(b->mu);
    tv = b->templ[r->template_index];
    if (!tv) return fail(ANTS_ERR_INVALID, "template %d is not set", r->template_index);
  }
  const ants_image_desc* fd = tv ? &tv->desc : &r->fixed;
  HostGrid gF, gM;
  ANTS_TRY(to_host_grid(fd, &gF));
  ANTS_TRY(to_host_grid(&r->moving, &gM));
  if (gF.D != r->params.dim || gM.D != r->params.dim) return fail(ANTS_ERR_INVALID, "dimension mismatch");
  if (!r->host_moving || (!tv && !r->host_fixed)) return fail(ANTS_ERR_INVALID, "null image");
  const int64_t fp = footprint_bytes(r->params, gF, gM);
  InitialXforms init;
  ANTS_TRY(compose_initial(r->params.dim, r->fixed_initial, r->n_fixed_initial, r->moving_initial,
                           r->n_moving_initial, &init));
  BatchItem it{0, *r, fp, now_s(), init, tv};
  {
    std::lock_guard<std::mutex> lk(b->mu);
    const int64_t tb = b->template_bytes.load();
    if (fp + tb > b->budget)
      return fail(ANTS_ERR_OOM, "request needs %lld bytes, the budget less resident templates is %lld",
                  (long long)fp, (long long)(b->budget - tb));
    it.id = b->next_id++;
    if (id) *id = it.id;
    b->queue.push_back(std::move(it));
    b->pending += 1;
  }
  b->cv_work.notify_all();
  return ANTS_OK;
}

int ants_batch_cancel(ants_batch* b, int64_t id) {
  if (!b) return fail(ANTS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(b->mu);
  for (auto q = b->queue.begin(); q != b->queue.end(); ++q) {
    if (q->id != id) continue;
    ants_request_result res{};
    res.id = id;
    res.status = ANTS_ERR_CANCELLED;
    res.slot = -1;
    res.user = q->req.user;
    res.t_submit = q->t_submit;
    res.t_start = res.t_end = now_s();
    res.footprint_bytes = q->footprint;
    std::snprintf(res.error, sizeof(res.error), "cancelled before it started");
    b->queue.erase(q);
    b->done.push_back(res);
    b->cv_done.notify_all();
    b->cv_work.notify_all();
    return ANTS_OK;
  }
  for (int k = 0; k < b->slots; ++k)
    if (b->run_id[k] == id) {
      b->cancel[k].store(1);
      return ANTS_OK;
    }
  return fail(ANTS_ERR_INVALID, "request %lld is neither queued nor running", static_cast<long long>(id));
}

int64_t ants_batch_device_bytes(ants_batch* b) {
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(b->mu);
  return b->committed();
}

static int take_results(ants_batch* b, ants_request_result* out, int max_results, int* n_out) {
  int n = 0;
  while (n < max_results && !b->done.empty()) {
    out[n++] = b->done.front();
    b->done.pop_front();
    b->pending -= 1;
  }
  if (n_out) *n_out = n;
  return ANTS_OK;
}

int ants_batch_poll(ants_batch* b, ants_request_result* out, int max_results, int* n_out) {
  if (!b || !out) return fail(ANTS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(b->mu);
  return take_results(b, out, max_results, n_out);
}

int ants_batch_wait(ants_batch* b, ants_request_result* out, int max_results, int* n_out) {
  if (!b || !out) return fail(ANTS_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(b->mu);
  b->cv_done.wait(lk, [&] { return !b->done.empty() || b->pending == 0; });
  return take_results(b, out, max_results, n_out);
}

int64_t ants_batch_pending(ants_batch* b) {
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(b->mu);
  return b->pending;
}

}  // extern "C"

// ===========================================================================
// Dense field-transform stages with the CC metric: -t GaussianDisplacementField
// and -t Exponential (Examples/itkantsRegistrationHelper.hxx:1733-1847,
// 3112-3277), each run by ImageRegistrationMethodv4.  Per level: the virtual
// domain is the shrunk fixed grid, the images are smoothed at full
// resolution, the transform's field (displacement u, or velocity v with
// u = exp(v)) is resampled onto the new grid.  Per iteration: the metric --
// fixed image at the virtual points through the fixed initial transforms,
// moving image at y = movingInit(x + u(x)) and its central-difference
// gradient there, CC box sums over the virtual grid, value -sum localCC / N,
// derivative c1 * A^T grad M(y) per voxel (local support: not divided by N)
// -- then the optimizer: GaussianDisplacementField runs the conjugate-
// gradient line search of every non-SyN stage (hxx:1290-1372; rate
// estimated at each level's first iteration as step / max |g|, scales 1 for
// a field in physical units) with the update u <- u + smooth(w, updVar),
// then smooth(u, totVar); Exponential runs plain gradient descent without
// a scales estimator (hxx:1405-1418): v <- v + smooth(step g, updVar),
// then smooth(v, velVar), u = exp(v) by scaling and squaring.
namespace {

// Recursive golden-section search over [a, c] with interior b (published
// recursive form; epsilon 0.2, at most 20 steps: the helper's line-search
// settings, hxx:1357-1362).
template <class F>
struct GoldenSection {
  F f;
  int it = 0;
  int rc = ANTS_OK;
  double search(double a, double b, double c, double fb, bool have_fb) {
    const double resphi = 2.0 - (1.0 + std::sqrt(5.0)) / 2.0;
    if (it > 20) return (c + a) / 2.0;
    ++it;
    const double x = (c - b > b - a) ? b + resphi * (c - b) : b - resphi * (b - a);
    if (std::fabs(c - a) < 0.2 * (std::fabs(b) + std::fabs(x))) return (c + a) / 2.0;
    double fx = 0.0;
    if ((rc = f(x, &fx)) != ANTS_OK) return b;
    if (!have_fb && (rc = f(b, &fb)) != ANTS_OK) return b;
    if (fx < fb) return (c - b > b - a) ? search(b, x, c, fx, true) : search(a, x, b, fx, true);
    return (c - b > b - a) ? search(a, b, x, fb, true) : search(x, b, c, fb, true);
  }
};

int validate_field_params(const ants_field_params* p) {
  if (!p) return fail(ANTS_ERR_INVALID, "null parameters");
  if (p->dim != 2 && p->dim != 3) return fail(ANTS_ERR_INVALID, "dim must be 2 or 3");
  if (p->model < ANTS_FIELD_GAUSSIAN_DISPLACEMENT || p->model > ANTS_FIELD_BSPLINE_EXPONENTIAL)
    return fail(ANTS_ERR_UNSUPPORTED,
                "field model must be GaussianDisplacementField, Exponential, BSplineDisplacementField or "
                "BSplineExponential");
  if (p->model >= ANTS_FIELD_BSPLINE_DISPLACEMENT) {
    if (p->spline_order < 1 || p->spline_order > 3) return fail(ANTS_ERR_UNSUPPORTED, "B-spline order must be 1..3");
    for (int d = 0; d < p->dim; ++d)
      if (p->update_mesh[d] < 0 || p->total_mesh[d] < 0 || p->update_mesh[d] > (1 << 12) ||
          p->total_mesh[d] > (1 << 12))
        return fail(ANTS_ERR_INVALID, "B-spline meshes must be in 0..4096 per axis");
  }
  if (p->n_levels < 1 || p->n_levels > ANTS_MAX_LEVELS) return fail(ANTS_ERR_INVALID, "n_levels must be 1..8");
  for (int l = 0; l < p->n_levels; ++l) {
    if (p->shrink[l] < 1) return fail(ANTS_ERR_INVALID, "shrink factors must be >= 1");
    if (!(p->sigma[l] >= 0.0)) return fail(ANTS_ERR_INVALID, "smoothing sigmas must be >= 0");
    if (p->iterations[l] < 0) return fail(ANTS_ERR_INVALID, "iterations must be >= 0");
  }
  if (p->convergence_window < 1) return fail(ANTS_ERR_INVALID, "convergence window must be >= 1");
  if (!(p->learning_rate > 0.0)) return fail(ANTS_ERR_INVALID, "gradient step must be > 0");
  if (!(p->update_variance >= 0.0) || !(p->total_variance >= 0.0))
    return fail(ANTS_ERR_INVALID, "field variances must be >= 0");
  if (p->integration_steps < 0 || p->integration_steps > 30)
    return fail(ANTS_ERR_INVALID, "integration steps must be 0..30");
  if (p->cc_radius < 1 || p->cc_radius > 16) return fail(ANTS_ERR_INVALID, "CC radius must be 1..16");
  return ANTS_OK;
}

struct FieldEngine {
  ants_field_params P{};
  int D = 3;
  hipStream_t s = nullptr;
  HostGrid gF{}, gM{}, gl{};
  GeomDev fgeo{}, mgeo{}, lgeo{};
  FieldMap fmap{}, mmap{};
  DevBuf F, M, Fs, Ms, tmp, minmax;
  DevBuf Fw, Mw, G;
  DevBuf state, u, g, last, conj, trial, upd, sm, work, extra, phi2, sdir;
  DevBuf ccpart, slot, dpart, dout, mbits;
  int64_t evaluations = 0;
  double minsp = 1.0;
  // B-spline forms: the level's fit tables of the update and total fields
  // (mesh << (level + 1); a zero mesh on any axis: no fit)
  bool fit_u = false, fit_t = false;
  BspLayout LU{}, LT{};
  BspDev TU{}, TT{};
  DevBuf tabU, tabT, lat, bpart;

  bool bspline() const { return P.model >= ANTS_FIELD_BSPLINE_DISPLACEMENT; }
  bool expo() const { return P.model == ANTS_FIELD_EXPONENTIAL || P.model == ANTS_FIELD_BSPLINE_EXPONENTIAL; }
  bool smooths_update() const { return bspline() ? fit_u : P.update_variance > 0.0; }
  bool smooths_total() const { return bspline() ? fit_t : P.total_variance > 0.0; }

  int level_tables(int l) {
    fit_u = fit_t = false;
    if (!bspline()) return ANTS_OK;
    int mu[3] = {1, 1, 1}, mt[3] = {1, 1, 1};
    bool zu = false, zt = false;
    for (int d = 0; d < D; ++d) {
      mu[d] = P.update_mesh[d] << (l + 1);
      mt[d] = P.total_mesh[d] << (l + 1);
      zu = zu || mu[d] == 0;
      zt = zt || mt[d] == 0;
    }
    int64_t lat_floats = 1;
    if (!zu) {
      std::vector<unsigned char> blob;
      ANTS_TRY(bsp_tables(gl, mu, P.spline_order, blob, &LU));
      ANTS_TRY(tabU.ensure(blob.size()));
      HIP_TRY(hipMemcpyAsync(tabU.p, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
      HIP_TRY(hipStreamSynchronize(s));  // blob is freed on return
      TU = bsp_dev(LU, tabU.p);
      lat_floats = std::max(lat_floats, LU.field_floats());
      fit_u = true;
    }
    if (!zt) {
      std::vector<unsigned char> blob;
      ANTS_TRY(bsp_tables(gl, mt, P.spline_order, blob, &LT));
      ANTS_TRY(tabT.ensure(blob.size()));
      HIP_TRY(hipMemcpyAsync(tabT.p, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
      HIP_TRY(hipStreamSynchronize(s));
      TT = bsp_dev(LT, tabT.p);
      lat_floats = std::max(lat_floats, LT.field_floats());
      fit_t = true;
    }
    ANTS_TRY(lat.ensure(static_cast<size_t>(lat_floats) * sizeof(float)));
    ANTS_TRY(bpart.ensure(static_cast<size_t>(bsp_partials_needed()) * sizeof(float)));
    return ANTS_OK;
  }

  // the update (or total / velocity) field's smoothing: Gaussian with the
  // stage's variance, or the B-spline fit of the level's mesh; in is scratch
  int smooth_field_of(bool total, float* in, float* out) {
    if (!bspline()) return smooth_into(in, out, total ? P.total_variance : P.update_variance);
    const float* i1[1] = {in};
    float* o1[1] = {out};
    float* b1[1] = {lat.f()};
    return bsp_fields_on(s, gl, total ? TT : TU, total ? LT : LU, 1, i1, o1, b1, nullptr, bpart.f());
  }

  int buf(DevBuf& b, int64_t floats) { return b.ensure(static_cast<size_t>(std::max<int64_t>(floats, 1)) * 4); }

  int init(const ants_field_params* p, const ants_image_desc* fd, const float* d_fixed, const ants_image_desc* md,
           const float* d_moving, const ants_transform* mi, int nmi, const ants_transform* fi, int nfi,
           void* stream) {
    ANTS_TRY(validate_field_params(p));
    P = *p;
    D = P.dim;
    s = static_cast<hipStream_t>(stream);
    if (!fd || !md || !d_fixed || !d_moving) return fail(ANTS_ERR_INVALID, "null image");
    if (fd->dim != D || md->dim != D) return fail(ANTS_ERR_INVALID, "image dimension differs from params");
    ANTS_TRY(to_host_grid(fd, &gF));
    ANTS_TRY(to_host_grid(md, &gM));
    ANTS_TRY(geom_from_desc(fd, &fgeo));
    ANTS_TRY(geom_from_desc(md, &mgeo));
    InitialXforms X;
    ANTS_TRY(compose_initial(D, fi, nfi, mi, nmi, &X));
    if (X.general[0] || X.general[1])
      return fail(ANTS_ERR_UNSUPPORTED, "field-transform stages take affine initial transforms only");
    for (int k = 0; k < 9; ++k) {
      fmap.a[k] = X.A[0][k];
      mmap.a[k] = X.A[1][k];
    }
    for (int k = 0; k < 3; ++k) {
      fmap.b[k] = X.o[0][k];
      mmap.b[k] = X.o[1][k];
      fmap.sp[k] = gF.sp[k];
      mmap.sp[k] = gM.sp[k];
    }
    const int64_t NF = gF.count(), NM = gM.count();
    ANTS_TRY(buf(F, NF));
    ANTS_TRY(buf(M, NM));
    ANTS_TRY(buf(Fs, NF));
    ANTS_TRY(buf(Ms, NM));
    ANTS_TRY(buf(tmp, std::max(NF, NM)));
    if (P.preprocess) {
      uint32_t pair[2];
      ANTS_TRY(preprocess_on(s, minmax, d_fixed, F.f(), NF, pair));
      ANTS_TRY(preprocess_on(s, minmax, d_moving, M.f(), NM, pair));
    } else {
      HIP_TRY(hipMemcpyAsync(F.p, d_fixed, NF * sizeof(float), hipMemcpyDeviceToDevice, s));
      HIP_TRY(hipMemcpyAsync(M.p, d_moving, NM * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    int64_t nmax = 1;
    for (int l = 0; l < P.n_levels; ++l) nmax = std::max(nmax, level_grid(gF, P.shrink[l]).count());
    const int64_t nf = nmax * D;
    ANTS_TRY(buf(Fw, nmax));
    ANTS_TRY(buf(Mw, nmax));
    ANTS_TRY(buf(G, nf));
    for (DevBuf* b : {&state, &u, &g, &last, &conj, &trial, &upd, &sm, &work, &extra, &phi2, &sdir})
      ANTS_TRY(buf(*b, nf));
    int64_t ncc = 1;
    for (int l = 0; l < P.n_levels; ++l)
      ncc = std::max(ncc, cc_partials_needed(to_dev(level_grid(gF, P.shrink[l])), P.cc_radius));
    ANTS_TRY(ccpart.ensure(static_cast<size_t>(ncc) * sizeof(Partial)));
    ANTS_TRY(slot.ensure(sizeof(RedSlot)));
    ANTS_TRY(dpart.ensure(static_cast<size_t>(field_dots_partials(nmax)) * 4 * sizeof(double)));
    ANTS_TRY(dout.ensure(4 * sizeof(double)));
    ANTS_TRY(mbits.ensure(16));
    return ANTS_OK;
  }

  int64_t nvox() const { return gl.count(); }

  // the level grid, its smoothed images and the fixed samples; `prev`: the
  // previous level's grid (the state field is resampled from it), or null
  int set_level(int l, const HostGrid* prev) {
    gl = level_grid(gF, P.shrink[l]);
    ants_image_desc ld;
    to_desc(gl, &ld);
    ANTS_TRY(geom_from_desc(&ld, &lgeo));
    minsp = gl.sp[0];
    for (int d = 1; d < D; ++d) minsp = std::min(minsp, gl.sp[d]);
    ANTS_TRY(smooth_image_on(s, gF, F.f(), Fs.f(), tmp.f(), P.sigma[l], P.sigma_in_mm != 0));
    ANTS_TRY(smooth_image_on(s, gM, M.f(), Ms.f(), tmp.f(), P.sigma[l], P.sigma_in_mm != 0));
    const int64_t nf = nvox() * D;
    if (!prev) {
      HIP_TRY(hipMemsetAsync(state.p, 0, nf * sizeof(float), s));
    } else {
      const float* in[1] = {state.f()};
      float* out[1] = {trial.f()};
      launch_resample_fields(to_dev(*prev), to_dev(gl), 1, in, out, s);
      std::swap(state.p, trial.p);
      std::swap(state.bytes, trial.bytes);
    }
    launch_field_sample(lgeo, fgeo, Fs.f(), nullptr, fmap, Fw.f(), nullptr, s);
    HIP_TRY(hipGetLastError());
    ANTS_TRY(level_tables(l));
    if (expo()) {
      int n = 0;
      ANTS_TRY(exponentiate(state.f(), u.f(), 1, P.integration_steps, &n));
    }
    return ANTS_OK;
  }

  int dots(const float* a, const float* b, double out[4]) {
    launch_field_dots(D, nvox(), a, b, static_cast<double*>(dpart.p), static_cast<double*>(dout.p), s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ANTS_OK;
  }

  // value (and with deriv the derivative into g) at displacement ufield
  // (null: 0), or at ufield + s2 u2 when u2 is given
  int metric(const float* ufield, bool deriv, double* value, const float* u2 = nullptr, float s2 = 0.f) {
    ++evaluations;
    launch_field_sample(lgeo, mgeo, Ms.f(), ufield, mmap, Mw.f(), deriv ? G.f() : nullptr, s, u2, s2);
    launch_cc(to_dev(gl), Fw.f(), Mw.f(), P.cc_radius, std::numeric_limits<float>::epsilon(),
              deriv ? g.f() : nullptr, nullptr, static_cast<Partial*>(ccpart.p), static_cast<RedSlot*>(slot.p), s,
              true, false, G.f());
    HIP_TRY(hipGetLastError());
    RedSlot r;
    HIP_TRY(hipMemcpyAsync(&r, slot.p, sizeof(r), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *value = -r.sum / static_cast<double>(nvox());
    return ANTS_OK;
  }

  // out = smooth(in, var) (GaussianSmoothDisplacementField), or a copy when
  // var <= 0; in is used as scratch by the separable path
  int smooth_into(float* in, float* out, double var) {
    if (!(var > 0.0)) {
      if (out != in) HIP_TRY(hipMemcpyAsync(out, in, nvox() * D * sizeof(float), hipMemcpyDeviceToDevice, s));
      return ANTS_OK;
    }
    float* i1[1] = {in};
    float* w1[1] = {work.f()};
    float* o1[1] = {out};
    float* e1[1] = {extra.f()};
    return smooth_fields_on(s, gl, 1, i1, w1, o1, var, static_cast<unsigned int*>(mbits.p), nullptr, e1);
  }

  // ExponentialDisplacementFieldImageFilter: steps given, or automatic (the
  // smallest n with max |v| / 2^n below half the smallest spacing,
  // floor(2 + log2(max |v| / min spacing) + 1), capped at the transform's
  // default of 10); phi = sign v / 2^n, then n times phi <- phi + phi(x + phi)
  int exponentiate(const float* v, float* out, int sign, int steps, int* used) {
    const int64_t nf = nvox() * D;
    int n = steps;
    if (n <= 0) {
      double q[4];
      ANTS_TRY(dots(v, nullptr, q));
      const double m2 = q[3] / (minsp * minsp);
      n = 0;
      if (m2 > 0.0) {
        const double nfl = 2.0 + 0.5 * std::log(m2) / std::log(2.0);
        if (nfl >= 0.0) n = std::min(static_cast<int>(nfl + 1.0), 10);
      }
    }
    *used = n;
    launch_field_axpy(nf, nullptr, static_cast<float>(sign) / static_cast<float>(1u << n), v, out, s);
    float* cur = out;
    float* nxt = phi2.f();
    const GridDev gd = to_dev(gl);
    for (int k = 0; k < n; ++k) {
      const float* dsp[1] = {cur};
      float* o1[1] = {nxt};
      launch_compose(gd, 1, dsp, dsp, o1, nullptr, 1.f, s);
      std::swap(cur, nxt);
    }
    if (cur != out) HIP_TRY(hipMemcpyAsync(out, cur, nf * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipGetLastError());
    return ANTS_OK;
  }

  // the smoothed search direction of this iteration, sdir = smooth(conj):
  // the update smoothing (Gaussian or B-spline fit) is linear, so the update
  // of a step s, smooth(s conj), is s sdir and one smoothing serves every
  // trial of the line search
  int smooth_direction() {
    const int64_t nf = nvox() * D;
    if (!smooths_update()) {
      HIP_TRY(hipMemcpyAsync(sdir.p, conj.p, nf * sizeof(float), hipMemcpyDeviceToDevice, s));
      return ANTS_OK;
    }
    HIP_TRY(hipMemcpyAsync(upd.p, conj.p, nf * sizeof(float), hipMemcpyDeviceToDevice, s));
    return smooth_field_of(false, upd.f(), sdir.f());
  }

  // displacement-form update with step s into dst: dst = state + s sdir,
  // then the total field's smoothing
  int gdf_update(double step, float* dst) {
    const int64_t nf = nvox() * D;
    launch_field_axpy(nf, state.f(), static_cast<float>(step), sdir.f(), dst, s);
    if (smooths_total()) {
      ANTS_TRY(smooth_field_of(true, dst, sm.f()));
      HIP_TRY(hipMemcpyAsync(dst, sm.p, nf * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipGetLastError());
    return ANTS_OK;
  }

  int run(std::vector<ants_linear_record>& trace, int* iterations_out) {
    HostGrid prev{};
    for (int l = 0; l < P.n_levels; ++l) {
      ANTS_TRY(set_level(l, l ? &prev : nullptr));
      prev = gl;
      const int64_t nf = nvox() * D;
      std::vector<double> energies;
      double lr = P.learning_rate;
      int it = 0;
      while (it < P.iterations[l]) {
        double val = 0.0;
        const bool gdf = !expo();
        ANTS_TRY(metric(gdf ? state.f() : u.f(), true, &val));
        energies.push_back(val);
        double conv = 0.0;
        ANTS_TRY(ants_syn_convergence_value(energies.data(), static_cast<int>(energies.size()),
                                            P.convergence_window, &conv));
        if (conv <= P.convergence_threshold && !P.force_iterations) break;
        if (gdf) {
          if (it == 0) {
            HIP_TRY(hipMemsetAsync(conj.p, 0, nf * sizeof(float), s));
            HIP_TRY(hipMemsetAsync(last.p, 0, nf * sizeof(float), s));
            double q[4];
            ANTS_TRY(dots(g.f(), nullptr, q));
            const double ss = std::sqrt(q[3]);
            lr = ss <= std::numeric_limits<double>::epsilon() ? 1.0 : P.learning_rate / ss;
          }
          double q[4];
          ANTS_TRY(dots(g.f(), last.f(), q));  // g.last, last.last, g.g
          double gamma = 0.0;
          if (q[1] > std::numeric_limits<double>::epsilon()) gamma = (q[2] - q[0]) / q[1];
          if (gamma < 0.0 || gamma > 5.0) gamma = 0.0;
          HIP_TRY(hipMemcpyAsync(last.p, g.p, nf * sizeof(float), hipMemcpyDeviceToDevice, s));
          launch_field_axpy(nf, g.f(), static_cast<float>(gamma), conj.f(), conj.f(), s);
          ANTS_TRY(smooth_direction());
          // a trial without total smoothing is sampled at state + s sdir
          // directly (the same float operations as the update writes)
          auto trial_value = [this](double step, double* v) -> int {
            if (!smooths_total()) return metric(state.f(), false, v, sdir.f(), static_cast<float>(step));
            ANTS_TRY(gdf_update(step, trial.f()));
            return metric(trial.f(), false, v);
          };
          GoldenSection<decltype(trial_value)> ls{trial_value};
          const double step = ls.search(lr * 0.0, lr, lr * 2.0, 0.0, false);
          if (ls.rc != ANTS_OK) return ls.rc;
          lr = step;
          ANTS_TRY(gdf_update(lr, trial.f()));
          std::swap(state.p, trial.p);
          std::swap(state.bytes, trial.bytes);
        } else {
          lr = P.learning_rate;
          launch_field_axpy(nf, nullptr, static_cast<float>(lr), g.f(), upd.f(), s);
          const float* src = upd.f();
          if (smooths_update()) {
            ANTS_TRY(smooth_field_of(false, upd.f(), sm.f()));
            src = sm.f();
          }
          launch_field_axpy(nf, state.f(), 1.f, src, trial.f(), s);
          if (smooths_total()) {
            ANTS_TRY(smooth_field_of(true, trial.f(), state.f()));
          } else {
            std::swap(state.p, trial.p);
            std::swap(state.bytes, trial.bytes);
          }
          int n = 0;
          ANTS_TRY(exponentiate(state.f(), u.f(), 1, P.integration_steps, &n));
        }
        trace.push_back(ants_linear_record{l, it + 1, val, conv, lr});
        ++it;
      }
      if (iterations_out) iterations_out[l] = it;
    }
    HIP_TRY(hipStreamSynchronize(s));
    return ANTS_OK;
  }
};

}  // namespace

extern "C" {

int ants_field_register(const ants_field_params* p, const ants_image_desc* fixed, const float* d_fixed,
                        const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                        int n_moving, const ants_transform* fixed_initial, int n_fixed, float* d_out_field,
                        float* d_out_inverse, float* d_out_velocity, ants_linear_record* trace,
                        int64_t trace_capacity, int64_t* trace_count, int* iterations_out, void* stream) {
  if (trace_capacity < 0 || (trace_capacity > 0 && !trace)) return fail(ANTS_ERR_INVALID, "bad trace buffer");
  FieldEngine e;
  ANTS_TRY(e.init(p, fixed, d_fixed, moving, d_moving, moving_initial, n_moving, fixed_initial, n_fixed, stream));
  std::vector<ants_linear_record> tr;
  ANTS_TRY(e.run(tr, iterations_out));
  const int64_t w = std::min<int64_t>(trace_capacity, static_cast<int64_t>(tr.size()));
  for (int64_t i = 0; i < w; ++i) trace[i] = tr[i];
  if (trace_count) *trace_count = static_cast<int64_t>(tr.size());
  const size_t bytes = static_cast<size_t>(e.nvox()) * e.D * sizeof(float);
  hipStream_t s = e.s;
  if (!e.expo()) {
    if (d_out_inverse || d_out_velocity)
      return fail(ANTS_ERR_INVALID, "a displacement-field stage has no inverse or velocity output");
    if (d_out_field) HIP_TRY(hipMemcpyAsync(d_out_field, e.state.p, bytes, hipMemcpyDeviceToDevice, s));
  } else {
    int n = 0;
    if (d_out_field) HIP_TRY(hipMemcpyAsync(d_out_field, e.u.p, bytes, hipMemcpyDeviceToDevice, s));
    if (d_out_inverse) ANTS_TRY(e.exponentiate(e.state.f(), d_out_inverse, -1, e.P.integration_steps, &n));
    if (d_out_velocity) HIP_TRY(hipMemcpyAsync(d_out_velocity, e.state.p, bytes, hipMemcpyDeviceToDevice, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  return ANTS_OK;
}

int ants_field_metric(const ants_field_params* p, int level, const ants_image_desc* fixed, const float* d_fixed,
                      const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                      int n_moving, const ants_transform* fixed_initial, int n_fixed, const float* d_field,
                      double* value, float* d_derivative, void* stream) {
  if (!value) return fail(ANTS_ERR_INVALID, "null value");
  FieldEngine e;
  ANTS_TRY(e.init(p, fixed, d_fixed, moving, d_moving, moving_initial, n_moving, fixed_initial, n_fixed, stream));
  if (level < 0 || level >= p->n_levels) return fail(ANTS_ERR_INVALID, "level out of range");
  e.gl = level_grid(e.gF, p->shrink[level]);
  ants_image_desc ld;
  to_desc(e.gl, &ld);
  ANTS_TRY(geom_from_desc(&ld, &e.lgeo));
  ANTS_TRY(smooth_image_on(e.s, e.gF, e.F.f(), e.Fs.f(), e.tmp.f(), p->sigma[level], p->sigma_in_mm != 0));
  ANTS_TRY(smooth_image_on(e.s, e.gM, e.M.f(), e.Ms.f(), e.tmp.f(), p->sigma[level], p->sigma_in_mm != 0));
  launch_field_sample(e.lgeo, e.fgeo, e.Fs.f(), nullptr, e.fmap, e.Fw.f(), nullptr, e.s);
  ANTS_TRY(e.metric(d_field, d_derivative != nullptr, value));
  if (d_derivative)
    HIP_TRY(hipMemcpyAsync(d_derivative, e.g.p, static_cast<size_t>(e.nvox()) * e.D * sizeof(float),
                           hipMemcpyDeviceToDevice, e.s));
  HIP_TRY(hipStreamSynchronize(e.s));
  return ANTS_OK;
}

int ants_op_exponential(const ants_image_desc* g, const float* d_v, float* d_out, int steps, int sign,
                        int* steps_used, void* stream) {
  if (!g || !d_v || !d_out) return fail(ANTS_ERR_INVALID, "null argument");
  if (sign != 1 && sign != -1) return fail(ANTS_ERR_INVALID, "sign must be 1 or -1");
  if (steps < 0 || steps > 30) return fail(ANTS_ERR_INVALID, "steps must be 0..30");
  FieldEngine e;
  e.s = static_cast<hipStream_t>(stream);
  e.D = g->dim;
  ANTS_TRY(to_host_grid(g, &e.gl));
  e.minsp = e.gl.sp[0];
  for (int d = 1; d < e.D; ++d) e.minsp = std::min(e.minsp, e.gl.sp[d]);
  const int64_t n = e.gl.count();
  ANTS_TRY(e.buf(e.phi2, n * e.D));
  ANTS_TRY(e.dpart.ensure(static_cast<size_t>(field_dots_partials(n)) * 4 * sizeof(double)));
  ANTS_TRY(e.dout.ensure(4 * sizeof(double)));
  int used = 0;
  ANTS_TRY(e.exponentiate(d_v, d_out, sign, steps, &used));
  HIP_TRY(hipStreamSynchronize(e.s));
  if (steps_used) *steps_used = used;
  return ANTS_OK;
}

}  // extern "C"
