// Host-side launch interface of the SyN kernels (gfx950).  All launches are
// asynchronous on the given stream; pointers are device pointers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ants {

// Level/image grid as the kernels see it: identity direction, x fastest.
struct GridDev {
  int D;
  int n[3];
  float sp[3];
  float isp[3];    // 1/spacing (rounded from double): index = (p - origin) * isp, as ITK's
                   // physical-point-to-index matrix does
  float org[3];
  double dsp[3];   // exact geometry for double-precision coordinate paths
  double dorg[3];
  // direction cosines (row-major 3x3, identity-padded): point = org +
  // dir * (index .* sp); `oblique` is 0 for the identity, which keeps the
  // per-axis arithmetic
  int oblique;
  float dir[9];
  double ddir[9];
  // unsigned 32-bit division by n[0] and n[1] as multiply-high + shifts
  // (Granlund & Montgomery, "Division by invariant integers using
  // multiplication", 1994, fig. 4.1); set by set_divisors().
  uint32_t div_m[2];
  uint32_t div_s1[2];
  uint32_t div_s2[2];
  __host__ __device__ int64_t count() const { return static_cast<int64_t>(n[0]) * n[1] * n[2]; }
};

// Fills GridDev::div_* for n[0] and n[1].
inline void set_divisors(GridDev& g) {
  for (int k = 0; k < 2; ++k) {
    const uint64_t d = static_cast<uint64_t>(g.n[k] > 0 ? g.n[k] : 1);
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;  // l = ceil(log2 d)
    g.div_m[k] = static_cast<uint32_t>((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    g.div_s1[k] = l < 1 ? l : 1;
    g.div_s2[k] = l > 1 ? l - 1 : 0;
  }
}

constexpr int kMaxTaps = 64;  // half-kernel length limit (radius <= 63)
struct Taps {
  float c[kMaxTaps];  // c[0..r]
  int r;
};

// Deterministic device reduction slot.  Blocks write plain per-block
// partials; a later launch (the next inversion step, or a one-block reduce
// kernel) sums them in block order, so the result does not depend on
// scheduling and no block waits on a device-wide fence or atomic.
struct RedSlot {
  double sum;
  unsigned int maxbits;
  unsigned int pad0;
  int computed;
  int pad;
};

// Per-block partial of a deterministic grid reduction.
struct Partial {
  double s;
  float m;
  float q;  // inversion: sampled lower bound of the next step's max norm
};

// Reducing elementwise kernels run grid-stride loops over at most this many
// blocks per problem, so the final (last-block) reduction reads few partials.
constexpr int kRedBlocks = 1024;

constexpr int kInvMaxIter = 20;            // InvertDisplacementFieldImageFilter MaximumNumberOfIterations (SyN)
constexpr int kInvSlots = kInvMaxIter + 1;
constexpr int kInvLaunchesDefault = 3;     // inversion steps per batch in an iteration's launch sequence

// Per-iteration device statistics block (zeroed by k_fill at iteration start).
struct IterStats {
  RedSlot energy;                // CC: sum of localCC over voxels
  unsigned int maxnorm[4];       // update-field max |u/spacing| (bits), per problem
  RedSlot inv[4][kInvSlots];     // inversion norms per problem per inner step
};

// Launch-shape tuning table (filled by the project's autotuner, see
// ants_amd/autotune.py): one value per knob and grid-size bucket
// floor(log2(voxels)); 0 means no entry, and the launcher's own model
// decides.  Knob values: inversion blocks per problem; CC and smoothing
// z-chunk depth in planes; CC tile (1: 16x16, 2: 16x8) and smoothing tile
// (1: 16x32, 2: 16x16).
enum TuneKnob : int {
  kTuneInvBlocks = 0,
  kTuneCCZChunk = 1,
  kTuneSmoothZChunk = 2,
  kTuneCCTile = 3,
  kTuneSmoothTile = 4,
  kTuneKnobCount = 5
};
constexpr int kTuneBuckets = 40;
int tune_get(int knob, int64_t voxels);
int tune_set(int knob, int bucket, int value);  // 0 on success, -1 on a bad knob, bucket or value
void tune_clear();
uint64_t tune_generation();  // changes whenever the table does

void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
void launch_minmax(const float* in, int64_t n, uint32_t* out2, hipStream_t s);  // out2: ordered bits (min, max)
void launch_window(const float* in, float* out, int64_t n, float lo, float hi, double scale, double shift,
                   hipStream_t s);

// Separable 1-D Gaussian correlation along `axis` with edge replication, ncomp interleaved components.
void launch_gauss_axis(const GridDev& g, int ncomp, const float* in, float* out, const Taps& t, int axis,
                       hipStream_t s);
// Same as the last pass of SyN field smoothing for `np` fields at once:
// along `axis`, then boundary voxels zeroed, blend w1*smooth + w2*orig,
// and the max norm |out/spacing| of each field into maxbits[k].
// boundary: zero the boundary voxels (SyN smoothing); false for the plain
// copy that variance <= 0 means (then only the max norm is taken).
void launch_gauss_field_final(const GridDev& g, int np, const float* const* in, const float* const* orig,
                              float* const* out, const Taps& t, int axis, float w1, float w2, bool boundary,
                              unsigned int* maxbits, hipStream_t s);

// The whole SyN field smoothing of np 3-D fields in one pass (all three axes,
// boundary zero, blend, max norm).  Returns false when this grid or kernel
// radius is not covered (then use launch_gauss_axis + launch_gauss_field_final).
bool launch_smooth3d(const GridDev& g, int np, const float* const* in, const float* const* orig, float* const* out,
                     const Taps& t, float w1, float w2, bool boundary, unsigned int* maxbits, hipStream_t s);
// The same fused 3-D pass for one scalar image with the same taps on every
// axis (edges replicated), radius 1..6, in != out; false when not applicable.
bool launch_smooth3d_image(const GridDev& g, const float* in, float* out, const Taps& t, hipStream_t s);

// Continuous index in an image grid of the level-grid voxel ix displaced by
// u and then mapped through an affine initial transform y = A p + o, folded
// into one map: c = G ix + Q u + h (built on the host in double).
struct WarpMap {
  float G[9];
  float Q[9];
  float h[3];
};

// Fw = Fs(x + uF(x)), Mw = Ms(x + uM(x)) on level grid gl (Fs on gF, Ms on gM).
// With `maps` (fixed, moving) the continuous indices come from the maps
// instead of the grids' own geometry.
void launch_warp_pair(const GridDev& gl, const GridDev& gF, const float* Fs, const float* uF, float* Fw,
                      const GridDev& gM, const float* Ms, const float* uM, float* Mw, hipStream_t s,
                      uint32_t* clear = nullptr, int nclear = 0, const WarpMap* maps = nullptr);
// Same with the smoothed images stored as BF16.
void launch_warp_pair_bf16(const GridDev& gl, const GridDev& gF, const uint16_t* Fs, const float* uF, float* Fw,
                           const GridDev& gM, const uint16_t* Ms, const float* uM, float* Mw, hipStream_t s,
                           uint32_t* clear = nullptr, int nclear = 0, const WarpMap* maps = nullptr);
// float -> BF16 (round to nearest even).
void launch_to_bf16(const float* in, uint16_t* out, int64_t n, hipStream_t s);
void launch_copy_calibration(const void* in, void* out, int64_t n4, hipStream_t s);
// Single image warp (u may be null).
void launch_warp_image(const GridDev& gi, const float* img, const GridDev& go, const float* u, float* out,
                       hipStream_t s);

// CC value + both derivatives.  `partials` must hold cc_partials_needed(g) entries.
int64_t cc_partials_needed(const GridDev& g, int radius, bool fp32_sums = false);
// Returns the number of partials written; with reduce = false the energy slot
// is left to launch_publish_stats.
// fp32_sums: float box sums over block-shifted samples (the BF16 pyramid
// form's precision) instead of double sums; ANTS_SYN_CC_SUMS=fp32|fp64
// overrides the choice.
// gext (field transforms): the moving image's physical gradient at each
// voxel's mapped point; then gF2M = a1 * gext (not written when gF2M is
// null: the value alone) and gM2M is not written.
int launch_cc(const GridDev& g, const float* F, const float* M, int radius, float eps, float* gF2M, float* gM2M,
              Partial* partials, RedSlot* slot, hipStream_t s, bool reduce = true, bool fp32_sums = false,
              const float* gext = nullptr);
int launch_cc_direct(const GridDev& g, const float* F, const float* M, int radius, float eps, float* gF2M,
                     float* gM2M, Partial* partials, RedSlot* slot, hipStream_t s, bool reduce = true,
                     const float* gext = nullptr);
// One block: reduce cc_n CC partials into *energy (skipped when cc_parts is
// null), copy nwords words of src to pinned host memory dst, then set *flag
// (pinned host memory) to 1 once the copy is visible to the host.
void launch_publish_stats(const Partial* cc_parts, int cc_n, RedSlot* energy, const uint32_t* src, uint32_t* dst,
                          int nwords, uint32_t* flag, hipStream_t s);

// B-spline fit of displacement fields (BSplineSyN's update and total field
// smoothing): one level of multilevel B-spline approximation over every
// voxel, boundary voxels as zero vectors of weight 1e10, evaluated back on
// the grid.  Separable in index space: the numerator lattice is three 1-D
// fit passes (axis tables fold in 1/sum_j b_j^2 and the interior mask), the
// denominator W_k = prod A + (1e10 - 1)(prod A - prod Ai) comes from 1-D sums,
// and the evaluation is three 1-D passes back.  Tables per axis (n voxels,
// m = spans + order control points):
struct BspAxisDev {
  const int* span;   // [n] span of each voxel
  const float* bw;   // [n][4] basis values b_j(t_i), j = 0..order
  const int* lo;     // [m] first voxel whose span lies in [k - order, k]
  const int* goff;   // [m + 1] control point k's fit weights are gw[goff[k] .. goff[k+1])
  const float* gw;   // fit weights b^3 / sum_j b_j^2 of voxels lo[k], lo[k] + 1, ... (zero at voxels 0, n-1)
  const double* A;   // [m] sum of b_{k-s}(t)^2 over all voxels of the range
  const double* Ai;  // [m] the same over the voxels 1..n-2
  int n, m;
  int gtot;          // goff[m] = (order + 1) n
};
struct BspDev {
  BspAxisDev ax[3];
  int D, q;
};
// Fit pass along `axis` of np D-component fields with input dims `dims`
// (the output has m control points on that axis).  `last`: the pass that
// completes the lattice, which it divides by W_k.
void launch_bsp_fit_axis(const BspDev& T, int np, const float* const* in, float* const* out, const int dims[3],
                         int axis, bool last, hipStream_t s);
// Evaluation pass along `axis` from input dims `dims` (m on that axis) to n
// voxels.  maxbits != nullptr (axis 0, the final pass): also each field's
// max |u / spacing| into maxbits[k] (atomic max of the float bits), through
// `part`, scratch of np * bsp_partials_needed() floats.
void launch_bsp_eval_axis(const BspDev& T, int np, const float* const* in, float* const* out, const int dims[3],
                          int axis, const float sp[3], unsigned int* maxbits, float* part, hipStream_t s);
int64_t bsp_partials_needed();
// Per-block sums of I, I x, I y, I z over an image (index coordinates),
// out[block * 4 + k]; returns the block count (at most kRedBlocks).
int launch_moments(const GridDev& g, const float* img, double* out, hipStream_t s);
// Fused evaluation along y and x from the y-unevaluated lattice (m0, m1,
// n2) to the field, with the max norm as above (false when the rows'
// tables exceed LDS; then use the separate passes).
bool launch_bsp_eval_yx(const BspDev& T, int np, const float* const* in, float* const* out, int n2,
                        const float sp[3], unsigned int* maxbits, float* part, hipStream_t s);

// out_k = warp_k + scale_k * disp_k(x + warp_k(x)), scale_k = rate / max_k when maxbits given (max>0), else 1.
void launch_compose(const GridDev& g, int np, const float* const* disp, const float* const* warp,
                    float* const* out, const unsigned int* maxbits, float rate, hipStream_t s);

// Inverse-consistency residual |fwd(x) + inv(x + fwd(x))| / spacing over the
// voxels at least `margin` from every face: slot->sum (double) and
// slot->maxbits (float max).  `partials` holds ic_partials_needed(g) entries.
int64_t ic_partials_needed(const GridDev& g);
void launch_ic_residual(const GridDev& g, const float* fwd, const float* inv, int margin, Partial* partials,
                        RedSlot* slot, hipStream_t s);

// Gate of an inversion launch: with step >= 0 the launch does nothing when
// any of slots[0..n) computed a norm at that step, i.e. when the inversions
// those slots track would still have run a further step.
struct InvGate {
  const RedSlot* slots[2];
  int n;
  int step;
};
constexpr InvGate kNoGate{{nullptr, nullptr}, 0, -1};

// One fused fixed-point inversion step j (0..20) for np problems (see
// kernels.hip): the pending updates of step j-1, the norms of step j and,
// on grids where inv_speculates(g), its speculative update.  j = -1 is the
// sample launch that bounds step 0's max norm from below (nothing is
// launched without speculation); launch it before step 0.  partials holds
// 2 x np x inv_partials_needed(g) entries and masks 2 x np x
// inv_mask_words(g) words (step j's are read by step j + 1).
int64_t inv_partials_needed(const GridDev& g);
// Words of one candidate mask of the inversion (one bit per voxel).
int64_t inv_mask_words(const GridDev& g);
bool inv_speculates(const GridDev& g);
void launch_inv_step(const GridDev& g, int np, const float* const* field, float* const* inv, RedSlot* const* slots,
                     Partial* partials, uint64_t* masks, int j, hipStream_t s, const InvGate& gate = kNoGate);

// Resample np fields from grid gi onto grid go (linear, 0 outside).
void launch_resample_fields(const GridDev& gi, const GridDev& go, int np, const float* const* in,
                            float* const* out, hipStream_t s);

// Grid geometry with direction cosines, in double: point = org + m * index,
// index = minv * (point - org), m = direction * diag(spacing).
struct GeomDev {
  int D;
  int n[3];
  double org[3];
  double m[9];
  double minv[9];
};
constexpr int kMaxTransforms = 8;
struct XformDev {
  int kind;       // 0 identity, 1 affine, 2 displacement field
  double a[9];    // affine: y = a * x + off
  double off[3];
  GeomDev fg;     // displacement field geometry
  const float* f;
};
struct ApplyDev {
  int nt;
  float defv;
  GeomDev ref, in;
  XformDev t[kMaxTransforms];
  const double* coef;  // B-spline: the input's interpolation coefficients (double)
};
// interp: 0 linear, 1 nearest neighbour, 2 generic label (linear indicators),
// 16 + q B-spline of order q (0..5) over P.coef, 32 + w windowed sinc of
// radius 3 (w: 0 cosine, 1 Hamming, 2 Lanczos, 3 Blackman, 4 Welch)
void launch_apply_transforms(const GridDev& go, const ApplyDev& P, int interp, const float* in, float* out,
                             hipStream_t s);
// CreateJacobianDeterminantImage: det of I + the direction-rotated
// fourth-order central-difference gradient of a displacement field, clamped
// at 0 (log(max(., 0.001)) with logj); grad (nullable): the D x D matrices.
struct JacobianDev {
  int D;
  int n[3];
  double isp[3];
  double R[9];
};
void launch_jacobian(const GridDev& g, const JacobianDev& J, const float* u, int logj, float* out, float* grad,
                     hipStream_t s);
// GeometricJacobianDeterminantImageFilter: |volume| of the voxel's simplex
// moved by the linearly interpolated field over `undisplaced` (its volume
// unmoved); log(max(., 0.001)) with logj.
void launch_jacobian_geometric(const GridDev& g, const JacobianDev& J, const float* u, int logj, double undisplaced,
                               float* out, hipStream_t s);
// CompositeTransform::TransformPoint of n points (D doubles each, in place
// allowed) through P's transform list.
void launch_transform_points(int D, const ApplyDev& P, const double* in, double* out, int64_t n, hipStream_t s);
// TransformToDisplacementFieldFilter: T(p) - p for every point p of grid go
// (P.ref), T = P's transform list; D interleaved components per voxel.
void launch_compose_field(const GridDev& go, const ApplyDev& P, float* out, hipStream_t s);
// BSplineDecompositionImageFilter: the B-spline interpolation coefficients of
// img (dims n, x fastest) for `order` 0..5 into coef (double), mirror
// boundaries, causal initialisation to tolerance 1e-10.
void launch_bspline_coefficients(const int n[3], int D, const float* img, double* coef, int order, hipStream_t s);

// ---- linear stages: metric sums over a level's virtual sample points ----
constexpr int kLinThreads = 256;
constexpr int kLinMaxBlocks = 512;
constexpr int kLinMaxParams = 12;
constexpr int kLinLdsBins = 64;              // joint histograms up to 64 x 64 bins sit in LDS
constexpr double kLinHistScale = 1099511627776.0;  // 2^40: Parzen weights as fixed-point integers
enum LinMode { kLinMS = 0, kLinGC = 1, kLinHist = 2, kLinHistDeriv = 3 };
// Stage Jacobian: dS/dp_p at x = L_p (x - c) + tau_p (row-major 3x3, 3).
struct LinJac {
  double L[kLinMaxParams][9];
  double tau[kLinMaxParams][3];
};
struct LinDev {
  GeomDev vg;               // virtual grid (points of a fully sampled level)
  const double* pts;        // or explicit physical points, npts x D
  int64_t npts;
  GeomDev fg, mg;           // fixed / moving image geometry
  const float* fimg;
  const float* mimg;
  double fa[9], fb[3];      // fixed point = fa x + fb
  double ya[9], yb[3];      // moving point = ya x + yb (stage, then the moving initial chain)
  double G[9];              // linear part of the moving initial chain
  double c[3];              // stage centre
  double msp[3];            // moving spacing (central differences)
  int bins;                 // Mattes
  double fbs, fnm, mbs, mnm;
  const double* wtab;       // Mattes derivative pass: bins x bins weights
};
// Sums per mode (all double): MS {n, sum d^2 [, sum 2 d dm_p]}; GC {n, sum f,
// sum m, sum f^2, sum m^2, sum f m [, sum dm_p, sum f dm_p, sum m dm_p]};
// Hist {n} plus the joint histogram (bins^2 fixed-point u64) and fixed
// counts (u32), both accumulated into zeroed buffers; HistDeriv {n, sum w dm_p}.
int lin_quantities(int mode, bool deriv, int P);
int64_t lin_blocks(int64_t npts);
// partials: lin_blocks(npts) x Q doubles; out: Q doubles.  False for an
// unsupported (D, P).
bool launch_lin_metric(const LinDev& L, const LinJac& J, int D, int P, int mode, bool deriv, double* partials,
                       double* out, unsigned long long* hist, unsigned int* fcnt, hipStream_t s);

// Dense field-transform stages: y = a (x + u(x)) + b maps a virtual-grid
// voxel into an image; sp = that image's spacing (central differences).
struct FieldMap {
  double a[9];
  double b[3];
  double sp[3];
};
// out = img(y) (0 outside), grad (optional) = a^T grad img(y) per voxel;
// with u2 the displacement is u + s2 u2 (formed as launch_field_axpy does).
void launch_field_sample(const GeomDev& vg, const GeomDev& ig, const float* img, const float* u, const FieldMap& map,
                         float* out, float* grad, hipStream_t s, const float* u2 = nullptr, float s2 = 0.f);
// out[0] = sum a.b, out[1] = sum b.b, out[2] = sum a.a, out[3] = max |a|^2
// over n D-vectors (b may be null); partials: field_dots_partials(n) x 4
// doubles.
int field_dots_partials(int64_t n);
void launch_field_dots(int D, int64_t n, const float* a, const float* b, double* partials, double* out,
                       hipStream_t s);
// out = a + sc * b over n floats (a may be null; out may alias a or b).
void launch_field_axpy(int64_t n, const float* a, float sc, const float* b, float* out, hipStream_t s);

void launch_labels_nn(const GridDev& go, const float* u, const GridDev& gl, const int32_t* labels, int32_t* out,
                      hipStream_t s);

}  // namespace ants
