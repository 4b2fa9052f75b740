// Internal declarations shared by the HIP translation units of the
// cone-beam library (gfx950).  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernelkit_hip.h"

namespace kk {

// Record an error message for kk_last_error() and return `code`.
int fail(int code, const char* fmt, ...);
// Record a HIP failure and return KK_ERR_HIP.
int hip_fail(hipError_t e, const char* what);

constexpr int kHaloLo = KK_HALO_LO;
constexpr int kHaloHi = KK_HALO_HI;
constexpr int kHalo = KK_HALO_LO + KK_HALO_HI;

// Number of XCDs on MI355X; consecutive workgroups are dealt round-robin
// over them, each XCD with its own L2.
constexpr int kNumXcd = 8;

// Forward-projection textures: for every ray axis a that occurs, one
// slice-major "frame texture" T_a[i][jb][jc] of float2 over the in-plane
// axes (b, c) = (y, z), (x, z) or (y, x) for a = x, y, z, with a zero halo
// of KK_HALO_LO / KK_HALO_HI along b and c (none along a).  Element
// (i, jb, jc) holds (V(i, jb-1, jc-1), V(i, jb, jc-1)): the value and its
// b-neighbour, so the four bilinear taps of padded coordinate (jb, jc) are
// one 16-byte load.  Frames are stored in the order x, y, z of those
// present.
constexpr int kFrameB[3] = {1, 0, 1};
constexpr int kFrameC[3] = {2, 2, 0};

// Per-angle forward-projection record in the ray frame (a, b, c) of that
// angle: a = ray (major) axis, (b, c) its frame's in-plane axes.  Geometry
// in float64: sample coordinates reach ~n + 1 (up to ~1025 voxels), where
// float32 would round them by ~1e-4 voxel.
struct alignas(16) FpAngle {
  double qa, qb, qc;        // centre of pixel (0,0) minus source
  double ua, ub, uc;        // column step U
  double va, vb, vc;        // row step V
  double sig, taub, tauc;   // slice i <-> in-plane coordinate offsets
  float sc1, sc2, oscale, _p0; // sqrt(1 + a1^2 sc1 + a2^2 sc2) * oscale
  int32_t na, nb, nc, axis;    // extents along a, b, c; world axis of a
  uint64_t tex_off;            // byte offset of the frame texture T_a
  uint32_t sab, sbb;           // slice / row byte strides of T_a
  int32_t colc;                // 1: V has no a/b part (a1, b1 row-free)
  int32_t _p1[3];
};
// Ray through pixel (r, c): D = q + c*U + r*V (relative to the source),
// a1 = D_b / D_a, a2 = D_c / D_a.  Padded coordinate along b at slice i:
// x_b(i) = a1 * (i + sig) + taub, with sig = 1/2 - n_a/2 - s_a and
// taub = s_b + n_b/2 - 1/2 + KK_HALO_LO, so x_b = j + f reads padded
// elements j and j + 1; 0 < x_b < n_b + 1 is the support of the volume.
static_assert(sizeof(FpAngle) == 160, "FpAngle layout");

// Padded extents of one frame texture (elements of float2).
inline int64_t frame_elems(const int32_t shape[3], int a) {
  return int64_t(shape[a]) * (shape[kFrameB[a]] + kHalo) *
         (shape[kFrameC[a]] + kHalo);
}

// Backprojection texture: T[a][jv][ju] of float2, a zero halo of
// KK_HALO_LO / KK_HALO_HI along rows and columns (none along angles);
// element (a, jv, ju) holds (P(a, jv-1, ju-1), P(a, jv, ju-1)), so the four
// taps at padded coordinate (jv, ju) are one 16-byte load.

// Per-angle backprojection record: affine forms of the voxel centre X
// (voxel frame, world order), float64.  u = nu(X)/den(X) and v = nv(X)/den(X)
// are padded texture coordinates on the detector (texel-centre 1/2 removed,
// KK_HALO_LO added), 1/den^2 is the adjoint weight.  fast = 1 when nu and
// den do not depend on z (circular orbit about z): then u and 1/den are
// shared by a z-run; fast = 2 when only den does not (detector tilted in
// its plane): 1/den is shared and u, v are affine along the run; fast = 0
// otherwise (per-voxel 1/den).
struct alignas(16) BpAngle {
  double nux, nuy, nuz, nuw;
  double nvx, nvy, nvz, nvw;
  double dx, dy, dz, dw;
  int32_t fast, _p0, _p1, _p2;
};
static_assert(sizeof(BpAngle) == 112, "BpAngle layout");

// One forward-projection problem of a launch: the plan's table read from
// the frame texture `tex`, written to `out` (element strides os) as
// mode 0: p, 1: += p, 2: (p - y) * R (SIRT residual), with an optional
// second write at out + out_dual.
struct FpSpec {
  const kk_plan_s* plan = nullptr;
  const float* tex = nullptr;
  float* out = nullptr;
  // optional device table of output bases, one per index along logical
  // axis ptr_axis (0 angle, 1 row, 2 col); then `out` is unused
  const int64_t* ptrs = nullptr;
  int32_t ptr_axis = 0;
  int64_t os[3] = {0, 0, 0};
  int64_t out_dual = 0;
  int32_t rows = 0, cols = 0, mode = 0;
  const float* y = nullptr;
  int64_t ys[3] = {0, 0, 0};
  const float* R = nullptr;
  int64_t rs[3] = {0, 0, 0};
};

// One backprojection problem: mode 0: b, 1: += b, 2: the SIRT update
// x <- clip(mask * (x - C * b)) with the FP frame textures of x refreshed.
struct BpSpec {
  const kk_plan_s* plan = nullptr;
  const float* tex = nullptr;
  int32_t n_angles = 0, rows = 0, cols = 0;
  float* out = nullptr;
  int32_t shape[3] = {0, 0, 0};
  int64_t vs[3] = {0, 0, 0};
  float scale = 1.f;
  int32_t mode = 0;
  const float* C = nullptr;
  int64_t cs[3] = {0, 0, 0};
  const float* mask = nullptr;
  int64_t ms[3] = {0, 0, 0};
  float lo = 0.f, hi = 0.f;
  int32_t clip_lo = 0, clip_hi = 0;
  float* xtex = nullptr;
  int32_t xtex_mask = 0;
};

// Device descriptors of many problems for one persistent launch.
struct FpBatch;
struct BpBatch;
int fp_batch_build(const std::vector<FpSpec>& specs, int32_t rows_per_thread,
                   FpBatch** out);
void fp_batch_free(FpBatch* b);
int fp_batch_launch(const FpBatch* b, int32_t grid_blocks, hipStream_t s);
int bp_batch_build(const std::vector<BpSpec>& specs, int32_t zrun,
                   BpBatch** out);
void bp_batch_free(BpBatch* b);
int bp_batch_launch(const BpBatch* b, int32_t grid_blocks, hipStream_t s);

// One texture build of a batch: the zero-halo float2 texture
// [n0][n1 + kHalo][n2 + kHalo] of a 3-D array addressed by element strides
// (an FP frame of a volume, or a BP texture of projections), optionally of
// the SIRT residual (src - y) * R.
struct TexJob {
  const float* src = nullptr;
  int64_t s0 = 0, s1 = 0, s2 = 0;
  const float* y = nullptr;  // nullptr: a plain copy
  int64_t y0 = 0, y1 = 0, y2 = 0;
  const float* R = nullptr;
  int64_t r0 = 0, r1 = 0, r2 = 0;
  float2* tex = nullptr;
  int32_t n0 = 0, n1 = 0, n2 = 0;
  int64_t first = 0;  // first tile of this job in the batch
};
struct TexJobs;
int tex_jobs_build(std::vector<TexJob> jobs, TexJobs** out);
void tex_jobs_free(TexJobs* j);
int tex_jobs_launch(const TexJobs* j, hipStream_t s);

// Fill a BpSpec from the C-ABI arguments of kk_cone_bp / kk_sirt_bp_update
// (validated when the launch arguments are built).
int bp_spec(BpSpec& S, const kk_plan_s* plan, const float* tex,
            int32_t n_angles, int32_t rows, int32_t cols, float* vol,
            const int32_t shape[3], const int64_t vs[3], float scale,
            int32_t additive);
int sirt_update_spec(BpSpec& S, const kk_plan_s* plan, const float* res_tex,
                     int32_t n_angles, int32_t rows, int32_t cols, float* x,
                     const int32_t shape[3], const int64_t xs[3],
                     const float* C, const int64_t cs[3], const float* mask,
                     const int64_t ms[3], float lo, float hi,
                     int32_t clip_lo, int32_t clip_hi, float* x_tex,
                     int32_t x_tex_mask, float scale);
// FpSpec of kk_cone_fp / kk_sirt_fp_residual.
int fp_spec(FpSpec& S, const kk_plan_s* plan, const float* tex, float* out,
            const int64_t os[3], int32_t rows, int32_t cols,
            int32_t additive);
int sirt_residual_spec(FpSpec& S, const kk_plan_s* plan, const float* x_tex,
                       const float* y, const int64_t ys[3], const float* R,
                       const int64_t rs[3], float* res_tex, int32_t n_angles,
                       int32_t rows, int32_t cols);

// Exact transpose of the forward projection (cone_fp_t.hip): the floats of
// workspace a launch needs (-1 on bad arguments), and the launch itself
// with the workspace given (validates, allocates nothing).
int64_t fp_transpose_work_floats(const kk_plan_s* plan, int32_t rows,
                                 int32_t cols);
int fp_transpose_launch(const kk_plan_s* plan, const float* proj,
                        const int64_t proj_strides[3], int32_t rows,
                        int32_t cols, float* vol, const int32_t vol_shape[3],
                        const int64_t vol_strides[3], int32_t additive,
                        float* work, hipStream_t stream);

}  // namespace kk

struct kk_plan_s {
  int device = 0;
  // forward projection table
  kk::FpAngle* fp = nullptr;
  int32_t fp_n = 0, fp_cap = 0;
  int32_t fp_shape[3] = {0, 0, 0};
  int32_t fp_mask = 0;  // bit a: frame texture of ray axis a is used
  bool fp_set = false;
  bool fp_all_colc = false;  // every angle column-coherent (FpAngle.colc)
  bool fp_any_colc = false;  // some angle column-coherent
  // angle ids grouped by ray axis: fr_idx[fr_start[a] .. fr_start[a+1])
  int32_t* fr_idx = nullptr;
  int32_t fr_cap = 0;
  int32_t fr_start[4] = {0, 0, 0, 0};
  // backprojection table
  kk::BpAngle* bp = nullptr;
  int32_t bp_n = 0, bp_cap = 0;
  bool bp_set = false;
  bool bp_all_circ = false;  // every angle on the circular path (fast == 1)
};

// Map a hardware workgroup id to a logical tile id such that, inside every
// chunk of kNumXcd * group consecutive ids, each XCD receives `group`
// consecutive logical tiles.  Tail chunks keep the identity mapping.
__device__ __forceinline__ int64_t kk_xcd_remap(int64_t bid, int64_t total,
                                                int32_t group) {
  if (group <= 1) return bid;
  const int64_t span = (int64_t)kk::kNumXcd * group;
  const int64_t chunk = bid / span;
  if ((chunk + 1) * span > total) return bid;
  const int64_t r = bid - chunk * span;
  return chunk * span + (r % kk::kNumXcd) * group + r / kk::kNumXcd;
}

__device__ __forceinline__ float kk_clamp(float x, float lo, float hi) {
  return __builtin_amdgcn_fmed3f(x, lo, hi);
}

// Loads at a wave-uniform base plus a 32-bit per-lane byte offset: the
// compiler emits the SGPR-base ("saddr") form of global loads, so no 64-bit
// address arithmetic is spent per lane.
// Four consecutive floats, 8-byte aligned: one global_load_dwordx4.  The
// address is stated to be global memory (textures are device allocations):
// for a base read from a descriptor in memory (the batch kernels) the
// compiler would otherwise fall back to flat loads, which cost more and
// also count against the LDS/scalar wait counter.
typedef float kk_float4u __attribute__((ext_vector_type(4), aligned(8)));
typedef const __attribute__((address_space(1))) kk_float4u* kk_gptr4;
__device__ __forceinline__ kk_float4u kk_ld4(const char* base, uint32_t off) {
  return *(kk_gptr4)(base + off);
}

// Copy of a wave-uniform, read-only record (an angle's parameters) read as
// constant memory, so it always comes through scalar loads into SGPRs,
// also when the table pointer was itself read from memory.
template <class T>
__device__ __forceinline__ T kk_uniform_load(const T* p) {
  static_assert(sizeof(T) % 16 == 0, "record of whole 16-byte words");
  typedef uint32_t word4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(4))) word4* cptr;
  const cptr q = (cptr)(p);
  word4 w[sizeof(T) / 16];
#pragma unroll
  for (size_t i = 0; i < sizeof(T) / 16; ++i) w[i] = q[i];
  return __builtin_bit_cast(T, w);
}

__device__ __forceinline__ int32_t kk_wave_min(int32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int32_t kk_wave_max(int32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}

// ---- forward-projection ray arithmetic, shared by the forward kernel and
// its exact transpose (both must compute identical samples) -------------

// Slopes a1, a2 and intercepts b1, b2 of the padded sample coordinates
// x_b(i) = a1 * i + b1, x_c(i) = a2 * i + b2 of the ray through pixel
// (row, col) of angle P, in float64; false for rays parallel to the slices.
__device__ __forceinline__ bool kk_fp_ray(const kk::FpAngle& P, int32_t row,
                                          int32_t col, double& a1,
                                          double& b1, double& a2,
                                          double& b2) {
  const double fr = static_cast<double>(row);
  const double fc = static_cast<double>(col);
  const double Da = fma(fr, P.va, fma(fc, P.ua, P.qa));
  const double Db = fma(fr, P.vb, fma(fc, P.ub, P.qb));
  const double Dc = fma(fr, P.vc, fma(fc, P.uc, P.qc));
  const double inv = 1.0 / Da;
  a1 = Db * inv;
  a2 = Dc * inv;
  b1 = fma(a1, P.sig, P.taub);
  b2 = fma(a2, P.sig, P.tauc);
  return isfinite(a1) && isfinite(a2) && isfinite(b1) && isfinite(b2);
}

// The ray's weight sqrt(1 + a1^2 sc1 + a2^2 sc2) * oscale.
__device__ __forceinline__ float kk_fp_weight(const kk::FpAngle& P,
                                              double a1, double a2) {
  const float f1 = static_cast<float>(a1), f2 = static_cast<float>(a2);
  return sqrtf(fmaf(f1 * f1, P.sc1, fmaf(f2 * f2, P.sc2, 1.f))) * P.oscale;
}

// (double)i for a wave-uniform 0 <= i < 2^31, assembled from the bits with
// scalar integer instructions (exponent from the leading-one position), so
// the vector unit does not convert the loop counter every slice.
__device__ __forceinline__ double kk_uniform_i2d(int32_t i) {
  const uint32_t u = static_cast<uint32_t>(i);
  const uint32_t e = 31u - static_cast<uint32_t>(__builtin_clz(u | 1u));
  const uint64_t bits = (static_cast<uint64_t>(u) << (52u - e)) +
                        (static_cast<uint64_t>(1022u + e) << 52);
  return u == 0u ? 0.0 : __builtin_bit_cast(double, bits);
}

// Split a coordinate x >= 0 into floor(x) and x - floor(x): for x >= 0 the
// truncating conversion is the floor, and v_fract_f64 returns x - floor(x)
// exactly (the difference is representable), one instruction fewer than
// floor + convert + subtract.
__device__ __forceinline__ void kk_split_nonneg(double x, uint32_t& idx,
                                                float& frac) {
  idx = static_cast<uint32_t>(x);
  frac = static_cast<float>(__builtin_amdgcn_fract(x));
}

// Sample coordinate a * i + b clamped into [0, n1] (the zero halo), split
// into the padded element index and the interpolation weight.  float64:
// coordinates reach ~1e3, where float32 would be off by ~1e-4.
__device__ __forceinline__ void kk_fp_coord(double a, double b, double fi,
                                            double n1, uint32_t& idx,
                                            float& frac) {
  kk_split_nonneg(fmin(fmax(fma(a, fi, b), 0.0), n1), idx, frac);
}

// kk_fp_coord for a coordinate known to lie in [0, n1]: the same index and
// weight without the clamp.
__device__ __forceinline__ void kk_fp_coord_in(double a, double b, double fi,
                                               uint32_t& idx, float& frac) {
  kk_split_nonneg(fma(a, fi, b), idx, frac);
}

