// Ray-driven (Joseph) cone-beam forward projection for gfx950.
//
// Behaviour follows cone_fp<DIR_X|Y|Z> of the reference
// (kernelkit/cuda/cone_fp.cu:73-207) as launched by
// kernelkit/kernels/cone_fp.py:195-262: for every angle a ray axis a is
// fixed; each ray through a pixel centre is sampled at the centre of every
// slice along a, bilinearly in the two other axes with zero outside the
// volume, and the sum is scaled by sqrt(1 + a1^2 sc1 + a2^2 sc2) * oscale.
//
// MI355X design:
// * one launch covers all angles and all slices (the reference issues one
//   launch per angle group and 16-slice chunk and accumulates in global
//   memory); each output pixel is written exactly once;
// * the volume is read from slice-major frame textures of (value,
//   b-neighbour) float2 pairs with a zero halo (kk_internal.h), so a sample
//   is: one FMA per coordinate, a clamp into the halo, v_cvt_u32 + v_fract
//   for the tap index and weight (the slice index itself is converted to
//   float64 with scalar integer instructions), and ONE 16-byte load for
//   all four taps (~15 VALU instructions per sample in the unclamped
//   interior; measured on MI355X at 512^3 / 720 angles the kernel keeps
//   the texture-address units ~85% busy, ~14 cycles per 64-lane gather,
//   and the VALU about as busy: it is close to both issue limits);
// * sample coordinates are evaluated directly, x_b(i) = a1*i + beta, instead
//   of the reference's running sums, so the per-sample error does not grow
//   along the ray;
// * lanes of a wavefront are 64 consecutive detector rows of one column and
//   every thread traces R rows 64 apart.  For rotation about z this makes a
//   wave sample consecutive z voxels of the same (x or y) row, contiguous
//   memory;
// * the slice loop is wave-uniform: it covers the union of the wave's ray
//   supports (samples outside a ray's support clamp into the zero halo), so
//   the slice base address lives in scalar registers and every tap is an
//   SGPR-base load with a 32-bit per-lane offset; it is software-pipelined
//   one slice deep, which hides part of the latency of the texture misses
//   that dominate at 512^3;
// * for a circular orbit the row step has no in-plane component, so the
//   b-coordinate, its tap row and weight are shared by the thread's R rows
//   ("column-coherent" angles, flagged per angle on the host);
// * per-angle parameters are wave-uniform (scalar loads); workgroups are
//   ordered in groups of kGroupAngles consecutive angles x kGroupCols
//   neighbouring column tiles (angle fastest), groups ordered angle group
//   -> column-tile group -> row band, and each group's ids go to one XCD:
//   neighbouring angles and columns read nearly the same texture lines, so
//   its L2 serves them (measured on MI355X at 512^3 / 720 angles: 86.9 ms
//   with 16 x 8, 91.0 ms with runs of 64 angles of one column tile, 110 ms
//   with column tiles fastest and workgroups dealt round-robin to the
//   XCDs; 16 x 4, 32 x 4 and 16 x 16 took 87.7, 88.1 and 90.5 ms).  In
//   batches (256-thread workgroups) groups of 16 x 16 took the batched
//   forward projection of the 64-problem workload from 104.8 to 98.1 ms.
#include <algorithm>
#include <new>
#include <vector>

#include "kk_internal.h"

namespace {

constexpr int kWave = 64;
// Detector columns per workgroup, one per wavefront.  Single problems: 16
// (1024 threads); the wavefronts of a workgroup share one CU's L1 and move
// through the slices nearly in step, so neighbouring columns' taps hit
// there (measured on MI355X at 512^3 / 720 angles: 3-8% faster than 8 or
// 4 columns).  Batches of small detectors: 4.
constexpr int kColsSingle = 16;
constexpr int kColsBatch = 4;
// The workgroups of kGroupAngles consecutive angles x kGroupCols (single
// problems; kGroupColsBatch in batches) neighbouring column tiles go to
// one XCD together.  An XCD holds 64 single-problem workgroups (32 CUs x
// 2 of 1024 threads) or 256 batch workgroups (32 x 8 of 256 threads);
// measured on MI355X: 16 x 8 and 16 x 16 were the fastest groups.  These
// are upper bounds: a group size is the one in [ceil(cap / 2), cap] that
// pads the angles or column tiles least (group_size).
constexpr int32_t kGroupAngles = 16;
constexpr int32_t kGroupCols = 8;
constexpr int32_t kGroupColsBatch = 16;

// Group size for n items, at most cap: the g in [ceil(cap / 2), cap] whose
// whole groups pad n least (the larger g on ties), or n if n <= cap.  Padded
// ids return at once, but the groups go to the XCDs whole, so an XCD dealt
// mostly padded groups idles while others work (160 detector columns = 10
// tiles in groups of 8: about a fifth more time at 128^3 / 180 angles than
// groups of 5).
int32_t group_size(int32_t n, int32_t cap) {
  if (n <= cap) return n;
  int32_t best = cap;
  int64_t best_pad = -1;
  for (int32_t g = cap; g >= (cap + 1) / 2; --g) {
    const int64_t pad = int64_t((n + g - 1) / g) * g - n;
    if (best_pad < 0 || pad < best_pad) {
      best = g;
      best_pad = pad;
    }
  }
  return best;
}

struct FpArgs {
  const char* tex;  // frame textures (byte base)
  float* out;
  // if non-null: the output base is out_ptrs[index along logical axis
  // ptr_axis] instead of out (whose stride along that axis is then 0)
  const int64_t* out_ptrs;
  int32_t ptr_axis;
  int64_t os0, os1, os2;  // output strides: angle, row, col
  int64_t out_dual;       // if nonzero: also write out[o + out_dual]
  const float* y;         // mode 2: measured projections
  int64_t ys0, ys1, ys2;
  const float* R;  // mode 2: ray weights
  int64_t rs0, rs1, rs2;
  int32_t rows, cols, n_angles;
  int32_t row_tiles, col_tiles;
  int32_t mode;  // 0: out = p, 1: out += p, 2: out = (p - y) * R
  int32_t xcd_group;
  // workgroup groups: grp_a angles x grp_c column tiles (ids past the
  // angles or column tiles, which pad the grid to whole groups, return)
  int32_t grp_a, grp_c;
  int64_t total_blocks;
};

// Narrow [lo, hi] (slice indices, initially [0, n_a - 1]) to the slices i
// with 0 < a*i + beta < n + 1 (padded coordinates).  For slopes below 1e-3
// the division is not trusted; the ray is then only dropped when it stays
// outside over the whole slice range.  Over-inclusion is harmless: clamped
// samples outside the volume read the zero halo.
__device__ __forceinline__ void clip_range(float a, float beta, float n1,
                                           float last, float& lo,
                                           float& hi) {
  if (fabsf(a) < 1e-3f) {
    const float e0 = beta, e1 = fmaf(a, last, beta);
    if (fmaxf(e0, e1) <= 0.f || fminf(e0, e1) >= n1) {
      lo = 3e38f;
      hi = -3e38f;
    }
    return;
  }
  const float inv = 1.f / a;
  const float t0 = (0.f - beta) * inv;
  const float t1 = (n1 - beta) * inv;
  lo = fmaxf(lo, fminf(t0, t1));
  hi = fminf(hi, fmaxf(t0, t1));
}

// Narrow [lo, hi] to the slices i with 1 <= a*i + beta <= n1 - 1: one
// voxel inside the clamp range [0, n1], which covers the float32 error of
// this estimate (~1e-4 voxel).  A coordinate is linear in i, so for flat
// slopes the end points decide.
__device__ __forceinline__ void inner_range(float a, float beta, float n1,
                                            float& lo, float& hi) {
  if (fabsf(a) < 1e-3f) {
    const float e0 = fmaf(a, lo, beta), e1 = fmaf(a, hi, beta);
    if (!(fminf(e0, e1) >= 1.f && fmaxf(e0, e1) <= n1 - 1.f)) {
      lo = 3e38f;
      hi = -3e38f;
    }
    return;
  }
  const float inv = 1.f / a;
  const float t0 = (1.f - beta) * inv;
  const float t1 = (n1 - 1.f - beta) * inv;
  lo = fmaxf(lo, fminf(t0, t1));
  hi = fminf(hi, fmaxf(t0, t1));
}

// Bilinear value from the 16-byte tap record (v[b][c], v[b+1][c],
// v[b][c+1], v[b+1][c+1]): lerp along b, then along c.
__device__ __forceinline__ float bilerp(kk_float4u t, float fb, float fc) {
  const float l0 = fmaf(fb, t.y - t.x, t.x);
  const float l1 = fmaf(fb, t.w - t.z, t.z);
  return fmaf(fc, l1 - l0, l0);
}

// The four taps and two weights of one slice for the R rays of a thread.
template <int R, int NB>
struct SliceTaps {
  kk_float4u t[R];
  float fb[NB];
  float fc[R];
};

// Coordinates, weights and tap loads of slice i for R rays.  COLC (NB = 1):
// the b slope/intercept are shared by the R rays.  CLAMP = false: every
// coordinate is known to lie inside the clamp range, which is then skipped.
template <int R, int NB, bool CLAMP>
__device__ __forceinline__ void fetch_slice(
    const char* __restrict__ frame, uint32_t sab, uint32_t sbb, double nb1,
    double nc1, const double (&a1)[NB], const double (&b1)[NB],
    const double (&a2)[R], const double (&b2)[R], int32_t i,
    SliceTaps<R, NB>& s) {
  constexpr bool COLC = NB == 1;
  const double fi = kk_uniform_i2d(i);
  const char* slice = frame + static_cast<size_t>(i) * sab;
  uint32_t offb = 0;
  if constexpr (COLC) {
    uint32_t ib;
    if constexpr (CLAMP)
      kk_fp_coord(a1[0], b1[0], fi, nb1, ib, s.fb[0]);
    else
      kk_fp_coord_in(a1[0], b1[0], fi, ib, s.fb[0]);
    offb = __umul24(ib, sbb);
    // one multiply for the R rays (not a multiply-add per ray)
    asm volatile("" : "+v"(offb));
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    uint32_t ob = offb;
    if constexpr (!COLC) {
      uint32_t ib;
      if constexpr (CLAMP)
        kk_fp_coord(a1[k], b1[k], fi, nb1, ib, s.fb[k]);
      else
        kk_fp_coord_in(a1[k], b1[k], fi, ib, s.fb[k]);
      ob = __umul24(ib, sbb);
    }
    uint32_t ic;
    if constexpr (CLAMP)
      kk_fp_coord(a2[k], b2[k], fi, nc1, ic, s.fc[k]);
    else
      kk_fp_coord_in(a2[k], b2[k], fi, ic, s.fc[k]);
    s.t[k] = kk_ld4(slice, ob + (ic << 3));
  }
}

template <int R, int NB>
__device__ __forceinline__ void add_slice(const SliceTaps<R, NB>& s,
                                          float (&acc)[R]) {
#pragma unroll
  for (int k = 0; k < R; ++k)
    acc[k] += bilerp(s.t[k], s.fb[NB == 1 ? 0 : k], s.fc[k]);
}

// One tracing pass over slices [i0, i1] (wave-uniform) for R rays of a
// thread.  PIPE software-pipelines the pass: the taps of slice i + 1 are
// requested before slice i is interpolated, so a wavefront keeps two
// slices of gathers in flight (two register buffers, the loop unrolled by
// two so that they swap roles without copies; the request past i1 repeats
// slice i1 and is discarded).  The scheduling barriers keep each request
// ahead of the interpolation (the scheduler would otherwise sink the loads
// next to their use).  Measured on MI355X at 512^3 / 720 angles: 5% faster
// than one slice in flight; three slices in flight (at 6 instead of 8
// wavefronts per SIMD, for the registers) are no faster than two.
template <int R, int NB, bool CLAMP, bool PIPE>
__device__ __forceinline__ void trace(const kk::FpAngle& P,
                                      const char* __restrict__ frame,
                                      const double (&a1)[NB],
                                      const double (&b1)[NB],
                                      const double (&a2)[R],
                                      const double (&b2)[R], int32_t i0,
                                      int32_t i1, float (&acc)[R]) {
  if (i0 > i1) return;
  const double nb1 = static_cast<double>(P.nb + 1);
  const double nc1 = static_cast<double>(P.nc + 1);
  const uint32_t sab = P.sab, sbb = P.sbb;
  auto fetch = [&](int32_t i, SliceTaps<R, NB>& s) {
    fetch_slice<R, NB, CLAMP>(frame, sab, sbb, nb1, nc1, a1, b1, a2, b2,
                              min(i, i1), s);
  };
  if constexpr (!PIPE) {
    for (int32_t i = i0; i <= i1; ++i) {
      SliceTaps<R, NB> A;
      fetch(i, A);
      add_slice(A, acc);
    }
  } else {
    SliceTaps<R, NB> A, B;
    fetch(i0, A);
    for (int32_t i = i0;;) {
      fetch(i + 1, B);
      __builtin_amdgcn_sched_barrier(0);
      add_slice(A, acc);
      if (++i > i1) break;
      fetch(i + 1, A);
      __builtin_amdgcn_sched_barrier(0);
      add_slice(B, acc);
      if (++i > i1) break;
    }
  }
}

// One workgroup's tile (row band rt, angle, column tile ct) of logical
// block id `bid`.  ALL_COLC: every angle of the launch is column-coherent,
// so a thread keeps one b slope/intercept for its R rays (fewer registers,
// more wavefronts in flight); otherwise the angle's flag picks the path.
template <int R, bool ALL_COLC, bool PIPE, int COLS, bool PTRS = false>
__device__ __forceinline__ void fp_tile_at(
    const kk::FpAngle* __restrict__ tab, int32_t ti, const FpArgs& A,
    int32_t ang, int32_t ct, int32_t rt) {
  constexpr int NB = ALL_COLC ? 1 : R;
  const int32_t lane = int32_t(threadIdx.x) & (kWave - 1);
  const int32_t col = ct * COLS + int32_t(threadIdx.x) / kWave;
  const int32_t row0 = rt * (kWave * R) + lane;

  const kk::FpAngle& P = tab[ti];
  const float nb1 = static_cast<float>(P.nb + 1);
  const float nc1 = static_cast<float>(P.nc + 1);
  const float last = static_cast<float>(P.na - 1);

  double a1[NB], b1[NB], a2[R], b2[R];
  float wscale[R];
  int32_t ilo = 0x7fffffff, ihi = -1;
  float jl = 0.f, jh = last;  // this thread's rays all inside
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int32_t row = row0 + kWave * k;
    // rays outside the detector, or parallel to the slices, contribute
    // nothing: they are parked at padded coordinate 1 (inside every
    // range, so also valid for unclamped passes) with weight 0 (for
    // column-coherent angles the b part of every ray of the thread equals
    // that of its first ray)
    double ra1, rb1;
    const bool ok = row < A.rows && col < A.cols &&
                    kk_fp_ray(P, row, col, ra1, rb1, a2[k], b2[k]);
    if (k < NB) {
      a1[k] = ok ? ra1 : 0.0;
      b1[k] = ok ? rb1 : 1.0;
    }
    wscale[k] = 0.f;
    if (!ok) {
      a2[k] = 0.0;
      b2[k] = 1.0;
      continue;
    }
    wscale[k] = kk_fp_weight(P, ra1, a2[k]);
    const float fa1 = static_cast<float>(ra1), fb1 = static_cast<float>(rb1);
    const float fa2 = static_cast<float>(a2[k]);
    const float fb2 = static_cast<float>(b2[k]);
    float lo = 0.f, hi = last;
    clip_range(fa1, fb1, nb1, last, lo, hi);
    clip_range(fa2, fb2, nc1, last, lo, hi);
    if (lo <= hi + 2.f) {
      ilo = min(ilo, max(0, static_cast<int32_t>(floorf(lo)) - 1));
      ihi = max(ihi, min(P.na - 1, static_cast<int32_t>(ceilf(hi)) + 1));
    }
    inner_range(fa1, fb1, nb1, jl, jh);
    inner_range(fa2, fb2, nc1, jl, jh);
  }
  ilo = kk_wave_min(ilo);
  ihi = kk_wave_max(ihi);
  // slices where every ray of the wave is inside: no clamps there
  int32_t jlo = 1, jhi = 0;
  if (jl <= jh) {
    jlo = max(0, static_cast<int32_t>(ceilf(jl)));
    jhi = min(P.na - 1, static_cast<int32_t>(floorf(jh)));
  }
  jlo = max(kk_wave_max(jlo), ilo);
  jhi = min(kk_wave_min(jhi), ihi);

  float acc[R];
#pragma unroll
  for (int k = 0; k < R; ++k) acc[k] = 0.f;
  const char* frame = A.tex + P.tex_off;
  if (ilo <= ihi) {
    // clamped head, unclamped interior, clamped tail (or one clamped pass)
    const bool split = jlo <= jhi;
    const int32_t h1 = split ? jlo - 1 : ihi;
    if constexpr (ALL_COLC) {
      trace<R, 1, true, PIPE>(P, frame, a1, b1, a2, b2, ilo, h1, acc);
      if (split) {
        trace<R, 1, false, PIPE>(P, frame, a1, b1, a2, b2, jlo, jhi, acc);
        trace<R, 1, true, PIPE>(P, frame, a1, b1, a2, b2, jhi + 1, ihi,
                                acc);
      }
    } else {
      if (P.colc) {
        const double a1c[1] = {a1[0]}, b1c[1] = {b1[0]};
        trace<R, 1, true, PIPE>(P, frame, a1c, b1c, a2, b2, ilo, ihi, acc);
      } else {
        trace<R, R, true, PIPE>(P, frame, a1, b1, a2, b2, ilo, h1, acc);
        if (split) {
          trace<R, R, false, PIPE>(P, frame, a1, b1, a2, b2, jlo, jhi,
                                   acc);
          trace<R, R, true, PIPE>(P, frame, a1, b1, a2, b2, jhi + 1, ihi,
                                  acc);
        }
      }
    }
  }

#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int32_t row = row0 + kWave * k;
    if (row >= A.rows || col >= A.cols) continue;
    float val = acc[k] * wscale[k];
    const int64_t o = ang * A.os0 + row * A.os1 + col * A.os2;
    float* out = A.out;
    if constexpr (PTRS)
      out = reinterpret_cast<float*>(
          A.out_ptrs[A.ptr_axis == 0 ? ang : (A.ptr_axis == 1 ? row : col)]);
    if (A.mode == 1) {
      out[o] += val;
      continue;
    }
    if (A.mode == 2) {
      const float yv = A.y[ang * A.ys0 + row * A.ys1 + col * A.ys2];
      const float rv = A.R[ang * A.rs0 + row * A.rs1 + col * A.rs2];
      val = (val - yv) * rv;
    }
    out[o] = val;
    if (A.out_dual) out[o + A.out_dual] = val;
  }
}

// Decode logical block id `bid` (groups of grp_a angles x grp_c column
// tiles, angle fastest inside a group; groups ordered angle group, then
// column-tile group, then row band) and trace that tile.  TABLE_IN_MEMORY: the angle table pointer was
// read from a descriptor (batch kernels); the record is then copied through
// constant-memory (scalar) loads first.  A table passed as a kernel
// argument is read in place (the compiler scalarizes it by itself; the
// explicit copy measured 5% slower at 512^3 / 720 angles).
template <int R, bool ALL_COLC, bool PIPE, int COLS, bool TABLE_IN_MEMORY,
          bool PTRS = false>
__device__ __forceinline__ void fp_tile(const kk::FpAngle* __restrict__ tab,
                                        const FpArgs& A, int64_t bid) {
  const int32_t ga = A.grp_a, gc = A.grp_c;
  const int64_t inner = bid % (int64_t(ga) * gc);
  const int64_t outer = bid / (int64_t(ga) * gc);
  const int32_t nag = (A.n_angles + ga - 1) / ga;
  const int32_t ncg = (A.col_tiles + gc - 1) / gc;
  const int32_t ang = static_cast<int32_t>(outer % nag) * ga +
                      static_cast<int32_t>(inner % ga);
  const int64_t t = outer / nag;
  const int32_t ct = static_cast<int32_t>(t % ncg) * gc +
                     static_cast<int32_t>(inner / ga);
  const int32_t rt = static_cast<int32_t>(t / ncg);
  if (ang >= A.n_angles || ct >= A.col_tiles) return;  // grid padding
  if constexpr (TABLE_IN_MEMORY) {
    const kk::FpAngle rec = kk_uniform_load(tab + ang);
    fp_tile_at<R, ALL_COLC, PIPE, COLS, PTRS>(&rec, 0, A, ang, ct, rt);
  } else {
    fp_tile_at<R, ALL_COLC, PIPE, COLS, PTRS>(tab, ang, A, ang, ct, rt);
  }
}

// Register caps: 8 wavefronts per SIMD up to two rows per thread, 6 at
// four (5 in the batch kernel), whose float64 ray data would otherwise
// spill.
constexpr int fp_min_blocks(int r) {
  return (r <= 2 ? 32 : 24) / kColsSingle;
}
constexpr int fp_batch_min_blocks(int r) {
  return (r <= 2 ? 32 : 20) / kColsBatch;
}
// Slice pipelining where the second tap buffer fits the register cap
// without spilling.
constexpr bool fp_pipe(int r) { return r <= 2; }
constexpr bool fp_batch_pipe(int r) { return r <= 1; }

// PTRS: output through the table of entry pointers (kk_cone_fp_ptrs); a
// separate instantiation, so the plain kernel's code is unchanged.
template <int R, bool ALL_COLC, bool PTRS>
__global__ __launch_bounds__(kColsSingle * kWave, fp_min_blocks(R)) void cone_fp_kernel(
    const kk::FpAngle* __restrict__ tab, FpArgs A) {
  fp_tile<R, ALL_COLC, fp_pipe(R), kColsSingle, false, PTRS>(
      tab, A, kk_xcd_remap(blockIdx.x, A.total_blocks, A.xcd_group));
}

}  // namespace

namespace {

// Rows per thread: 2 unless the detector has at most 64 rows (measured on
// MI355X: 2 beats 4 by 9-30% at 256^3..512^3, the float64 coordinates of
// four rays costing occupancy).
int auto_rows_per_thread(int32_t rows, int32_t rpt) {
  if (rpt == 1 || rpt == 2 || rpt == 4) return rpt;
  return rows > 64 ? 2 : 1;
}

// Validate one forward-projection problem and fill its kernel arguments
// for `rpt` rows per thread.
int make_fp_args(const kk::FpSpec& S, int32_t rpt, int32_t cols_per_block,
                 FpArgs& A) {
  using kk::fail;
  if (S.plan == nullptr || !S.plan->fp_set)
    return fail(KK_ERR_NOT_SET,
                "Please set the projection geometry with `set_params()`.");
  if (S.tex == nullptr || (S.out == nullptr && S.ptrs == nullptr))
    return fail(KK_ERR_INVALID, "kk_cone_fp: null argument");
  if (S.ptrs != nullptr &&
      (S.mode == 2 || S.out_dual != 0 || S.ptr_axis < 0 || S.ptr_axis > 2))
    return fail(KK_ERR_INVALID,
                "kk_cone_fp_ptrs: a pointer table takes plain or additive "
                "output along axis 0, 1 or 2");
  if (S.rows <= 0 || S.cols <= 0)
    return fail(KK_ERR_INVALID, "kk_cone_fp: bad detector size");
  if (S.mode == 2 && (S.y == nullptr || S.R == nullptr))
    return fail(KK_ERR_INVALID, "kk_cone_fp: residual mode needs y and R");
  if (S.mode == 1 && S.out_dual != 0)
    return fail(KK_ERR_INVALID, "kk_cone_fp: dual output is write-only");
  A = FpArgs{};
  A.tex = reinterpret_cast<const char*>(S.tex);
  A.out = S.out;
  A.out_ptrs = S.ptrs;
  A.ptr_axis = S.ptr_axis;
  A.os0 = S.ptrs != nullptr && S.ptr_axis == 0 ? 0 : S.os[0];
  A.os1 = S.ptrs != nullptr && S.ptr_axis == 1 ? 0 : S.os[1];
  A.os2 = S.ptrs != nullptr && S.ptr_axis == 2 ? 0 : S.os[2];
  A.out_dual = S.out_dual;
  if (S.mode == 2) {
    A.y = S.y;
    A.ys0 = S.ys[0];
    A.ys1 = S.ys[1];
    A.ys2 = S.ys[2];
    A.R = S.R;
    A.rs0 = S.rs[0];
    A.rs1 = S.rs[1];
    A.rs2 = S.rs[2];
  }
  A.rows = S.rows;
  A.cols = S.cols;
  A.n_angles = S.plan->fp_n;
  A.row_tiles = (S.rows + kWave * rpt - 1) / (kWave * rpt);
  A.col_tiles = (S.cols + cols_per_block - 1) / cols_per_block;
  A.mode = S.mode;
  A.grp_a = group_size(A.n_angles, kGroupAngles);
  A.grp_c = group_size(
      A.col_tiles,
      cols_per_block == kColsSingle ? kGroupCols : kGroupColsBatch);
  A.xcd_group = A.grp_a * A.grp_c;
  A.total_blocks = int64_t(A.row_tiles) *
                   ((A.n_angles + A.grp_a - 1) / A.grp_a) * A.grp_a *
                   ((A.col_tiles + A.grp_c - 1) / A.grp_c) * A.grp_c;
  if (A.total_blocks >= (int64_t(1) << 31))
    return fail(KK_ERR_UNSUPPORTED, "kk_cone_fp: grid too large");
  return KK_OK;
}

// Many problems, one launch: a persistent grid walks the concatenated
// workgroup tiles of all problems; a tile finds its problem by binary
// search over the problems' first tile ids (wave-uniform).
struct FpItem {
  const kk::FpAngle* tab;
  FpArgs args;
};

template <int R, bool ALL_COLC>
__global__ __launch_bounds__(kColsBatch * kWave, fp_batch_min_blocks(R)) void cone_fp_batch_kernel(
    const FpItem* __restrict__ items, const int64_t* __restrict__ first,
    int32_t n, int64_t total) {
  for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (first[mid] <= g)
        lo = mid;
      else
        hi = mid - 1;
    }
    const FpItem& it = items[lo];
    const int64_t local = g - first[lo];
    fp_tile<R, ALL_COLC, fp_batch_pipe(R), kColsBatch, true>(it.tab, it.args,
               kk_xcd_remap(local, it.args.total_blocks, it.args.xcd_group));
  }
}

template <bool C, bool P>
void launch_single(int32_t rpt, dim3 grid, hipStream_t stream,
                   const kk::FpAngle* tab, const FpArgs& A) {
  const dim3 block(kColsSingle * kWave);
  switch (rpt) {
    case 1:
      hipLaunchKernelGGL((cone_fp_kernel<1, C, P>), grid, block, 0, stream,
                         tab, A);
      break;
    case 2:
      hipLaunchKernelGGL((cone_fp_kernel<2, C, P>), grid, block, 0, stream,
                         tab, A);
      break;
    default:
      hipLaunchKernelGGL((cone_fp_kernel<4, C, P>), grid, block, 0, stream,
                         tab, A);
      break;
  }
}

template <bool C>
void launch_batch(int32_t rpt, dim3 grid, hipStream_t stream,
                  const FpItem* items, const int64_t* first, int32_t n,
                  int64_t total) {
  const dim3 block(kColsBatch * kWave);
  switch (rpt) {
    case 1:
      hipLaunchKernelGGL((cone_fp_batch_kernel<1, C>), grid, block, 0, stream,
                         items, first, n, total);
      break;
    case 2:
      hipLaunchKernelGGL((cone_fp_batch_kernel<2, C>), grid, block, 0, stream,
                         items, first, n, total);
      break;
    default:
      hipLaunchKernelGGL((cone_fp_batch_kernel<4, C>), grid, block, 0, stream,
                         items, first, n, total);
      break;
  }
}

}  // namespace

namespace kk {

struct FpBatch {
  FpItem* items = nullptr;
  int64_t* first = nullptr;
  int32_t n = 0;
  int64_t total = 0;
  int32_t rpt = 4;
  bool all_colc = false;
};

int fp_batch_build(const std::vector<FpSpec>& specs, int32_t rows_per_thread,
                   FpBatch** out) {
  if (specs.empty()) return fail(KK_ERR_INVALID, "empty forward batch");
  int32_t max_rows = 0;
  for (const FpSpec& S : specs) max_rows = S.rows > max_rows ? S.rows : max_rows;
  const int32_t rpt = auto_rows_per_thread(max_rows, rows_per_thread);
  std::vector<FpItem> items(specs.size());
  std::vector<int64_t> first(specs.size());
  int64_t total = 0;
  for (size_t p = 0; p < specs.size(); ++p) {
    const int rc = make_fp_args(specs[p], rpt, kColsBatch, items[p].args);
    if (rc != KK_OK) return rc;
    items[p].tab = specs[p].plan->fp;
    first[p] = total;
    total += items[p].args.total_blocks;
  }
  auto* b = new (std::nothrow) FpBatch();
  if (b == nullptr) return fail(KK_ERR_INVALID, "out of host memory");
  b->n = static_cast<int32_t>(specs.size());
  b->total = total;
  b->rpt = rpt;
  b->all_colc = true;
  for (const FpSpec& S : specs) b->all_colc = b->all_colc && S.plan->fp_all_colc;
  hipError_t e = hipMalloc(&b->items, sizeof(FpItem) * items.size());
  if (e == hipSuccess) e = hipMalloc(&b->first, sizeof(int64_t) * first.size());
  if (e == hipSuccess)
    e = hipMemcpy(b->items, items.data(), sizeof(FpItem) * items.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->first, first.data(), sizeof(int64_t) * first.size(),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fp_batch_free(b);
    return hip_fail(e, "forward batch descriptors");
  }
  *out = b;
  return KK_OK;
}

void fp_batch_free(FpBatch* b) {
  if (b == nullptr) return;
  if (b->items) hipFree(b->items);
  if (b->first) hipFree(b->first);
  delete b;
}

int fp_batch_launch(const FpBatch* b, int32_t grid_blocks,
                    hipStream_t stream) {
  if (b == nullptr) return fail(KK_ERR_NOT_SET, "forward batch not built");
  int64_t g = grid_blocks > 0 ? grid_blocks : b->total;
  if (g > b->total) g = b->total;
  if (g <= 0) return KK_OK;
  const dim3 grid(static_cast<unsigned>(g));
  if (b->all_colc)
    launch_batch<true>(b->rpt, grid, stream, b->items, b->first, b->n,
                       b->total);
  else
    launch_batch<false>(b->rpt, grid, stream, b->items, b->first, b->n,
                        b->total);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "cone_fp_batch_kernel");
  return KK_OK;
}

}  // namespace kk

namespace kk {

int fp_spec(FpSpec& S, const kk_plan* plan, const float* tex, float* out,
            const int64_t os[3], int32_t rows, int32_t cols,
            int32_t additive) {
  if (os == nullptr) return fail(KK_ERR_INVALID, "kk_cone_fp: null strides");
  S = FpSpec{};
  S.plan = plan;
  S.tex = tex;
  S.out = out;
  for (int t = 0; t < 3; ++t) S.os[t] = os[t];
  S.rows = rows;
  S.cols = cols;
  S.mode = additive ? 1 : 0;
  return KK_OK;
}

int sirt_residual_spec(FpSpec& S, const kk_plan* plan, const float* x_tex,
                       const float* y, const int64_t ys[3], const float* R,
                       const int64_t rs[3], float* res_tex, int32_t n_angles,
                       int32_t rows, int32_t cols) {
  if (plan == nullptr || !plan->fp_set)
    return fail(KK_ERR_NOT_SET, "kk_sirt_fp_residual: no parameters");
  if (n_angles != plan->fp_n)
    return fail(KK_ERR_SHAPE, "kk_sirt_fp_residual: angle count");
  if (res_tex == nullptr || ys == nullptr || rs == nullptr)
    return fail(KK_ERR_INVALID, "kk_sirt_fp_residual: null argument");
  // write the residual straight into the backprojection texture: P(a, r, c)
  // is component 0 of element (a, r+1, c+1) and component 1 of (a, r, c+1)
  const int64_t tc = cols + kHalo, tr = rows + kHalo;
  S = FpSpec{};
  S.plan = plan;
  S.tex = x_tex;
  S.out = res_tex + 2 * (kHaloLo * tc + kHaloLo);
  S.os[0] = 2 * tr * tc;
  S.os[1] = 2 * tc;
  S.os[2] = 2;
  S.out_dual = 1 - 2 * tc;
  S.rows = rows;
  S.cols = cols;
  S.mode = 2;
  S.y = y;
  S.R = R;
  for (int t = 0; t < 3; ++t) {
    S.ys[t] = ys[t];
    S.rs[t] = rs[t];
  }
  return KK_OK;
}

}  // namespace kk

namespace {

int launch_fp_spec(const kk::FpSpec& S, int32_t rows_per_thread,
                   hipStream_t stream) {
  const int32_t rpt = auto_rows_per_thread(S.rows, rows_per_thread);
  FpArgs A;
  const int rc = make_fp_args(S, rpt, kColsSingle, A);
  if (rc != KK_OK) return rc;
  const dim3 grid(static_cast<unsigned>(A.total_blocks));
  const bool c = S.plan->fp_all_colc, p = S.ptrs != nullptr;
  if (c && !p)
    launch_single<true, false>(rpt, grid, stream, S.plan->fp, A);
  else if (!c && !p)
    launch_single<false, false>(rpt, grid, stream, S.plan->fp, A);
  else if (c)
    launch_single<true, true>(rpt, grid, stream, S.plan->fp, A);
  else
    launch_single<false, true>(rpt, grid, stream, S.plan->fp, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "cone_fp_kernel");
  return KK_OK;
}

}  // namespace

extern "C" int kk_cone_fp(const kk_plan* plan, const float* vol_tex,
                          float* proj, const int64_t ps[3], int32_t rows,
                          int32_t cols, int32_t additive, void* stream) {
  kk::FpSpec S;
  const int rc = kk::fp_spec(S, plan, vol_tex, proj, ps, rows, cols,
                             additive);
  if (rc != KK_OK) return rc;
  return launch_fp_spec(S, 0, static_cast<hipStream_t>(stream));
}

extern "C" int kk_cone_fp_ptrs(const kk_plan* plan, const float* vol_tex,
                               const int64_t* ptrs, int32_t ptr_axis,
                               const int64_t ps[3], int32_t rows,
                               int32_t cols, int32_t additive, void* stream) {
  if (ptrs == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_cone_fp_ptrs: null pointer table");
  kk::FpSpec S;
  const int rc = kk::fp_spec(S, plan, vol_tex, nullptr, ps, rows, cols,
                             additive);
  if (rc != KK_OK) return rc;
  S.ptrs = ptrs;
  S.ptr_axis = ptr_axis;
  return launch_fp_spec(S, 0, static_cast<hipStream_t>(stream));
}

extern "C" int kk_sirt_fp_residual(const kk_plan* plan, const float* x_tex,
                                   const float* y, const int64_t ys[3],
                                   const float* R, const int64_t rs[3],
                                   float* res_tex, int32_t n_angles,
                                   int32_t rows, int32_t cols, void* stream) {
  kk::FpSpec S;
  const int rc = kk::sirt_residual_spec(S, plan, x_tex, y, ys, R, rs,
                                        res_tex, n_angles, rows, cols);
  if (rc != KK_OK) return rc;
  return launch_fp_spec(S, 0, static_cast<hipStream_t>(stream));
}
