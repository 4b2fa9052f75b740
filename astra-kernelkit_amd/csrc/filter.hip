// FDK pre-weighting and row-wise ramp filtering for gfx950.
//
// Behaviour follows kernelkit/processing.py of the reference:
// * preweight (:149-188): p[a, r, c] *= |d - s| / |corner(r, c) - s| with
//   corner(r, c) = d_min + c * U + r * V (pixel corners, world units);
// * filter (:96-126): every detector row is zero-padded to N (a power of two,
//   N >= 2 * cols), transformed, multiplied by the real response H, inverse
//   transformed, and its first `cols` entries kept.  The reference does this
//   in float64 with cuFFT, one Python-level loop iteration per angle.
//
// MI355X design: one workgroup filters two detector rows at once (radix-4
// passes in registers, twiddles staged in LDS).  They are
// packed as the real and imaginary parts of one complex signal; because H
// is real and even, the real part of the result is the filtered first row
// and the imaginary part the filtered second row.  The transform runs in
// LDS: decimation-in-frequency passes (natural -> bit-reversed order;
// radix 4, plus one radix-2 stage when log2 N is odd), the multiplication
// by H stored in bit-reversed order with 1/N folded in, and
// decimation-in-time inverse passes (bit-reversed -> natural), so no
// explicit permutation is needed.  Twiddles and H come from
// float64 host tables rounded once to float32.  The pre-weight can be fused
// into the load, and the output can be written anywhere through strides
// (e.g. straight into the backprojector's padded texture).
#include <cmath>
#include <vector>

#include "kk_internal.h"

namespace {

struct PreweightAngle {  // world units, float32 after an origin shift
  float qx, qy, qz;      // d_min - s
  float ux, uy, uz;      // column step u * pixel_width
  float vx, vy, vz;      // row step v * pixel_height
  float central;         // |d - s|
  float _pad;
};

__device__ __forceinline__ float preweight_factor(const PreweightAngle& P,
                                                  int32_t r, int32_t c) {
  const float fr = static_cast<float>(r), fc = static_cast<float>(c);
  const float x = fmaf(fr, P.vx, fmaf(fc, P.ux, P.qx));
  const float y = fmaf(fr, P.vy, fmaf(fc, P.uy, P.qy));
  const float z = fmaf(fr, P.vz, fmaf(fc, P.uz, P.qz));
  return P.central / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
}

__global__ __launch_bounds__(256) void preweight_kernel(
    float* __restrict__ p, int64_t s0, int64_t s1, int64_t s2, int32_t rows,
    int32_t cols, const PreweightAngle* __restrict__ tab) {
  const int32_t c = blockIdx.x * 256 + threadIdx.x;
  const int32_t r = blockIdx.y;
  const int32_t a = blockIdx.z;
  if (c >= cols) return;
  const int64_t o = a * s0 + r * s1 + c * s2;
  p[o] *= preweight_factor(tab[a], r, c);
}

struct FilterArgs {
  const float* in;
  int64_t is0, is1, is2;
  float* out;
  int64_t os0, os1, os2;
  int64_t out_dual;  // if nonzero: also write out[o + out_dual]
  int64_t total_rows;
  int32_t rows, cols, n;
  const PreweightAngle* pw;  // optional fused pre-weight
  const float2* tw;          // exp(-2 pi i k / N), k < N/2
  const float* hb;           // H[bitrev(k)] / N
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}

// TW_LDS: the twiddles are copied into LDS after the signal (when both fit
// in 64 KiB); otherwise they are read from global memory.
template <bool TW_LDS>
__global__ __launch_bounds__(256) void filter_rows_kernel(FilterArgs A) {
  extern __shared__ float2 buf[];
  const int32_t N = A.n;
  const int64_t g0 = 2 * int64_t(blockIdx.x);
  const int64_t g1 = g0 + 1;
  const bool has1 = g1 < A.total_rows;
  const int32_t a0 = static_cast<int32_t>(g0 / A.rows);
  const int32_t r0 = static_cast<int32_t>(g0 % A.rows);
  const int32_t a1 = has1 ? static_cast<int32_t>(g1 / A.rows) : a0;
  const int32_t r1 = has1 ? static_cast<int32_t>(g1 % A.rows) : r0;
  const float* in0 = A.in + a0 * A.is0 + r0 * A.is1;
  const float* in1 = A.in + a1 * A.is0 + r1 * A.is1;

  for (int32_t c = threadIdx.x; c < N; c += 256) {
    float re = 0.f, im = 0.f;
    if (c < A.cols) {
      re = in0[c * A.is2];
      if (has1) im = in1[c * A.is2];
      if (A.pw) {
        re *= preweight_factor(A.pw[a0], r0, c);
        if (has1) im *= preweight_factor(A.pw[a1], r1, c);
      }
    }
    buf[c] = make_float2(re, im);
  }
  __syncthreads();

  const int32_t half = N >> 1;
  const float2* tw = A.tw;
  if constexpr (TW_LDS) {
    float2* tl = buf + N;
    for (int32_t k = threadIdx.x; k < half; k += 256) tl[k] = A.tw[k];
    __syncthreads();
    tw = tl;
  }
  const int32_t quarter = N >> 2;
  // forward transform, decimation in frequency: natural -> bit-reversed.
  // Pairs of radix-2 stages (spans m and m / 2) run as one radix-4 pass
  // in registers; a last span-1 stage remains when log2(N) is odd.
  int32_t m = half, ts = 1;
  for (; m >= 2; m >>= 2, ts <<= 2) {
    const int32_t q = m >> 1;
    for (int32_t g = threadIdx.x; g < quarter; g += 256) {
      const int32_t k = g & (q - 1);
      const int32_t ia = ((g - k) << 2) + k;
      const float2 xa = buf[ia], xb = buf[ia + q];
      const float2 xc = buf[ia + m], xd = buf[ia + m + q];
      // span m: (a, c) with twiddle k, (b, d) with twiddle k + q
      const float2 a1 = make_float2(xa.x + xc.x, xa.y + xc.y);
      const float2 c1 =
          cmul(make_float2(xa.x - xc.x, xa.y - xc.y), tw[k * ts]);
      const float2 b1 = make_float2(xb.x + xd.x, xb.y + xd.y);
      const float2 d1 =
          cmul(make_float2(xb.x - xd.x, xb.y - xd.y), tw[(k + q) * ts]);
      // span q: (a, b) and (c, d), both with twiddle k of the next stage
      const float2 w2 = tw[k * (ts << 1)];
      buf[ia] = make_float2(a1.x + b1.x, a1.y + b1.y);
      buf[ia + q] = cmul(make_float2(a1.x - b1.x, a1.y - b1.y), w2);
      buf[ia + m] = make_float2(c1.x + d1.x, c1.y + d1.y);
      buf[ia + m + q] = cmul(make_float2(c1.x - d1.x, c1.y - d1.y), w2);
    }
    __syncthreads();
  }
  if (m == 1) {  // span 1: twiddle 1
    for (int32_t j = threadIdx.x; j < half; j += 256) {
      const float2 x0 = buf[2 * j], x1 = buf[2 * j + 1];
      buf[2 * j] = make_float2(x0.x + x1.x, x0.y + x1.y);
      buf[2 * j + 1] = make_float2(x0.x - x1.x, x0.y - x1.y);
    }
    __syncthreads();
  }
  // multiply by the real response (bit-reversed order, 1/N folded in)
  for (int32_t c = threadIdx.x; c < N; c += 256) {
    const float h = A.hb[c];
    const float2 x = buf[c];
    buf[c] = make_float2(x.x * h, x.y * h);
  }
  __syncthreads();
  // inverse transform, decimation in time: bit-reversed -> natural; a
  // first span-1 stage when log2(N) is odd, then radix-4 passes (spans
  // m and 2 m)
  m = 1;
  ts = half;
  if (__builtin_ctz(static_cast<uint32_t>(N)) & 1) {  // log2(N) odd
    for (int32_t j = threadIdx.x; j < half; j += 256) {
      const float2 x0 = buf[2 * j], x1 = buf[2 * j + 1];
      buf[2 * j] = make_float2(x0.x + x1.x, x0.y + x1.y);
      buf[2 * j + 1] = make_float2(x0.x - x1.x, x0.y - x1.y);
    }
    __syncthreads();
    m = 2;
    ts = half >> 1;
  }
  for (; m < N; m <<= 2, ts >>= 2) {
    for (int32_t g = threadIdx.x; g < quarter; g += 256) {
      const int32_t k = g & (m - 1);
      const int32_t ia = ((g - k) << 2) + k;
      const float2 xa = buf[ia], xb = buf[ia + m];
      const float2 xc = buf[ia + 2 * m], xd = buf[ia + 3 * m];
      // span m: (a, b) and (c, d), both with twiddle k
      const float2 w1 = tw[k * ts];
      const float2 tb = cmulc(xb, w1), td = cmulc(xd, w1);
      const float2 a1 = make_float2(xa.x + tb.x, xa.y + tb.y);
      const float2 b1 = make_float2(xa.x - tb.x, xa.y - tb.y);
      const float2 c1 = make_float2(xc.x + td.x, xc.y + td.y);
      const float2 d1 = make_float2(xc.x - td.x, xc.y - td.y);
      // span 2 m: (a, c) with twiddle k, (b, d) with twiddle k + m
      const float2 tc = cmulc(c1, tw[k * (ts >> 1)]);
      const float2 tdd = cmulc(d1, tw[(k + m) * (ts >> 1)]);
      buf[ia] = make_float2(a1.x + tc.x, a1.y + tc.y);
      buf[ia + 2 * m] = make_float2(a1.x - tc.x, a1.y - tc.y);
      buf[ia + m] = make_float2(b1.x + tdd.x, b1.y + tdd.y);
      buf[ia + 3 * m] = make_float2(b1.x - tdd.x, b1.y - tdd.y);
    }
    __syncthreads();
  }
  float* out0 = A.out + a0 * A.os0 + r0 * A.os1;
  float* out1 = A.out + a1 * A.os0 + r1 * A.os1;
  for (int32_t c = threadIdx.x; c < A.cols; c += 256) {
    const float2 x = buf[c];
    out0[c * A.os2] = x.x;
    if (has1) out1[c * A.os2] = x.y;
    if (A.out_dual) {
      out0[c * A.os2 + A.out_dual] = x.x;
      if (has1) out1[c * A.os2 + A.out_dual] = x.y;
    }
  }
}

}  // namespace

// Longest row transform: the signal stays in LDS (8 bytes per point,
// 128 KiB at 16384 points, within gfx950's 160 KiB per workgroup), so
// detectors up to 8192 columns filter in one pass.
constexpr int32_t kMaxFilterN = 16384;

struct kk_filter_s {
  int device = 0;
  int32_t n = 0;
  float2* tw = nullptr;
  float* hb = nullptr;
  PreweightAngle* pw = nullptr;
  int32_t pw_n = 0, pw_cap = 0;
};

extern "C" int kk_filter_create(kk_filter** out) {
  if (out == nullptr) return kk::fail(KK_ERR_INVALID, "kk_filter_create");
  auto* f = new (std::nothrow) kk_filter();
  if (f == nullptr) return kk::fail(KK_ERR_INVALID, "out of host memory");
  hipError_t e = hipGetDevice(&f->device);
  if (e != hipSuccess) {
    delete f;
    return kk::hip_fail(e, "hipGetDevice");
  }
  *out = f;
  return KK_OK;
}

extern "C" int kk_filter_destroy(kk_filter* f) {
  if (f == nullptr) return KK_OK;
  if (f->tw) hipFree(f->tw);
  if (f->hb) hipFree(f->hb);
  if (f->pw) hipFree(f->pw);
  delete f;
  return KK_OK;
}

extern "C" int kk_filter_set_response(kk_filter* f, const double* H,
                                      int32_t n) {
  if (f == nullptr || H == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_filter_set_response: null");
  if (n < 2 || (n & (n - 1)) != 0 || n > kMaxFilterN)
    return kk::fail(KK_ERR_INVALID,
                    "filter length must be a power of two in [2, %d]",
                    kMaxFilterN);
  int log2n = 0;
  while ((1 << log2n) < n) ++log2n;
  std::vector<float2> tw(n / 2);
  const double pi = 3.14159265358979323846;
  for (int32_t k = 0; k < n / 2; ++k) {
    const double ang = -2.0 * pi * double(k) / double(n);
    tw[k] = make_float2(float(std::cos(ang)), float(std::sin(ang)));
  }
  // Only the even part He[k] = (H[k] + H[-k]) / 2 is kept.  For a real row
  // p, Re(IDFT(DFT(p) H)) = IDFT(DFT(p) He): the odd part of a real H only
  // adds an imaginary component, which the real part drops.  Windows such
  // as fftshift(hanning(N)) are not even (hanning is symmetric about
  // (N-1)/2), so this matters; with an even response the kernel may pack
  // two real rows into one complex transform.
  std::vector<float> hb(n);
  for (int32_t k = 0; k < n; ++k) {
    int32_t rev = 0;
    for (int b = 0; b < log2n; ++b) rev |= ((k >> b) & 1) << (log2n - 1 - b);
    const double he = 0.5 * (H[rev] + H[(n - rev) & (n - 1)]);
    hb[k] = float(he / double(n));
  }
  if (f->tw) hipFree(f->tw);
  if (f->hb) hipFree(f->hb);
  f->tw = nullptr;
  f->hb = nullptr;
  f->n = 0;
  hipError_t e = hipMalloc(&f->tw, sizeof(float2) * tw.size());
  if (e != hipSuccess) return kk::hip_fail(e, "hipMalloc twiddles");
  e = hipMalloc(&f->hb, sizeof(float) * hb.size());
  if (e != hipSuccess) return kk::hip_fail(e, "hipMalloc response");
  e = hipMemcpy(f->tw, tw.data(), sizeof(float2) * tw.size(),
                hipMemcpyHostToDevice);
  if (e != hipSuccess) return kk::hip_fail(e, "hipMemcpy twiddles");
  e = hipMemcpy(f->hb, hb.data(), sizeof(float) * hb.size(),
                hipMemcpyHostToDevice);
  if (e != hipSuccess) return kk::hip_fail(e, "hipMemcpy response");
  f->n = n;
  return KK_OK;
}

extern "C" int kk_filter_set_preweight(kk_filter* f, const double* p,
                                       int32_t n) {
  if (f == nullptr || p == nullptr || n <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_filter_set_preweight: bad args");
  std::vector<PreweightAngle> tab(n);
  for (int32_t k = 0; k < n; ++k) {
    const double* q = p + 10 * int64_t(k);
    PreweightAngle& e = tab[k];
    e.qx = float(q[0]);
    e.qy = float(q[1]);
    e.qz = float(q[2]);
    e.ux = float(q[3]);
    e.uy = float(q[4]);
    e.uz = float(q[5]);
    e.vx = float(q[6]);
    e.vy = float(q[7]);
    e.vz = float(q[8]);
    e.central = float(q[9]);
    e._pad = 0.f;
  }
  if (n > f->pw_cap) {
    if (f->pw) hipFree(f->pw);
    f->pw = nullptr;
    f->pw_cap = 0;
    hipError_t e = hipMalloc(&f->pw, sizeof(PreweightAngle) * n);
    if (e != hipSuccess) return kk::hip_fail(e, "hipMalloc preweight");
    f->pw_cap = n;
  }
  hipError_t e = hipMemcpy(f->pw, tab.data(), sizeof(PreweightAngle) * n,
                           hipMemcpyHostToDevice);
  if (e != hipSuccess) return kk::hip_fail(e, "hipMemcpy preweight");
  f->pw_n = n;
  return KK_OK;
}

extern "C" int kk_preweight(const kk_filter* f, float* proj, int32_t n,
                            int32_t rows, int32_t cols, const int64_t st[3],
                            void* stream) {
  if (f == nullptr || proj == nullptr || st == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_preweight: null argument");
  if (f->pw == nullptr || f->pw_n != n)
    return kk::fail(KK_ERR_NOT_SET, "kk_preweight: pre-weights not set for "
                                    "%d angles",
                    n);
  if (rows <= 0 || cols <= 0 || rows > 65535 || n > 65535)
    return kk::fail(KK_ERR_INVALID, "kk_preweight: bad sizes");
  const dim3 grid((cols + 255) / 256, rows, n);
  hipLaunchKernelGGL(preweight_kernel, grid, dim3(256), 0,
                     static_cast<hipStream_t>(stream), proj, st[0], st[1],
                     st[2], rows, cols, f->pw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "preweight_kernel");
  return KK_OK;
}

static int filter_rows(const kk_filter* f, const float* in,
                       const int64_t is[3], float* out, const int64_t os[3],
                       int64_t out_dual, int32_t n, int32_t rows,
                       int32_t cols, int32_t fuse_preweight,
                       hipStream_t stream) {
  if (f == nullptr || in == nullptr || out == nullptr || is == nullptr ||
      os == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_filter_rows: null argument");
  if (f->n == 0) return kk::fail(KK_ERR_NOT_SET, "filter response not set");
  if (n <= 0 || rows <= 0 || cols <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_filter_rows: bad sizes");
  if (cols > f->n)
    return kk::fail(KK_ERR_INVALID,
                    "kk_filter_rows: %d columns exceed the filter length %d",
                    cols, f->n);
  if (fuse_preweight && (f->pw == nullptr || f->pw_n != n))
    return kk::fail(KK_ERR_NOT_SET, "kk_filter_rows: pre-weights not set");
  FilterArgs A{};
  A.in = in;
  A.is0 = is[0];
  A.is1 = is[1];
  A.is2 = is[2];
  A.out = out;
  A.os0 = os[0];
  A.os1 = os[1];
  A.os2 = os[2];
  A.out_dual = out_dual;
  A.total_rows = int64_t(n) * rows;
  A.rows = rows;
  A.cols = cols;
  A.n = f->n;
  A.pw = fuse_preweight ? f->pw : nullptr;
  A.tw = f->tw;
  A.hb = f->hb;
  const int64_t pairs = (A.total_rows + 1) / 2;
  if (pairs >= (int64_t(1) << 31))
    return kk::fail(KK_ERR_UNSUPPORTED, "kk_filter_rows: too many rows");
  const size_t lds_tw = sizeof(float2) * (size_t(f->n) + size_t(f->n) / 2);
  const size_t lds = sizeof(float2) * size_t(f->n);
  if (lds_tw <= 65536) {
    hipLaunchKernelGGL(filter_rows_kernel<true>, dim3(unsigned(pairs)),
                       dim3(256), lds_tw, stream, A);
  } else {
    if (lds > 65536) {
      // beyond 64 KiB of dynamic LDS: check the device limit, and raise
      // the kernel's
      int dev = 0, limit = 0;
      hipError_t e = hipGetDevice(&dev);
      if (e == hipSuccess)
        e = hipDeviceGetAttribute(
            &limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
      if (e != hipSuccess) return kk::hip_fail(e, "LDS limit");
      if (size_t(limit) < lds)
        return kk::fail(KK_ERR_UNSUPPORTED,
                        "kk_filter_rows: a %d-point row needs %zu bytes of "
                        "LDS, the device allows %d",
                        f->n, lds, limit);
      e = hipFuncSetAttribute(
          reinterpret_cast<const void*>(&filter_rows_kernel<false>),
          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
      if (e != hipSuccess) return kk::hip_fail(e, "hipFuncSetAttribute");
    }
    hipLaunchKernelGGL(filter_rows_kernel<false>, dim3(unsigned(pairs)),
                       dim3(256), lds, stream, A);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "filter_rows_kernel");
  return KK_OK;
}

extern "C" int kk_filter_rows(const kk_filter* f, const float* in,
                              const int64_t is[3], float* out,
                              const int64_t os[3], int32_t n, int32_t rows,
                              int32_t cols, int32_t fuse_preweight,
                              void* stream) {
  return filter_rows(f, in, is, out, os, 0, n, rows, cols, fuse_preweight,
                     static_cast<hipStream_t>(stream));
}

extern "C" int kk_filter_rows_to_bp_texture(const kk_filter* f,
                                            const float* in,
                                            const int64_t is[3], float* tex,
                                            int32_t n, int32_t rows,
                                            int32_t cols,
                                            int32_t fuse_preweight,
                                            void* stream) {
  if (tex == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_filter_rows_to_bp_texture: null");
  // P(a, r, c) is component 0 of texture element (a, r+1, c+1) and
  // component 1 of (a, r, c+1)
  const int64_t tc = cols + kk::kHalo, tr = rows + kk::kHalo;
  const int64_t os[3] = {2 * tr * tc, 2 * tc, 2};
  float* out = tex + 2 * (kk::kHaloLo * tc + kk::kHaloLo);
  return filter_rows(f, in, is, out, os, 1 - 2 * tc, n, rows, cols,
                     fuse_preweight, static_cast<hipStream_t>(stream));
}
