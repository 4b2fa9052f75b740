// Host side of the C ABI: error reporting, plans and the per-angle device
// tables derived (in float64) from the reference's kernel parameters.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kk_internal.h"

namespace {
thread_local std::string g_last_error;
}

namespace kk {

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(KK_ERR_HIP, "HIP error %d (%s) in %s", static_cast<int>(e),
              hipGetErrorString(e), what);
}

}  // namespace kk

#define KK_HIP(call)                                          \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return kk::hip_fail(e_, #call);     \
  } while (0)

extern "C" const char* kk_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* kk_version(void) { return "kernelkit-hip 0.1 gfx950"; }

extern "C" int kk_plan_create(kk_plan** out) {
  if (out == nullptr) return kk::fail(KK_ERR_INVALID, "kk_plan_create: null");
  auto* p = new (std::nothrow) kk_plan();
  if (p == nullptr) return kk::fail(KK_ERR_INVALID, "out of host memory");
  hipError_t e = hipGetDevice(&p->device);
  if (e != hipSuccess) {
    delete p;
    return kk::hip_fail(e, "hipGetDevice");
  }
  *out = p;
  return KK_OK;
}

extern "C" int kk_plan_destroy(kk_plan* plan) {
  if (plan == nullptr) return KK_OK;
  int dev = 0;
  hipGetDevice(&dev);
  hipSetDevice(plan->device);
  if (plan->fp) hipFree(plan->fp);
  if (plan->bp) hipFree(plan->bp);
  if (plan->fr_idx) hipFree(plan->fr_idx);
  hipSetDevice(dev);
  delete plan;
  return KK_OK;
}

namespace {

template <typename T>
int upload_table(T** dev_ptr, int32_t* cap, const std::vector<T>& host) {
  const int32_t n = static_cast<int32_t>(host.size());
  if (n > *cap) {
    if (*dev_ptr) KK_HIP(hipFree(*dev_ptr));
    *dev_ptr = nullptr;
    *cap = 0;
    KK_HIP(hipMalloc(reinterpret_cast<void**>(dev_ptr), sizeof(T) * n));
    *cap = n;
  }
  KK_HIP(hipMemcpy(*dev_ptr, host.data(), sizeof(T) * n,
                   hipMemcpyHostToDevice));
  return KK_OK;
}

}  // namespace

// Host-only construction of the forward-projection table.  `axes`
// (nullable) gives the per-angle ray axis computed by the caller from the
// detector centres; when null the axis is taken from |source - detector
// centre| with the centre rebuilt from the corner, which needs rows/cols.
// `mask_out` receives the set of frame textures the table refers to.
static int build_fp_table(const double* params, int32_t n,
                          const int32_t* axes, int32_t rows, int32_t cols,
                          const int32_t shape[3], const double voxel_size[3],
                          std::vector<kk::FpAngle>& tab, int32_t* mask_out) {
  if (params == nullptr || shape == nullptr || voxel_size == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_set_params_fp: null argument");
  if (n <= 0) return kk::fail(KK_ERR_INVALID, "kk_set_params_fp: no angles");
  for (int k = 0; k < 3; ++k) {
    if (shape[k] <= 0)
      return kk::fail(KK_ERR_INVALID, "kk_set_params_fp: bad volume shape");
    if (!(voxel_size[k] > 0))
      return kk::fail(KK_ERR_INVALID, "kk_set_params_fp: bad voxel size");
  }
  if (axes == nullptr && (rows <= 0 || cols <= 0))
    return kk::fail(KK_ERR_INVALID,
                    "kk_set_params_fp: rows/cols needed to find ray axes");

  // Output scale table (kernels/cone_fp.py:287-300), reproduced verbatim.
  const double* v = voxel_size;
  double osc[3][3];
  const double eps = 1e-4;
  if (std::fabs(v[0] / v[1] - 1) > eps || std::fabs(v[0] / v[2] - 1) > eps) {
    const double t[3][3] = {
        {(v[1] / v[0]) * (v[1] / v[0]), (v[2] / v[0]) * (v[2] / v[0]),
         v[0] * v[0]},
        {(v[0] / v[1]) * (v[0] / v[1]), (v[2] / v[1]) * (v[2] / v[1]),
         v[0] * v[1]},
        {(v[0] / v[2]) * (v[0] / v[2]), (v[1] / v[2]) * (v[1] / v[2]),
         v[0] * v[2]}};
    std::memcpy(osc, t, sizeof(t));
  } else {
    for (int a = 0; a < 3; ++a) {
      osc[a][0] = 1.0;
      osc[a][1] = 1.0;
      osc[a][2] = v[0];
    }
  }

  std::vector<int> ray_axis(n);
  int32_t mask = 0;
  for (int32_t k = 0; k < n; ++k) {
    int m = 0;
    if (axes) {
      m = axes[k];
      if (m < 0 || m > 2)
        return kk::fail(KK_ERR_INVALID, "kk_set_params_fp: bad ray axis");
    } else {
      double best = -1.0;
      for (int t = 0; t < 3; ++t) {
        const double s = params[t * n + k];
        const double dc = params[(3 + t) * n + k] +
                          params[(6 + t) * n + k] * (cols / 2.0) +
                          params[(9 + t) * n + k] * (rows / 2.0);
        const double dist = std::fabs(s - dc);
        if (dist > best) {
          best = dist;
          m = t;
        }
      }
    }
    ray_axis[k] = m;
    mask |= 1 << m;
  }
  // frame textures in order x, y, z of those present
  uint64_t frame_off[3] = {0, 0, 0};
  uint64_t off = 0;
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    frame_off[a] = off;
    off += uint64_t(kk::frame_elems(shape, a)) * 8;
  }
  tab.assign(n, kk::FpAngle{});
  for (int32_t k = 0; k < n; ++k) {
    double s[3], d0[3], U[3], V[3];
    for (int t = 0; t < 3; ++t) {
      s[t] = params[(0 + t) * n + k];
      d0[t] = params[(3 + t) * n + k];
      U[t] = params[(6 + t) * n + k];
      V[t] = params[(9 + t) * n + k];
    }
    const int a = ray_axis[k], b = kk::kFrameB[a], c = kk::kFrameC[a];
    double q[3];
    for (int t = 0; t < 3; ++t) q[t] = d0[t] + 0.5 * U[t] + 0.5 * V[t] - s[t];
    kk::FpAngle& e = tab[k];
    std::memset(&e, 0, sizeof(e));
    e.qa = q[a];
    e.qb = q[b];
    e.qc = q[c];
    e.ua = U[a];
    e.ub = U[b];
    e.uc = U[c];
    e.va = V[a];
    e.vb = V[b];
    e.vc = V[c];
    e.sig = 0.5 - 0.5 * shape[a] - s[a];
    e.taub = s[b] + 0.5 * shape[b] - 0.5 + kk::kHaloLo;
    e.tauc = s[c] + 0.5 * shape[c] - 0.5 + kk::kHaloLo;
    e.sc1 = float(osc[a][0]);
    e.sc2 = float(osc[a][1]);
    e.oscale = float(osc[a][2]);
    e.na = shape[a];
    e.nb = shape[b];
    e.nc = shape[c];
    e.axis = a;
    e.tex_off = frame_off[a];
    const int64_t sbb = int64_t(shape[c] + kk::kHalo) * 8;
    const int64_t sab = int64_t(shape[b] + kk::kHalo) * sbb;
    // in-slice byte offsets must fit 32 bits, row strides 24 bits
    // (v_mul_u32_u24)
    if (sab >= (int64_t(1) << 32) || sbb >= (int64_t(1) << 24))
      return kk::fail(KK_ERR_UNSUPPORTED,
                      "volume slice too large for 32-bit offsets");
    e.sab = uint32_t(sab);
    e.sbb = uint32_t(sbb);
    e.colc = (V[a] == 0.0 && V[b] == 0.0) ? 1 : 0;
  }
  if (mask_out) *mask_out = mask;
  return KK_OK;
}

extern "C" int kk_set_params_fp(kk_plan* plan, const double* params,
                                int32_t n, const int32_t* axes, int32_t rows,
                                int32_t cols, const int32_t shape[3],
                                const double voxel_size[3]) {
  if (plan == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_set_params_fp: null plan");
  std::vector<kk::FpAngle> tab;
  int32_t mask = 0;
  const int brc = build_fp_table(params, n, axes, rows, cols, shape,
                                 voxel_size, tab, &mask);
  if (brc != KK_OK) return brc;
  int dev = 0;
  KK_HIP(hipGetDevice(&dev));
  KK_HIP(hipSetDevice(plan->device));
  // angle ids grouped by ray axis (the transpose runs one pass per axis)
  std::vector<int32_t> idx;
  idx.reserve(n);
  int32_t start[4] = {0, 0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    start[a] = static_cast<int32_t>(idx.size());
    for (int32_t k = 0; k < n; ++k)
      if (tab[k].axis == a) idx.push_back(k);
  }
  start[3] = n;
  int rc = upload_table(&plan->fp, &plan->fp_cap, tab);
  if (rc == KK_OK) rc = upload_table(&plan->fr_idx, &plan->fr_cap, idx);
  hipSetDevice(dev);
  if (rc != KK_OK) return rc;
  for (int a = 0; a < 4; ++a) plan->fr_start[a] = start[a];
  plan->fp_all_colc = true;
  plan->fp_any_colc = false;
  for (int32_t k = 0; k < n; ++k) {
    plan->fp_all_colc = plan->fp_all_colc && tab[k].colc != 0;
    plan->fp_any_colc = plan->fp_any_colc || tab[k].colc != 0;
  }
  plan->fp_n = n;
  for (int t = 0; t < 3; ++t) plan->fp_shape[t] = shape[t];
  plan->fp_mask = mask;
  plan->fp_set = true;
  return KK_OK;
}

static int build_bp_table(const double* params, int32_t n,
                          std::vector<kk::BpAngle>& tab) {
  if (params == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_set_params_bp: null argument");
  if (n <= 0) return kk::fail(KK_ERR_INVALID, "kk_set_params_bp: no angles");
  tab.assign(n, kk::BpAngle{});
  for (int32_t k = 0; k < n; ++k) {
    const double* p = params + 12 * int64_t(k);
    // numerators shifted by (halo - 1/2) * den: texture coordinate ->
    // padded array coordinate
    double nu[4], nv[4], de[4];
    const double shift = kk::kHaloLo - 0.5;
    for (int t = 0; t < 4; ++t) {
      de[t] = p[8 + t];
      nu[t] = p[0 + t] + shift * de[t];
      nv[t] = p[4 + t] + shift * de[t];
    }
    for (int t = 0; t < 12; ++t)
      if (!std::isfinite(p[t]))
        return kk::fail(KK_ERR_INVALID,
                        "kk_set_params_bp: non-finite parameter at angle %d",
                        k);
    kk::BpAngle& e = tab[k];
    std::memset(&e, 0, sizeof(e));
    e.nux = nu[0];
    e.nuy = nu[1];
    e.nuz = nu[2];
    e.nuw = nu[3];
    e.nvx = nv[0];
    e.nvy = nv[1];
    e.nvz = nv[2];
    e.nvw = nv[3];
    e.dx = de[0];
    e.dy = de[1];
    e.dz = de[2];
    e.dw = de[3];
    // z-invariant weight (and column coordinate): negligible (relative
    // 1e-13) z coefficients are treated as zero.
    const double du = std::fabs(nu[0]) + std::fabs(nu[1]) + std::fabs(nu[3]);
    const double dd = std::fabs(de[0]) + std::fabs(de[1]) + std::fabs(de[3]);
    const bool den_flat = std::fabs(de[2]) <= 1e-13 * dd;
    const bool u_flat = std::fabs(nu[2]) <= 1e-13 * du;
    e.fast = den_flat ? (u_flat ? 1 : 2) : 0;
    if (den_flat) e.dz = 0.0;
    if (e.fast == 1) e.nuz = 0.0;
  }
  return KK_OK;
}

extern "C" int kk_set_params_bp(kk_plan* plan, const double* params,
                                int32_t n) {
  if (plan == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_set_params_bp: null plan");
  std::vector<kk::BpAngle> tab;
  const int brc = build_bp_table(params, n, tab);
  if (brc != KK_OK) return brc;
  int dev = 0;
  KK_HIP(hipGetDevice(&dev));
  KK_HIP(hipSetDevice(plan->device));
  int rc = upload_table(&plan->bp, &plan->bp_cap, tab);
  hipSetDevice(dev);
  if (rc != KK_OK) return rc;
  plan->bp_n = n;
  plan->bp_all_circ = true;
  for (int32_t k = 0; k < n; ++k)
    plan->bp_all_circ = plan->bp_all_circ && tab[k].fast == 1;
  plan->bp_set = true;
  return KK_OK;
}

extern "C" int32_t kk_fp_frame_mask(const kk_plan* plan) {
  return (plan && plan->fp_set) ? plan->fp_mask : 0;
}

extern "C" int64_t kk_fp_texture_elems(const int32_t s[3], int32_t mask) {
  if (s == nullptr || s[0] <= 0 || s[1] <= 0 || s[2] <= 0 || mask <= 0 ||
      mask > 7)
    return -1;
  int64_t n = 0;
  for (int a = 0; a < 3; ++a)
    if (mask & (1 << a)) n += 2 * kk::frame_elems(s, a);
  return n;
}

extern "C" int64_t kk_bp_texture_elems(int32_t n, int32_t rows,
                                       int32_t cols) {
  if (n <= 0 || rows <= 0 || cols <= 0) return -1;
  return 2 * int64_t(n) * (rows + kk::kHalo) * (cols + kk::kHalo);
}

extern "C" int kk_fp_table(const double* params, int32_t n,
                           const int32_t* axes, int32_t rows, int32_t cols,
                           const int32_t shape[3], const double voxel_size[3],
                           void* out, int32_t* mask) {
  if (out == nullptr) return kk::fail(KK_ERR_INVALID, "kk_fp_table: null");
  std::vector<kk::FpAngle> tab;
  const int rc = build_fp_table(params, n, axes, rows, cols, shape,
                                voxel_size, tab, mask);
  if (rc != KK_OK) return rc;
  std::memcpy(out, tab.data(), sizeof(kk::FpAngle) * tab.size());
  return KK_OK;
}

extern "C" int kk_bp_table(const double* params, int32_t n, void* out) {
  if (out == nullptr) return kk::fail(KK_ERR_INVALID, "kk_bp_table: null");
  std::vector<kk::BpAngle> tab;
  const int rc = build_bp_table(params, n, tab);
  if (rc != KK_OK) return rc;
  std::memcpy(out, tab.data(), sizeof(kk::BpAngle) * tab.size());
  return KK_OK;
}
