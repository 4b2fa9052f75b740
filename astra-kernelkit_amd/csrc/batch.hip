// Batches of independent cone-beam problems: one launch per operator for
// all of them.
//
// The reference handles one geometry per projector (projector.py:160-183)
// and its autograd operator loops over samples in Python
// (torch_support.py:19-20).  Many small problems (64^3..256^3 volumes, a
// few hundred angles) underfill an MI355X one launch at a time and pay a
// launch and a host round trip each.  Here a batch collects problem
// descriptions (the arguments of kk_cone_fp / kk_cone_bp /
// kk_sirt_fp_residual / kk_sirt_bp_update), builds device descriptors once
// for a chosen tile shape, and launches every forward (or backward) problem
// in one grid that walks the problems' concatenated workgroup tiles with a
// grid-stride loop: one workgroup per tile, or a persistent grid of a few
// workgroups per CU (the tuner times both; on the batching workload one
// per tile wins).  Texture builds (SIRT residuals into backprojection textures,
// updated volumes into forward-projection frames) are batched the same way.
// Launches allocate nothing and can be captured in a HIP graph (a SIRT
// iteration of the whole batch is four launches: forward projections,
// residuals, backprojections with the update, frame rebuilds).
#include <new>
#include <vector>

#include "kk_internal.h"

struct kk_batch_s {
  int device = 0;
  std::vector<kk::FpSpec> fp_specs;
  std::vector<kk::BpSpec> bp_specs;
  // texture builds: residuals into BP textures, volumes into FP frames
  std::vector<kk::TexJob> res_jobs, vol_jobs;
  int32_t n_res = 0, n_vol = 0;  // problems added (a volume has 1-3 jobs)
  kk::FpBatch* fp = nullptr;
  kk::BpBatch* bp = nullptr;
  kk::TexJobs* res = nullptr;
  kk::TexJobs* vol = nullptr;
};

namespace {

void release(kk_batch* b) {
  kk::fp_batch_free(b->fp);
  kk::bp_batch_free(b->bp);
  kk::tex_jobs_free(b->res);
  kk::tex_jobs_free(b->vol);
  b->fp = nullptr;
  b->bp = nullptr;
  b->res = nullptr;
  b->vol = nullptr;
}

int check_same_device(const kk_batch* b) {
  int dev = 0;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return kk::hip_fail(e, "hipGetDevice");
  if (dev != b->device)
    return kk::fail(KK_ERR_INVALID, "batch used on another device");
  return KK_OK;
}

}  // namespace

extern "C" int kk_batch_create(kk_batch** out) {
  if (out == nullptr) return kk::fail(KK_ERR_INVALID, "kk_batch_create");
  auto* b = new (std::nothrow) kk_batch();
  if (b == nullptr) return kk::fail(KK_ERR_INVALID, "out of host memory");
  const hipError_t e = hipGetDevice(&b->device);
  if (e != hipSuccess) {
    delete b;
    return kk::hip_fail(e, "hipGetDevice");
  }
  *out = b;
  return KK_OK;
}

extern "C" int kk_batch_destroy(kk_batch* b) {
  if (b == nullptr) return KK_OK;
  release(b);
  delete b;
  return KK_OK;
}

extern "C" int32_t kk_batch_size(const kk_batch* b, int32_t which) {
  if (b == nullptr) return -1;
  switch (which) {
    case 0: return static_cast<int32_t>(b->fp_specs.size());
    case 1: return static_cast<int32_t>(b->bp_specs.size());
    case 2: return b->n_res;
    case 3: return b->n_vol;
    default: return -1;
  }
}

extern "C" int kk_batch_add_fp(kk_batch* b, const kk_plan* plan,
                               const float* vol_tex, float* proj,
                               const int64_t proj_strides[3], int32_t rows,
                               int32_t cols, int32_t additive) {
  if (b == nullptr) return kk::fail(KK_ERR_INVALID, "kk_batch_add_fp: null");
  kk::FpSpec S;
  const int rc =
      kk::fp_spec(S, plan, vol_tex, proj, proj_strides, rows, cols, additive);
  if (rc != KK_OK) return rc;
  release(b);
  b->fp_specs.push_back(S);
  return KK_OK;
}

extern "C" int kk_batch_add_bp(kk_batch* b, const kk_plan* plan,
                               const float* proj_tex, int32_t n_angles,
                               int32_t rows, int32_t cols, float* vol,
                               const int32_t vol_shape[3],
                               const int64_t vol_strides[3],
                               float output_scale, int32_t additive) {
  if (b == nullptr) return kk::fail(KK_ERR_INVALID, "kk_batch_add_bp: null");
  kk::BpSpec S;
  const int rc = kk::bp_spec(S, plan, proj_tex, n_angles, rows, cols, vol,
                             vol_shape, vol_strides, output_scale, additive);
  if (rc != KK_OK) return rc;
  release(b);
  b->bp_specs.push_back(S);
  return KK_OK;
}

extern "C" int kk_batch_add_sirt_fp(kk_batch* b, const kk_plan* plan,
                                    const float* x_tex, const float* y,
                                    const int64_t y_strides[3],
                                    const float* R,
                                    const int64_t R_strides[3],
                                    float* res_tex, int32_t n_angles,
                                    int32_t rows, int32_t cols) {
  if (b == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_batch_add_sirt_fp: null");
  kk::FpSpec S;
  const int rc = kk::sirt_residual_spec(S, plan, x_tex, y, y_strides, R,
                                        R_strides, res_tex, n_angles, rows,
                                        cols);
  if (rc != KK_OK) return rc;
  release(b);
  b->fp_specs.push_back(S);
  return KK_OK;
}

extern "C" int kk_batch_add_sirt_bp(
    kk_batch* b, const kk_plan* plan, const float* res_tex, int32_t n_angles,
    int32_t rows, int32_t cols, float* x, const int32_t vol_shape[3],
    const int64_t x_strides[3], const float* C, const int64_t C_strides[3],
    const float* mask, const int64_t mask_strides[3], float lo, float hi,
    int32_t clip_lo, int32_t clip_hi, float* x_tex, int32_t x_tex_mask,
    float output_scale) {
  if (b == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_batch_add_sirt_bp: null");
  kk::BpSpec S;
  const int rc = kk::sirt_update_spec(
      S, plan, res_tex, n_angles, rows, cols, x, vol_shape, x_strides, C,
      C_strides, mask, mask_strides, lo, hi, clip_lo, clip_hi, x_tex,
      x_tex_mask, output_scale);
  if (rc != KK_OK) return rc;
  release(b);
  b->bp_specs.push_back(S);
  return KK_OK;
}

extern "C" int kk_batch_add_residual(kk_batch* b, const float* proj,
                                     const int64_t ps[3], const float* y,
                                     const int64_t ys[3], const float* R,
                                     const int64_t rs[3], int32_t n_angles,
                                     int32_t rows, int32_t cols, float* tex) {
  if (b == nullptr || proj == nullptr || ps == nullptr || y == nullptr ||
      ys == nullptr || R == nullptr || rs == nullptr || tex == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_batch_add_residual: null");
  if (n_angles <= 0 || rows <= 0 || cols <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_batch_add_residual: bad sizes");
  kk::TexJob J;
  J.src = proj;
  J.s0 = ps[0], J.s1 = ps[1], J.s2 = ps[2];
  J.y = y;
  J.y0 = ys[0], J.y1 = ys[1], J.y2 = ys[2];
  J.R = R;
  J.r0 = rs[0], J.r1 = rs[1], J.r2 = rs[2];
  J.tex = reinterpret_cast<float2*>(tex);
  J.n0 = n_angles, J.n1 = rows, J.n2 = cols;
  release(b);
  b->res_jobs.push_back(J);
  ++b->n_res;
  return KK_OK;
}

extern "C" int kk_batch_add_volume_texture(kk_batch* b, const float* vol,
                                           const int32_t shape[3],
                                           const int64_t st[3], int32_t mask,
                                           float* tex) {
  if (b == nullptr || vol == nullptr || shape == nullptr || st == nullptr ||
      tex == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_batch_add_volume_texture: null");
  if (shape[0] <= 0 || shape[1] <= 0 || shape[2] <= 0 || mask <= 0 ||
      mask > 7)
    return kk::fail(KK_ERR_INVALID, "kk_batch_add_volume_texture: bad args");
  // one job per frame, at the offsets of kk_volume_to_fp_texture
  std::vector<kk::TexJob> jobs;
  int64_t off = 0;  // in float2 elements
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    const int fb = kk::kFrameB[a], fc = kk::kFrameC[a];
    kk::TexJob J;
    J.src = vol;
    J.s0 = st[a], J.s1 = st[fb], J.s2 = st[fc];
    J.tex = reinterpret_cast<float2*>(tex) + off;
    J.n0 = shape[a], J.n1 = shape[fb], J.n2 = shape[fc];
    jobs.push_back(J);
    off += kk::frame_elems(shape, a);
  }
  release(b);
  b->vol_jobs.insert(b->vol_jobs.end(), jobs.begin(), jobs.end());
  ++b->n_vol;
  return KK_OK;
}

extern "C" int kk_batch_configure(kk_batch* b, int32_t fp_rows_per_thread,
                                  int32_t bp_zrun) {
  if (b == nullptr) return kk::fail(KK_ERR_INVALID, "kk_batch_configure");
  int rc = check_same_device(b);
  if (rc != KK_OK) return rc;
  release(b);
  if (!b->fp_specs.empty()) {
    rc = kk::fp_batch_build(b->fp_specs, fp_rows_per_thread, &b->fp);
    if (rc != KK_OK) return rc;
  }
  if (!b->bp_specs.empty()) {
    rc = kk::bp_batch_build(b->bp_specs, bp_zrun, &b->bp);
    if (rc != KK_OK) {
      release(b);
      return rc;
    }
  }
  if (!b->res_jobs.empty()) {
    rc = kk::tex_jobs_build(b->res_jobs, &b->res);
    if (rc != KK_OK) {
      release(b);
      return rc;
    }
  }
  if (!b->vol_jobs.empty()) {
    rc = kk::tex_jobs_build(b->vol_jobs, &b->vol);
    if (rc != KK_OK) {
      release(b);
      return rc;
    }
  }
  return KK_OK;
}

extern "C" int kk_batch_residual(const kk_batch* b, void* stream) {
  if (b == nullptr) return kk::fail(KK_ERR_INVALID, "kk_batch_residual: null");
  if (b->res_jobs.empty()) return KK_OK;
  if (b->res == nullptr)
    return kk::fail(KK_ERR_NOT_SET,
                    "kk_batch_residual: call kk_batch_configure");
  return kk::tex_jobs_launch(b->res, static_cast<hipStream_t>(stream));
}

extern "C" int kk_batch_volume_textures(const kk_batch* b, void* stream) {
  if (b == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_batch_volume_textures: null");
  if (b->vol_jobs.empty()) return KK_OK;
  if (b->vol == nullptr)
    return kk::fail(KK_ERR_NOT_SET,
                    "kk_batch_volume_textures: call kk_batch_configure");
  return kk::tex_jobs_launch(b->vol, static_cast<hipStream_t>(stream));
}

extern "C" int kk_batch_fp(const kk_batch* b, int32_t grid_blocks,
                           void* stream) {
  if (b == nullptr) return kk::fail(KK_ERR_INVALID, "kk_batch_fp: null");
  if (b->fp_specs.empty()) return KK_OK;
  if (b->fp == nullptr)
    return kk::fail(KK_ERR_NOT_SET, "kk_batch_fp: call kk_batch_configure");
  return kk::fp_batch_launch(b->fp, grid_blocks,
                             static_cast<hipStream_t>(stream));
}

extern "C" int kk_batch_bp(const kk_batch* b, int32_t grid_blocks,
                           void* stream) {
  if (b == nullptr) return kk::fail(KK_ERR_INVALID, "kk_batch_bp: null");
  if (b->bp_specs.empty()) return KK_OK;
  if (b->bp == nullptr)
    return kk::fail(KK_ERR_NOT_SET, "kk_batch_bp: call kk_batch_configure");
  return kk::bp_batch_launch(b->bp, grid_blocks,
                             static_cast<hipStream_t>(stream));
}
