// Batches of independent cone-beam problems: one persistent launch per
// operator for all of them.
//
// The reference handles one geometry per projector (projector.py:160-183)
// and its autograd operator loops over samples in Python
// (torch_support.py:19-20).  Many small problems (64^3..256^3 volumes, a
// few hundred angles) underfill an MI355X one launch at a time and pay a
// launch and a host round trip each.  Here a batch collects problem
// descriptions (the arguments of kk_cone_fp / kk_cone_bp /
// kk_sirt_fp_residual / kk_sirt_bp_update), builds device descriptors once
// for a chosen tile shape, and launches every forward (or backward) problem
// in one persistent grid that walks the problems' concatenated workgroup
// tiles.  Launches allocate nothing and can be captured in a HIP graph (a
// SIRT iteration of the whole batch is two launches).
#include <new>
#include <vector>

#include "kk_internal.h"

struct kk_batch_s {
  int device = 0;
  std::vector<kk::FpSpec> fp_specs;
  std::vector<kk::BpSpec> bp_specs;
  kk::FpBatch* fp = nullptr;
  kk::BpBatch* bp = nullptr;
};

namespace {

void release(kk_batch* b) {
  kk::fp_batch_free(b->fp);
  kk::bp_batch_free(b->bp);
  b->fp = nullptr;
  b->bp = nullptr;
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
  return static_cast<int32_t>(which == 0 ? b->fp_specs.size()
                                         : b->bp_specs.size());
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
  return KK_OK;
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
