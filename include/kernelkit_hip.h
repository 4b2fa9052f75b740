/*
 * kernelkit_hip.h -- C ABI of the MI355X (gfx950) cone-beam operators.
 *
 * Plain C: pointers, sizes and status codes; no torch or C++ types.  Every
 * device-side call is asynchronous on the given HIP stream (a hipStream_t
 * passed as void*; NULL = the null stream), allocates nothing and never
 * synchronises, so it can be captured in a HIP graph.  Only the
 * kk_set_params_*, kk_filter_set_* and kk_batch_configure calls allocate (a
 * small device table) and synchronise; scratch that a launch needs is passed
 * in by the caller (kk_cone_fp_transpose's work).
 *
 * Errors: every int-returning function returns KK_OK (0) or a negative
 * KK_ERR_* code, and records a message for kk_last_error() (thread-local).
 * Validation happens before any launch.
 *
 * Conventions (mirroring the reference, paths relative to its repository):
 *   world axes (x, y, z); volumes are addressed by world strides;
 *   projections by (angle, row, col) strides, in elements (floats);
 *   kernel parameters are the reference's kernel parameters, in float64,
 *   in the volume's voxel frame (kernelkit/geom/__init__.py:305-330).
 */
#ifndef KERNELKIT_HIP_H
#define KERNELKIT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KK_OK 0
#define KK_ERR_INVALID (-1)     /* bad argument (ValueError)                */
#define KK_ERR_HIP (-2)         /* HIP runtime failure                       */
#define KK_ERR_NOT_SET (-3)     /* parameters not set (KernelMemoryUnprepared) */
#define KK_ERR_SHAPE (-4)       /* shape mismatch (ValueError)               */
#define KK_ERR_UNSUPPORTED (-5) /* unsupported case (NotImplementedError)    */

/* Textures.  The reference samples CUDA texture objects with border
 * addressing (zero outside, kernelkit/kernel.py:22-32).  Here a "texture" is
 * a library-defined zero-halo layout of float2 pairs (value, neighbour),
 * KK_HALO_LO elements before and KK_HALO_HI after along the interpolated
 * axes, built from any strided array by the kk_*_to_*_texture calls:
 *  - FP texture: for each ray axis a present in the geometry (bit a of the
 *    frame mask), a slice-major frame [n_a][n_b+3][n_c+3] of float2 with
 *    (b, c) = (y, z), (x, z), (y, x) for a = x, y, z; frames in x, y, z order;
 *  - BP texture: [angle][rows+3][cols+3] of float2 (value, next row).
 * Sizes are given in floats by kk_fp_texture_elems / kk_bp_texture_elems. */
#define KK_HALO_LO 1
#define KK_HALO_HI 2

typedef struct kk_plan_s kk_plan; /* opaque per-geometry device tables */

/* Last error message of the calling thread ("" if none). */
const char* kk_last_error(void);
/* Library version string. */
const char* kk_version(void);

/* Create/destroy a plan (replaces the NVRTC module holding __constant__
 * geometry, kernelkit/kernel.py:122-147, :188-226).  The plan binds to the
 * current HIP device. */
int kk_plan_create(kk_plan** out);
int kk_plan_destroy(kk_plan* plan);

/* RayDrivenConeFP.set_params (kernelkit/kernels/cone_fp.py:341-363).
 * params: 12 x n_angles float64, structure of arrays, in the order returned
 * by RayDrivenConeFP.geoms2params (kernels/cone_fp.py:336-339):
 * src.xyz, det_corner.xyz, U.xyz (column step), V.xyz (row step), in voxel
 * units.  ray_axes (nullable): per-angle ray axis 0/1/2, the reference's
 * argmax |source - detector centre| (kernels/cone_fp.py:264-285); when NULL
 * it is computed here from the corner plus rows/cols.  vol_shape and
 * voxel_size are world (x, y, z); they fix the texture layout and the
 * output scale (kernels/cone_fp.py:287-300). */
int kk_set_params_fp(kk_plan* plan, const double* params, int32_t n_angles,
                     const int32_t* ray_axes, int32_t rows, int32_t cols,
                     const int32_t vol_shape[3], const double voxel_size[3]);

/* Frame mask (bit a: ray axis a occurs) of the plan's FP parameters. */
int32_t kk_fp_frame_mask(const kk_plan* plan);

/* VoxelDrivenConeBP.set_params (kernelkit/kernels/cone_bp.py:168-190).
 * params: n_angles x 12 float64 as returned by VoxelDrivenConeBP.geoms2params
 * (kernels/cone_bp.py:401-421): per angle numU.xyzw, numV.xyzw, den.xyzw. */
int kk_set_params_bp(kk_plan* plan, const double* params, int32_t n_angles);

/* Texture sizes in floats (-1 on bad arguments). */
int64_t kk_fp_texture_elems(const int32_t vol_shape[3], int32_t frame_mask);
int64_t kk_bp_texture_elems(int32_t n_angles, int32_t rows, int32_t cols);

/* Volume -> FP texture (the volume setter's copy_to_texture,
 * kernelkit/projector.py:266-283).  vol_strides: element strides of world
 * x, y, z in the source (any volume_axes permutation or pitch).  Writes
 * every element of the frames in frame_mask, halo included. */
int kk_volume_to_fp_texture(const float* vol, const int32_t vol_shape[3],
                            const int64_t vol_strides[3], int32_t frame_mask,
                            float* tex, void* stream);

/* Projections -> BP texture (the projections setter, projector.py:503-556).
 * strides: element strides of (angle, row, col) in the source (any
 * projection_axes permutation or pitch). */
int kk_projections_to_bp_texture(const float* proj, int32_t n_angles,
                                 int32_t rows, int32_t cols,
                                 const int64_t strides[3], float* tex,
                                 void* stream);

/* value-fill of n floats (a kernel, not hipMemsetAsync: graph-safe). */
int kk_fill(float* ptr, int64_t n, float value, void* stream);

/* ---- device memory without PyTorch ---------------------------------- */

/* The reference's arrays are CuPy arrays (kernelkit/kernel.py:47-119,
 * kernelkit/projector.py:266-283): CuPy allocates, copies and synchronises
 * for it.  These calls give the drop-in (kernelkit.DeviceArray) and C
 * callers the same without PyTorch.  kk_memcpy is asynchronous on the
 * stream (host memory must stay valid until the stream is synchronised);
 * kk_copy_strided copies a 3-D index space of shape[0] x shape[1] x
 * shape[2] elements of elem_bytes (4 or 8) between any element strides, on
 * the device. */
#define KK_COPY_H2D 1
#define KK_COPY_D2H 2
#define KK_COPY_D2D 3
int kk_device_count(int32_t* count);
int kk_get_device(int32_t* device);
int kk_set_device(int32_t device);
int kk_malloc(void** ptr, int64_t bytes);
int kk_free(void* ptr);
int kk_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind,
              void* stream);
int kk_stream_synchronize(void* stream);
int kk_copy_strided(const void* src, const int64_t src_strides[3], void* dst,
                    const int64_t dst_strides[3], const int64_t shape[3],
                    int32_t elem_bytes, void* stream);

/* ---- streams, events, graphs (callers without PyTorch) -------------- */

/* The reference records projector calls in a CUDA graph through CuPy's
 * stream capture (example/cuda/graph.py:46-51).  These calls give the
 * drop-in and C callers streams (default priority, non-blocking), timing
 * events and stream capture without PyTorch.  kk_graph_end_capture
 * instantiates the captured work and returns an executable graph, replayed
 * with kk_graph_launch on any stream; the null stream cannot capture.
 * kk_device_info: the current device's compute-unit count and name, or
 * its architecture name where the runtime reports no marketing name
 * (name_len bytes including the terminator; either output may be null). */
int kk_device_info(int32_t* cu_count, char* name, int32_t name_len);
int kk_stream_create(void** stream);
int kk_stream_destroy(void* stream);
int kk_stream_wait_event(void* stream, void* event);
int kk_event_create(void** event);
int kk_event_destroy(void* event);
int kk_event_record(void* event, void* stream);
int kk_event_synchronize(void* event);
int kk_event_elapsed_ms(void* start, void* stop, float* ms);
int kk_graph_begin_capture(void* stream);
int kk_graph_end_capture(void* stream, void** graph_exec);
int kk_graph_launch(void* graph_exec, void* stream);
int kk_graph_destroy(void* graph_exec);

/* Ray-driven Joseph forward projection, all angles in one launch
 * (cone_fp<DIR_X|Y|Z>, kernelkit/cuda/cone_fp.cu:73-207, launched by
 * kernels/cone_fp.py:195-262).  vol_tex: FP texture holding at least the
 * plan's frames.  proj: output addressed by (angle, row, col) element strides
 * (any projection_axes or pitch).  additive != 0: proj += result; else
 * proj = result (the reference zero-fills first, projector.py:337-343). */
int kk_cone_fp(const kk_plan* plan, const float* vol_tex, float* proj,
               const int64_t proj_strides[3], int32_t rows, int32_t cols,
               int32_t additive, void* stream);

/* kk_cone_fp writing through a table of base pointers: the reference's
 * `float** projections` argument (kernelkit/cuda/cone_fp.cu:76, :199-204),
 * an int64 device array with one base address per index of the first
 * projection array axis (kernels/cone_fp.py:173, built by
 * projector.py:301-314 from a list of per-angle, per-row or per-column
 * arrays).  ptrs: device array of int64 addresses; ptr_axis: the logical
 * axis (0 angle, 1 row, 2 col) the table is indexed by; proj_strides: the
 * element strides of the other two logical axes inside every entry (the
 * ptr_axis stride is ignored), e.g. for the reference's pitched entries
 * stride(projection_axes[1]) = pitch of projection_axes[2] and
 * stride(projection_axes[2]) = 1.  Plain or additive output as kk_cone_fp.
 * Allocates nothing: the table stays valid in a captured graph. */
int kk_cone_fp_ptrs(const kk_plan* plan, const float* vol_tex,
                    const int64_t* ptrs, int32_t ptr_axis,
                    const int64_t proj_strides[3], int32_t rows,
                    int32_t cols, int32_t additive, void* stream);

/* Voxel-driven backprojection, all angles in one launch (cone_bp,
 * kernelkit/cuda/cone_bp.cu:35-159, kernels/cone_bp.py:292-302).
 * proj_tex: BP texture; n_angles must equal the parameter count.  vol:
 * output addressed by world (x, y, z) element strides; output_scale: the
 * voxel volume (kernels/cone_bp.py:294).  additive != 0: vol += result;
 * else vol = result (projector.py:586-587). */
int kk_cone_bp(const kk_plan* plan, const float* proj_tex, int32_t n_angles,
               int32_t rows, int32_t cols, float* vol,
               const int32_t vol_shape[3], const int64_t vol_strides[3],
               float output_scale, int32_t additive, void* stream);

/* Exact transpose of kk_cone_fp (matched adjoint, for gradients): every
 * (ray, slice) sample of the forward projection adds proj * scale *
 * bilinear weight to its in-volume taps, with the forward kernel's own
 * coordinate and weight arithmetic (bit-identical weights), so
 * <A x, y> = <x, A^T y> up to rounding.  No counterpart in
 * the reference, whose autograd backward uses the voxel-driven
 * backprojector (kernelkit/torch_support.py:29-33); it replaces that call
 * where an exact gradient is wanted.  proj: (angle, row, col) strides; vol:
 * world strides, shape equal to the parameters' volume shape.  additive !=
 * 0: vol += A^T proj, else vol = A^T proj.  work: caller-owned device
 * scratch of kk_cone_fp_transpose_work_elems(plan, rows, cols) floats
 * (overwritten; the weighted projections in column-major order with zero
 * rows around each column).  rows * cols < 2^31 and
 * cols * (rows + 32) < 2^30 (one angle's workspace below 4 GiB), else
 * KK_ERR_UNSUPPORTED.  The angles of each ray axis run in launches of as
 * many angles as have 192 MiB of workspace, the later ones adding to vol. */
int kk_cone_fp_transpose(const kk_plan* plan, const float* proj,
                         const int64_t proj_strides[3], int32_t rows,
                         int32_t cols, float* vol, const int32_t vol_shape[3],
                         const int64_t vol_strides[3], int32_t additive,
                         float* work, void* stream);
/* Floats of kk_cone_fp_transpose's workspace for the plan's angles and a
 * rows x cols detector (host-only; -1 on bad arguments). */
int64_t kk_cone_fp_transpose_work_elems(const kk_plan* plan, int32_t rows,
                                        int32_t cols);

/* kk_cone_fp_transpose without the caller workspace: it allocates nothing
 * (no hipMalloc, no library allocation).  No counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The workspace is a static device array of
 * kk_cone_fp_transpose_static_capacity() floats (64 Mi floats, 256 MiB) that
 * every device the library runs on holds until the process ends; problems
 * whose kk_cone_fp_transpose_work_elems is above it are rejected with
 * KK_ERR_UNSUPPORTED.  All calls on one device share that array: calls that
 * overlap on one device (other streams or host threads) race.  Arguments,
 * results and errors are otherwise those of kk_cone_fp_transpose. */
int kk_cone_fp_transpose_static(const kk_plan* plan, const float* proj,
                                const int64_t proj_strides[3], int32_t rows,
                                int32_t cols, float* vol,
                                const int32_t vol_shape[3],
                                const int64_t vol_strides[3], int32_t additive,
                                void* stream);
/* Floats of kk_cone_fp_transpose_static's workspace (host-only). */
int64_t kk_cone_fp_transpose_static_capacity(void);

/* ---- batches of independent problems (one launch per operator) ------ */

/* No counterpart in the reference, which runs one geometry per projector
 * (kernelkit/projector.py:160-183) and loops over samples in Python
 * (kernelkit/torch_support.py:19-20).  A batch records problems with the
 * arguments of kk_cone_fp, kk_cone_bp, kk_sirt_fp_residual and
 * kk_sirt_bp_update (pointers are kept, not copied: they must stay valid),
 * kk_batch_configure builds device descriptors for one tile shape
 * (fp_rows_per_thread 1/2/4 or 0 = automatic; bp_zrun 4/8/16), and
 * kk_batch_fp / kk_batch_bp run all forward / backward problems in one
 * launch of grid_blocks workgroups walking all tiles in a grid-stride loop
 * (0 = one workgroup per tile; fewer make the grid persistent).  Launches
 * allocate nothing and are graph-capturable; a SIRT iteration of the batch
 * is kk_batch_fp, kk_batch_residual, kk_batch_bp, kk_batch_volume_textures.
 * Adding a problem drops the descriptors (configure again). */
typedef struct kk_batch_s kk_batch;
int kk_batch_create(kk_batch** out);
int kk_batch_destroy(kk_batch* batch);
/* Number of forward (which = 0) or backward (which = 1) problems, of
 * residual texture builds (2) or of volume texture builds (3). */
int32_t kk_batch_size(const kk_batch* batch, int32_t which);
int kk_batch_add_fp(kk_batch* batch, const kk_plan* plan,
                    const float* vol_tex, float* proj,
                    const int64_t proj_strides[3], int32_t rows, int32_t cols,
                    int32_t additive);
int kk_batch_add_bp(kk_batch* batch, const kk_plan* plan,
                    const float* proj_tex, int32_t n_angles, int32_t rows,
                    int32_t cols, float* vol, const int32_t vol_shape[3],
                    const int64_t vol_strides[3], float output_scale,
                    int32_t additive);
int kk_batch_add_sirt_fp(kk_batch* batch, const kk_plan* plan,
                         const float* x_tex, const float* y,
                         const int64_t y_strides[3], const float* R,
                         const int64_t R_strides[3], float* res_tex,
                         int32_t n_angles, int32_t rows, int32_t cols);
int kk_batch_add_sirt_bp(kk_batch* batch, const kk_plan* plan,
                         const float* res_tex, int32_t n_angles, int32_t rows,
                         int32_t cols, float* x, const int32_t vol_shape[3],
                         const int64_t x_strides[3], const float* C,
                         const int64_t C_strides[3], const float* mask,
                         const int64_t mask_strides[3], float lo, float hi,
                         int32_t clip_lo, int32_t clip_hi, float* x_tex,
                         int32_t x_tex_mask, float output_scale);
/* Texture builds of a batch, each kind in one launch: the SIRT residual
 * texture of kk_residual_to_bp_texture ((proj - y) * R into tex), and the
 * forward-projection frames of kk_volume_to_fp_texture (vol into tex for
 * the frames in mask); same arguments, same results bit for bit. */
int kk_batch_add_residual(kk_batch* batch, const float* proj,
                          const int64_t proj_strides[3], const float* y,
                          const int64_t y_strides[3], const float* R,
                          const int64_t R_strides[3], int32_t n_angles,
                          int32_t rows, int32_t cols, float* tex);
int kk_batch_add_volume_texture(kk_batch* batch, const float* vol,
                                const int32_t vol_shape[3],
                                const int64_t vol_strides[3], int32_t mask,
                                float* tex);
int kk_batch_configure(kk_batch* batch, int32_t fp_rows_per_thread,
                       int32_t bp_zrun);
int kk_batch_fp(const kk_batch* batch, int32_t grid_blocks, void* stream);
int kk_batch_bp(const kk_batch* batch, int32_t grid_blocks, void* stream);
int kk_batch_residual(const kk_batch* batch, void* stream);
int kk_batch_volume_textures(const kk_batch* batch, void* stream);

/* ---- host-only introspection (no device calls) ---------------------- */

/* The per-angle device records that kk_set_params_fp / kk_set_params_bp
 * would upload, written to `out` (n_angles x 160 bytes for the FP table,
 * n_angles x 112 bytes for the BP table; layouts FpAngle / BpAngle of
 * csrc/kk_internal.h); frame_mask (nullable) receives the FP frame mask.
 * Used by CPU tests to check the table arithmetic. */
int kk_fp_table(const double* params, int32_t n_angles,
                const int32_t* ray_axes, int32_t rows, int32_t cols,
                const int32_t vol_shape[3], const double voxel_size[3],
                void* out, int32_t* frame_mask);
int kk_bp_table(const double* params, int32_t n_angles, void* out);

/* ---- SIRT (kernelkit/algo.py:226-368) ---------------------------------- */

/* res = (FP(x) - y) * R for every ray: the residual step of one SIRT
 * iteration (algo.py:339-344) fused into the forward projection epilogue,
 * written straight into the BP texture res_tex (whose halo must be zero).
 * y, R: (angle, row, col) element strides.  Not on the product path (nor is
 * its batch form kk_batch_add_sirt_fp): the SIRT loops use kk_cone_fp plus
 * kk_residual_to_bp_texture, which is faster on MI355X; it is kept, and
 * tested to write the same texture bit for bit, for callers that want one
 * launch instead of two. */
int kk_sirt_fp_residual(const kk_plan* plan, const float* x_tex,
                        const float* y, const int64_t y_strides[3],
                        const float* R, const int64_t R_strides[3],
                        float* res_tex, int32_t n_angles, int32_t rows,
                        int32_t cols, void* stream);

/* res = (proj - y) * R written as the BP texture res_tex (halo included):
 * the same residual step of a SIRT iteration (algo.py:339-344) when the
 * forward projection went to a plain buffer proj first (kk_cone_fp).  One
 * coalesced pass; proj, y, R: (angle, row, col) element strides.  This is
 * what algo.sirt uses per problem: the fused epilogue above reads y and R
 * one detector row per lane, which costs more than this pass. */
int kk_residual_to_bp_texture(const float* proj, const int64_t proj_strides[3],
                              const float* y, const int64_t y_strides[3],
                              const float* R, const int64_t R_strides[3],
                              int32_t n_angles, int32_t rows, int32_t cols,
                              float* res_tex, void* stream);

/* x = clip(mask * (x - C * BP(res))) (algo.py:356-362), fused into the
 * backprojection epilogue.  mask (nullable), C, x: world strides; clip_lo /
 * clip_hi select the min/max constraints.  x_tex (nullable): the FP texture
 * of x with frames x_tex_mask, refreshed with the new x so the next forward
 * projection needs no copy. */
int kk_sirt_bp_update(const kk_plan* plan, const float* res_tex,
                      int32_t n_angles, int32_t rows, int32_t cols, float* x,
                      const int32_t vol_shape[3], const int64_t x_strides[3],
                      const float* C, const int64_t C_strides[3],
                      const float* mask, const int64_t mask_strides[3],
                      float lo, float hi, int32_t clip_lo, int32_t clip_hi,
                      float* x_tex, int32_t x_tex_mask, float output_scale,
                      void* stream);

/* p = 1/p with +inf mapped to 0, n contiguous floats (algo.py:324-332). */
int kk_recip_or_zero(float* ptr, int64_t n, void* stream);

/* ---- FDK processing (kernelkit/processing.py) ------------------------- */

typedef struct kk_filter_s kk_filter; /* opaque filter tables */

int kk_filter_create(kk_filter** out);
int kk_filter_destroy(kk_filter* filter);

/* Frequency response H (float64, length n, a power of two <= 8192), the
 * output of processing._filter (processing.py:44-93). */
int kk_filter_set_response(kk_filter* filter, const double* H, int32_t n);

/* Per-angle pre-weight geometry, n_angles x 10 float64 in world units:
 * (d_min - s).xyz, (u * pixel_width).xyz, (v * pixel_height).xyz, |d - s|
 * (processing.py:170-188). */
int kk_filter_set_preweight(kk_filter* filter, const double* params,
                            int32_t n_angles);

/* In-place FDK cosine pre-weight (processing.preweight, :149-188). */
int kk_preweight(const kk_filter* filter, float* proj, int32_t n_angles,
                 int32_t rows, int32_t cols, const int64_t strides[3],
                 void* stream);

/* Row filter (processing.filter, :96-126): zero-pad each detector row to
 * the response length, multiply by H in frequency space, keep the first
 * cols values.  in/out: (angle, row, col) strides, may alias exactly (in
 * place).  fuse_preweight != 0 applies kk_preweight's factor on load. */
int kk_filter_rows(const kk_filter* filter, const float* in,
                   const int64_t in_strides[3], float* out,
                   const int64_t out_strides[3], int32_t n_angles,
                   int32_t rows, int32_t cols, int32_t fuse_preweight,
                   void* stream);

/* Same, writing the filtered projections straight into a BP texture (FDK:
 * algo.py:94-126 without the intermediate copy); the texture's halo must
 * be zero. */
int kk_filter_rows_to_bp_texture(const kk_filter* filter, const float* in,
                                 const int64_t in_strides[3], float* tex,
                                 int32_t n_angles, int32_t rows,
                                 int32_t cols, int32_t fuse_preweight,
                                 void* stream);

/* ---- measurement probe -------------------------------------------- */

/* No counterpart in the reference.  Streams float4 reads over buf (n_floats
 * floats, n_floats / 4 a power of two, 16-byte aligned) with `blocks`
 * workgroups of 256 threads, each thread reading 8 * iters float4 values
 * (addresses wrap around the buffer); writes one float per thread to out
 * (blocks * 256 floats).  A buffer that fits each XCD's L2 but not a CU's
 * L1 measures the aggregate L2 read rate (bytes = blocks * 256 * 8 * iters
 * * 16). */
int kk_probe_l2_read(const float* buf, int64_t n_floats, int32_t blocks,
                     int32_t iters, float* out, void* stream);

/* No counterpart in the reference.  The projectors' access shape: each
 * thread of `blocks` workgroups of 256 issues 8 * iters 16-byte loads at
 * 8-byte alignment (two neighbouring float2); the 64 lanes of a wave form
 * an 8 x 8 tile with lane steps sx and sy (float2 units, in [0, 2^20]), and
 * every load moves the tile to a pseudo-random base inside the first
 * 2^footprint_log2 float2 of buf (footprint_log2 in [6, 30]; n_floats >=
 * 2 * (2^footprint_log2 + 1), buf 8-byte aligned).  Writes one float per
 * thread to out (blocks * 256 floats).  Bytes read = blocks * 256 * 8 *
 * iters * 16. */
int kk_probe_gather(const float* buf, int64_t n_floats,
                    int32_t footprint_log2, int32_t sx, int32_t sy,
                    int32_t blocks, int32_t iters, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KERNELKIT_HIP_H */
