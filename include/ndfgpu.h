/*
 * ndfgpu — MI355X (gfx950) NaN-aware window selection: median and rank
 * filters over size/footprint/origin windows with reflect/constant/nearest/
 * mirror/wrap boundaries, float32 and float64, bit-exact with ndfiltrs' CPU
 * semantics.
 *
 * C ABI: plain pointers and sizes only. Host pointers are caller-owned and
 * never retained after a call returns. Error messages (err, err_len) are
 * NUL-terminated and byte-identical to the reference's texts where the
 * reference has one.
 */
#ifndef NDFGPU_H
#define NDFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1..5 mirror ndfiltrs' Error enum (src/error.rs:6-18) and its
 * Python mapping (src/pyapi/filter.rs:37-45): MEMORY -> MemoryError,
 * OVERFLOW -> OverflowError, the rest -> ValueError. DEVICE has no reference
 * counterpart (a HIP runtime failure -> RuntimeError). */
typedef int32_t ndf_status;
enum {
    NDF_OK = 0,
    NDF_ERR_EMPTY_AXIS = 1,      /* "cannot map an empty axis" */
    NDF_ERR_INVALID = 2,         /* InvalidConfiguration(message) */
    NDF_ERR_LENGTH_MISMATCH = 3, /* "input length does not match the configured shape" */
    NDF_ERR_MEMORY = 4,          /* "unable to allocate filtering workspace" */
    NDF_ERR_OVERFLOW = 5,        /* "shape or index arithmetic overflowed" */
    NDF_ERR_DEVICE = 6,          /* HIP runtime error */
    NDF_ERR_KIND = 7,            /* "unsupported window filter kind" */
    NDF_ERR_MODE = 8             /* "unsupported boundary mode" */
};

/* Boundary modes, in the order of BoundaryMode (src/boundary.rs:16-27). */
enum { NDF_MODE_REFLECT = 0, NDF_MODE_CONSTANT = 1, NDF_MODE_NEAREST = 2, NDF_MODE_MIRROR = 3, NDF_MODE_WRAP = 4 };

/* Window-filter kinds on this path (src/pyapi/filter.rs:289-301). All three
 * median kinds and both rank kinds give identical bits; the histogram kinds
 * keep their extra validation (rectangular only, support overflow). */
enum {
    NDF_KIND_MEDIAN = 0,
    NDF_KIND_MEDIAN_NETWORK = 1,
    NDF_KIND_MEDIAN_HISTOGRAM = 2,
    NDF_KIND_RANK = 3,
    NDF_KIND_RANK_HISTOGRAM = 4,
    /* Minimum / Maximum in IEEE total order (src/reduction.rs:43-46,
     * 110-119); a rectangle runs the separable extrema engine, as the
     * reference does (src/plan.rs:270-281). */
    NDF_KIND_MIN = 5,
    NDF_KIND_MAX = 6
};

/* Device engines. AUTO picks by geometry (EXTREMA for rectangular minimum /
 * maximum plans that need no retained count; STACK for box windows with
 * compiled slice networks: 2-D 3x3, 5x5, 7x7, 9x9 and 3-D 3x3x3, 3x5x5,
 * 3x7x7, float64 up to 49 taps; NETWORK for other windows of at most 160
 * taps in float32 or 64 in float64; else SLIDE where the tile can be staged,
 * else GENERIC); the others force one engine and exist so tests can check
 * engines against each other. */
enum {
    NDF_ENGINE_AUTO = 0,
    NDF_ENGINE_GENERIC = 1,
    NDF_ENGINE_SLIDE = 2,
    NDF_ENGINE_NETWORK = 3,
    NDF_ENGINE_EXTREMA = 4,
    NDF_ENGINE_STACK = 5
};

/* One-shot filter over host buffers.
 * Replaces _core.window_filter(array, kind, size, footprint, origin, mode,
 * cval, omit_nan, ignore_inf, min_count, parameter)
 * (src/pyapi/filter.rs:265-314, python/ndfiltrs/_core.pyi:34-46).
 * dtype_bytes: 4 (float32) or 8 (float64). shape/size/origin/modes have
 * `rank` entries; footprint is NULL (rectangle of `size`) or prod(size)
 * bytes, nonzero = selected, C order. `cval` is cast to the dtype.
 * omit_nan = 0 is InvalidPolicy::Propagate; otherwise Omit{ignore_inf,
 * min_count} (src/reduction.rs:6-18). `parameter` is the zero-based rank
 * for the rank kinds. */
ndf_status ndf_window_filter(int32_t dtype_bytes, const void *input, void *output, size_t rank,
                             const size_t *shape, const size_t *size, const uint8_t *footprint,
                             const ptrdiff_t *origin, const int32_t *modes, double cval, int32_t kind,
                             int32_t omit_nan, int32_t ignore_inf, size_t min_count, size_t parameter,
                             char *err, size_t err_len);

/* Robust-filter kinds (src/pyapi/robust.rs:110-124); only the percentile is
 * on this path. */
enum { NDF_ROBUST_PERCENTILE = 0 };

/* One-shot robust filter over host buffers: replaces
 * _core.robust_filter(array, size, footprint, origin, mode, cval, kind,
 * parameter, omit_nan, ignore_inf, min_count, ...) (src/pyapi/robust.rs:
 * 86-158) for kind "percentile": RobustReduction::Percentile{q, policy}
 * (src/robust.rs:145-151), the linear-interpolated percentile of the
 * retained samples (src/robust.rs:320-367), q = `parameter` in [0, 100].
 * Errors in the reference's order: "ignore_inf and min_count require NaN
 * omission", "unsupported robust filter kind", the window's, then
 * "percentile must be finite and between zero and 100" and "min_count must
 * be between one and the number of window elements". */
ndf_status ndf_robust_filter(int32_t dtype_bytes, const void *input, void *output, size_t rank, const size_t *shape,
                             const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin,
                             const int32_t *modes, double cval, int32_t kind, double parameter, int32_t omit_nan,
                             int32_t ignore_inf, size_t min_count, char *err, size_t err_len);

/* Reusable plan for one shape: replaces WindowFilter::<T>::new
 * (src/plan.rs:260-324). Validation order and messages follow
 * Window::new (src/window.rs:44-137), FilterConfig::from_window
 * (src/filter_nd.rs:32-55), validate_rectangular (src/plan.rs:154-197) and
 * validate_support (src/histogram.rs:74-88). */
typedef struct ndf_plan ndf_plan;

ndf_status ndf_plan_create(ndf_plan **plan, int32_t dtype_bytes, size_t rank, const size_t *shape,
                           const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin,
                           const int32_t *modes, double cval, int32_t kind, int32_t omit_nan,
                           int32_t ignore_inf, size_t min_count, size_t parameter, char *err,
                           size_t err_len);

/* Plan form of ndf_robust_filter: replaces RobustConfig::from_window
 * (src/robust.rs:698-722) for the percentile; apply it with the ndf_plan_*
 * calls. */
ndf_status ndf_robust_plan_create(ndf_plan **plan, int32_t dtype_bytes, size_t rank, const size_t *shape,
                                  const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin,
                                  const int32_t *modes, double cval, int32_t kind, double parameter, int32_t omit_nan,
                                  int32_t ignore_inf, size_t min_count, char *err, size_t err_len);

/* Filter host buffers of `len` elements each: replaces WindowFilter::apply_into
 * (src/plan.rs:367-427); LENGTH_MISMATCH when len differs from prod(shape).
 * Copies to the device, runs, copies back; blocking. */
ndf_status ndf_plan_apply(ndf_plan *plan, const void *input, void *output, size_t len, char *err,
                          size_t err_len);

/* ndf_plan_apply in two halves, so a caller can copy (snapshot) the input
 * while it holds its own lock and release that lock for the compute, as the
 * reference does with the GIL (src/pyapi/filter.rs:219-242).
 * ndf_plan_upload copies `len` host elements to the plan's device buffer;
 * ndf_plan_compute_download filters that buffer and copies the result into
 * `output`. */
ndf_status ndf_plan_upload(ndf_plan *plan, const void *input, size_t len, char *err, size_t err_len);
ndf_status ndf_plan_compute_download(ndf_plan *plan, void *output, size_t len, char *err, size_t err_len);

/* Filter device-resident buffers on `stream` (a hipStream_t, NULL = default
 * stream). Asynchronous: returns after the kernels are enqueued. The output
 * must not overlap the input (INVALID otherwise). No reference counterpart
 * (the reference has no device; it always writes a new array). */
ndf_status ndf_plan_apply_device(ndf_plan *plan, const void *d_input, void *d_output, size_t len,
                                 void *stream, char *err, size_t err_len);

/* No-allocation form of ndf_plan_apply_device. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc, hipFree or other allocating routine: the plan's
 * device tables and every scratch buffer of the launch (error flag, learned-
 * bracket block moments, tile features, key brackets and fallback list) live
 * in one static device array of 32 MiB per GPU, held until the process ends.
 * A plan whose tables and scratch need more returns INVALID with the sizes
 * (the learned bracket needs 0.375 bytes per pixel, so 8192 x 8192 fits).
 * The tables are copied in on every call. Every call on one GPU shares that
 * GPU's array and it keeps the last call's contents: calls of this function
 * that overlap on one GPU (on other streams or threads) race, and the caller
 * orders them. Results are bit-identical to ndf_plan_apply_device.
 * Asynchronous on `stream`. */
ndf_status ndf_plan_apply_device_static(ndf_plan *plan, const void *d_input, void *d_output, size_t len,
                                        void *stream, char *err, size_t err_len);

/* Launches of one plan are ordered: each waits for the plan's previous
 * launch (on whichever stream it ran) before it starts, because launches
 * share the plan's device scratch (error flag, bracket buffers). */

/* Waits for the plan's last launch and returns DEVICE if one of its kernels
 * reported an internal inconsistency (a rank walk that left its bounds),
 * NDF_OK otherwise. Device-resident callers use it after
 * ndf_plan_apply_device; the host-buffer calls check it themselves. No
 * reference counterpart. */
ndf_status ndf_plan_last_error(ndf_plan *plan, char *err, size_t err_len);

/* Test hook (no reference counterpart): when enabled, host-buffer launches
 * (ndf_plan_apply, ndf_plan_compute_download) first fill the plan's device
 * output with `bits`, so an output element that no kernel writes keeps a
 * pattern no filter produces (for example a non-canonical NaN payload). */
ndf_status ndf_plan_set_output_fill(ndf_plan *plan, int32_t enabled, uint64_t bits);

/* Force an engine (NDF_ENGINE_*); returns INVALID when the engine cannot run
 * this geometry. */
ndf_status ndf_plan_set_engine(ndf_plan *plan, int32_t engine, char *err, size_t err_len);

/* Writes a short JSON description of the chosen engine and tile shape. */
ndf_status ndf_plan_describe(const ndf_plan *plan, char *buffer, size_t buffer_len);

void ndf_plan_destroy(ndf_plan *plan);

/* Mask-directed median repair (fixpix(method="median")).
 * Replaces _core.fixpix_median(array, mask, size, footprint, origin, mode,
 * cval, ignore_inf, min_count) (src/pyapi/repair.rs:125-183) and
 * fixpix_median (src/repair.rs:107-126) with its repair_median
 * (src/filter_nd.rs:174-215): every output whose mask byte is nonzero is the
 * NaN-omitting median (Omit{ignore_inf, min_count}) of its window over a
 * snapshot of `input` in which every masked element is NaN; every other
 * output equals the input. `mask` has mask_len bytes (LENGTH_MISMATCH unless
 * mask_len == prod(shape)); the other arguments are those of
 * ndf_window_filter. Blocking; host buffers. */
ndf_status ndf_fixpix_median(int32_t dtype_bytes, const void *input, void *output, const uint8_t *mask,
                             size_t mask_len, size_t rank, const size_t *shape, const size_t *size,
                             const uint8_t *footprint, const ptrdiff_t *origin, const int32_t *modes, double cval,
                             int32_t ignore_inf, size_t min_count, char *err, size_t err_len);

/* ndf_fixpix_median in two halves on a median plan (omit_nan = 1), so a
 * caller can snapshot under its own lock and release it for the compute, as
 * the reference does with the GIL (src/pyapi/repair.rs:23-83).
 * ndf_fixpix_upload copies `len` elements of input and `len` mask bytes to
 * the device (the mask's nonzero positions are compacted there); the
 * caller's buffers may change once it returns. ndf_fixpix_compute_download
 * repairs the uploaded data and writes all `len` elements of the result to
 * `output`. Blocking. */
ndf_status ndf_fixpix_upload(ndf_plan *plan, const void *input, const uint8_t *mask, size_t len, char *err,
                             size_t err_len);
ndf_status ndf_fixpix_compute_download(ndf_plan *plan, void *output, size_t len, char *err, size_t err_len);

/* Repair paths: SPARSE evaluates only the listed positions (one wave each
 * for windows of up to 4096 taps, else one thread each), DENSE runs the
 * plan's engine over the whole snapshot and keeps the listed outputs, AUTO
 * picks the one with the lower estimated cost for this window and number of
 * positions. All give identical bits. */
enum { NDF_REPAIR_AUTO = 0, NDF_REPAIR_SPARSE = 1, NDF_REPAIR_DENSE = 2 };

/* Device-resident form of the repair for any plan (fixpix uses a median
 * plan with omit_nan = 1): d_output = d_input with the plan's statistic over
 * the snapshot at the `npos` flat indices in d_positions (int64). d_snapshot
 * is caller scratch of `len` elements that must not alias d_input or
 * d_output (INVALID otherwise); d_output may equal d_input (in place). Positions may come in any order;
 * one outside [0, len) gives LENGTH_MISMATCH. Blocking on `stream`. No
 * reference counterpart (the reference has no device). */
ndf_status ndf_plan_repair_device(ndf_plan *plan, const void *d_input, void *d_output, void *d_snapshot,
                                  const int64_t *d_positions, size_t npos, size_t len, int32_t path, void *stream,
                                  char *err, size_t err_len);
/* Times `iterations` repairs on one stream with HIP events (mean ms per
 * repair: snapshot, estimates and the result copy). */
ndf_status ndf_plan_time_repair_device(ndf_plan *plan, const void *d_input, void *d_output, void *d_snapshot,
                                       const int64_t *d_positions, size_t npos, size_t len, int32_t path,
                                       int32_t warmup, int32_t iterations, double *ms_per_call, char *err,
                                       size_t err_len);

/* Learned bracket (2-D images on the slide engine; no reference
 * counterpart). A tiny MLP (6 -> 32 -> 32 -> 2, ReLU, fp32; NDF_BRACKET_PARAMS
 * floats: W1[6][32], b1[32], W2[32][32], b2[32], W3[32][2], b3[2]) proposes
 * a value bracket per tile from box-sum features of its support; only
 * in-range samples are sorted (sort_len entries per pass: 1024, 2048, 4096
 * or 8192). A tile whose range holds more samples than the sort splits it,
 * and one whose window targets leave the ranges adds adjacent ranges, in
 * further passes (at most 8); a tile that needs more is recomputed by the
 * plain kernel. Needs a 2-D image and a rectangular window at most 64 wide.
 * Results are bit-identical with and without a bracket. weights = NULL
 * disables the bracket. */
#define NDF_BRACKET_PARAMS 1346
#define NDF_BRACKET_FEATURES 6
ndf_status ndf_plan_set_bracket(ndf_plan *plan, const float *weights, size_t n_weights, int32_t sort_len,
                                char *err, size_t err_len);
/* Tiles of the plan and tiles that fell back in the last launch (synchronises). */
ndf_status ndf_plan_bracket_report(ndf_plan *plan, int64_t *tiles, int64_t *fallback_tiles);
size_t ndf_plan_bracket_tiles(const ndf_plan *plan);
/* Training data for the bracket model: per-tile features (tiles x 6) of the
 * device input and per-tile targets (tiles x 2: smallest and largest result
 * of the tile, standardised by the tile's mean and std) from a device output
 * of the plain filter; weight is 0 for tiles without a result. Host arrays. */
ndf_status ndf_plan_bracket_training_data(ndf_plan *plan, const void *d_input, const void *d_output, float *features,
                                          float *targets, float *weight, char *err, size_t err_len);
/* Full-batch Adam on the pinball losses (quantile tau_lo of the lower
 * target, tau_hi of the upper) with the hand-written HIP kernels; weights
 * are updated in place; losses receives `steps` values (may be NULL). */
ndf_status ndf_mlp_train(float *weights, size_t n_weights, const float *features, const float *targets,
                         const float *weight, size_t count, int32_t steps, float learning_rate, float tau_lo,
                         float tau_hi, float *losses, char *err, size_t err_len);

/* Batches of 2-D images (no reference counterpart: the reference filters one
 * array per call, src/pyapi/filter.rs:265-314, so a caller with many cutouts
 * makes one call each). Image i has shape shapes[2i..2i+1], window
 * sizes[2i..2i+1] (rectangle, centred) and modes[2i..2i+1]; cval, kind and
 * the policy are shared. Each image is validated exactly as
 * ndf_plan_create would validate it (messages prefixed "batch image i: ").
 * The images are stored back to back in one device buffer (image i starts
 * at the sum of the earlier images' element counts; ndf_batch_length gives
 * the total). One launch runs every tile of every image through a work
 * queue, and is captured once per input/output pair in a hipGraph and
 * replayed (use_graph = 0 launches directly). */
typedef struct ndf_batch ndf_batch;
ndf_status ndf_batch_create(ndf_batch **batch, int32_t dtype_bytes, size_t count, const size_t *shapes,
                            const size_t *sizes, const int32_t *modes, double cval, int32_t kind, int32_t omit_nan,
                            int32_t ignore_inf, size_t min_count, size_t parameter, char *err, size_t err_len);
size_t ndf_batch_length(const ndf_batch *batch);
/* Tile sort length (1024 .. 16384) and cap on tile rows (multiple of 4). */
ndf_status ndf_batch_configure(ndf_batch *batch, int32_t sort_len, int32_t max_tile_rows, int32_t use_graph,
                               char *err, size_t err_len);
/* Times every (sort length, tile-row cap) candidate on these buffers
 * (`repeats` launches each) and keeps the fastest. */
ndf_status ndf_batch_autotune(ndf_batch *batch, const void *d_input, void *d_output, size_t len, int32_t repeats,
                              char *err, size_t err_len);
/* Asynchronous on `stream` (a hipStream_t, NULL = default stream). The
 * output must not overlap the input (INVALID otherwise). */
ndf_status ndf_batch_apply_device(ndf_batch *batch, const void *d_input, void *d_output, size_t len, void *stream,
                                  char *err, size_t err_len);
/* Mean milliseconds per launch over `iterations` (HIP events). */
ndf_status ndf_batch_time_device(ndf_batch *batch, const void *d_input, void *d_output, size_t len, int32_t warmup,
                                 int32_t iterations, double *ms_per_launch, char *err, size_t err_len);
/* JSON: configuration, work tiles, LDS and workgroups. */
ndf_status ndf_batch_describe(ndf_batch *batch, char *buffer, size_t buffer_len);
void ndf_batch_destroy(ndf_batch *batch);

/* Median filter of mesh-background grids: replaces the grid stage of the
 * reference's mesh background, MeshBackground::filter
 * (src/background.rs:797-855) with MeshFilterEdge (:102-111) and
 * MeshThreshold (:113-123). background / rms are float64 grids of
 * prod(mesh_shape) values (C order); every mesh takes the median of the
 * meshes in its filter_size window (NaN if one is NaN; an even count
 * averages the middle two, src/background.rs:940-963), the window shrunk to
 * stay centred (SHRINK) or truncated at the grid (TRUNCATE). ALL replaces
 * every mesh, DEVIATION meshes with |median - background| >= threshold,
 * ABOVE meshes with background > threshold; a replaced mesh takes the RMS
 * median too. A filter of 1 on every axis copies the grids. Host buffers;
 * blocking. "filter threshold must be finite" for a nonfinite threshold
 * with DEVIATION or ABOVE. */
enum { NDF_MESH_SHRINK = 0, NDF_MESH_TRUNCATE = 1 };
enum { NDF_MESH_ALL = 0, NDF_MESH_DEVIATION = 1, NDF_MESH_ABOVE = 2 };
ndf_status ndf_mesh_median_filter(const double *background, const double *rms, double *out_background,
                                  double *out_rms, size_t ndim, const size_t *mesh_shape, const size_t *filter_size,
                                  int32_t edge, int32_t threshold_kind, double threshold, char *err, size_t err_len);

/* Replaces get_num_threads (src/pyapi/filter.rs:377-380): the host thread
 * count a CPU caller would use (RAYON_NUM_THREADS, else the CPU count). */
size_t ndf_get_num_threads(void);

/* Device plumbing for benchmarks and tests (no reference counterpart). */
ndf_status ndf_device_count(int32_t *count);
ndf_status ndf_device_malloc(void **ptr, size_t bytes);
ndf_status ndf_device_free(void *ptr);
ndf_status ndf_memcpy_to_device(void *d_dst, const void *h_src, size_t bytes);
ndf_status ndf_memcpy_to_host(void *h_dst, const void *d_src, size_t bytes);
/* Device-to-device copy (a plain copy kernel; used to calibrate HBM counters). */
ndf_status ndf_memcpy_device(void *d_dst, const void *d_src, size_t bytes);
/* Fill `count` elements of `element_bytes` (4 or 8) at d_ptr with `bits`
 * (a fill kernel; blocking). */
ndf_status ndf_device_fill(void *d_ptr, size_t count, int32_t element_bytes, uint64_t bits);
ndf_status ndf_device_synchronize(void);
/* Times `iterations` launches of ndf_plan_apply_device with HIP events on
 * one stream; writes the mean milliseconds per launch (every kernel of the
 * launch) and, when ms_main_kernel is not NULL, the mean duration of the
 * launch's main kernel (the slide, bracket or generic kernel), measured with
 * events recorded around it on the same stream. */
ndf_status ndf_plan_time_device(ndf_plan *plan, const void *d_input, void *d_output, size_t len,
                                int32_t warmup, int32_t iterations, double *ms_per_launch, double *ms_main_kernel,
                                char *err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif
