/*
 * CPU oracle for the NaN-aware window-selection path (test infrastructure).
 *
 * This library is the checker, never the product: only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it. It restates
 * the reference's CPU semantics in plain C:
 *
 *   - boundary mapping            src/boundary.rs:29-71
 *   - window taps / gather        src/window.rs:44-137, 211-294
 *   - NaN/inf policy + selection  src/reduction.rs:78-130
 *   - percentile interpolation    src/robust.rs:320-367, src/convolution.rs:203-384
 *   - separable extrema           src/extrema.rs:125-436
 *   - mesh-grid median filter     src/background.rs:797-855, 940-963
 *   - total-order keys            src/median_network.rs:9-37
 *   - even-count midpoint         docs/semantics.md:74-85, tests/_reference.py:36-47
 *   - mask-directed median repair src/repair.rs:14-29, 107-126,
 *                                 src/filter_nd.rs:174-215
 *   - block rank histogram        src/histogram.rs:30-631 (used as the fast
 *                                 multi-threaded CPU baseline and as a second,
 *                                 independent algorithm for cross-checking)
 */
#ifndef NDF_ORACLE_H
#define NDF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; identical numbering to include/ndfgpu.h. */
enum {
    NDFO_OK = 0,
    NDFO_EMPTY_AXIS = 1,
    NDFO_INVALID = 2,
    NDFO_LENGTH_MISMATCH = 3,
    NDFO_MEMORY = 4,
    NDFO_OVERFLOW = 5
};

/* Boundary modes in the order of src/boundary.rs:16-27. */
enum { NDFO_REFLECT = 0, NDFO_CONSTANT = 1, NDFO_NEAREST = 2, NDFO_MIRROR = 3, NDFO_WRAP = 4 };

/* Statistics: median, fixed rank, the linear-interpolated percentile of
 * src/robust.rs:320-367, and the total-order minimum / maximum of
 * src/reduction.rs:110-119. */
enum { NDFO_STAT_MEDIAN = 0, NDFO_STAT_RANK = 1, NDFO_STAT_PERCENTILE = 2, NDFO_STAT_MIN = 3, NDFO_STAT_MAX = 4 };
enum { NDFO_METHOD_DIRECT = 0, NDFO_METHOD_HISTOGRAM = 1 };

/* Maps `index` on an axis of `length`; *mapped = -1 marks a constant
 * boundary sample. Returns NDFO_EMPTY_AXIS for length 0. */
int ndfo_map_index(int64_t index, uint64_t length, int32_t mode, int64_t *mapped);

/* Even-count midpoint of two f64 values (the reducers contract). */
double ndfo_midpoint(double lower, double upper);

/* Unsigned total-order keys and their inverses. */
uint32_t ndfo_key32(uint32_t bits);
uint64_t ndfo_key64(uint64_t bits);
uint32_t ndfo_unkey32(uint32_t key);
uint64_t ndfo_unkey64(uint64_t key);

/* Full filter. `dtype_bytes` is 4 or 8; buffers are C-ordered with
 * prod(shape) elements. `footprint` is NULL (rectangle) or prod(size) bytes.
 * statistic: NDFO_STAT_MEDIAN or NDFO_STAT_RANK (with `parameter` = rank).
 * omit_nan = 0 selects the propagate policy. method: direct or histogram.
 * nthreads <= 0 uses one thread. On error, a message is written to err. */
int ndfo_filter(int dtype_bytes, const void *input, void *output, size_t rank,
                const size_t *shape, const size_t *size, const uint8_t *footprint,
                const ptrdiff_t *origin, const int32_t *modes, double cval,
                int32_t statistic, size_t parameter, int32_t omit_nan, int32_t ignore_inf,
                size_t min_count, int32_t method, int32_t nthreads, char *err,
                size_t err_len);

/* ndfo_filter with the percentile q (0..100) of NDFO_STAT_PERCENTILE, whose
 * validation follows RobustReduction::validate (src/robust.rs:189-241). Both
 * methods serve every statistic; the histogram method needs a rectangle. */
int ndfo_filter_q(int dtype_bytes, const void *input, void *output, size_t rank,
                  const size_t *shape, const size_t *size, const uint8_t *footprint,
                  const ptrdiff_t *origin, const int32_t *modes, double cval,
                  int32_t statistic, size_t parameter, double q, int32_t omit_nan, int32_t ignore_inf,
                  size_t min_count, int32_t method, int32_t nthreads, char *err,
                  size_t err_len);

/* interpolate() of src/robust.rs:320-356. */
double ndfo_interpolate(double lower, double upper, double fraction);

/* Rectangular minimum (maximum != 0: maximum) by the reference's exact
 * separable van Herk / Gil-Werman passes (src/extrema.rs:125-436), a second
 * algorithm beside the direct NDFO_STAT_MIN / NDFO_STAT_MAX. */
int ndfo_extrema(int dtype_bytes, const void *input, void *output, size_t rank, const size_t *shape,
                 const size_t *size, const ptrdiff_t *origin, const int32_t *modes, double cval, int32_t maximum,
                 int32_t omit_nan, int32_t ignore_inf, size_t min_count, char *err, size_t err_len);

/* Median filter of a mesh background grid and its RMS grid
 * (MeshBackground::filter, src/background.rs:797-855): each mesh takes the
 * median (NaN if any sample is NaN; an even count averages the middle two)
 * of its window, shrunk to stay centred (edge 0) or truncated at the grid
 * (edge 1); threshold_kind 0 replaces every mesh, 1 meshes whose median
 * deviates by at least `threshold`, 2 meshes whose background is above it. */
int ndfo_mesh_filter(const double *background, const double *rms, double *out_background, double *out_rms,
                     size_t ndim, const size_t *mesh_shape, const size_t *filter_size, int32_t edge,
                     int32_t threshold_kind, double threshold, char *err, size_t err_len);

/* Mask-directed median repair (fixpix(method="median")): outputs whose
 * mask byte is nonzero become the Omit median of their window over the input
 * with every masked element set to NaN; the rest copy the input. */
int ndfo_fixpix_median(int dtype_bytes, const void *input, void *output, const uint8_t *mask, size_t mask_len,
                       size_t rank, const size_t *shape, const size_t *size, const uint8_t *footprint,
                       const ptrdiff_t *origin, const int32_t *modes, double cval, int32_t ignore_inf,
                       size_t min_count, int32_t nthreads, char *err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif
