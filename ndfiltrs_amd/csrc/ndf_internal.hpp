// Internal interfaces between the host planner (ndf_host.hip) and the device
// engines (ndf_generic.hip, ndf_slide.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ndf {

constexpr int MAX_RANK = 32;

// Per-axis boundary maps hold premultiplied input offsets:
//   >= 0  offset of the mapped coordinate (index * stride)
//   MAP_CVAL   the sample is cval (constant mode, outside the axis)
//   MAP_PAD    the coordinate only feeds outputs beyond the array (tile
//              padding); the sample is omitted and never reaches a result.
constexpr int64_t MAP_CVAL = -1;
constexpr int64_t MAP_PAD = -2;

// Engine 1: one thread per output, any rank and footprint. Samples are
// gathered from global memory through the maps; selection is a bitwise
// binary search over the key (count of keys below a candidate, one bit per
// pass). Used for ranks above 3, geometries the tiled engine cannot stage,
// and as an independent cross-check of the tiled engine.
struct GenericArgs {
    const void *input;
    void *output;
    int64_t len;
    int32_t rank;
    int32_t count;
    int64_t stride[MAX_RANK];   // output/input strides in elements (C order)
    int64_t map_base[MAX_RANK]; // start of axis a's map in `maps`
    const int64_t *maps;        // entry (map_base[a] + coord + rel) for tap rel
    const int32_t *taps;        // count * rank entries: tap offset - low offset
    uint64_t cval_bits;
    int32_t ignore_inf;
    int32_t stat;
    int64_t min_count;
    int64_t rank_param;
    double q100;  // percentile / 100 (STAT_PERCENTILE)
    // Sparse mode (mask-directed repair): when `positions` is set, `len`
    // counts positions and thread i computes the output at flat index
    // positions[i] of an array of `extent` elements. A position outside
    // [0, extent) is skipped and sets *error_flag.
    const int64_t *positions;
    int64_t extent;
    int32_t *error_flag;
    // Rank-2 rectangular windows: rows and columns of the box (0 otherwise);
    // tap t is then (t / rect_kx, t % rect_kx) relative to the box corner.
    int32_t rect_ky, rect_kx;
};

// Engine 2: LDS-staged tile, block sort of the support, per-lane column slide
// with a rank walk. Canonical 3-D geometry (nz planes of ny x nx); a tile is
// 64 output columns x ty rows of one plane; wave w of the block slides down
// rows [w*tw, (w+1)*tw). Support positions are packed u16 (row << 8 | x).
struct SlideArgs {
    const void *input;
    void *output;
    int64_t nz, ny, nx;
    int64_t in_sz, in_sy;  // input strides of z and y (x stride is 1)
    const int64_t *mapz;   // index z0 + z, z in [0, kz)
    const int64_t *mapy;   // index y0 + y, y in [0, sy)
    const int64_t *mapx;   // index x0 + x, x in [0, sx)
    int32_t kz, ky, kx;    // window bounding box
    int32_t sx, sy;        // support row length (64 + kx - 1) and rows per plane (ty + ky - 1)
    int32_t zrows, zshift; // row stride between planes (a power of two 1 << zshift when kz > 1)
    int32_t rows, s;       // support rows ((kz - 1) * zrows + sy, <= 254) and s = rows * sx
    int32_t sort_e;        // sort elements per thread; sort length 256 * sort_e >= s
    int32_t hist_stride;   // u16 entries of one wave's histogram region
    int32_t ty, tw;        // tile rows and rows per wave
    int32_t tiles_x, tiles_y;
    int32_t count;         // taps in the window
    const uint32_t *taps;  // count support offsets (dz*zrows + dy)*sx + dx (footprints)
    int32_t nedge;
    const uint32_t *edge_remove;  // offsets relative to row y-1
    const uint32_t *edge_add;     // offsets relative to row y
    const uint8_t *mask;          // kz*ky*kx footprint, nullptr for a rectangle
    uint64_t cval_bits;
    int32_t ignore_inf;
    int32_t stat;
    int32_t min_count;
    int32_t rank_param;
    double q100;          // percentile / 100 (STAT_PERCENTILE)
    // Rectangles with a high min_count: count every window's retained
    // samples before the sort and skip it when no window has a result.
    int32_t count_gate;
    int32_t *error_flag;  // set when a rank walk leaves its bounds (plain mode)
    // Tile selection: tiles 0 .. tiles_x*tiles_y*nz - 1, or tile_list[0 .. *tile_count).
    const int32_t *tile_list;
    const int32_t *tile_count;
    // Bracket mode: per-tile key bracket [lo, hi]; a tile whose targets leave
    // it is appended to fallback_tiles (count in fallback_count).
    const void *bracket_lo;
    const void *bracket_hi;
    int32_t *fallback_tiles;
    int32_t *fallback_count;
    // Bracket mode: BRACKET_QUEUES tile counters (zero at launch), one per
    // XCD-contiguous chunk of tiles, from which the persistent workgroups take
    // their tiles.
    int32_t *tile_queue;
};

constexpr int BRACKET_QUEUES = 8;

// Engine 3: small windows (count <= 160 for f32, 64 for f64). A tile of 64
// columns x ty rows of one plane; its support (kz planes of sy = ty + ky - 1
// rows of sx = 64 + kx - 1 samples) is staged in LDS as classified keys; each
// thread sorts its window in registers with a sorting network of `live`
// wires and reads the statistic. taps[t] = (dz * sy + dy) * sx + dx.
struct NetworkArgs {
    const void *input;
    void *output;
    int64_t nz, ny, nx;
    const int64_t *mapz, *mapy, *mapx;
    int32_t sx, sy, rows, s;
    int32_t ty, tiles_x, tiles_y;
    int32_t count, live;
    // Two-row step (f32): the windows at rows r and r + 1 share pair_c core
    // taps and have pair_e taps each beyond them; 0 for one output per step.
    int32_t pair_c, pair_e;
    int32_t pad;  // sentinel keys after the support (taps past count read there)
    const uint32_t *taps;
    uint64_t cval_bits;
    int32_t ignore_inf;
    int32_t stat;
    int32_t min_count;
    int32_t rank_param;
    double q100;  // percentile / 100 (STAT_PERCENTILE)
};

// Engine 5: box windows of K0 slices along their slowest axis (3-D: kz
// planes of ky x kx; 2-D: ky rows of 1 x kx), one sorted slice built per
// output step and K0 sorted slices merged per output (ndf_stack.hip). A wave
// owns a unit: 64 columns of one row of the row axis and a run of `run`
// positions along the slide axis. Map entry i of an axis is coordinate
// i + low of its window (the row axis of a 2-D window has no window: its
// entry r is row r's offset).
struct StackArgs {
    const void *input;
    void *output;
    const int64_t *maps;  // slide axis: entries [0, ns + k0 - 1)
    const int64_t *mapr;  // row axis: entries [0, nr + sh - 1)
    const int64_t *mapx;  // columns: entries [0, tiles_x * 64 + sw - 1)
    int64_t ns, nr, nx;
    int64_t out_ss, out_sr;  // output strides of the slide and row axes
    int64_t units;           // tiles_x * nr * ceil(ns / run)
    int32_t k0, sh, sw;
    int32_t sx;              // 64 + sw - 1 staged columns
    int32_t tiles_x;
    int32_t run;
    uint64_t cval_bits;
    int32_t ignore_inf;
    int32_t stat;
    int32_t min_count;
    int32_t rank_param;
    double q100;  // percentile / 100 (STAT_PERCENTILE)
};

// Whether a (k0, sh, sw) box has compiled networks for this dtype.
bool stack_supported(int dtype_bytes, int k0, int sh, int sw);
size_t stack_lds_bytes(int dtype_bytes, int sh, int sx);
int64_t stack_max_units();
hipError_t launch_stack(int dtype_bytes, const StackArgs &args, hipStream_t stream);
// Batched stack kernel for images sharing the box (k0, sh, sw) and the
// statistic of `shape`: `images` one StackArgs per image (input / output
// unused), `work` nwork (image, block of 4 units) pairs taken through
// `counter`.
hipError_t stack_batch_grid(int dtype_bytes, const StackArgs &shape, int *grid);
hipError_t launch_stack_batch(int dtype_bytes, const StackArgs &shape, const StackArgs *images,
                              const int64_t *offsets, const int32_t *work, int nwork, int32_t *counter,
                              const void *input, void *output, int grid, hipStream_t stream);

// Engine 4: rectangular minimum / maximum (rank <= 3) by separable passes
// over an LDS tile of tx columns x ty rows of one plane; its support is kz
// planes of sy = ty + ky - 1 rows of sx = tx + kx - 1 samples. Maps as for
// the slide engine (index i <-> coordinate i + low of the window).
struct ExtremaArgs {
    const void *input;
    void *output;
    int64_t nz, ny, nx;
    const int64_t *mapz, *mapy, *mapx;
    int32_t kz, ky, kx;
    int32_t tx, ty, sx, sy;
    int32_t tiles_x, tiles_y;
    uint64_t cval_bits;
    int32_t ignore_inf;
    int32_t omit;     // 0: a NaN in the window gives NaN (propagate)
    int32_t maximum;  // 0: minimum
};

size_t extrema_lds_bytes(int dtype_bytes, int kz, int sy, int sx, int tx);
int extrema_max_support_columns();
hipError_t launch_extrema(int dtype_bytes, const ExtremaArgs &args, size_t lds_bytes, hipStream_t stream);

// Network size (live wires) for a window of `count` taps, 0 if too large.
int network_live(int dtype_bytes, int count);
// Two-row step class (*c, *e) for windows whose rows r and r + 1 share
// `core` taps and have `extra` taps each beyond them: true when a compiled
// class holds them and costs fewer comparators per output than the
// one-output network of `live` wires (median or full statistics).
bool network_pair_class(int dtype_bytes, int core, int extra, int live, bool median, int *c, int *e);
size_t network_lds_bytes(int dtype_bytes, int s, int pad, int live);
hipError_t launch_network(int dtype_bytes, const NetworkArgs &args, size_t lds_bytes, hipStream_t stream);
// Batched network kernel for images whose windows have at most `live` taps:
// `images` one NetworkArgs per image (input/output unused; taps in global
// memory), `work` nwork (image, tile) pairs taken through `counter`.
hipError_t network_batch_grid(int dtype_bytes, int live, size_t lds_bytes, int *grid);
hipError_t launch_network_batch(int dtype_bytes, int live, const NetworkArgs *images, const int64_t *offsets,
                                const int32_t *work, int nwork, int32_t *counter, const void *input, void *output,
                                size_t lds_bytes, int grid, hipStream_t stream);

constexpr int SLIDE_WAVES = 4;
constexpr int SLIDE_TILE_X = 64;
constexpr uint32_t SLIDE_NO_POS = 0xFFFFFFFFu;

hipError_t launch_generic(int dtype_bytes, const GenericArgs &args, hipStream_t stream);
// Sparse repair with one wave per position (args.positions set): the window's
// retained keys are staged in LDS, so args.count <= REPAIR_WAVE_MAX_TAPS.
constexpr int REPAIR_WAVE_MAX_TAPS = 4096;
size_t repair_wave_lds_bytes(int dtype_bytes, int count);
hipError_t launch_repair_wave(int dtype_bytes, const GenericArgs &args, hipStream_t stream);
hipError_t launch_slide(int dtype_bytes, const SlideArgs &args, size_t lds_bytes, int max_tiles, hipStream_t stream);
hipError_t launch_fill(int dtype_bytes, void *output, int64_t len, uint64_t bits, hipStream_t stream);
// Nonzero bytes of a mask (added to *count), and their flat indices
// (written from *cursor on, in no particular order). The mask is 16-byte
// aligned.
hipError_t launch_mask_count(const uint8_t *mask, int64_t len, unsigned long long *count, hipStream_t stream);
hipError_t launch_mask_positions(const uint8_t *mask, int64_t len, int64_t *positions, unsigned long long *cursor,
                                 hipStream_t stream);
// out[positions[i]] = bits for i < n; positions outside [0, extent) are
// skipped and set *error_flag.
hipError_t launch_scatter_fill(int dtype_bytes, void *output, const int64_t *positions, int64_t n, int64_t extent,
                               uint64_t bits, int32_t *error_flag, hipStream_t stream);
// dst[positions[i]] = src[positions[i]] for i < n (same bounds rule).
hipError_t launch_scatter_copy(int dtype_bytes, void *dst, const void *src, const int64_t *positions, int64_t n,
                               int64_t extent, int32_t *error_flag, hipStream_t stream);
size_t slide_lds_bytes(int dtype_bytes, int s, int sort_e, int hist_stride, int mask_bytes);
// Bracket kernel (2-D rectangles, support rows <= 128 samples): ranks only
// the samples inside each tile's key bracket; tiles it cannot finish are
// appended to args.fallback_tiles for the plain kernel.
hipError_t launch_bracket_slide(int dtype_bytes, const SlideArgs &args, size_t lds_bytes, int max_tiles,
                                hipStream_t stream);
size_t bracket_lds_bytes(int dtype_bytes, int s, int rows, int sort_e);
// Batched plain kernel (2-D rectangles): `images` holds one SlideArgs per
// image (input/output fields unused), `offsets` its first element in the
// batch buffers, `work` nwork (image, tile) pairs taken in order through
// `counter` (zero before the launch). `grid` from slide_batch_grid.
hipError_t slide_batch_grid(int dtype_bytes, int sort_e, size_t lds_bytes, int *grid);
hipError_t launch_slide_batch(int dtype_bytes, int sort_e, const SlideArgs *images, const int64_t *offsets,
                              const int32_t *work, int nwork, int32_t *counter, const void *input, void *output,
                              int32_t *error_flag, size_t lds_bytes, int grid, hipStream_t stream);

// Pinned double buffers for large host <-> device copies (ndf_transfer.hip).
constexpr size_t STAGING_CHUNK = (size_t)16 << 20;
constexpr size_t STAGING_MIN_BYTES = (size_t)8 << 20;
struct Staging {
    void *pinned[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    Staging() = default;
    Staging(const Staging &) = delete;
    Staging &operator=(const Staging &) = delete;
    ~Staging();
    hipError_t ensure();
};
// Copies host -> device; returns once the host buffer may change.
hipError_t staged_upload(Staging &s, void *d_dst, const void *h_src, size_t bytes, hipStream_t stream);
// Copies device -> host after the work already enqueued on `stream`; blocking.
hipError_t staged_download(Staging &s, void *h_dst, const void *d_src, size_t bytes, hipStream_t stream);

}  // namespace ndf
