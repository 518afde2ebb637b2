// Host side of libndfgpu: validation, planning, device tables and the C ABI.
//
// Validation restates, in order and with the same messages:
//   Window::new                 src/window.rs:44-137
//   FilterConfig::from_window   src/filter_nd.rs:32-55
//   WindowFilter::new           src/plan.rs:260-324 (histogram kinds:
//                               validate_rectangular :154-197 and
//                               validate_support src/histogram.rs:74-88)
//   _core.window_filter         src/pyapi/filter.rs:265-314 (kind dispatch)
#include <hip/hip_runtime.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ndfgpu.h"
#include "ndf_bracket.hpp"
#include "ndf_device.hpp"
#include "ndf_internal.hpp"

namespace {

using namespace ndf;

const char *const MSG_RANK0 = "filtering requires at least one dimension";
const char *const MSG_SIZE0 = "window sizes must be positive";
const char *const MSG_ORIGIN = "origin must satisfy -(size // 2) <= origin <= (size - 1) // 2";
const char *const MSG_EMPTY_FOOTPRINT = "footprint must select at least one element";
const char *const MSG_MIN_COUNT = "min_count must be between 1 and the number of window elements";
const char *const MSG_HIST_FOOTPRINT = "histogram reductions require rectangular size, not footprint";
const char *const MSG_EMPTY_AXIS = "cannot map an empty axis";
const char *const MSG_LENGTH = "input length does not match the configured shape";
const char *const MSG_MEMORY = "unable to allocate filtering workspace";
const char *const MSG_OVERFLOW = "shape or index arithmetic overflowed";
const char *const MSG_KIND = "unsupported window filter kind";
const char *const MSG_MODE = "unsupported boundary mode";

ndf_status fail(ndf_status code, const char *message, char *err, size_t err_len)
{
    if (err && err_len) {
        snprintf(err, err_len, "%s", message);
    }
    return code;
}

ndf_status fail_code(ndf_status code, char *err, size_t err_len)
{
    switch (code) {
    case NDF_ERR_EMPTY_AXIS: return fail(code, MSG_EMPTY_AXIS, err, err_len);
    case NDF_ERR_LENGTH_MISMATCH: return fail(code, MSG_LENGTH, err, err_len);
    case NDF_ERR_MEMORY: return fail(code, MSG_MEMORY, err, err_len);
    case NDF_ERR_OVERFLOW: return fail(code, MSG_OVERFLOW, err, err_len);
    case NDF_ERR_KIND: return fail(code, MSG_KIND, err, err_len);
    case NDF_ERR_MODE: return fail(code, MSG_MODE, err, err_len);
    default: return fail(code, "invalid configuration", err, err_len);
    }
}

ndf_status fail_hip(hipError_t e, char *err, size_t err_len)
{
    char buffer[256];
    snprintf(buffer, sizeof(buffer), "HIP error %d: %s", (int)e, hipGetErrorString(e));
    return fail(NDF_ERR_DEVICE, buffer, err, err_len);
}

bool checked_mul(size_t a, size_t b, size_t *out)
{
    if (a != 0 && b > SIZE_MAX / a) {
        return false;
    }
    *out = a * b;
    return true;
}

// map_index (src/boundary.rs:29-71); returns -1 for a constant-mode sample.
int64_t map_coordinate(int64_t index, int64_t length, int32_t mode)
{
    if (index >= 0 && index < length) {
        return index;
    }
    auto rem = [](int64_t v, int64_t m) {
        const int64_t r = v % m;
        return r < 0 ? r + m : r;
    };
    switch (mode) {
    case NDF_MODE_CONSTANT: return -1;
    case NDF_MODE_NEAREST: return index < 0 ? 0 : length - 1;
    case NDF_MODE_WRAP: return rem(index, length);
    case NDF_MODE_REFLECT: {
        if (length == 1) return 0;
        const int64_t period = 2 * length;
        const int64_t r = rem(index, period);
        return r < length ? r : period - 1 - r;
    }
    default: {  // mirror
        if (length == 1) return 0;
        const int64_t period = 2 * (length - 1);
        const int64_t r = rem(index, period);
        return r < length ? r : period - r;
    }
    }
}

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need)
    {
        if (need <= bytes && ptr) return hipSuccess;
        if (ptr) {
            (void)hipFree(ptr);
            ptr = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&ptr, need < 16 ? 16 : need);
        if (e == hipSuccess) bytes = need < 16 ? 16 : need;
        return e;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

template <typename V>
hipError_t upload(DeviceBuffer &buffer, const std::vector<V> &values)
{
    hipError_t e = buffer.ensure(values.size() * sizeof(V));
    if (e != hipSuccess || values.empty()) return e;
    return hipMemcpy(buffer.ptr, values.data(), values.size() * sizeof(V), hipMemcpyHostToDevice);
}

// LDS per workgroup on gfx950 (160 KiB); a present device may report less.
constexpr int GFX950_LDS_BYTES = 163840;

int lds_limit()
{
    static int limit = -1;
    if (limit < 0) {
        int count = 0;
        int device = 0;
        int value = 0;
        limit = GFX950_LDS_BYTES;
        if (hipGetDeviceCount(&count) == hipSuccess && count > 0 && hipGetDevice(&device) == hipSuccess &&
            hipDeviceGetAttribute(&value, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
            value > 0 && value < limit) {
            limit = value;
        }
    }
    return limit;
}

}  // namespace

// Plan kind of the robust percentile (ndf_robust_plan_create); not a
// window-filter kind of the public ABI.
constexpr int32_t KIND_PERCENTILE = 64;

struct ndf_plan {
    int32_t dbytes = 4;
    size_t rank = 0;
    std::vector<size_t> shape, size;
    std::vector<uint8_t> footprint;  // empty = rectangle
    std::vector<ptrdiff_t> origin;
    std::vector<int32_t> modes;
    double cval = 0.0;
    uint64_t cval_bits = 0;
    int32_t kind = 0;
    bool omit = false, ignore_inf = false;
    size_t min_count = 1, parameter = 0;
    int32_t stat = STAT_MEDIAN;
    double q100 = 0.0;  // STAT_PERCENTILE: percentile / 100 (0 min, 1 max)
    size_t len = 0, count = 0;
    std::vector<int64_t> taps;  // count * rank offsets
    std::vector<int64_t> low, high;
    // Effective policy: propagate == omit{ignore_inf=false, min_count=count}.
    bool eff_ignore_inf = false;
    int64_t eff_min_count = 1;
    bool all_nan = false;  // rank >= count: every output is NaN
    int32_t engine = NDF_ENGINE_GENERIC;
    bool slide_ok = false;

    // Network geometry (host side).
    bool network_ok = false;
    NetworkArgs network{};
    size_t network_lds = 0;
    std::vector<int64_t> net_mapz, net_mapy, net_mapx;
    std::vector<uint32_t> net_taps;       // one output per step: `live` entries
    std::vector<uint32_t> net_pair_taps;  // two-row step: core, extra of row r, extra of row r + 1

    // Stack geometry (box windows with compiled slice networks).
    bool stack_ok = false;
    StackArgs stack{};
    std::vector<int64_t> stk_maps, stk_mapr, stk_mapx;
    DeviceBuffer k_maps, k_mapr, k_mapx;

    // Extrema geometry (rectangular minimum / maximum).
    bool extrema_ok = false;
    ExtremaArgs extrema{};
    size_t extrema_lds = 0;
    std::vector<int64_t> ext_mapz, ext_mapy, ext_mapx;
    DeviceBuffer e_mapz, e_mapy, e_mapx;

    // Slide geometry (host side).
    SlideArgs slide{};
    size_t slide_lds = 0;
    std::vector<int64_t> slide_mapz, slide_mapy, slide_mapx;
    std::vector<uint32_t> slide_taps, slide_rm, slide_add;
    std::vector<uint8_t> slide_mask;

    // Bracket mode (2-D images on the slide engine).
    bool bracket_on = false;
    bool bracket_ready = false;
    bool bracket_launched = false;  // the last launch took the bracket path
    int bracket_e = 16;
    int bracket_ty = 0;  // tile rows of the bracket kernel (0: no bracket for this plan)
    SlideArgs slide_fb{};  // plain kernel on the bracket tiles (fallback)
    size_t slide_fb_lds = 0;
    std::vector<float> bracket_weights;
    SlideArgs slide_b{};
    size_t slide_b_lds = 0;
    BracketArgs bargs{};
    DeviceBuffer b_weights, b_blocks, b_features, b_moments, b_lo, b_hi, b_pred, b_fallback, b_fcount;

    // When set, launch() records these events around its main kernel (the
    // slide, bracket or generic kernel) for ndf_plan_time_device.
    hipEvent_t main_begin = nullptr, main_end = nullptr;

    // Launch order: every launch waits for the plan's previous launch (on
    // whatever stream that ran) and records `done` at its end, so launches
    // of one plan never overlap on the plan's scratch (error flag, bracket
    // buffers, fallback list, repair flag).
    hipEvent_t done = nullptr;
    bool done_recorded = false;

    // Test hook: host-buffer launches first fill the device output with
    // these bits, so an output a kernel fails to write keeps them.
    bool output_fill = false;
    uint64_t output_fill_bits = 0;

    // Device state.
    std::mutex mutex;
    bool tables_ready = false;
    bool generic_ready = false;
    bool uploaded = false;
    DeviceBuffer g_maps, g_taps, s_mapz, s_mapy, s_mapx, s_taps, s_rm, s_add, s_mask, d_in, d_out, d_error;
    DeviceBuffer r_flag, r_snap, r_pos;  // mask-directed repair: position flag, snapshot, positions
    DeviceBuffer r_mask, r_count;        // host-buffer repair: uploaded mask, position counter
    size_t fix_npos = 0;                 // positions of the uploaded mask
    bool fix_uploaded = false;
    Staging staging;                     // pinned buffers of the host-buffer calls
    DeviceBuffer n_mapz, n_mapy, n_mapx, n_taps;
    GenericArgs generic{};
    // Host copies of the generic engine's tables (maps, tap offsets).
    bool generic_host_ready = false;
    std::vector<int64_t> gen_maps;
    std::vector<int32_t> gen_taps;
    // Fallback counter of the last bracket launch (plan or static scratch).
    const int32_t *report_fcount = nullptr;

    ~ndf_plan()
    {
        if (done) (void)hipEventDestroy(done);
        for (DeviceBuffer *b : {&n_mapz, &n_mapy, &n_mapx, &n_taps, &e_mapz, &e_mapy, &e_mapx, &k_maps, &k_mapr, &k_mapx}) {
            b->release();
        }
        for (DeviceBuffer *b : {&g_maps, &g_taps, &s_mapz, &s_mapy, &s_mapx, &s_taps, &s_rm, &s_add, &s_mask, &d_in,
                                &d_out, &d_error, &r_flag, &r_snap, &r_pos, &r_mask, &r_count, &b_weights, &b_blocks, &b_features, &b_moments, &b_lo, &b_hi,
                                &b_pred, &b_fallback, &b_fcount}) {
            b->release();
        }
    }
};

namespace {

// Slide arguments of the same 2-D plan (one plane, kz = 1) with tiles of
// `ty` rows.
SlideArgs with_tile_rows(const SlideArgs &s, int ty)
{
    SlideArgs r = s;
    r.ty = ty;
    r.tw = ty / SLIDE_WAVES;
    r.sy = ty + s.ky - 1;
    r.zrows = r.sy;
    r.rows = r.sy;
    r.s = r.rows * s.sx;
    int e = 4;
    while (256 * e < r.s) e <<= 1;
    r.sort_e = e;
    r.tiles_y = (int32_t)((s.ny + ty - 1) / ty);
    return r;
}

// Plan the network engine (small windows); false when it does not apply.
bool plan_network(ndf_plan *p)
{
    if (p->rank > 3 || p->len == 0) return false;
    const int live = network_live(p->dbytes, (int)p->count);
    if (live == 0) return false;
    int64_t dims[3] = {1, 1, 1}, lo[3] = {0, 0, 0}, box[3] = {1, 1, 1};
    int32_t modes[3] = {NDF_MODE_NEAREST, NDF_MODE_NEAREST, NDF_MODE_NEAREST};
    const size_t pad = 3 - p->rank;
    for (size_t a = 0; a < p->rank; ++a) {
        dims[pad + a] = (int64_t)p->shape[a];
        lo[pad + a] = p->low[a];
        box[pad + a] = p->high[a] - p->low[a] + 1;
        modes[pad + a] = p->modes[a];
    }
    const int64_t kz = box[0], ky = box[1], kx = box[2];
    const int64_t sx = 64 + kx - 1;
    // Tile rows: 32 when the staged support stays within 32 KB, else 16, 8, 4.
    int64_t ty = 32;
    while (ty > 4 && kz * (ty + ky - 1) * sx * p->dbytes > 32768) ty >>= 1;
    const int64_t sy = ty + ky - 1;
    const int64_t rows = kz * sy;
    const int64_t s = rows * sx;
    // Sentinel run read by the taps past the count: window bases span
    // (ty - 1) * sx + 64 keys.
    int64_t pad_keys = (int64_t)p->count < live ? (ty - 1) * sx + 64 : 0;
    const size_t lds = network_lds_bytes(p->dbytes, (int)s, (int)pad_keys, live);
    if (lds > (size_t)lds_limit() || s > (1 << 24)) return false;
    const int64_t tiles_x = (dims[2] + 63) / 64;
    const int64_t tiles_y = (dims[1] + ty - 1) / ty;
    if (tiles_x * tiles_y * dims[0] > 0x7FFFFFFF) return false;
    const int64_t in_sy = dims[2], in_sz = dims[1] * dims[2];
    auto build = [](std::vector<int64_t> &map, int64_t length, int64_t n, int64_t low, int64_t k, int32_t mode,
                    int64_t stride) {
        map.assign((size_t)length, MAP_PAD);
        for (int64_t i = 0; i < length && i < n + k - 1; ++i) {
            const int64_t m = map_coordinate(i + low, n, mode);
            map[(size_t)i] = m < 0 ? MAP_CVAL : m * stride;
        }
    };
    build(p->net_mapz, dims[0] + kz - 1, dims[0], lo[0], kz, modes[0], in_sz);
    build(p->net_mapy, tiles_y * ty + ky - 1, dims[1], lo[1], ky, modes[1], in_sy);
    build(p->net_mapx, tiles_x * 64 + kx - 1, dims[2], lo[2], kx, modes[2], 1);
    p->net_taps.clear();
    for (size_t t = 0; t < p->count; ++t) {
        int64_t d[3] = {0, 0, 0};
        for (size_t a = 0; a < p->rank; ++a) d[pad + a] = p->taps[t * p->rank + a] - p->low[a];
        p->net_taps.push_back((uint32_t)((d[0] * sy + d[1]) * sx + d[2]));
    }
    // The kernels read `live` taps; those past the count read the sentinel
    // run that starts at offset s.
    p->net_taps.resize((size_t)live, (uint32_t)s);
    // Two-row step: rows r and r + 1 of a tile share the taps of
    // T ∩ (T + one row) (the core); row r has T \ (T + one row) of its own and
    // row r + 1 (T + one row) \ T, all as offsets from row r.
    p->net_pair_taps.clear();
    int pair_c = 0, pair_e = 0;
    {
        std::vector<std::array<int64_t, 3>> rel;
        for (size_t t = 0; t < p->count; ++t) {
            std::array<int64_t, 3> d = {0, 0, 0};
            for (size_t a = 0; a < p->rank; ++a) d[pad + a] = p->taps[t * p->rank + a] - p->low[a];
            rel.push_back(d);
        }
        std::sort(rel.begin(), rel.end());
        auto has = [&](const std::array<int64_t, 3> &d) { return std::binary_search(rel.begin(), rel.end(), d); };
        std::vector<uint32_t> core, extra0, extra1;
        auto offset = [&](int64_t dz, int64_t dy, int64_t dx) { return (uint32_t)((dz * sy + dy) * sx + dx); };
        for (const auto &d : rel) {
            const std::array<int64_t, 3> up = {d[0], d[1] - 1, d[2]};  // the same sample seen from row r + 1
            if (d[1] >= 1 && has(up)) {
                core.push_back(offset(d[0], d[1], d[2]));
            } else {
                extra0.push_back(offset(d[0], d[1], d[2]));
            }
            const std::array<int64_t, 3> down = {d[0], d[1] + 1, d[2]};
            if (!has(down)) extra1.push_back(offset(d[0], d[1] + 1, d[2]));
        }
        if (extra0.size() == extra1.size() &&
            network_pair_class(p->dbytes, (int)core.size(), (int)extra0.size(), live, p->stat == STAT_MEDIAN,
                               &pair_c, &pair_e)) {
            core.resize((size_t)pair_c, (uint32_t)s);
            extra0.resize((size_t)pair_e, (uint32_t)s);
            extra1.resize((size_t)pair_e, (uint32_t)s);
            // Padded entries read the sentinel run after the support.
            pad_keys = (ty - 1) * sx + 64;
            p->net_pair_taps = core;
            p->net_pair_taps.insert(p->net_pair_taps.end(), extra0.begin(), extra0.end());
            p->net_pair_taps.insert(p->net_pair_taps.end(), extra1.begin(), extra1.end());
        } else {
            pair_c = pair_e = 0;
        }
    }
    NetworkArgs &n = p->network;
    n = NetworkArgs{};
    n.nz = dims[0];
    n.ny = dims[1];
    n.nx = dims[2];
    n.sx = (int32_t)sx;
    n.sy = (int32_t)sy;
    n.rows = (int32_t)rows;
    n.s = (int32_t)s;
    n.ty = (int32_t)ty;
    n.tiles_x = (int32_t)tiles_x;
    n.tiles_y = (int32_t)tiles_y;
    n.count = (int32_t)p->count;
    n.live = live;
    n.pair_c = pair_c;
    n.pair_e = pair_e;
    n.pad = (int32_t)pad_keys;
    n.cval_bits = p->cval_bits;
    n.ignore_inf = p->eff_ignore_inf;
    n.stat = p->stat;
    n.min_count = (int32_t)p->eff_min_count;
    n.rank_param = (int32_t)std::min(p->parameter, p->count);
    n.q100 = p->q100;
    p->network_lds = pair_c ? network_lds_bytes(p->dbytes, (int)s, (int)pad_keys, pair_c + 2 * pair_e) : lds;
    if (p->network_lds > (size_t)lds_limit()) return false;
    return true;
}

void build_map(std::vector<int64_t> &map, int64_t length, int64_t n, int64_t low, int64_t k, int32_t mode,
               int64_t stride);

// Plan the stack engine (box windows of rank 2 or 3 with compiled networks);
// false when it does not apply. A window with more than one plane slides
// along z (slices ky x kx, rows = y); a single-plane window slides along y
// (slices 1 x kx, rows = z planes).
bool plan_stack(ndf_plan *p)
{
    if (p->rank < 2 || p->rank > 3 || p->len == 0) return false;
    int64_t dims[3] = {1, 1, 1}, lo[3] = {0, 0, 0}, box[3] = {1, 1, 1};
    int32_t modes[3] = {NDF_MODE_NEAREST, NDF_MODE_NEAREST, NDF_MODE_NEAREST};
    const size_t pad = 3 - p->rank;
    int64_t volume = 1;
    for (size_t a = 0; a < p->rank; ++a) {
        dims[pad + a] = (int64_t)p->shape[a];
        lo[pad + a] = p->low[a];
        box[pad + a] = p->high[a] - p->low[a] + 1;
        modes[pad + a] = p->modes[a];
        volume *= box[pad + a];
    }
    // Every position of the taps' bounding box must be a tap.
    if (volume != (int64_t)p->count) return false;
    const bool planes = box[0] > 1;
    const int k0 = (int)(planes ? box[0] : box[1]);
    const int sh = planes ? (int)box[1] : 1;
    const int sw = (int)box[2];
    if (!stack_supported(p->dbytes, k0, sh, sw)) return false;
    const int64_t ns = planes ? dims[0] : dims[1];
    const int64_t nr = planes ? dims[1] : dims[0];
    const int64_t tiles_x = (dims[2] + 63) / 64;
    const int64_t in_sy = dims[2], in_sz = dims[1] * dims[2];
    // Runs along the slide axis: whole columns when that gives 8192 waves,
    // else shorter runs (at least 2 * k0 steps, so the k0 - 1 slices each run
    // builds ahead of its first output stay a small share).
    const int64_t columns = tiles_x * nr;
    int64_t run = ns;
    if (columns < 8192) {
        const int64_t wanted = (8192 + columns - 1) / columns;
        run = std::max<int64_t>((ns + wanted - 1) / wanted, 2 * k0);
        run = std::min(run, ns);
    }
    const int64_t units = columns * ((ns + run - 1) / run);
    if (units > stack_max_units() || run > INT32_MAX) return false;
    if (planes) {
        build_map(p->stk_maps, ns + k0 - 1, dims[0], lo[0], k0, modes[0], in_sz);
        build_map(p->stk_mapr, nr + sh - 1, dims[1], lo[1], sh, modes[1], in_sy);
    } else {
        build_map(p->stk_maps, ns + k0 - 1, dims[1], lo[1], k0, modes[1], in_sy);
        build_map(p->stk_mapr, nr, dims[0], 0, 1, NDF_MODE_NEAREST, in_sz);
    }
    build_map(p->stk_mapx, tiles_x * 64 + sw - 1, dims[2], lo[2], sw, modes[2], 1);
    StackArgs &s = p->stack;
    s = StackArgs{};
    s.ns = ns;
    s.nr = nr;
    s.nx = dims[2];
    s.out_ss = planes ? dims[1] * dims[2] : dims[2];
    s.out_sr = planes ? dims[2] : dims[1] * dims[2];
    s.units = units;
    s.k0 = k0;
    s.sh = sh;
    s.sw = sw;
    s.sx = 64 + sw - 1;
    s.tiles_x = (int32_t)tiles_x;
    s.run = (int32_t)run;
    s.cval_bits = p->cval_bits;
    s.ignore_inf = p->eff_ignore_inf;
    s.stat = p->stat;
    s.min_count = (int32_t)p->eff_min_count;
    s.rank_param = (int32_t)std::min(p->parameter, p->count);
    s.q100 = p->q100;
    return true;
}

// The engine a plan runs unless one is forced: extrema for rectangular
// minimum / maximum, then the stack engine for box windows it has networks
// for, the network engine for other small windows, the slide engine for
// large ones, the generic engine for the rest.
int32_t auto_engine(const ndf_plan *p)
{
    return p->extrema_ok   ? NDF_ENGINE_EXTREMA
           : p->stack_ok   ? NDF_ENGINE_STACK
           : p->network_ok ? NDF_ENGINE_NETWORK
                           : (p->slide_ok ? NDF_ENGINE_SLIDE : NDF_ENGINE_GENERIC);
}

// Maps for the tiled engines: entry i is the input offset of coordinate
// i + low along an axis of n (MAP_CVAL outside under constant mode); entries
// past the last window coordinate (n + k - 1) only feed padded outputs.
void build_map(std::vector<int64_t> &map, int64_t length, int64_t n, int64_t low, int64_t k, int32_t mode,
               int64_t stride)
{
    map.assign((size_t)length, MAP_PAD);
    for (int64_t i = 0; i < length && i < n + k - 1; ++i) {
        const int64_t m = map_coordinate(i + low, n, mode);
        map[(size_t)i] = m < 0 ? MAP_CVAL : m * stride;
    }
}

// Plan the extrema engine (rectangular minimum / maximum of rank <= 3 that
// needs no retained count); false when it does not apply or the tile does
// not fit LDS.
bool plan_extrema(ndf_plan *p)
{
    if (p->kind != NDF_KIND_MIN && p->kind != NDF_KIND_MAX) return false;
    if (!p->footprint.empty() || p->rank > 3 || p->len == 0) return false;
    if (p->omit && p->min_count > 1) return false;
    int64_t dims[3] = {1, 1, 1}, lo[3] = {0, 0, 0}, box[3] = {1, 1, 1};
    int32_t modes[3] = {NDF_MODE_NEAREST, NDF_MODE_NEAREST, NDF_MODE_NEAREST};
    const size_t pad = 3 - p->rank;
    for (size_t a = 0; a < p->rank; ++a) {
        dims[pad + a] = (int64_t)p->shape[a];
        lo[pad + a] = p->low[a];
        box[pad + a] = p->high[a] - p->low[a] + 1;
        modes[pad + a] = p->modes[a];
    }
    const int64_t kz = box[0], ky = box[1], kx = box[2];
    const int limit = lds_limit();
    // The tile with the least support per output that fits LDS; ties keep
    // the larger tile. A tile above half of LDS allows one workgroup per CU
    // (no second workgroup to hide the staging latency): it must save a
    // fifth of the support to be chosen.
    int best_tx = 0, best_ty = 0;
    double best_ratio = 0.0;
    const int64_t rows_cap = dims[1];
    for (int tx : {256, 128, 64}) {
        if (tx > 64 && (int64_t)(tx / 2) >= dims[2]) continue;
        for (int ty : {64, 48, 32, 24, 16, 12, 8, 6, 4, 2, 1}) {
            if (ty > rows_cap && ty > 1) continue;
            const int64_t sx = tx + kx - 1, sy = ty + ky - 1;
            if (sx > extrema_max_support_columns() || sy > (1 << 20)) continue;
            const size_t lds = extrema_lds_bytes(p->dbytes, (int)kz, (int)sy, (int)sx, tx);
            if (lds > (size_t)limit) continue;
            const double ratio = (double)(kz * sy * sx) / (double)(ty * tx) * (2 * lds > (size_t)limit ? 1.2 : 1.0);
            if (best_tx == 0 || ratio < best_ratio * 0.999) {
                best_tx = tx;
                best_ty = ty;
                best_ratio = ratio;
            }
        }
    }
    if (best_tx == 0) return false;
    const int64_t tiles_x = (dims[2] + best_tx - 1) / best_tx;
    const int64_t tiles_y = (dims[1] + best_ty - 1) / best_ty;
    if (tiles_x * tiles_y * dims[0] > 0x7FFFFFFF) return false;
    const int64_t in_sy = dims[2], in_sz = dims[1] * dims[2];
    build_map(p->ext_mapz, dims[0] + kz - 1, dims[0], lo[0], kz, modes[0], in_sz);
    build_map(p->ext_mapy, tiles_y * best_ty + ky - 1, dims[1], lo[1], ky, modes[1], in_sy);
    build_map(p->ext_mapx, tiles_x * best_tx + kx - 1, dims[2], lo[2], kx, modes[2], 1);
    ExtremaArgs &e = p->extrema;
    e = ExtremaArgs{};
    e.nz = dims[0];
    e.ny = dims[1];
    e.nx = dims[2];
    e.kz = (int32_t)kz;
    e.ky = (int32_t)ky;
    e.kx = (int32_t)kx;
    e.tx = best_tx;
    e.ty = best_ty;
    e.sx = (int32_t)(best_tx + kx - 1);
    e.sy = (int32_t)(best_ty + ky - 1);
    e.tiles_x = (int32_t)tiles_x;
    e.tiles_y = (int32_t)tiles_y;
    e.cval_bits = p->cval_bits;
    e.ignore_inf = p->eff_ignore_inf;
    e.omit = p->omit;
    e.maximum = p->kind == NDF_KIND_MAX;
    p->extrema_lds = extrema_lds_bytes(p->dbytes, e.kz, e.sy, e.sx, e.tx);
    return true;
}

// Plan the slide engine; false when the geometry cannot be staged.
// force_e > 0 fixes the sort (256 * force_e entries) and max_ty caps the tile
// rows (the batch autotuner's knobs); 0 / 256 let the planner choose.
bool plan_slide(ndf_plan *p, int force_e = 0, int max_ty = 256)
{
    if (p->rank > 3 || p->count > 65534 || p->len == 0) return false;
    // Canonical 3-D view: leading singleton axes for rank 1 and 2.
    int64_t dims[3] = {1, 1, 1}, lo[3] = {0, 0, 0}, box[3] = {1, 1, 1};
    int32_t modes[3] = {NDF_MODE_NEAREST, NDF_MODE_NEAREST, NDF_MODE_NEAREST};
    const size_t pad = 3 - p->rank;
    for (size_t a = 0; a < p->rank; ++a) {
        dims[pad + a] = (int64_t)p->shape[a];
        lo[pad + a] = p->low[a];
        box[pad + a] = p->high[a] - p->low[a] + 1;
        modes[pad + a] = p->modes[a];
    }
    const int64_t kz = box[0], ky = box[1], kx = box[2];
    const int64_t sx = SLIDE_TILE_X + kx - 1;
    if (sx > 254 || ky > 254 || kz > 254) return false;
    // Footprint mask over the tight box.
    std::vector<uint8_t> mask((size_t)(kz * ky * kx), 0);
    for (size_t t = 0; t < p->count; ++t) {
        int64_t d[3] = {0, 0, 0};
        for (size_t a = 0; a < p->rank; ++a) d[pad + a] = p->taps[t * p->rank + a] - p->low[a];
        mask[(size_t)((d[0] * ky + d[1]) * kx + d[2])] = 1;
    }
    const bool rect = p->count == mask.size();
    const int mask_bytes = rect ? 0 : (int)mask.size();
    if (mask_bytes > 16384) return false;
    // Rows per wave: as many as LDS allows, but no more than the image needs.
    const int limit = lds_limit();
    // Tile rows: the support (kz x (ty + ky - 1) x sx samples) must fit the
    // sort (256 * E <= 16384 entries, u16 positions) and LDS. Larger tiles
    // amortise the sort and the per-wave start; a support much larger than
    // the window lengthens every rank walk, so it is capped at 32 windows.
    const int64_t hist_stride = rect ? (int64_t)18 * (sx + 1) : 0;
    // Rows of the support: planes of ty + ky - 1 rows; with several planes
    // the plane stride is a power of two so positions decode with shifts.
    auto zrows_of = [&](int64_t ty) {
        int64_t r = ty + ky - 1;
        if (kz == 1) return r;
        int64_t pow2 = 1;
        while (pow2 < r) pow2 <<= 1;
        return pow2;
    };
    auto rows_of = [&](int64_t ty) { return (kz - 1) * zrows_of(ty) + ty + ky - 1; };
    auto support_of = [&](int64_t ty) { return rows_of(ty) * sx; };
    auto sort_e_of = [](int64_t s) {
        int e = 4;
        while ((int64_t)256 * e < s) e <<= 1;
        return e;
    };
    auto fits = [&](int64_t ty) {
        const int64_t s = support_of(ty);
        if (rows_of(ty) > 254 || s > 16384) return false;
        if (force_e > 0 && (int64_t)256 * force_e < s) return false;
        const int e = force_e > 0 ? force_e : sort_e_of(s);
        return e <= 64 && slide_lds_bytes(p->dbytes, (int)s, e, (int)hist_stride, mask_bytes) <= (size_t)limit;
    };
    const int64_t rows_needed = ((dims[1] + SLIDE_WAVES - 1) / SLIDE_WAVES) * SLIDE_WAVES;
    int64_t ty = 0;
    for (int64_t cand = SLIDE_WAVES; cand <= max_ty; cand += SLIDE_WAVES) {
        if (!fits(cand)) break;
        if (force_e == 0 && cand > SLIDE_WAVES && support_of(cand) > 32 * (int64_t)p->count) break;
        ty = cand;
        if (cand >= rows_needed) break;
    }
    if (ty == 0) return false;
    const int tw = (int)(ty / SLIDE_WAVES);
    const int sy = (int)(ty + ky - 1);
    const int zrows = (int)zrows_of(ty);
    int zshift = 8;
    if (kz > 1) {
        zshift = 0;
        while ((1 << zshift) < zrows) ++zshift;
    }
    const int rows = (int)rows_of(ty);
    const int s = (int)support_of(ty);
    const int sort_e = force_e > 0 ? force_e : sort_e_of(s);
    const int64_t tiles_x = (dims[2] + SLIDE_TILE_X - 1) / SLIDE_TILE_X;
    const int64_t tiles_y = (dims[1] + ty - 1) / ty;
    if (tiles_x * tiles_y * dims[0] > 0x7FFFFFFF) return false;

    const int64_t in_sy = dims[2];
    const int64_t in_sz = dims[1] * dims[2];
    // Maps: index i <-> coordinate i + lo; entries past the last valid
    // window coordinate only feed padded outputs.
    auto build = [](std::vector<int64_t> &map, int64_t length, int64_t n, int64_t low, int64_t k, int32_t mode,
                    int64_t stride) {
        map.assign((size_t)length, MAP_PAD);
        for (int64_t i = 0; i < length; ++i) {
            if (i >= n + k - 1) break;
            const int64_t m = map_coordinate(i + low, n, mode);
            map[(size_t)i] = m < 0 ? MAP_CVAL : m * stride;
        }
    };
    build(p->slide_mapz, dims[0] + kz - 1, dims[0], lo[0], kz, modes[0], in_sz);
    // Long enough for any tile height up to 256 rows (the bracket tiles).
    build(p->slide_mapy, std::max(tiles_y * ty, dims[1] + 256) + ky - 1, dims[1], lo[1], ky, modes[1], in_sy);
    build(p->slide_mapx, tiles_x * SLIDE_TILE_X + kx - 1, dims[2], lo[2], kx, modes[2], 1);

    p->slide_taps.clear();
    p->slide_rm.clear();
    p->slide_add.clear();
    for (int64_t dz = 0; dz < kz; ++dz) {
        for (int64_t dy = 0; dy < ky; ++dy) {
            for (int64_t dx = 0; dx < kx; ++dx) {
                auto at = [&](int64_t yy) { return mask[(size_t)((dz * ky + yy) * kx + dx)] != 0; };
                if (!at(dy)) continue;
                const uint32_t off = (uint32_t)((dz * zrows + dy) * sx + dx);
                p->slide_taps.push_back(off);
                if (dy == 0 || !at(dy - 1)) p->slide_rm.push_back(off);
                if (dy == ky - 1 || !at(dy + 1)) p->slide_add.push_back(off);
            }
        }
    }
    if (p->slide_rm.size() != p->slide_add.size()) return false;
    p->slide_mask = rect ? std::vector<uint8_t>() : mask;

    SlideArgs &s_args = p->slide;
    s_args = SlideArgs{};
    s_args.nz = dims[0];
    s_args.ny = dims[1];
    s_args.nx = dims[2];
    s_args.in_sz = in_sz;
    s_args.in_sy = in_sy;
    s_args.kz = (int32_t)kz;
    s_args.ky = (int32_t)ky;
    s_args.kx = (int32_t)kx;
    s_args.sx = (int32_t)sx;
    s_args.sy = sy;
    s_args.zrows = zrows;
    s_args.zshift = zshift;
    s_args.rows = rows;
    s_args.s = s;
    s_args.sort_e = sort_e;
    s_args.hist_stride = (int32_t)hist_stride;
    s_args.ty = (int32_t)ty;
    s_args.tw = tw;
    s_args.tiles_x = (int32_t)tiles_x;
    s_args.tiles_y = (int32_t)tiles_y;
    s_args.count = (int32_t)p->count;
    s_args.nedge = (int32_t)p->slide_rm.size();
    s_args.cval_bits = p->cval_bits;
    s_args.ignore_inf = p->eff_ignore_inf;
    s_args.stat = p->stat;
    s_args.min_count = (int32_t)p->eff_min_count;
    s_args.rank_param = (int32_t)std::min(p->parameter, p->count);
    s_args.q100 = p->q100;
    // A min_count above half the window (Propagate: all of it) leaves most
    // windows of a sky with NaNs without a result; such tiles skip the sort.
    s_args.count_gate = rect && sx <= 128 && 2 * p->eff_min_count > p->count ? 1 : 0;
    p->slide_lds = slide_lds_bytes(p->dbytes, s, sort_e, (int)hist_stride, mask_bytes);

    // Bracket tiles (2-D rectangles with support rows of at most 128
    // samples): the tallest tile of 24 to 32 rows among those that keep the
    // most bracket workgroups resident per CU (at most 4 counted for float32,
    // 3 for float64); shorter tiles only when none of those fits.
    p->bracket_ty = 0;
    if (force_e == 0 && rect && kz == 1 && dims[0] == 1 && kx <= 64 && p->rank >= 2) {
        int best_per_cu = 0;
        for (int cand = 32; cand >= SLIDE_WAVES && (best_per_cu == 0 || cand >= 24); cand -= SLIDE_WAVES) {
            const SlideArgs v = with_tile_rows(s_args, cand);
            if (v.rows > 254 || v.s > 16384 || v.sort_e > 64 ||
                slide_lds_bytes(p->dbytes, v.s, v.sort_e, (int)hist_stride, 0) > (size_t)limit) {
                continue;
            }
            const size_t need = bracket_lds_bytes(p->dbytes, v.s, v.rows, 16);
            const size_t granule = 512;
            // The bracket kernel's launch bounds allow 4 workgroups per CU
            // for float32 and 3 for float64 (register budget).
            const int by_registers = p->dbytes == 4 ? 4 : 3;
            const int per_cu =
                std::min(by_registers, (int)((size_t)limit / ((need + 16 + granule - 1) / granule * granule)));
            if (per_cu > best_per_cu) {
                best_per_cu = per_cu;
                p->bracket_ty = cand;
            }
        }
    }
    return true;
}

ndf_status prepare_bracket(ndf_plan *p, char *err, size_t err_len);
ndf_status prepare_engine_tables(ndf_plan *p, char *err, size_t err_len);

ndf_status prepare_tables(ndf_plan *p, char *err, size_t err_len)
{
    ndf_status st = prepare_engine_tables(p, err, err_len);
    if (st == NDF_OK && p->bracket_on && p->engine == NDF_ENGINE_SLIDE && !p->bracket_ready) {
        st = prepare_bracket(p, err, err_len);
    }
    return st;
}

// Generic-engine tables (per-axis maps and tap offsets) on the host, and
// the generic launch arguments without device pointers.
void generic_host_tables(ndf_plan *p)
{
    if (p->generic_host_ready) return;
    std::vector<int64_t> &maps = p->gen_maps;
    maps.clear();
    GenericArgs &g = p->generic;
    g = GenericArgs{};
    g.rank = (int32_t)p->rank;
    g.count = (int32_t)p->count;
    g.len = (int64_t)p->len;
    int64_t stride = 1;
    for (size_t a = p->rank; a-- > 0;) {
        g.stride[a] = stride;
        stride *= (int64_t)p->shape[a];
    }
    for (size_t a = 0; a < p->rank; ++a) {
        g.map_base[a] = (int64_t)maps.size();
        const int64_t n = (int64_t)p->shape[a];
        const int64_t length = n + p->high[a] - p->low[a];
        for (int64_t i = 0; i < length; ++i) {
            const int64_t m = map_coordinate(i + p->low[a], n, p->modes[a]);
            maps.push_back(m < 0 ? MAP_CVAL : m * g.stride[a]);
        }
    }
    p->gen_taps.assign(p->count * p->rank, 0);
    for (size_t t = 0; t < p->count; ++t) {
        for (size_t a = 0; a < p->rank; ++a) {
            p->gen_taps[t * p->rank + a] = (int32_t)(p->taps[t * p->rank + a] - p->low[a]);
        }
    }
    g.cval_bits = p->cval_bits;
    g.ignore_inf = p->eff_ignore_inf;
    g.stat = p->stat;
    g.min_count = p->eff_min_count;
    g.rank_param = (int64_t)std::min(p->parameter, p->count);
    g.q100 = p->q100;
    if (p->rank == 2 && p->footprint.empty()) {
        g.rect_ky = (int32_t)p->size[0];
        g.rect_kx = (int32_t)p->size[1];
    }
    p->generic_host_ready = true;
}

// Generic-engine tables on the device. Also used by the sparse repair path
// whatever engine the plan runs for whole images.
ndf_status prepare_generic_tables(ndf_plan *p, char *err, size_t err_len)
{
    if (p->generic_ready) return NDF_OK;
    generic_host_tables(p);
    hipError_t e;
    if ((e = upload(p->g_maps, p->gen_maps)) != hipSuccess || (e = upload(p->g_taps, p->gen_taps)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    p->generic.maps = static_cast<const int64_t *>(p->g_maps.ptr);
    p->generic.taps = static_cast<const int32_t *>(p->g_taps.ptr);
    p->generic_ready = true;
    return NDF_OK;
}

ndf_status prepare_engine_tables(ndf_plan *p, char *err, size_t err_len)
{
    if (p->tables_ready) return NDF_OK;
    const bool fresh = p->d_error.ptr == nullptr;
    hipError_t e = p->d_error.ensure(sizeof(int32_t));
    if (e == hipSuccess && fresh) e = hipMemset(p->d_error.ptr, 0, sizeof(int32_t));
    if (e != hipSuccess) return fail_hip(e, err, err_len);
    if (p->engine == NDF_ENGINE_GENERIC) {
        const ndf_status st = prepare_generic_tables(p, err, err_len);
        if (st != NDF_OK) return st;
    } else if (p->engine == NDF_ENGINE_EXTREMA) {
        if ((e = upload(p->e_mapz, p->ext_mapz)) != hipSuccess || (e = upload(p->e_mapy, p->ext_mapy)) != hipSuccess ||
            (e = upload(p->e_mapx, p->ext_mapx)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        ExtremaArgs &x = p->extrema;
        x.mapz = static_cast<const int64_t *>(p->e_mapz.ptr);
        x.mapy = static_cast<const int64_t *>(p->e_mapy.ptr);
        x.mapx = static_cast<const int64_t *>(p->e_mapx.ptr);
    } else if (p->engine == NDF_ENGINE_NETWORK) {
        if ((e = upload(p->n_mapz, p->net_mapz)) != hipSuccess || (e = upload(p->n_mapy, p->net_mapy)) != hipSuccess ||
            (e = upload(p->n_mapx, p->net_mapx)) != hipSuccess ||
            (e = upload(p->n_taps, p->network.pair_c ? p->net_pair_taps : p->net_taps)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        NetworkArgs &n = p->network;
        n.mapz = static_cast<const int64_t *>(p->n_mapz.ptr);
        n.mapy = static_cast<const int64_t *>(p->n_mapy.ptr);
        n.mapx = static_cast<const int64_t *>(p->n_mapx.ptr);
        n.taps = static_cast<const uint32_t *>(p->n_taps.ptr);
    } else if (p->engine == NDF_ENGINE_STACK) {
        if ((e = upload(p->k_maps, p->stk_maps)) != hipSuccess || (e = upload(p->k_mapr, p->stk_mapr)) != hipSuccess ||
            (e = upload(p->k_mapx, p->stk_mapx)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        StackArgs &k = p->stack;
        k.maps = static_cast<const int64_t *>(p->k_maps.ptr);
        k.mapr = static_cast<const int64_t *>(p->k_mapr.ptr);
        k.mapx = static_cast<const int64_t *>(p->k_mapx.ptr);
    } else if (p->engine == NDF_ENGINE_SLIDE) {
        if ((e = upload(p->s_mapz, p->slide_mapz)) != hipSuccess ||
            (e = upload(p->s_mapy, p->slide_mapy)) != hipSuccess ||
            (e = upload(p->s_mapx, p->slide_mapx)) != hipSuccess ||
            (e = upload(p->s_taps, p->slide_taps)) != hipSuccess ||
            (e = upload(p->s_rm, p->slide_rm)) != hipSuccess || (e = upload(p->s_add, p->slide_add)) != hipSuccess ||
            (e = upload(p->s_mask, p->slide_mask)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        SlideArgs &s = p->slide;
        s.mapz = static_cast<const int64_t *>(p->s_mapz.ptr);
        s.mapy = static_cast<const int64_t *>(p->s_mapy.ptr);
        s.mapx = static_cast<const int64_t *>(p->s_mapx.ptr);
        s.taps = static_cast<const uint32_t *>(p->s_taps.ptr);
        s.edge_remove = static_cast<const uint32_t *>(p->s_rm.ptr);
        s.edge_add = static_cast<const uint32_t *>(p->s_add.ptr);
        s.mask = p->slide_mask.empty() ? nullptr : static_cast<const uint8_t *>(p->s_mask.ptr);
        s.error_flag = static_cast<int32_t *>(p->d_error.ptr);
    }
    p->tables_ready = true;
    return NDF_OK;
}

bool bracket_supported(const ndf_plan *p) { return p->slide_ok && p->bracket_ty > 0; }

// Tiles of the bracket geometry (bracket_ty rows each).
int64_t bracket_tiles(const ndf_plan *p)
{
    return (int64_t)p->slide.tiles_x * ((p->slide.ny + p->bracket_ty - 1) / p->bracket_ty);
}

// Bracket launch geometry (host only): the plain fallback kernel on the
// bracket tiles, the bracket kernel and the proposal arguments, without
// device pointers.
void bracket_geometry_into(const ndf_plan *p, const SlideArgs &slide, SlideArgs &fb, size_t &fb_lds, SlideArgs &sb,
                           size_t &sb_lds, BracketArgs &b)
{
    const int64_t bx = (slide.nx + 7) / 8;
    fb = with_tile_rows(slide, p->bracket_ty);
    fb_lds = slide_lds_bytes(p->dbytes, fb.s, fb.sort_e, fb.hist_stride, 0);
    b = BracketArgs{};
    b.ny = slide.ny;
    b.nx = slide.nx;
    b.bx = bx;
    b.ty = fb.ty;
    b.tiles_y = fb.tiles_y;
    b.tiles_x = fb.tiles_x;
    const size_t ay = p->rank - 2, ax = p->rank - 1;
    b.low_y = (int32_t)p->low[ay];
    b.high_y = (int32_t)p->high[ay];
    b.low_x = (int32_t)p->low[ax];
    b.high_x = (int32_t)p->high[ax];
    b.quantile = p->stat == STAT_RANK
                     ? (float)((double)std::min(p->parameter, p->count) / (double)(p->count > 1 ? p->count - 1 : 1))
                 : p->stat == STAT_PERCENTILE ? (float)p->q100
                                              : 0.5f;
    sb = fb;
    sb.sort_e = p->bracket_e;
    sb_lds = bracket_lds_bytes(p->dbytes, fb.s, fb.rows, p->bracket_e);
}

void bracket_geometry(ndf_plan *p)
{
    bracket_geometry_into(p, p->slide, p->slide_fb, p->slide_fb_lds, p->slide_b, p->slide_b_lds, p->bargs);
}

// Bytes of the bracket's per-launch buffers, in carve order: block moments
// (double[3] per 8x8 block), features, tile moments, lo / hi keys,
// predictions, fallback list, counters (the fallback count, then the
// bracket kernel's tile queues).
void bracket_buffer_sizes(const ndf_plan *p, size_t sizes[8])
{
    const int64_t tiles = bracket_tiles(p);
    const int64_t by = (p->slide.ny + 7) / 8, bx = (p->slide.nx + 7) / 8;
    sizes[0] = (size_t)(by * bx * 3) * sizeof(double);
    sizes[1] = (size_t)tiles * FEATURES * sizeof(float);
    sizes[2] = (size_t)tiles * 3 * sizeof(double);
    sizes[3] = (size_t)tiles * (size_t)p->dbytes;
    sizes[4] = (size_t)tiles * (size_t)p->dbytes;
    sizes[5] = (size_t)tiles * 2 * sizeof(float);
    sizes[6] = (size_t)tiles * sizeof(int32_t);
    sizes[7] = (1 + BRACKET_QUEUES) * sizeof(int32_t);
}

// Points the bracket arguments at its buffers (in bracket_buffer_sizes order).
void bracket_bind(SlideArgs &slide_b, BracketArgs &b, void *const buffers[8])
{
    b.features = static_cast<float *>(buffers[1]);
    b.moments = static_cast<double *>(buffers[2]);
    b.lo_keys = buffers[3];
    b.hi_keys = buffers[4];
    b.predicted = static_cast<float *>(buffers[5]);
    b.reset = static_cast<int32_t *>(buffers[7]);
    slide_b.bracket_lo = buffers[3];
    slide_b.bracket_hi = buffers[4];
    slide_b.fallback_tiles = static_cast<int32_t *>(buffers[6]);
    slide_b.fallback_count = static_cast<int32_t *>(buffers[7]);
    slide_b.tile_queue = static_cast<int32_t *>(buffers[7]) + 1;
    b.reset_count = 1 + BRACKET_QUEUES;
}

// Allocates the bracket buffers and fills the bracket launch arguments.
ndf_status prepare_bracket(ndf_plan *p, char *err, size_t err_len)
{
    if (p->bracket_ready) return NDF_OK;
    size_t sizes[8];
    bracket_buffer_sizes(p, sizes);
    DeviceBuffer *buffers[8] = {&p->b_blocks, &p->b_features, &p->b_moments, &p->b_lo,
                                &p->b_hi,     &p->b_pred,     &p->b_fallback, &p->b_fcount};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 8 && e == hipSuccess; ++i) e = buffers[i]->ensure(sizes[i]);
    if (e != hipSuccess) return fail_hip(e, err, err_len);
    if (!p->bracket_weights.empty() && (e = upload(p->b_weights, p->bracket_weights)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    // The plain fallback kernel shares the plan's device tables (p->slide).
    bracket_geometry(p);
    void *ptrs[8];
    for (int i = 0; i < 8; ++i) ptrs[i] = buffers[i]->ptr;
    bracket_bind(p->slide_b, p->bargs, ptrs);
    p->bracket_ready = true;
    return NDF_OK;
}

// Device pointers of one launch: the launch arguments of every engine with
// their tables bound, and the bracket and error buffers. A plan launch takes
// them from the plan's own device buffers; the no-allocation launch from the
// static scratch array.
struct LaunchSet {
    SlideArgs slide{}, slide_b{}, slide_fb{};
    ExtremaArgs extrema{};
    NetworkArgs network{};
    StackArgs stack{};
    GenericArgs generic{};
    BracketArgs bargs{};
    size_t slide_b_lds = 0, slide_fb_lds = 0;
    double *blocks = nullptr;
    const float *weights = nullptr;
    int32_t *fallback = nullptr;
    int32_t *fcount = nullptr;
    int32_t *error = nullptr;
};

LaunchSet plan_set(const ndf_plan *p)
{
    LaunchSet t;
    t.slide = p->slide;
    t.slide_b = p->slide_b;
    t.slide_fb = p->slide_fb;
    t.network = p->network;
    t.stack = p->stack;
    t.extrema = p->extrema;
    t.generic = p->generic;
    t.bargs = p->bargs;
    t.slide_b_lds = p->slide_b_lds;
    t.slide_fb_lds = p->slide_fb_lds;
    t.blocks = static_cast<double *>(p->b_blocks.ptr);
    t.weights = static_cast<const float *>(p->b_weights.ptr);
    t.fallback = static_cast<int32_t *>(p->b_fallback.ptr);
    t.fcount = static_cast<int32_t *>(p->b_fcount.ptr);
    t.error = static_cast<int32_t *>(p->d_error.ptr);
    return t;
}

// Enqueues the plan's kernels with the pointers of `t` on `stream`.
hipError_t enqueue(ndf_plan *p, const LaunchSet &t, const void *d_in, void *d_out, hipStream_t stream)
{
    hipError_t e = hipSuccess;
    const uint64_t nan_bits = p->dbytes == 4 ? 0x7FC00000ull : 0x7FF8000000000000ull;
    p->bracket_launched = false;
    auto mark = [&](hipEvent_t ev) {
        if (ev && e == hipSuccess) e = hipEventRecord(ev, stream);
    };
    // The slide kernels report a rank walk that left its bounds through the
    // error flag; it is cleared on the launch stream, ahead of them.
    if (p->engine == NDF_ENGINE_SLIDE && !p->all_nan) e = launch_fill(4, t.error, 1, 0, stream);
    if (e != hipSuccess) return e;
    if (p->all_nan) {
        mark(p->main_begin);
        e = launch_fill(p->dbytes, d_out, (int64_t)p->len, nan_bits, stream);
        mark(p->main_end);
    } else if (p->engine == NDF_ENGINE_SLIDE && p->bracket_on) {
        // Learned bracket: propose (block moments, tile features, MLP),
        // filter in-bracket samples only, then recompute every fallback tile
        // with the plain kernel.
        const int tiles = (int)bracket_tiles(p);
        p->bracket_launched = true;
        p->report_fcount = t.fcount;
        // The proposal launch also clears the fallback count (t.bargs.reset).
        e = launch_block_stats(p->dbytes, d_in, p->slide.ny, p->slide.nx, p->eff_ignore_inf, t.blocks, stream);
        if (e == hipSuccess) e = launch_bracket_proposal(p->dbytes, t.bargs, t.blocks, t.weights, tiles, stream);
        if (e == hipSuccess) {
            SlideArgs args = t.slide_b;
            args.input = d_in;
            args.output = d_out;
            mark(p->main_begin);
            if (e == hipSuccess) e = launch_bracket_slide(p->dbytes, args, t.slide_b_lds, tiles, stream);
            mark(p->main_end);
        }
        if (e == hipSuccess) {
            SlideArgs args = t.slide_fb;
            args.input = d_in;
            args.output = d_out;
            args.tile_list = t.fallback;
            args.tile_count = t.fcount;
            e = launch_slide(p->dbytes, args, t.slide_fb_lds, tiles, stream);
        }
    } else if (p->engine == NDF_ENGINE_EXTREMA) {
        ExtremaArgs args = t.extrema;
        args.input = d_in;
        args.output = d_out;
        mark(p->main_begin);
        if (e == hipSuccess) e = launch_extrema(p->dbytes, args, p->extrema_lds, stream);
        mark(p->main_end);
    } else if (p->engine == NDF_ENGINE_NETWORK) {
        NetworkArgs args = t.network;
        args.input = d_in;
        args.output = d_out;
        mark(p->main_begin);
        if (e == hipSuccess) e = launch_network(p->dbytes, args, p->network_lds, stream);
        mark(p->main_end);
    } else if (p->engine == NDF_ENGINE_STACK) {
        StackArgs args = t.stack;
        args.input = d_in;
        args.output = d_out;
        mark(p->main_begin);
        if (e == hipSuccess) e = launch_stack(p->dbytes, args, stream);
        mark(p->main_end);
    } else if (p->engine == NDF_ENGINE_SLIDE) {
        SlideArgs args = t.slide;
        args.input = d_in;
        args.output = d_out;
        mark(p->main_begin);
        if (e == hipSuccess) {
            e = launch_slide(p->dbytes, args, p->slide_lds, args.tiles_x * args.tiles_y * (int)args.nz, stream);
        }
        mark(p->main_end);
    } else {
        GenericArgs args = t.generic;
        args.input = d_in;
        args.output = d_out;
        mark(p->main_begin);
        if (e == hipSuccess) e = launch_generic(p->dbytes, args, stream);
        mark(p->main_end);
    }
    return e;
}

// Orders a launch of plan p on `stream` after the plan's previous launch.
hipError_t plan_begin(ndf_plan *p, hipStream_t stream)
{
    if (!p->done) {
        hipError_t e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
        if (e != hipSuccess) {
            p->done = nullptr;
            return e;
        }
    }
    return p->done_recorded ? hipStreamWaitEvent(stream, p->done, 0) : hipSuccess;
}

hipError_t plan_end(ndf_plan *p, hipStream_t stream)
{
    hipError_t e = hipEventRecord(p->done, stream);
    if (e == hipSuccess) p->done_recorded = true;
    return e;
}

ndf_status launch(ndf_plan *p, const void *d_in, void *d_out, hipStream_t stream, char *err, size_t err_len)
{
    hipError_t e = plan_begin(p, stream);
    if (e == hipSuccess) e = enqueue(p, plan_set(p), d_in, d_out, stream);
    if (e == hipSuccess) e = plan_end(p, stream);
    return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
}

// Validation + tap construction shared by plan creation.
ndf_status build_plan(ndf_plan *p, int32_t dtype_bytes, size_t rank, const size_t *shape, const size_t *size,
                      const uint8_t *footprint, const ptrdiff_t *origin, const int32_t *modes, double cval,
                      int32_t kind, int32_t omit_nan, int32_t ignore_inf, size_t min_count, size_t parameter,
                      double q, char *err, size_t err_len)
{
    if (dtype_bytes != 4 && dtype_bytes != 8) {
        return fail(NDF_ERR_INVALID, "array dtype must be float32 or float64", err, err_len);
    }
    const bool robust = kind == KIND_PERCENTILE;
    if (!robust && (kind < NDF_KIND_MEDIAN || kind > NDF_KIND_MAX)) return fail_code(NDF_ERR_KIND, err, err_len);
    for (size_t a = 0; a < rank; ++a) {
        if (modes[a] < NDF_MODE_REFLECT || modes[a] > NDF_MODE_WRAP) return fail_code(NDF_ERR_MODE, err, err_len);
    }
    if (rank > (size_t)MAX_RANK) return fail(NDF_ERR_INVALID, "array rank exceeds 32", err, err_len);
    const bool histogram = kind == NDF_KIND_MEDIAN_HISTOGRAM || kind == NDF_KIND_RANK_HISTOGRAM;
    const bool extrema = kind == NDF_KIND_MIN || kind == NDF_KIND_MAX;
    // Rectangular histogram kinds and rectangular extrema are validated by
    // validate_rectangular first (src/plan.rs:154-197, 270-312).
    const bool rectangular_plan = histogram || (extrema && !footprint);
    p->dbytes = dtype_bytes;
    p->rank = rank;
    p->shape.assign(shape, shape + rank);
    p->size.assign(size, size + rank);
    p->origin.assign(origin, origin + rank);
    p->modes.assign(modes, modes + rank);
    p->kind = kind;
    p->omit = omit_nan != 0;
    p->ignore_inf = ignore_inf != 0;
    p->min_count = min_count;
    p->parameter = parameter;
    if (kind == NDF_KIND_RANK || kind == NDF_KIND_RANK_HISTOGRAM) {
        p->stat = STAT_RANK;
    } else if (extrema || robust) {
        // The minimum and maximum are the 0th and 100th percentiles: rank 0
        // and n - 1 of the retained samples, with a zero fraction.
        p->stat = STAT_PERCENTILE;
        p->q100 = robust ? q / 100.0 : (kind == NDF_KIND_MAX ? 1.0 : 0.0);
    } else {
        p->stat = STAT_MEDIAN;
    }
    if (dtype_bytes == 4) {
        const float f = (float)cval;
        uint32_t b;
        memcpy(&b, &f, 4);
        p->cval_bits = b;
    } else {
        memcpy(&p->cval_bits, &cval, 8);
    }
    p->cval = cval;

    if (rectangular_plan) {
        // src/plan.rs:294-312, validate_rectangular, validate_support.
        if (footprint) return fail(NDF_ERR_INVALID, MSG_HIST_FOOTPRINT, err, err_len);
        if (rank == 0) return fail(NDF_ERR_INVALID, MSG_RANK0, err, err_len);
        for (size_t a = 0; a < rank; ++a) {
            if (size[a] == 0) return fail(NDF_ERR_INVALID, MSG_SIZE0, err, err_len);
        }
        size_t len = 1, volume = 1;
        for (size_t a = 0; a < rank; ++a) {
            if (!checked_mul(len, shape[a], &len)) return fail_code(NDF_ERR_OVERFLOW, err, err_len);
        }
        for (size_t a = 0; a < rank; ++a) {
            if (!checked_mul(volume, size[a], &volume)) return fail_code(NDF_ERR_OVERFLOW, err, err_len);
        }
        for (size_t a = 0; a < rank; ++a) {
            if (size[a] > (size_t)INT64_MAX) return fail_code(NDF_ERR_OVERFLOW, err, err_len);
            const int64_t length = (int64_t)size[a];
            const int64_t shift = (int64_t)origin[a];
            if (shift < -(length / 2) || shift > (length - 1) / 2) {
                return fail(NDF_ERR_INVALID, MSG_ORIGIN, err, err_len);
            }
        }
        if (p->omit && (min_count == 0 || min_count > volume)) {
            return fail(NDF_ERR_INVALID, MSG_MIN_COUNT, err, err_len);
        }
        // validate_support: block lengths of max(size, 32) per axis.
        size_t support = 1;
        for (size_t a = 0; a < rank && histogram; ++a) {
            const size_t window = std::max<size_t>(size[a], 32);
            size_t blocks = shape[a] / window;
            if (blocks < 1) blocks = 1;
            size_t block = (shape[a] + blocks - 1) / blocks;
            if (block < 1) block = 1;
            const size_t length = block + size[a] - 1;
            if (length < block || !checked_mul(support, length, &support)) {
                return fail_code(NDF_ERR_OVERFLOW, err, err_len);
            }
        }
        if (support > 0xFFFFFFFEull) return fail_code(NDF_ERR_OVERFLOW, err, err_len);
    }

    // Window::new (src/window.rs:44-137).
    if (rank == 0) return fail(NDF_ERR_INVALID, MSG_RANK0, err, err_len);
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] == 0) return fail(NDF_ERR_INVALID, MSG_SIZE0, err, err_len);
    }
    size_t volume = 1;
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(volume, size[a], &volume)) return fail_code(NDF_ERR_OVERFLOW, err, err_len);
    }
    std::vector<int64_t> anchors(rank);
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] > (size_t)INT64_MAX) return fail_code(NDF_ERR_OVERFLOW, err, err_len);
        const int64_t length = (int64_t)size[a];
        const int64_t shift = (int64_t)origin[a];
        if (shift < -(length / 2) || shift > (length - 1) / 2) {
            return fail(NDF_ERR_INVALID, MSG_ORIGIN, err, err_len);
        }
        anchors[a] = length / 2 + shift;
    }
    size_t len = 1;
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(len, shape[a], &len)) return fail_code(NDF_ERR_OVERFLOW, err, err_len);
    }
    size_t count = 0;
    if (footprint) {
        for (size_t i = 0; i < volume; ++i) count += footprint[i] != 0;
    } else {
        count = volume;
    }
    if (count == 0) return fail(NDF_ERR_INVALID, MSG_EMPTY_FOOTPRINT, err, err_len);
    if (robust) {
        // RobustReduction::validate (src/robust.rs:189-241).
        if (!(q >= 0.0 && q <= 100.0)) {
            return fail(NDF_ERR_INVALID, "percentile must be finite and between zero and 100", err, err_len);
        }
        if (p->omit && (min_count == 0 || min_count > count)) {
            return fail(NDF_ERR_INVALID, "min_count must be between one and the number of window elements", err,
                        err_len);
        }
    }
    if (p->omit && (min_count == 0 || min_count > count)) {
        return fail(NDF_ERR_INVALID, MSG_MIN_COUNT, err, err_len);
    }
    if (count > (size_t)INT32_MAX / (rank ? rank : 1)) return fail_code(NDF_ERR_MEMORY, err, err_len);
    try {
        p->taps.resize(count * rank);
        p->low.assign(rank, INT64_MAX);
        p->high.assign(rank, INT64_MIN);
        if (footprint) p->footprint.assign(footprint, footprint + volume);
    } catch (const std::bad_alloc &) {
        return fail_code(NDF_ERR_MEMORY, err, err_len);
    }
    std::vector<size_t> index(rank, 0);
    size_t t = 0;
    for (size_t element = 0; element < volume; ++element) {
        if (!footprint || footprint[element]) {
            for (size_t a = 0; a < rank; ++a) {
                const int64_t offset = (int64_t)index[a] - anchors[a];
                p->taps[t * rank + a] = offset;
                p->low[a] = std::min(p->low[a], offset);
                p->high[a] = std::max(p->high[a], offset);
            }
            ++t;
        }
        for (size_t a = rank; a-- > 0;) {
            if (++index[a] < size[a]) break;
            index[a] = 0;
        }
    }
    p->len = len;
    p->count = count;
    p->eff_ignore_inf = p->omit && p->ignore_inf;
    p->eff_min_count = p->omit ? (int64_t)min_count : (int64_t)count;
    p->all_nan = p->stat == STAT_RANK && parameter >= count;
    p->slide_ok = plan_slide(p);
    p->network_ok = plan_network(p);
    p->stack_ok = plan_stack(p);
    p->extrema_ok = plan_extrema(p);
    p->engine = auto_engine(p);
    return NDF_OK;
}

// True when the len-element input and output ranges share a byte: the
// filters read every input sample of a tile's support after other tiles have
// written their outputs, so they need disjoint buffers.
bool ranges_overlap(const void *d_input, const void *d_output, size_t bytes)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(d_input), b = reinterpret_cast<uintptr_t>(d_output);
    return a < b + bytes && b < a + bytes;
}

bool buffers_overlap(const ndf_plan *plan, const void *d_input, const void *d_output, size_t len)
{
    return ranges_overlap(d_input, d_output, len * (size_t)plan->dbytes);
}

}  // namespace

extern "C" {

ndf_status ndf_plan_create(ndf_plan **plan, int32_t dtype_bytes, size_t rank, const size_t *shape,
                           const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin,
                           const int32_t *modes, double cval, int32_t kind, int32_t omit_nan, int32_t ignore_inf,
                           size_t min_count, size_t parameter, char *err, size_t err_len)
{
    if (!plan) return fail(NDF_ERR_INVALID, "plan pointer is null", err, err_len);
    *plan = nullptr;
    ndf_plan *p = new (std::nothrow) ndf_plan();
    if (!p) return fail_code(NDF_ERR_MEMORY, err, err_len);
    const ndf_status st = build_plan(p, dtype_bytes, rank, shape, size, footprint, origin, modes, cval, kind,
                                     omit_nan, ignore_inf, min_count, parameter, 0.0, err, err_len);
    if (st != NDF_OK) {
        delete p;
        return st;
    }
    *plan = p;
    return NDF_OK;
}

ndf_status ndf_robust_plan_create(ndf_plan **plan, int32_t dtype_bytes, size_t rank, const size_t *shape,
                                  const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin,
                                  const int32_t *modes, double cval, int32_t kind, double parameter, int32_t omit_nan,
                                  int32_t ignore_inf, size_t min_count, char *err, size_t err_len)
{
    if (!plan) return fail(NDF_ERR_INVALID, "plan pointer is null", err, err_len);
    *plan = nullptr;
    // robust_filter (src/pyapi/robust.rs:98-116): the omission options need
    // NaN omission, then the kind.
    if (!omit_nan && (ignore_inf || min_count != 1)) {
        return fail(NDF_ERR_INVALID, "ignore_inf and min_count require NaN omission", err, err_len);
    }
    if (kind != NDF_ROBUST_PERCENTILE) return fail(NDF_ERR_KIND, "unsupported robust filter kind", err, err_len);
    ndf_plan *p = new (std::nothrow) ndf_plan();
    if (!p) return fail_code(NDF_ERR_MEMORY, err, err_len);
    const ndf_status st = build_plan(p, dtype_bytes, rank, shape, size, footprint, origin, modes, cval,
                                     KIND_PERCENTILE, omit_nan, ignore_inf, omit_nan ? min_count : 1, 0, parameter,
                                     err, err_len);
    if (st != NDF_OK) {
        delete p;
        return st;
    }
    *plan = p;
    return NDF_OK;
}

void ndf_plan_destroy(ndf_plan *plan) { delete plan; }

ndf_status ndf_plan_set_engine(ndf_plan *plan, int32_t engine, char *err, size_t err_len)
{
    std::lock_guard<std::mutex> guard(plan->mutex);
    int32_t chosen;
    if (engine == NDF_ENGINE_AUTO) {
        chosen = auto_engine(plan);
    } else if (engine == NDF_ENGINE_STACK) {
        if (!plan->stack_ok) {
            return fail(NDF_ERR_INVALID,
                        "the stack engine needs a box window of rank 2 or 3 with compiled slice networks "
                        "(2-D 3x3, 5x5, 7x7, 9x9; 3-D 3x3x3, 3x5x5, 3x7x7)",
                        err, err_len);
        }
        chosen = NDF_ENGINE_STACK;
    } else if (engine == NDF_ENGINE_EXTREMA) {
        if (!plan->extrema_ok) {
            return fail(NDF_ERR_INVALID,
                        "the extrema engine needs a rectangular minimum or maximum of rank 3 or less without a "
                        "min_count above 1",
                        err, err_len);
        }
        chosen = NDF_ENGINE_EXTREMA;
    } else if (engine == NDF_ENGINE_GENERIC) {
        chosen = NDF_ENGINE_GENERIC;
    } else if (engine == NDF_ENGINE_SLIDE) {
        if (!plan->slide_ok) return fail(NDF_ERR_INVALID, "the slide engine cannot stage this geometry", err, err_len);
        chosen = NDF_ENGINE_SLIDE;
    } else if (engine == NDF_ENGINE_NETWORK) {
        if (!plan->network_ok) {
            return fail(NDF_ERR_INVALID, "the network engine needs a window of at most 160 taps (64 for float64)",
                        err, err_len);
        }
        chosen = NDF_ENGINE_NETWORK;
    } else {
        return fail(NDF_ERR_INVALID, "unknown engine", err, err_len);
    }
    if (chosen != plan->engine) {
        plan->engine = chosen;
        plan->tables_ready = false;
    }
    return NDF_OK;
}

ndf_status ndf_plan_describe(const ndf_plan *plan, char *buffer, size_t buffer_len)
{
    const char *engine = plan->all_nan                       ? "fill"
                         : plan->engine == NDF_ENGINE_SLIDE   ? "slide"
                         : plan->engine == NDF_ENGINE_NETWORK ? "network"
                         : plan->engine == NDF_ENGINE_STACK   ? "stack"
                         : plan->engine == NDF_ENGINE_EXTREMA ? "extrema"
                                                              : "generic";
    if (plan->engine == NDF_ENGINE_EXTREMA) {
        snprintf(buffer, buffer_len,
                 "{\"engine\": \"%s\", \"count\": %zu, \"tile\": [%d, %d], \"support\": [%d, %d, %d], "
                 "\"lds_bytes\": %zu, \"tiles\": %lld}",
                 engine, plan->count, plan->extrema.ty, plan->extrema.tx, plan->extrema.kz, plan->extrema.sy,
                 plan->extrema.sx, plan->extrema_lds,
                 (long long)plan->extrema.tiles_x * plan->extrema.tiles_y * plan->extrema.nz);
        return NDF_OK;
    }
    if (plan->engine == NDF_ENGINE_STACK) {
        const StackArgs &k = plan->stack;
        snprintf(buffer, buffer_len,
                 "{\"engine\": \"%s\", \"count\": %zu, \"slices\": [%d, %d, %d], \"run\": %d, \"units\": %lld, "
                 "\"lds_bytes\": %zu, \"bracket_tile_rows\": %d}",
                 engine, plan->count, k.k0, k.sh, k.sw, k.run, (long long)k.units,
                 stack_lds_bytes(plan->dbytes, k.sh, k.sx), plan->bracket_ty);
        return NDF_OK;
    }
    if (plan->engine == NDF_ENGINE_NETWORK) {
        snprintf(buffer, buffer_len,
                 "{\"engine\": \"%s\", \"count\": %zu, \"tile\": [%d, 64], \"network_wires\": %d, "
                 "\"row_pair\": [%d, %d], \"lds_bytes\": %zu, \"tiles\": %lld, \"bracket_tile_rows\": %d}",
                 engine, plan->count, plan->network.ty, plan->network.live, plan->network.pair_c,
                 plan->network.pair_e, plan->network_lds,
                 (long long)plan->network.tiles_x * plan->network.tiles_y * plan->network.nz, plan->bracket_ty);
    } else if (plan->engine == NDF_ENGINE_SLIDE) {
        snprintf(buffer, buffer_len,
                 "{\"engine\": \"%s\", \"count\": %zu, \"tile\": [%d, %d], \"support\": %d, \"sort_len\": %d, "
                 "\"lds_bytes\": %zu, \"edge_taps\": %d, \"tiles\": %lld, \"bracket_tile_rows\": %d, "
                 "\"bracket\": %s}",
                 engine, plan->count, plan->slide.ty, SLIDE_TILE_X, plan->slide.s, 256 * plan->slide.sort_e, plan->slide_lds,
                 plan->slide.nedge,
                 (long long)plan->slide.tiles_x * plan->slide.tiles_y * plan->slide.nz, plan->bracket_ty,
                 plan->bracket_on ? "true" : "false");
    } else {
        snprintf(buffer, buffer_len, "{\"engine\": \"%s\", \"count\": %zu}", engine, plan->count);
    }
    return NDF_OK;
}

ndf_status ndf_plan_apply_device(ndf_plan *plan, const void *d_input, void *d_output, size_t len, void *stream,
                                 char *err, size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (len == 0) return NDF_OK;
    if (buffers_overlap(plan, d_input, d_output, len)) {
        return fail(NDF_ERR_INVALID, "the output buffer must not overlap the input buffer", err, err_len);
    }
    std::lock_guard<std::mutex> guard(plan->mutex);
    const ndf_status st = prepare_tables(plan, err, err_len);
    if (st != NDF_OK) return st;
    return launch(plan, d_input, d_output, static_cast<hipStream_t>(stream), err, err_len);
}

namespace {

ndf_status read_error_flag(ndf_plan *plan, char *err, size_t err_len);

}  // namespace

ndf_status ndf_plan_last_error(ndf_plan *plan, char *err, size_t err_len)
{
    std::lock_guard<std::mutex> guard(plan->mutex);
    return read_error_flag(plan, err, err_len);
}

ndf_status ndf_plan_set_output_fill(ndf_plan *plan, int32_t enabled, uint64_t bits)
{
    std::lock_guard<std::mutex> guard(plan->mutex);
    plan->output_fill = enabled != 0;
    plan->output_fill_bits = bits;
    return NDF_OK;
}

namespace {

// Waits for the plan's last launch and reports a rank walk that left its
// bounds (the slide kernels' internal consistency check).
ndf_status read_error_flag(ndf_plan *plan, char *err, size_t err_len)
{
    hipError_t e = hipSuccess;
    if (plan->done_recorded) e = hipEventSynchronize(plan->done);
    int32_t flag = 0;
    if (e == hipSuccess && plan->d_error.ptr) {
        e = hipMemcpy(&flag, plan->d_error.ptr, sizeof(flag), hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return fail_hip(e, err, err_len);
    if (flag) return fail(NDF_ERR_DEVICE, "internal error: rank walk left its bounds", err, err_len);
    return NDF_OK;
}

ndf_status upload_locked(ndf_plan *plan, const void *input, char *err, size_t err_len)
{
    ndf_status st = prepare_tables(plan, err, err_len);
    if (st != NDF_OK) return st;
    const size_t bytes = plan->len * (size_t)plan->dbytes;
    hipError_t e;
    if ((e = plan->d_in.ensure(bytes)) != hipSuccess || (e = plan->d_out.ensure(bytes)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    if ((e = staged_upload(plan->staging, plan->d_in.ptr, input, bytes, nullptr)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    plan->uploaded = true;
    return NDF_OK;
}

ndf_status compute_download_locked(ndf_plan *plan, void *output, char *err, size_t err_len)
{
    if (!plan->uploaded) return fail(NDF_ERR_INVALID, "no input has been uploaded to this plan", err, err_len);
    plan->uploaded = false;
    const size_t bytes = plan->len * (size_t)plan->dbytes;
    hipError_t e;
    if (plan->output_fill &&
        (e = launch_fill(plan->dbytes, plan->d_out.ptr, (int64_t)plan->len, plan->output_fill_bits, nullptr)) !=
            hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    ndf_status st = launch(plan, plan->d_in.ptr, plan->d_out.ptr, nullptr, err, err_len);
    if (st != NDF_OK) return st;
    if ((e = staged_download(plan->staging, output, plan->d_out.ptr, bytes, nullptr)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    return read_error_flag(plan, err, err_len);
}

}  // namespace

ndf_status ndf_plan_upload(ndf_plan *plan, const void *input, size_t len, char *err, size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (len == 0) return NDF_OK;
    std::lock_guard<std::mutex> guard(plan->mutex);
    return upload_locked(plan, input, err, err_len);
}

ndf_status ndf_plan_compute_download(ndf_plan *plan, void *output, size_t len, char *err, size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (len == 0) return NDF_OK;
    std::lock_guard<std::mutex> guard(plan->mutex);
    return compute_download_locked(plan, output, err, err_len);
}

ndf_status ndf_plan_apply(ndf_plan *plan, const void *input, void *output, size_t len, char *err, size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (len == 0) return NDF_OK;
    std::lock_guard<std::mutex> guard(plan->mutex);
    const ndf_status st = upload_locked(plan, input, err, err_len);
    if (st != NDF_OK) return st;
    return compute_download_locked(plan, output, err, err_len);
}

ndf_status ndf_window_filter(int32_t dtype_bytes, const void *input, void *output, size_t rank,
                             const size_t *shape, const size_t *size, const uint8_t *footprint,
                             const ptrdiff_t *origin, const int32_t *modes, double cval, int32_t kind,
                             int32_t omit_nan, int32_t ignore_inf, size_t min_count, size_t parameter, char *err,
                             size_t err_len)
{
    ndf_plan *plan = nullptr;
    ndf_status st = ndf_plan_create(&plan, dtype_bytes, rank, shape, size, footprint, origin, modes, cval, kind,
                                    omit_nan, ignore_inf, min_count, parameter, err, err_len);
    if (st != NDF_OK) return st;
    st = ndf_plan_apply(plan, input, output, plan->len, err, err_len);
    ndf_plan_destroy(plan);
    return st;
}

ndf_status ndf_robust_filter(int32_t dtype_bytes, const void *input, void *output, size_t rank, const size_t *shape,
                             const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin,
                             const int32_t *modes, double cval, int32_t kind, double parameter, int32_t omit_nan,
                             int32_t ignore_inf, size_t min_count, char *err, size_t err_len)
{
    ndf_plan *plan = nullptr;
    ndf_status st = ndf_robust_plan_create(&plan, dtype_bytes, rank, shape, size, footprint, origin, modes, cval,
                                           kind, parameter, omit_nan, ignore_inf, min_count, err, err_len);
    if (st != NDF_OK) return st;
    st = ndf_plan_apply(plan, input, output, plan->len, err, err_len);
    ndf_plan_destroy(plan);
    return st;
}

ndf_status ndf_plan_set_bracket(ndf_plan *plan, const float *weights, size_t n_weights, int32_t sort_len, char *err,
                                size_t err_len)
{
    std::lock_guard<std::mutex> guard(plan->mutex);
    if (weights == nullptr) {
        plan->bracket_on = false;
        return NDF_OK;
    }
    if (n_weights != (size_t)PARAMS) return fail(NDF_ERR_INVALID, "bracket model must have 1346 parameters", err, err_len);
    if (!bracket_supported(plan)) {
        return fail(NDF_ERR_INVALID,
                    "the learned bracket needs a 2-D image, a rectangular window at most 64 wide and the slide engine",
                    err, err_len);
    }
    int e = 4;
    while (256 * e < sort_len && e < 64) e <<= 1;
    if (sort_len <= 0 || 256 * e != sort_len || e > 32) {
        return fail(NDF_ERR_INVALID, "bracket sort length must be 1024, 2048, 4096 or 8192", err, err_len);
    }
    const SlideArgs geometry = with_tile_rows(plan->slide, plan->bracket_ty);
    if (bracket_lds_bytes(plan->dbytes, geometry.s, geometry.rows, e) > (size_t)lds_limit()) {
        return fail(NDF_ERR_INVALID, "bracket sort length does not fit LDS", err, err_len);
    }
    plan->bracket_weights.assign(weights, weights + n_weights);
    plan->bracket_e = e;
    plan->bracket_on = true;
    plan->bracket_ready = false;
    // The bracket is a mode of the slide engine (its fallback is the plain
    // slide kernel on the same tiles).
    if (plan->engine != NDF_ENGINE_SLIDE) {
        plan->engine = NDF_ENGINE_SLIDE;
        plan->tables_ready = false;
    }
    return NDF_OK;
}

ndf_status ndf_plan_bracket_report(ndf_plan *plan, int64_t *tiles, int64_t *fallback_tiles)
{
    std::lock_guard<std::mutex> guard(plan->mutex);
    *tiles = bracket_supported(plan) ? bracket_tiles(plan) : 0;
    *fallback_tiles = 0;
    if (!plan->bracket_launched || !plan->report_fcount) return NDF_OK;
    int32_t count = 0;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&count, plan->report_fcount, sizeof(count), hipMemcpyDeviceToHost) != hipSuccess) {
        return NDF_ERR_DEVICE;
    }
    *fallback_tiles = count;
    return NDF_OK;
}

size_t ndf_plan_bracket_tiles(const ndf_plan *plan)
{
    return bracket_supported(plan) ? (size_t)bracket_tiles(plan) : 0;
}

ndf_status ndf_plan_bracket_training_data(ndf_plan *plan, const void *d_input, const void *d_output, float *features,
                                          float *targets, float *weight, char *err, size_t err_len)
{
    std::lock_guard<std::mutex> guard(plan->mutex);
    if (!bracket_supported(plan)) {
        return fail(NDF_ERR_INVALID,
                    "the learned bracket needs a 2-D image, a rectangular window at most 64 wide and the slide engine",
                    err, err_len);
    }
    ndf_status st = prepare_tables(plan, err, err_len);
    if (st != NDF_OK) return st;
    const bool was_on = plan->bracket_on;
    plan->bracket_on = false;
    st = prepare_bracket(plan, err, err_len);
    plan->bracket_on = was_on;
    if (st != NDF_OK) return st;
    const int64_t tiles = bracket_tiles(plan);
    DeviceBuffer d_targets, d_weight;
    hipError_t e;
    if ((e = d_targets.ensure((size_t)tiles * 2 * sizeof(float))) != hipSuccess ||
        (e = d_weight.ensure((size_t)tiles * sizeof(float))) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    // Features only: block moments, then per-tile features.
    e = launch_block_stats(plan->dbytes, d_input, plan->slide.ny, plan->slide.nx, plan->eff_ignore_inf,
                           static_cast<double *>(plan->b_blocks.ptr), nullptr);
    if (e == hipSuccess) {
        e = launch_tile_features(plan->bargs, static_cast<const double *>(plan->b_blocks.ptr), tiles, nullptr);
    }
    if (e == hipSuccess) {
        e = launch_tile_labels(plan->dbytes, plan->bargs, d_output, static_cast<float *>(d_targets.ptr),
                               static_cast<float *>(d_weight.ptr), tiles, nullptr);
    }
    if (e == hipSuccess) {
        e = hipMemcpy(features, plan->b_features.ptr, (size_t)tiles * FEATURES * sizeof(float), hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess) {
        e = hipMemcpy(targets, d_targets.ptr, (size_t)tiles * 2 * sizeof(float), hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess) {
        e = hipMemcpy(weight, d_weight.ptr, (size_t)tiles * sizeof(float), hipMemcpyDeviceToHost);
    }
    d_targets.release();
    d_weight.release();
    return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
}

ndf_status ndf_mlp_train(float *weights, size_t n_weights, const float *features, const float *targets,
                         const float *weight, size_t count, int32_t steps, float learning_rate, float tau_lo,
                         float tau_hi, float *losses, char *err, size_t err_len)
{
    if (n_weights != (size_t)PARAMS) return fail(NDF_ERR_INVALID, "bracket model must have 1346 parameters", err, err_len);
    if (count == 0 || steps <= 0) return fail(NDF_ERR_INVALID, "training needs samples and steps", err, err_len);
    DeviceBuffer w, m1, m2, x, t, wt, partial, loss_partial, loss;
    const int blocks = mlp_grad_blocks((int64_t)count);
    std::vector<float> zeros(PARAMS, 0.0f);
    hipError_t e;
    if ((e = w.ensure(PARAMS * sizeof(float))) != hipSuccess || (e = m1.ensure(PARAMS * sizeof(float))) != hipSuccess ||
        (e = m2.ensure(PARAMS * sizeof(float))) != hipSuccess ||
        (e = x.ensure(count * FEATURES * sizeof(float))) != hipSuccess ||
        (e = t.ensure(count * 2 * sizeof(float))) != hipSuccess ||
        (e = wt.ensure(count * sizeof(float))) != hipSuccess ||
        (e = partial.ensure((size_t)blocks * PARAMS * sizeof(float))) != hipSuccess ||
        (e = loss_partial.ensure((size_t)blocks * sizeof(float))) != hipSuccess ||
        (e = loss.ensure((size_t)steps * sizeof(float))) != hipSuccess ||
        (e = hipMemcpy(w.ptr, weights, PARAMS * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(m1.ptr, zeros.data(), PARAMS * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(m2.ptr, zeros.data(), PARAMS * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(x.ptr, features, count * FEATURES * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(t.ptr, targets, count * 2 * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(wt.ptr, weight, count * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    for (int step = 1; step <= steps && e == hipSuccess; ++step) {
        e = launch_mlp_step(static_cast<float *>(w.ptr), static_cast<float *>(m1.ptr), static_cast<float *>(m2.ptr),
                            static_cast<const float *>(x.ptr), static_cast<const float *>(t.ptr),
                            static_cast<const float *>(wt.ptr), nullptr, (int64_t)count, tau_lo, tau_hi,
                            learning_rate, step, static_cast<float *>(partial.ptr),
                            static_cast<float *>(loss_partial.ptr), static_cast<float *>(loss.ptr) + (step - 1),
                            nullptr);
    }
    if (e == hipSuccess) e = hipMemcpy(weights, w.ptr, PARAMS * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && losses) {
        e = hipMemcpy(losses, loss.ptr, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost);
    }
    return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
}

// ---------------------------------------------------------------------------
// Batches of 2-D images, each with its own window size and boundary modes
// (no reference counterpart: the reference filters one array per call). One
// persistent launch of the batched plain kernel covers every tile of every
// image through a work queue; the launch is captured once in a hipGraph and
// replayed. ndf_batch_autotune picks the sort length and the tile-row cap.
// ---------------------------------------------------------------------------

}  // extern "C"

struct ndf_batch {
    int32_t dbytes = 4;
    std::vector<ndf_plan *> plans;
    std::vector<int64_t> offsets;  // first element of each image
    int64_t total = 0;
    int sort_e = 8, max_ty = 16;
    bool use_graph = true;
    bool tuned = false;
    // Device state for the current configuration.
    bool ready = false;
    std::vector<SlideArgs> images;
    std::vector<int32_t> fill_images;  // images whose every output is NaN
    DeviceBuffer d_images, d_offsets, d_work, d_counter, d_error, d_maps, d_taps;
    int nwork = 0;
    size_t lds = 0;
    int grid = 0;
    // Images with small windows run on the batched network kernel, one
    // group (and launch) per network size.
    struct NetworkGroup {
        int live = 0;
        std::vector<NetworkArgs> args;
        std::vector<int64_t> offsets;
        std::vector<int32_t> pairs;
        DeviceBuffer d_args, d_offsets, d_work;
        size_t lds = 0;
        int grid = 0;
        int nwork = 0;
    };
    std::vector<NetworkGroup> groups;
    // Images with box windows the stack engine has networks for: one group
    // (and launch) per box and statistic.
    struct StackGroup {
        StackArgs shape{};
        std::vector<StackArgs> args;
        std::vector<int64_t> offsets;
        std::vector<int32_t> pairs;
        DeviceBuffer d_args, d_offsets, d_work;
        int grid = 0;
        int nwork = 0;
    };
    std::vector<StackGroup> sgroups;
    // Captured launch (for one input/output pair).
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *graph_in = nullptr;
    void *graph_out = nullptr;
    std::mutex mutex;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_in = nullptr;
        graph_out = nullptr;
    }
    ~ndf_batch()
    {
        drop_graph();
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        for (ndf_plan *p : plans) ndf_plan_destroy(p);
        for (DeviceBuffer *b : {&d_images, &d_offsets, &d_work, &d_counter, &d_error, &d_maps, &d_taps}) b->release();
        for (NetworkGroup &g : groups) {
            for (DeviceBuffer *b : {&g.d_args, &g.d_offsets, &g.d_work}) b->release();
        }
        for (StackGroup &g : sgroups) {
            for (DeviceBuffer *b : {&g.d_args, &g.d_offsets, &g.d_work}) b->release();
        }
    }
};

namespace {

// Plans every image for (sort_e, max_ty), pools the maps, builds the work
// queue (largest estimated tile cost first) and uploads it all.
ndf_status batch_prepare(ndf_batch *b, char *err, size_t err_len)
{
    if (b->ready) return NDF_OK;
    b->drop_graph();
    std::vector<int64_t> maps;
    std::vector<std::array<int64_t, 3>> map_at;
    b->images.assign(b->plans.size(), SlideArgs{});
    b->fill_images.clear();
    size_t lds = 0;
    struct Work {
        int64_t cost;
        int32_t image, tile;
    };
    std::vector<Work> work;
    for (ndf_batch::NetworkGroup &g : b->groups) {
        for (DeviceBuffer *buffer : {&g.d_args, &g.d_offsets, &g.d_work}) buffer->release();
    }
    b->groups.clear();
    for (ndf_batch::StackGroup &g : b->sgroups) {
        for (DeviceBuffer *buffer : {&g.d_args, &g.d_offsets, &g.d_work}) buffer->release();
    }
    b->sgroups.clear();
    std::vector<std::array<int64_t, 4>> stack_at;  // image, maps offset, group, index in group
    std::vector<uint32_t> taps;
    std::vector<std::array<int64_t, 5>> net_at;  // image, maps offset, taps offset, group, index in group
    for (size_t i = 0; i < b->plans.size(); ++i) {
        ndf_plan *p = b->plans[i];
        map_at.push_back({0, 0, 0});
        if (p->len == 0) continue;
        if (p->all_nan) {
            b->fill_images.push_back((int32_t)i);
            continue;
        }
        if (p->stack_ok) {
            const StackArgs &k = p->stack;
            size_t gi = 0;
            while (gi < b->sgroups.size() &&
                   !(b->sgroups[gi].shape.k0 == k.k0 && b->sgroups[gi].shape.sh == k.sh &&
                     b->sgroups[gi].shape.sw == k.sw && b->sgroups[gi].shape.stat == k.stat)) {
                ++gi;
            }
            if (gi == b->sgroups.size()) {
                b->sgroups.emplace_back();
                b->sgroups.back().shape = k;
            }
            ndf_batch::StackGroup &g = b->sgroups[gi];
            stack_at.push_back({(int64_t)i, (int64_t)maps.size(), (int64_t)gi, (int64_t)g.args.size()});
            maps.insert(maps.end(), p->stk_maps.begin(), p->stk_maps.end());
            maps.insert(maps.end(), p->stk_mapr.begin(), p->stk_mapr.end());
            maps.insert(maps.end(), p->stk_mapx.begin(), p->stk_mapx.end());
            const int index = (int)g.args.size();
            g.args.push_back(k);
            g.offsets.push_back(b->offsets[i]);
            const int64_t blocks = (k.units + 3) / 4;
            for (int64_t t = 0; t < blocks; ++t) {
                g.pairs.push_back(index);
                g.pairs.push_back((int32_t)t);
            }
            continue;
        }
        if (p->network_ok) {
            const int live = p->network.live;
            size_t gi = 0;
            while (gi < b->groups.size() && b->groups[gi].live != live) ++gi;
            if (gi == b->groups.size()) {
                b->groups.emplace_back();
                b->groups.back().live = live;
            }
            ndf_batch::NetworkGroup &g = b->groups[gi];
            net_at.push_back({(int64_t)i, (int64_t)maps.size(), (int64_t)taps.size(), (int64_t)gi,
                              (int64_t)g.args.size()});
            maps.insert(maps.end(), p->net_mapz.begin(), p->net_mapz.end());
            maps.insert(maps.end(), p->net_mapy.begin(), p->net_mapy.end());
            maps.insert(maps.end(), p->net_mapx.begin(), p->net_mapx.end());
            taps.insert(taps.end(), p->net_taps.begin(), p->net_taps.end());
            const int index = (int)g.args.size();
            g.args.push_back(p->network);
            g.offsets.push_back(b->offsets[i]);
            g.lds = std::max(g.lds, p->network_lds);
            const int tiles = p->network.tiles_x * p->network.tiles_y * (int)p->network.nz;
            for (int t = 0; t < tiles; ++t) {
                g.pairs.push_back(index);
                g.pairs.push_back(t);
            }
            continue;
        }
        if (!plan_slide(p, b->sort_e, b->max_ty)) {
            char msg[160];
            snprintf(msg, sizeof msg, "batch image %zu: its window does not fit a %d-entry tile sort", i,
                     256 * b->sort_e);
            return fail(NDF_ERR_INVALID, msg, err, err_len);
        }
        map_at.back() = {(int64_t)maps.size(), (int64_t)(maps.size() + p->slide_mapz.size()),
                         (int64_t)(maps.size() + p->slide_mapz.size() + p->slide_mapy.size())};
        maps.insert(maps.end(), p->slide_mapz.begin(), p->slide_mapz.end());
        maps.insert(maps.end(), p->slide_mapy.begin(), p->slide_mapy.end());
        maps.insert(maps.end(), p->slide_mapx.begin(), p->slide_mapx.end());
        b->images[i] = p->slide;
        lds = std::max(lds, p->slide_lds);
        const SlideArgs &a = p->slide;
        const int tiles = a.tiles_x * a.tiles_y;
        const int64_t cost = (int64_t)a.ty * SLIDE_TILE_X * (2 * a.kx + 8) + (int64_t)a.s * 12;
        for (int t = 0; t < tiles; ++t) work.push_back({cost, (int32_t)i, (int32_t)t});
    }
    std::stable_sort(work.begin(), work.end(), [](const Work &x, const Work &y) { return x.cost > y.cost; });
    std::vector<int32_t> pairs;
    pairs.reserve(work.size() * 2);
    for (const Work &w : work) {
        pairs.push_back(w.image);
        pairs.push_back(w.tile);
    }
    hipError_t e;
    if ((e = upload(b->d_maps, maps)) != hipSuccess || (e = upload(b->d_taps, taps)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    const int64_t *base = static_cast<const int64_t *>(b->d_maps.ptr);
    for (const std::array<int64_t, 5> &at : net_at) {
        const ndf_plan *p = b->plans[(size_t)at[0]];
        NetworkArgs &n = b->groups[(size_t)at[3]].args[(size_t)at[4]];
        n.mapz = base + at[1];
        n.mapy = n.mapz + p->net_mapz.size();
        n.mapx = n.mapy + p->net_mapy.size();
        n.taps = static_cast<const uint32_t *>(b->d_taps.ptr) + at[2];
        n.input = nullptr;
        n.output = nullptr;
    }
    for (const std::array<int64_t, 4> &at : stack_at) {
        const ndf_plan *p = b->plans[(size_t)at[0]];
        StackArgs &k = b->sgroups[(size_t)at[2]].args[(size_t)at[3]];
        k.maps = base + at[1];
        k.mapr = k.maps + p->stk_maps.size();
        k.mapx = k.mapr + p->stk_mapr.size();
        k.input = nullptr;
        k.output = nullptr;
    }
    for (ndf_batch::StackGroup &g : b->sgroups) {
        g.nwork = (int)(g.pairs.size() / 2);
        if ((e = upload(g.d_args, g.args)) != hipSuccess || (e = upload(g.d_offsets, g.offsets)) != hipSuccess ||
            (e = upload(g.d_work, g.pairs)) != hipSuccess ||
            (e = stack_batch_grid(b->dbytes, g.shape, &g.grid)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        g.grid = std::max(1, std::min(g.grid, g.nwork));
    }
    for (ndf_batch::NetworkGroup &g : b->groups) {
        g.nwork = (int)(g.pairs.size() / 2);
        if ((e = upload(g.d_args, g.args)) != hipSuccess || (e = upload(g.d_offsets, g.offsets)) != hipSuccess ||
            (e = upload(g.d_work, g.pairs)) != hipSuccess ||
            (e = network_batch_grid(b->dbytes, g.live, g.lds, &g.grid)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        g.grid = std::max(1, std::min(g.grid, g.nwork));
    }
    for (size_t i = 0; i < b->images.size(); ++i) {
        SlideArgs &a = b->images[i];
        if (b->plans[i]->len == 0 || b->plans[i]->all_nan) continue;
        a.mapz = base + map_at[i][0];
        a.mapy = base + map_at[i][1];
        a.mapx = base + map_at[i][2];
        a.taps = nullptr;
        a.edge_remove = nullptr;
        a.edge_add = nullptr;
        a.mask = nullptr;
        a.input = nullptr;
        a.output = nullptr;
    }
    if ((e = upload(b->d_images, b->images)) != hipSuccess || (e = upload(b->d_offsets, b->offsets)) != hipSuccess ||
        (e = upload(b->d_work, pairs)) != hipSuccess ||
        (e = b->d_counter.ensure(sizeof(int32_t) * (1 + b->groups.size() + b->sgroups.size()))) != hipSuccess ||
        (e = b->d_error.ensure(sizeof(int32_t))) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    b->nwork = (int)work.size();
    b->lds = lds;
    b->grid = 1;
    if (b->nwork > 0 && (e = slide_batch_grid(b->dbytes, b->sort_e, lds, &b->grid)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    b->grid = std::max(1, std::min(b->grid, b->nwork));
    b->ready = true;
    return NDF_OK;
}

// Enqueues the batch on `stream`: NaN fills for all-NaN images, the queue
// counter reset and the batched kernel.
hipError_t batch_enqueue(ndf_batch *b, const void *d_in, void *d_out, hipStream_t stream)
{
    const uint64_t nan_bits = b->dbytes == 4 ? 0x7FC00000ull : 0x7FF8000000000000ull;
    hipError_t e = hipSuccess;
    for (int32_t i : b->fill_images) {
        if (e != hipSuccess) break;
        void *out = static_cast<char *>(d_out) + b->offsets[(size_t)i] * b->dbytes;
        e = launch_fill(b->dbytes, out, (int64_t)b->plans[(size_t)i]->len, nan_bits, stream);
    }
    if (e == hipSuccess) {
        e = launch_fill(4, b->d_counter.ptr, (int64_t)(1 + b->groups.size() + b->sgroups.size()), 0, stream);
    }
    for (size_t gi = 0; gi < b->sgroups.size() && e == hipSuccess; ++gi) {
        const ndf_batch::StackGroup &g = b->sgroups[gi];
        if (g.nwork == 0) continue;
        e = launch_stack_batch(b->dbytes, g.shape, static_cast<const StackArgs *>(g.d_args.ptr),
                               static_cast<const int64_t *>(g.d_offsets.ptr),
                               static_cast<const int32_t *>(g.d_work.ptr), g.nwork,
                               static_cast<int32_t *>(b->d_counter.ptr) + 1 + b->groups.size() + gi, d_in, d_out,
                               g.grid, stream);
    }
    for (size_t gi = 0; gi < b->groups.size() && e == hipSuccess; ++gi) {
        const ndf_batch::NetworkGroup &g = b->groups[gi];
        if (g.nwork == 0) continue;
        e = launch_network_batch(b->dbytes, g.live, static_cast<const NetworkArgs *>(g.d_args.ptr),
                                 static_cast<const int64_t *>(g.d_offsets.ptr),
                                 static_cast<const int32_t *>(g.d_work.ptr), g.nwork,
                                 static_cast<int32_t *>(b->d_counter.ptr) + 1 + gi, d_in, d_out, g.lds, g.grid,
                                 stream);
    }
    if (e == hipSuccess && b->nwork > 0) {
        e = launch_slide_batch(b->dbytes, b->sort_e, static_cast<const SlideArgs *>(b->d_images.ptr),
                               static_cast<const int64_t *>(b->d_offsets.ptr),
                               static_cast<const int32_t *>(b->d_work.ptr), b->nwork,
                               static_cast<int32_t *>(b->d_counter.ptr), d_in, d_out,
                               static_cast<int32_t *>(b->d_error.ptr), b->lds, b->grid, stream);
    }
    return e;
}

// Runs the batch once: replays the captured graph (capturing it first for a
// new input/output pair) or enqueues the kernels directly.
ndf_status batch_launch(ndf_batch *b, const void *d_in, void *d_out, hipStream_t stream, char *err, size_t err_len)
{
    ndf_status st = batch_prepare(b, err, err_len);
    if (st != NDF_OK) return st;
    hipError_t e = hipSuccess;
    if (!b->use_graph) {
        e = batch_enqueue(b, d_in, d_out, stream);
        return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
    }
    if (!b->exec || b->graph_in != d_in || b->graph_out != d_out) {
        b->drop_graph();
        if (!b->capture_stream && (e = hipStreamCreate(&b->capture_stream)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        if ((e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
        const hipError_t enqueue = batch_enqueue(b, d_in, d_out, b->capture_stream);
        e = hipStreamEndCapture(b->capture_stream, &b->graph);
        if (enqueue != hipSuccess) e = enqueue;
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            b->drop_graph();
            return fail_hip(e, err, err_len);
        }
        b->graph_in = d_in;
        b->graph_out = d_out;
    }
    e = hipGraphLaunch(b->exec, stream);
    return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
}

// Mean milliseconds per batch launch over `iterations` (HIP events).
ndf_status batch_time(ndf_batch *b, const void *d_in, void *d_out, int warmup, int iterations, double *ms, char *err,
                      size_t err_len)
{
    ndf_status st = batch_prepare(b, err, err_len);
    if (st != NDF_OK) return st;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    hipError_t e;
    if ((e = hipStreamCreate(&stream)) != hipSuccess) return fail_hip(e, err, err_len);
    if ((e = hipEventCreate(&start)) != hipSuccess || (e = hipEventCreate(&stop)) != hipSuccess) {
        st = fail_hip(e, err, err_len);
    }
    if (st == NDF_OK && (e = launch_fill(4, b->d_error.ptr, 1, 0, stream)) != hipSuccess) st = fail_hip(e, err, err_len);
    for (int i = 0; st == NDF_OK && i < warmup; ++i) st = batch_launch(b, d_in, d_out, stream, err, err_len);
    if (st == NDF_OK && (e = hipEventRecord(start, stream)) != hipSuccess) st = fail_hip(e, err, err_len);
    for (int i = 0; st == NDF_OK && i < iterations; ++i) st = batch_launch(b, d_in, d_out, stream, err, err_len);
    if (st == NDF_OK && (e = hipEventRecord(stop, stream)) != hipSuccess) st = fail_hip(e, err, err_len);
    if (st == NDF_OK && (e = hipEventSynchronize(stop)) != hipSuccess) st = fail_hip(e, err, err_len);
    float elapsed = 0.0f;
    if (st == NDF_OK && (e = hipEventElapsedTime(&elapsed, start, stop)) != hipSuccess) st = fail_hip(e, err, err_len);
    if (st == NDF_OK) {
        *ms = iterations > 0 ? (double)elapsed / iterations : 0.0;
        int32_t flag = 0;
        if ((e = hipMemcpy(&flag, b->d_error.ptr, sizeof(flag), hipMemcpyDeviceToHost)) != hipSuccess) {
            st = fail_hip(e, err, err_len);
        } else if (flag) {
            st = fail(NDF_ERR_DEVICE, "internal error: rank walk left its bounds", err, err_len);
        }
    }
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
    (void)hipStreamDestroy(stream);
    return st;
}

}  // namespace

extern "C" {

ndf_status ndf_batch_create(ndf_batch **batch, int32_t dtype_bytes, size_t count, const size_t *shapes,
                            const size_t *sizes, const int32_t *modes, double cval, int32_t kind, int32_t omit_nan,
                            int32_t ignore_inf, size_t min_count, size_t parameter, char *err, size_t err_len)
{
    if (!batch) return fail(NDF_ERR_INVALID, "batch pointer is null", err, err_len);
    *batch = nullptr;
    if (count == 0) return fail(NDF_ERR_INVALID, "a batch needs at least one image", err, err_len);
    ndf_batch *b = new (std::nothrow) ndf_batch();
    if (!b) return fail_code(NDF_ERR_MEMORY, err, err_len);
    b->dbytes = dtype_bytes;
    for (size_t i = 0; i < count; ++i) {
        ndf_plan *p = new (std::nothrow) ndf_plan();
        if (!p) {
            delete b;
            return fail_code(NDF_ERR_MEMORY, err, err_len);
        }
        b->plans.push_back(p);
        const ptrdiff_t origin[2] = {0, 0};
        char inner[200] = {0};
        const ndf_status st = build_plan(p, dtype_bytes, 2, shapes + 2 * i, sizes + 2 * i, nullptr, origin,
                                         modes + 2 * i, cval, kind, omit_nan, ignore_inf, min_count, parameter,
                                         0.0, inner, sizeof inner);
        if (st != NDF_OK) {
            char msg[260];
            snprintf(msg, sizeof msg, "batch image %zu: %s", i, inner);
            delete b;
            return fail(st, msg, err, err_len);
        }
        b->offsets.push_back(b->total);
        b->total += (int64_t)p->len;
    }
    *batch = b;
    return NDF_OK;
}

size_t ndf_batch_length(const ndf_batch *batch) { return batch ? (size_t)batch->total : 0; }

ndf_status ndf_batch_configure(ndf_batch *batch, int32_t sort_len, int32_t max_tile_rows, int32_t use_graph,
                               char *err, size_t err_len)
{
    std::lock_guard<std::mutex> guard(batch->mutex);
    int e = 4;
    while (256 * e < sort_len && e < 64) e <<= 1;
    if (sort_len <= 0 || 256 * e != sort_len) {
        return fail(NDF_ERR_INVALID, "batch sort length must be 1024, 2048, 4096, 8192 or 16384", err, err_len);
    }
    if (max_tile_rows < SLIDE_WAVES || max_tile_rows > 256 || max_tile_rows % SLIDE_WAVES != 0) {
        return fail(NDF_ERR_INVALID, "batch tile rows must be a multiple of 4 in [4, 256]", err, err_len);
    }
    if (e != batch->sort_e || max_tile_rows != batch->max_ty) batch->ready = false;
    batch->sort_e = e;
    batch->max_ty = max_tile_rows;
    batch->use_graph = use_graph != 0;
    if (!batch->use_graph) batch->drop_graph();
    return NDF_OK;
}

ndf_status ndf_batch_apply_device(ndf_batch *batch, const void *d_input, void *d_output, size_t len, void *stream,
                                  char *err, size_t err_len)
{
    if (len != (size_t)batch->total) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (len == 0) return NDF_OK;
    if (ranges_overlap(d_input, d_output, len * (size_t)batch->dbytes)) {
        return fail(NDF_ERR_INVALID, "the output buffer must not overlap the input buffer", err, err_len);
    }
    std::lock_guard<std::mutex> guard(batch->mutex);
    return batch_launch(batch, d_input, d_output, static_cast<hipStream_t>(stream), err, err_len);
}

ndf_status ndf_batch_time_device(ndf_batch *batch, const void *d_input, void *d_output, size_t len, int32_t warmup,
                                 int32_t iterations, double *ms_per_launch, char *err, size_t err_len)
{
    if (len != (size_t)batch->total) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    std::lock_guard<std::mutex> guard(batch->mutex);
    return batch_time(batch, d_input, d_output, warmup, iterations, ms_per_launch, err, err_len);
}

ndf_status ndf_batch_autotune(ndf_batch *batch, const void *d_input, void *d_output, size_t len, int32_t repeats,
                              char *err, size_t err_len)
{
    if (len != (size_t)batch->total) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    std::lock_guard<std::mutex> guard(batch->mutex);
    const int sort_es[] = {4, 8, 16};
    const int rows[] = {4, 8, 16, 32, 64};
    double best = -1.0;
    int best_e = batch->sort_e, best_rows = batch->max_ty;
    char inner[200];
    for (int e : sort_es) {
        for (int r : rows) {
            batch->sort_e = e;
            batch->max_ty = r;
            batch->ready = false;
            double ms = 0.0;
            // A configuration some image cannot use is skipped.
            if (batch_time(batch, d_input, d_output, 1, repeats > 0 ? repeats : 3, &ms, inner, sizeof inner) !=
                NDF_OK) {
                continue;
            }
            if (best < 0.0 || ms < best) {
                best = ms;
                best_e = e;
                best_rows = r;
            }
        }
    }
    batch->sort_e = best_e;
    batch->max_ty = best_rows;
    batch->ready = false;
    if (best < 0.0) return fail(NDF_ERR_INVALID, "no batch configuration fits every image", err, err_len);
    batch->tuned = true;
    return batch_prepare(batch, err, err_len);
}

ndf_status ndf_batch_describe(ndf_batch *batch, char *buffer, size_t buffer_len)
{
    std::lock_guard<std::mutex> guard(batch->mutex);
    char err[200];
    const ndf_status st = batch_prepare(batch, err, sizeof err);
    snprintf(buffer, buffer_len,
             "{\"images\": %zu, \"elements\": %lld, \"sort_len\": %d, \"max_tile_rows\": %d, \"graph\": %s, "
             "\"tuned\": %s, \"work_tiles\": %d, \"lds_bytes\": %zu, \"workgroups\": %d, \"ok\": %s, "
             "\"network_groups\": [",
             batch->plans.size(), (long long)batch->total, 256 * batch->sort_e, batch->max_ty,
             batch->use_graph ? "true" : "false", batch->tuned ? "true" : "false", batch->nwork, batch->lds,
             batch->grid, st == NDF_OK ? "true" : "false");
    for (size_t gi = 0; gi < batch->groups.size(); ++gi) {
        const size_t used = strlen(buffer);
        if (used + 96 >= buffer_len) break;
        snprintf(buffer + used, buffer_len - used, "%s{\"wires\": %d, \"images\": %zu, \"tiles\": %d}",
                 gi ? ", " : "", batch->groups[gi].live, batch->groups[gi].args.size(), batch->groups[gi].nwork);
    }
    size_t used = strlen(buffer);
    if (used + 24 < buffer_len) snprintf(buffer + used, buffer_len - used, "], \"stack_groups\": [");
    for (size_t gi = 0; gi < batch->sgroups.size(); ++gi) {
        used = strlen(buffer);
        if (used + 96 >= buffer_len) break;
        const ndf_batch::StackGroup &g = batch->sgroups[gi];
        snprintf(buffer + used, buffer_len - used, "%s{\"slices\": [%d, %d, %d], \"images\": %zu, \"blocks\": %d}",
                 gi ? ", " : "", g.shape.k0, g.shape.sh, g.shape.sw, g.args.size(), g.nwork);
    }
    used = strlen(buffer);
    if (used + 3 < buffer_len) snprintf(buffer + used, buffer_len - used, "]}");
    return st;
}

void ndf_batch_destroy(ndf_batch *batch) { delete batch; }

size_t ndf_get_num_threads(void)
{
    const char *env = getenv("RAYON_NUM_THREADS");
    if (env && *env) {
        const long v = strtol(env, nullptr, 10);
        if (v > 0) return (size_t)v;
    }
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) {
        const int n = CPU_COUNT(&set);
        if (n > 0) return (size_t)n;
    }
    const long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (size_t)n : 1;
}

ndf_status ndf_device_count(int32_t *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return e == hipSuccess ? NDF_OK : NDF_ERR_DEVICE;
}

ndf_status ndf_device_malloc(void **ptr, size_t bytes)
{
    return hipMalloc(ptr, bytes) == hipSuccess ? NDF_OK : NDF_ERR_DEVICE;
}

ndf_status ndf_device_free(void *ptr) { return hipFree(ptr) == hipSuccess ? NDF_OK : NDF_ERR_DEVICE; }

ndf_status ndf_memcpy_to_device(void *d_dst, const void *h_src, size_t bytes)
{
    return hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice) == hipSuccess ? NDF_OK : NDF_ERR_DEVICE;
}

ndf_status ndf_memcpy_to_host(void *h_dst, const void *d_src, size_t bytes)
{
    return hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? NDF_OK : NDF_ERR_DEVICE;
}

ndf_status ndf_memcpy_device(void *d_dst, const void *d_src, size_t bytes)
{
    return hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? NDF_OK : NDF_ERR_DEVICE;
}

ndf_status ndf_device_fill(void *d_ptr, size_t count, int32_t element_bytes, uint64_t bits)
{
    if ((element_bytes != 4 && element_bytes != 8) || (count > 0 && !d_ptr)) return NDF_ERR_INVALID;
    if (count == 0) return NDF_OK;
    if (launch_fill(element_bytes, d_ptr, (int64_t)count, bits, nullptr) != hipSuccess) return NDF_ERR_DEVICE;
    return hipDeviceSynchronize() == hipSuccess ? NDF_OK : NDF_ERR_DEVICE;
}

ndf_status ndf_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? NDF_OK : NDF_ERR_DEVICE; }

ndf_status ndf_plan_time_device(ndf_plan *plan, const void *d_input, void *d_output, size_t len, int32_t warmup,
                                int32_t iterations, double *ms_per_launch, double *ms_main_kernel, char *err,
                                size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (iterations < 0 || warmup < 0) return fail(NDF_ERR_INVALID, "iterations must be nonnegative", err, err_len);
    std::lock_guard<std::mutex> guard(plan->mutex);
    ndf_status st = prepare_tables(plan, err, err_len);
    if (st != NDF_OK) return st;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    std::vector<hipEvent_t> marks((size_t)iterations * 2, nullptr);
    hipError_t e;
    if ((e = hipStreamCreate(&stream)) != hipSuccess) return fail_hip(e, err, err_len);
    if ((e = hipEventCreate(&start)) != hipSuccess || (e = hipEventCreate(&stop)) != hipSuccess) {
        st = fail_hip(e, err, err_len);
    }
    for (size_t i = 0; st == NDF_OK && i < marks.size(); ++i) {
        if ((e = hipEventCreate(&marks[i])) != hipSuccess) st = fail_hip(e, err, err_len);
    }
    for (int i = 0; st == NDF_OK && i < warmup; ++i) st = launch(plan, d_input, d_output, stream, err, err_len);
    if (st == NDF_OK && (e = hipEventRecord(start, stream)) != hipSuccess) st = fail_hip(e, err, err_len);
    for (int i = 0; st == NDF_OK && i < iterations; ++i) {
        plan->main_begin = marks[(size_t)i * 2];
        plan->main_end = marks[(size_t)i * 2 + 1];
        st = launch(plan, d_input, d_output, stream, err, err_len);
    }
    plan->main_begin = plan->main_end = nullptr;
    if (st == NDF_OK && (e = hipEventRecord(stop, stream)) != hipSuccess) st = fail_hip(e, err, err_len);
    if (st == NDF_OK && (e = hipEventSynchronize(stop)) != hipSuccess) st = fail_hip(e, err, err_len);
    float ms = 0.0f;
    if (st == NDF_OK && (e = hipEventElapsedTime(&ms, start, stop)) != hipSuccess) st = fail_hip(e, err, err_len);
    double main_ms = 0.0;
    for (int i = 0; st == NDF_OK && i < iterations; ++i) {
        float part = 0.0f;
        if ((e = hipEventElapsedTime(&part, marks[(size_t)i * 2], marks[(size_t)i * 2 + 1])) != hipSuccess) {
            st = fail_hip(e, err, err_len);
        }
        main_ms += part;
    }
    if (st == NDF_OK) {
        *ms_per_launch = iterations > 0 ? (double)ms / iterations : 0.0;
        if (ms_main_kernel) *ms_main_kernel = iterations > 0 ? main_ms / iterations : 0.0;
        st = read_error_flag(plan, err, err_len);
    }
    for (hipEvent_t ev : marks) {
        if (ev) (void)hipEventDestroy(ev);
    }
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
    (void)hipStreamDestroy(stream);
    return st;
}

}  // extern "C"

namespace {

// Mask-directed repair of one array (the reference's repair_median,
// src/filter_nd.rs:174-215): the snapshot is the input with every listed
// position set to NaN; each listed output is the plan's statistic over the
// snapshot, every other output is the input. All estimates read the
// snapshot only, so their order (and any duplicate position) cannot change
// a result.
//
// Two device forms give identical bits:
//   sparse  one wave per position (k_repair_wave: the window staged in LDS,
//           a wave-wide bitwise select), or one generic-engine thread per
//           position for windows above REPAIR_WAVE_MAX_TAPS taps
//           (cost ~ positions);
//   dense   the plan's engine over the whole snapshot, then the estimates
//           are scattered back (cost ~ array).
// Auto compares estimated costs (measured on MI355X, 8192^2 float32, 1 %
// of the pixels listed; the paths give identical bits, so the choice only
// affects time). Sparse, per position, by the keys each lane holds: 1.0 ns
// (<= 64 taps), 2 ns (<= 256), 4.3 ns (<= 1024), 24 ns (<= 4096), and a
// thread per position beyond. Dense, per element: the slide engine about
// 0.2 ns at any window size (31x31 and 63x63 alike), 0.07 ns with a learned
// bracket (63x63: 4.6 ms for 8192^2), the network engine growing with the
// taps W (0.023 ns at 25 taps, 0.11 ns at 68, 0.37 ns at 147). Plans without
// a tiled engine always run sparse.
double repair_sparse_ns(size_t count)
{
    return count <= 64 ? 1.0 : count <= 256 ? 2.0 : count <= 1024 ? 4.3 : count <= 4096 ? 24.0 : 1000.0;
}

double repair_dense_ns(const ndf_plan *p)
{
    const double w = (double)p->count;
    const double scale = p->dbytes == 8 ? 1.5 : 1.0;
    if (p->engine == NDF_ENGINE_NETWORK) return scale * 0.0009 * w * (1.0 + w / 60.0);
    if (p->engine == NDF_ENGINE_STACK) return scale * 0.0009 * w * (1.0 + w / 150.0);
    return scale * (p->bracket_on ? 0.07 : 0.2);
}

bool repair_sparse(const ndf_plan *p, size_t npos, int32_t path)
{
    if (path == NDF_REPAIR_SPARSE) return true;
    if (path == NDF_REPAIR_DENSE) return false;
    if (p->engine == NDF_ENGINE_GENERIC) return true;
    return (double)npos * repair_sparse_ns(p->count) <= (double)p->len * repair_dense_ns(p);
}

// Enqueues the repair on `stream`; positions outside the array set
// p->r_flag. The caller holds p->mutex.
ndf_status repair_launch(ndf_plan *p, const void *d_in, void *d_out, void *d_snap, const int64_t *d_pos,
                         size_t npos, int32_t path, hipStream_t stream, char *err, size_t err_len)
{
    const size_t bytes = p->len * (size_t)p->dbytes;
    const uint64_t nan_bits = p->dbytes == 4 ? 0x7FC00000ull : 0x7FF8000000000000ull;
    int32_t *flag = static_cast<int32_t *>(p->r_flag.ptr);
    const int64_t extent = (int64_t)p->len;
    const bool sparse = repair_sparse(p, npos, path);
    ndf_status st = sparse ? prepare_generic_tables(p, err, err_len) : prepare_tables(p, err, err_len);
    if (st != NDF_OK) return st;
    // Both flags are cleared on the repair stream, after the plan's previous
    // launch, so no earlier or concurrent call can clear a flag this one set.
    hipError_t e = plan_begin(p, stream);
    if (e == hipSuccess) e = launch_fill(4, flag, 1, 0, stream);
    if (e == hipSuccess) e = launch_fill(4, p->d_error.ptr, 1, 0, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_snap, d_in, bytes, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = launch_scatter_fill(p->dbytes, d_snap, d_pos, (int64_t)npos, extent, nan_bits, flag, stream);
    if (e != hipSuccess) return fail_hip(e, err, err_len);
    if (sparse) {
        if (d_out != d_in) e = hipMemcpyAsync(d_out, d_in, bytes, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess && p->all_nan) {
            e = launch_scatter_fill(p->dbytes, d_out, d_pos, (int64_t)npos, extent, nan_bits, flag, stream);
        } else if (e == hipSuccess) {
            GenericArgs args = p->generic;
            args.input = d_snap;
            args.output = d_out;
            args.len = (int64_t)npos;
            args.positions = d_pos;
            args.extent = extent;
            args.error_flag = flag;
            e = p->count <= (size_t)REPAIR_WAVE_MAX_TAPS ? launch_repair_wave(p->dbytes, args, stream)
                                                        : launch_generic(p->dbytes, args, stream);
        }
        if (e == hipSuccess) e = plan_end(p, stream);
        return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
    }
    // Whole-image statistic of the snapshot into d_out, estimates back into
    // the snapshot (whose unlisted entries are the input), then the result.
    st = launch(p, d_snap, d_out, stream, err, err_len);
    if (st != NDF_OK) return st;
    e = launch_scatter_copy(p->dbytes, d_snap, d_out, d_pos, (int64_t)npos, extent, flag, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_out, d_snap, bytes, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = plan_end(p, stream);
    return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
}

// Waits for the stream and reports a position outside the array or a rank
// walk that left its bounds.
ndf_status repair_finish(ndf_plan *p, hipStream_t stream, char *err, size_t err_len)
{
    hipError_t e = hipStreamSynchronize(stream);
    int32_t flags[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpy(&flags[0], p->r_flag.ptr, sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&flags[1], p->d_error.ptr, sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, err, err_len);
    if (flags[0]) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (flags[1]) return fail(NDF_ERR_DEVICE, "internal error: rank walk left its bounds", err, err_len);
    return NDF_OK;
}

ndf_status repair_buffers(ndf_plan *p, char *err, size_t err_len)
{
    hipError_t e;
    if ((e = p->r_flag.ensure(sizeof(int32_t))) != hipSuccess || (e = p->d_error.ensure(sizeof(int32_t))) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    return NDF_OK;
}

}  // namespace

extern "C" {

ndf_status ndf_plan_repair_device(ndf_plan *plan, const void *d_input, void *d_output, void *d_snapshot,
                                  const int64_t *d_positions, size_t npos, size_t len, int32_t path, void *stream,
                                  char *err, size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (path < NDF_REPAIR_AUTO || path > NDF_REPAIR_DENSE) return fail(NDF_ERR_INVALID, "unknown repair path", err, err_len);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (len == 0) return NDF_OK;
    std::lock_guard<std::mutex> guard(plan->mutex);
    if (d_snapshot == d_input || d_snapshot == d_output) {
        return fail(NDF_ERR_INVALID, "the repair snapshot must not alias the input or the output", err, err_len);
    }
    if (npos == 0) {
        hipError_t e = hipSuccess;
        if (d_output != d_input) {
            e = hipMemcpyAsync(d_output, d_input, len * (size_t)plan->dbytes, hipMemcpyDeviceToDevice, s);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) return NDF_OK;
        return fail_hip(e, err, err_len);
    }
    ndf_status st = repair_buffers(plan, err, err_len);
    if (st == NDF_OK) st = repair_launch(plan, d_input, d_output, d_snapshot, d_positions, npos, path, s, err, err_len);
    if (st == NDF_OK) st = repair_finish(plan, s, err, err_len);
    return st;
}

ndf_status ndf_plan_time_repair_device(ndf_plan *plan, const void *d_input, void *d_output, void *d_snapshot,
                                       const int64_t *d_positions, size_t npos, size_t len, int32_t path,
                                       int32_t warmup, int32_t iterations, double *ms_per_call, char *err,
                                       size_t err_len)
{
    if (len != plan->len || npos == 0) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (d_snapshot == d_input || d_snapshot == d_output) {
        return fail(NDF_ERR_INVALID, "the repair snapshot must not alias the input or the output", err, err_len);
    }
    if (path < NDF_REPAIR_AUTO || path > NDF_REPAIR_DENSE) return fail(NDF_ERR_INVALID, "unknown repair path", err, err_len);
    std::lock_guard<std::mutex> guard(plan->mutex);
    ndf_status st = repair_buffers(plan, err, err_len);
    if (st != NDF_OK) return st;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&start);
    if (e == hipSuccess) e = hipEventCreate(&stop);
    if (e != hipSuccess) st = fail_hip(e, err, err_len);
    for (int i = 0; st == NDF_OK && i < warmup; ++i) {
        st = repair_launch(plan, d_input, d_output, d_snapshot, d_positions, npos, path, stream, err, err_len);
    }
    if (st == NDF_OK && (e = hipEventRecord(start, stream)) != hipSuccess) st = fail_hip(e, err, err_len);
    for (int i = 0; st == NDF_OK && i < iterations; ++i) {
        st = repair_launch(plan, d_input, d_output, d_snapshot, d_positions, npos, path, stream, err, err_len);
    }
    if (st == NDF_OK && (e = hipEventRecord(stop, stream)) != hipSuccess) st = fail_hip(e, err, err_len);
    if (st == NDF_OK) st = repair_finish(plan, stream, err, err_len);
    if (st == NDF_OK) {
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, start, stop)) != hipSuccess) {
            st = fail_hip(e, err, err_len);
        } else {
            *ms_per_call = iterations > 0 ? (double)ms / iterations : 0.0;
        }
    }
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
    if (stream) (void)hipStreamDestroy(stream);
    return st;
}

namespace {

// Snapshot half of the host-buffer repair: the data and the mask go to the
// device, and the mask's flat indices are compacted there. The caller's
// buffers are free to change once it returns.
ndf_status fixpix_upload_locked(ndf_plan *p, const void *input, const uint8_t *mask, char *err, size_t err_len)
{
    const size_t bytes = p->len * (size_t)p->dbytes;
    hipError_t e;
    p->fix_uploaded = false;
    if ((e = p->d_in.ensure(bytes)) != hipSuccess || (e = p->d_out.ensure(bytes)) != hipSuccess ||
        (e = p->r_mask.ensure(p->len)) != hipSuccess ||
        (e = p->r_count.ensure(2 * sizeof(unsigned long long))) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    if ((e = staged_upload(p->staging, p->d_in.ptr, input, bytes, nullptr)) != hipSuccess ||
        (e = staged_upload(p->staging, p->r_mask.ptr, mask, p->len, nullptr)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    unsigned long long *counter = static_cast<unsigned long long *>(p->r_count.ptr);
    unsigned long long count = 0;
    if ((e = launch_fill(8, counter, 2, 0, nullptr)) != hipSuccess ||
        (e = launch_mask_count(static_cast<const uint8_t *>(p->r_mask.ptr), (int64_t)p->len, counter, nullptr)) !=
            hipSuccess ||
        (e = hipMemcpy(&count, counter, sizeof(count), hipMemcpyDeviceToHost)) != hipSuccess) {
        return fail_hip(e, err, err_len);
    }
    p->fix_npos = (size_t)count;
    if (count > 0) {
        if ((e = p->r_pos.ensure((size_t)count * sizeof(int64_t))) != hipSuccess ||
            (e = launch_mask_positions(static_cast<const uint8_t *>(p->r_mask.ptr), (int64_t)p->len,
                                       static_cast<int64_t *>(p->r_pos.ptr), counter + 1, nullptr)) != hipSuccess) {
            return fail_hip(e, err, err_len);
        }
    }
    p->fix_uploaded = true;
    return NDF_OK;
}

// Compute half: repair the uploaded data at the mask's positions and copy
// the result into `output`.
ndf_status fixpix_compute_locked(ndf_plan *p, void *output, char *err, size_t err_len)
{
    if (!p->fix_uploaded) return fail(NDF_ERR_INVALID, "no input has been uploaded to this plan", err, err_len);
    p->fix_uploaded = false;
    const size_t bytes = p->len * (size_t)p->dbytes;
    const void *result = p->d_in.ptr;
    hipError_t e;
    if (p->fix_npos > 0) {
        if ((e = p->r_snap.ensure(bytes)) != hipSuccess) return fail_hip(e, err, err_len);
        ndf_status st = repair_buffers(p, err, err_len);
        if (st == NDF_OK) {
            st = repair_launch(p, p->d_in.ptr, p->d_out.ptr, p->r_snap.ptr, static_cast<const int64_t *>(p->r_pos.ptr),
                               p->fix_npos, NDF_REPAIR_AUTO, nullptr, err, err_len);
        }
        if (st == NDF_OK) st = repair_finish(p, nullptr, err, err_len);
        if (st != NDF_OK) return st;
        result = p->d_out.ptr;
    }
    if ((e = staged_download(p->staging, output, result, bytes, nullptr)) != hipSuccess) return fail_hip(e, err, err_len);
    return NDF_OK;
}

}  // namespace

ndf_status ndf_fixpix_upload(ndf_plan *plan, const void *input, const uint8_t *mask, size_t len, char *err,
                             size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    std::lock_guard<std::mutex> guard(plan->mutex);
    if (len == 0) {
        plan->fix_uploaded = true;
        plan->fix_npos = 0;
        return NDF_OK;
    }
    return fixpix_upload_locked(plan, input, mask, err, err_len);
}

ndf_status ndf_fixpix_compute_download(ndf_plan *plan, void *output, size_t len, char *err, size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    std::lock_guard<std::mutex> guard(plan->mutex);
    if (len == 0) {
        plan->fix_uploaded = false;
        return NDF_OK;
    }
    return fixpix_compute_locked(plan, output, err, err_len);
}

ndf_status ndf_fixpix_median(int32_t dtype_bytes, const void *input, void *output, const uint8_t *mask,
                             size_t mask_len, size_t rank, const size_t *shape, const size_t *size,
                             const uint8_t *footprint, const ptrdiff_t *origin, const int32_t *modes, double cval,
                             int32_t ignore_inf, size_t min_count, char *err, size_t err_len)
{
    ndf_plan *plan = nullptr;
    ndf_status st = ndf_plan_create(&plan, dtype_bytes, rank, shape, size, footprint, origin, modes, cval,
                                    NDF_KIND_MEDIAN, 1, ignore_inf, min_count, 0, err, err_len);
    if (st != NDF_OK) return st;
    std::unique_ptr<ndf_plan> owner(plan);
    // positions() in src/repair.rs:14-29: the mask must cover the array.
    if (mask_len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    st = ndf_fixpix_upload(plan, input, mask, plan->len, err, err_len);
    if (st == NDF_OK) st = ndf_fixpix_compute_download(plan, output, plan->len, err, err_len);
    return st;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// No-allocation launches.
//
// ndf_plan_apply_device_static runs the plan's kernels with every device
// table (boundary maps, tap offsets, edge lists, footprint mask, bracket
// weights) and every scratch buffer (error flag, block moments, tile
// features and key brackets, fallback list and count) carved out of one
// file-scope device array, so it never calls hipMalloc or hipFree. Each GPU
// holds its own copy of the array for the life of the process; its address
// on a device comes from hipGetSymbolAddress, taken once per device with the
// device current. All calls on one GPU share that copy and it keeps the
// last call's contents: the tables are copied in on every call, and the
// counters that must start at zero are cleared by fill kernels on the call's
// stream. Calls that overlap on one GPU (on other streams or threads) race
// on the array; the caller orders them.
// ---------------------------------------------------------------------------

namespace {

constexpr size_t STATIC_SCRATCH_BYTES = (size_t)32 << 20;
constexpr int STATIC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

std::mutex g_scratch_mutex;
void *g_scratch_address[STATIC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device).
hipError_t scratch_address(void **out)
{
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= STATIC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_address[device] = ptr;
    }
    *out = g_scratch_address[device];
    return hipSuccess;
}

// Sequential 256-byte aligned regions of the scratch array. With a null base
// it only counts bytes.
struct Carve {
    char *base;
    size_t used = 0;
    void *take(size_t bytes)
    {
        const size_t at = (used + 255) & ~(size_t)255;
        used = at + bytes;
        return base ? base + at : nullptr;
    }
};

// Lays out one launch of plan p in the scratch array at `base` (null: size
// only) and copies the plan's host tables there on `stream`. *needed
// receives the bytes the layout uses. Leaves the plan's own launch
// arguments untouched.
hipError_t static_layout(ndf_plan *p, char *base, hipStream_t stream, LaunchSet *t, size_t *needed)
{
    Carve c{base};
    hipError_t e = hipSuccess;
    auto put = [&](const auto &values) -> void * {
        const size_t bytes = values.size() * sizeof(values[0]);
        void *dst = c.take(bytes);
        if (base && bytes && e == hipSuccess) {
            e = hipMemcpyAsync(dst, values.data(), bytes, hipMemcpyHostToDevice, stream);
        }
        return dst;
    };
    *t = LaunchSet{};
    t->error = static_cast<int32_t *>(c.take(sizeof(int32_t)));
    if (p->all_nan) {
        // A fill kernel only.
    } else if (p->engine == NDF_ENGINE_GENERIC) {
        generic_host_tables(p);
        t->generic = p->generic;
        t->generic.maps = static_cast<const int64_t *>(put(p->gen_maps));
        t->generic.taps = static_cast<const int32_t *>(put(p->gen_taps));
    } else if (p->engine == NDF_ENGINE_EXTREMA) {
        ExtremaArgs &x = t->extrema;
        x = p->extrema;
        x.mapz = static_cast<const int64_t *>(put(p->ext_mapz));
        x.mapy = static_cast<const int64_t *>(put(p->ext_mapy));
        x.mapx = static_cast<const int64_t *>(put(p->ext_mapx));
    } else if (p->engine == NDF_ENGINE_STACK) {
        StackArgs &k = t->stack;
        k = p->stack;
        k.maps = static_cast<const int64_t *>(put(p->stk_maps));
        k.mapr = static_cast<const int64_t *>(put(p->stk_mapr));
        k.mapx = static_cast<const int64_t *>(put(p->stk_mapx));
    } else if (p->engine == NDF_ENGINE_NETWORK) {
        NetworkArgs &n = t->network;
        n = p->network;
        n.mapz = static_cast<const int64_t *>(put(p->net_mapz));
        n.mapy = static_cast<const int64_t *>(put(p->net_mapy));
        n.mapx = static_cast<const int64_t *>(put(p->net_mapx));
        n.taps = static_cast<const uint32_t *>(put(p->network.pair_c ? p->net_pair_taps : p->net_taps));
    } else {
        SlideArgs &sl = t->slide;
        sl = p->slide;
        sl.mapz = static_cast<const int64_t *>(put(p->slide_mapz));
        sl.mapy = static_cast<const int64_t *>(put(p->slide_mapy));
        sl.mapx = static_cast<const int64_t *>(put(p->slide_mapx));
        sl.taps = static_cast<const uint32_t *>(put(p->slide_taps));
        sl.edge_remove = static_cast<const uint32_t *>(put(p->slide_rm));
        sl.edge_add = static_cast<const uint32_t *>(put(p->slide_add));
        const void *mask = put(p->slide_mask);
        sl.mask = p->slide_mask.empty() ? nullptr : static_cast<const uint8_t *>(mask);
        sl.error_flag = t->error;
        if (p->bracket_on) {
            bracket_geometry_into(p, sl, t->slide_fb, t->slide_fb_lds, t->slide_b, t->slide_b_lds, t->bargs);
            size_t sizes[8];
            bracket_buffer_sizes(p, sizes);
            void *ptrs[8];
            for (int i = 0; i < 8; ++i) ptrs[i] = c.take(sizes[i]);
            bracket_bind(t->slide_b, t->bargs, ptrs);
            t->blocks = static_cast<double *>(ptrs[0]);
            t->fallback = static_cast<int32_t *>(ptrs[6]);
            t->fcount = static_cast<int32_t *>(ptrs[7]);
            t->weights = static_cast<const float *>(put(p->bracket_weights));
        }
    }
    *needed = c.used;
    return e;
}

}  // namespace

extern "C" {

ndf_status ndf_plan_apply_device_static(ndf_plan *plan, const void *d_input, void *d_output, size_t len,
                                        void *stream, char *err, size_t err_len)
{
    if (len != plan->len) return fail_code(NDF_ERR_LENGTH_MISMATCH, err, err_len);
    if (len == 0) return NDF_OK;
    if (buffers_overlap(plan, d_input, d_output, len)) {
        return fail(NDF_ERR_INVALID, "the output buffer must not overlap the input buffer", err, err_len);
    }
    std::lock_guard<std::mutex> guard(plan->mutex);
    hipStream_t s = static_cast<hipStream_t>(stream);
    LaunchSet t;
    size_t needed = 0;
    (void)static_layout(plan, nullptr, nullptr, &t, &needed);
    if (needed > STATIC_SCRATCH_BYTES) {
        char message[192];
        snprintf(message, sizeof message,
                 "the no-allocation launch needs %zu bytes of device scratch for this plan; the limit is %zu",
                 needed, STATIC_SCRATCH_BYTES);
        return fail(NDF_ERR_INVALID, message, err, err_len);
    }
    void *base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e == hipSuccess) e = static_layout(plan, static_cast<char *>(base), s, &t, &needed);
    if (e == hipSuccess) e = enqueue(plan, t, d_input, d_output, s);
    return e == hipSuccess ? NDF_OK : fail_hip(e, err, err_len);
}

}  // extern "C"
