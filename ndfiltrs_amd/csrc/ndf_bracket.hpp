// Learned-bracket proposal: layout of the MLP parameters and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ndf {

constexpr int FEATURES = 6;
constexpr int HIDDEN = 32;
// Flat float parameter layout: W1[6][32], b1[32], W2[32][32], b2[32],
// W3[32][2], b3[2] (row-major, input index first).
constexpr int W1_OFF = 0;
constexpr int B1_OFF = W1_OFF + FEATURES * HIDDEN;
constexpr int W2_OFF = B1_OFF + HIDDEN;
constexpr int B2_OFF = W2_OFF + HIDDEN * HIDDEN;
constexpr int W3_OFF = B2_OFF + HIDDEN;
constexpr int B3_OFF = W3_OFF + HIDDEN * 2;
constexpr int PARAMS = B3_OFF + 2;

struct BracketArgs {
    int64_t ny, nx;  // 2-D image
    int64_t bx;      // 8x8 blocks per image row
    int32_t ty, tiles_y, tiles_x;
    int32_t low_y, high_y, low_x, high_x;  // tap offset extents of the window
    float quantile;                         // target quantile of the statistic
    float *features;                        // tiles x FEATURES
    double *moments;                        // tiles x 3: level c, scale sigma, retained count
    void *lo_keys;                          // tiles keys (uint32 / uint64)
    void *hi_keys;
    float *predicted;                       // tiles x 2 standardised offsets
    int32_t *reset;                         // reset[0 .. reset_count) cleared by the proposal launch, or null
    int32_t reset_count;                    // (the fallback count and the bracket kernel's tile queues)
};

hipError_t launch_block_stats(int dtype_bytes, const void *image, int64_t ny, int64_t nx, int ignore_inf,
                              double *blocks, hipStream_t stream);
hipError_t launch_tile_features(const BracketArgs &args, const double *blocks, int64_t tiles, hipStream_t stream);
// Features, MLP and key bracket of every tile in one launch; also clears
// args.reset[0 .. args.reset_count) when it is set.
hipError_t launch_bracket_proposal(int dtype_bytes, const BracketArgs &args, const double *blocks,
                                   const float *weights, int64_t tiles, hipStream_t stream);
hipError_t launch_tile_labels(int dtype_bytes, const BracketArgs &args, const void *output, float *targets,
                              float *weight, int64_t tiles, hipStream_t stream);
int mlp_grad_blocks(int64_t count);
hipError_t launch_mlp_step(float *weights, float *m1, float *m2, const float *features, const float *targets,
                           const float *weight, const int32_t *index, int64_t count, float tau_lo, float tau_hi,
                           float lr, int step, float *partial, float *loss_partial, float *loss_out,
                           hipStream_t stream);

}  // namespace ndf
