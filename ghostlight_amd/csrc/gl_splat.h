// gl_splat.h — deposit records and the tiled resolve of the ghost splat.
#pragma once

#include <hip/hip_runtime.h>

struct gl_ctx;

namespace gl {

struct SplatGeom {
    int width, height;
    int tile_shift;          // tile side = 1 << tile_shift pixels
    int tiles_x, tiles_y;
    int frames;              // output frames (one per source in frame mode, else 1)
    int tiles_per_frame;
    int n_tiles;             // tiles_per_frame * frames
    int lds_hist;            // 1: per-workgroup LDS histograms in the binning passes
};

struct SplatBufs {
    float4* rec;                          // {px, py, power, lambda index bits}
    unsigned long long rec_cap;
    unsigned long long* rec_count;        // records produced (may exceed rec_cap: overflow)
    float4* binned;                       // records copied into their tiles' ranges
    unsigned long long bin_cap;
    unsigned long long* bin_total;
    unsigned* tile_count;
    unsigned long long* tile_start;
    unsigned long long* tile_cursor;
    unsigned* chunk_start;                // n_tiles + 1
    unsigned* queue;                      // resolve work-queue head
    unsigned* tile_mag;                   // per tile: bits of the largest |power x colour weight|
    unsigned* mc_slot;                    // per tile with several chunks: its staging slot
    unsigned* mc_tile;                    // per staging slot: the tile
    unsigned* mc_count;                   // tiles with several chunks
    unsigned long long* stage;            // int64 fixed-point planes of those tiles (kept zero between renders)
    unsigned* blk_hist;                   // LDS form: each binning workgroup's tile counts of its record slice
    unsigned stage_tiles;                 // staging slots allocated
};

SplatGeom splat_geom(int width, int height, int frames = 1);
int splat_prepare(gl_ctx* c, const SplatGeom& g, unsigned long long rec_cap, SplatBufs& B);
int splat_clear(hipStream_t s, const SplatGeom& g, const SplatBufs& B);
int splat_resolve(hipStream_t s, const SplatGeom& g, const SplatBufs& B, const float4* spec, float* out_r,
                  float* out_g, float* out_b);

}  // namespace gl
