// mpmw_sort.hpp -- particle binning into cell tiles.
//
// Particles are kept ordered by tile (a block of TX x TY x TZ cells; along
// each axis the tiles start at the cell of the AABB-min particle, found with
// the same FP64 expression the kernels use, so the tiles line up with the
// workpiece) and,
// inside a tile, slot-major: the first particle of every cell, then the
// second of every cell, and so on.  A 64-lane wave then holds particles of 64
// different cells, so its LDS scatter for one corner offset touches 64
// different nodes, while global loads stay contiguous.
//
// Key space: tile-major cell keys [0, nkeys), then key nkeys for particles
// whose cell is outside the tiles or has a corner node outside the grid
// ("loose" particles, handled by a per-particle pass), then key nkeys + 1 for
// INACTIVE particles.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpmw {

constexpr int TX = 8, TY = 8, TZ = 4;                 // cells per tile
constexpr int kTileCells = TX * TY * TZ;              // 256 = one cell per thread
constexpr int WX = TX + 1, WY = TY + 1, WZ = TZ + 1;  // node window of a tile
constexpr int WN = WX * WY * WZ;                      // 405 nodes
// node window of a tile for the quadratic (APIC) stencil: one node before
// the tile's first cell, two after its last
constexpr int QX = TX + 3, QY = TY + 3, QZ = TZ + 3;
constexpr int QN = QX * QY * QZ;                      // 847 nodes

struct ScratchBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// Tile geometry over the current grid.
struct TileGeom {
    int32_t ox = 4, oy = 4, oz = 4;    // first tiled cell along x, y, z
    int32_t ntx = 0, nty = 0, ntz = 0;
    int32_t ntiles = 0;
    int32_t nkeys = 0;                 // ntiles * kTileCells
};

// Stable device radix sort of (key, value) pairs over the low `nbits` bits.
hipError_t radix_sort_pairs(ScratchBuf &temp, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                            int32_t *vout, int n, int nbits, hipStream_t st);
// Exclusive prefix sum of n int32 values.
hipError_t exclusive_scan(ScratchBuf &temp, const int32_t *in, int32_t *out, int n, hipStream_t st);
// The same scan into caller-provided temporary storage (no allocation):
// exclusive_scan_temp_bytes gives the bytes it needs for n values.
hipError_t exclusive_scan_temp_bytes(int n, size_t *bytes);
hipError_t exclusive_scan_buf(void *temp, size_t bytes, const int32_t *in, int32_t *out, int n, hipStream_t st);
void free_scratch(ScratchBuf &t);

}  // namespace mpmw

struct SortState {
    bool valid = false;                // particles are binned on the current grid
    mpmw::TileGeom geom;
    int32_t key_cap = 0, n_cap = 0;
    uint32_t *keys = nullptr, *keys_sorted = nullptr;
    int32_t *idx = nullptr, *idx_sorted = nullptr, *perm = nullptr;
    int32_t *counts = nullptr;         // nkeys + 2
    int32_t *cell_start = nullptr;     // nkeys + 2 (exclusive scan of counts)
    int32_t *tile_start = nullptr;     // ntiles + 1
    int32_t tile_cap = 0;
    // device counters: [0] particle visits outside their tile window since
    // the last rebuild, [1] particles of the last deterministic P2G pass that
    // were not in their binned cell
    int32_t *oow = nullptr;
    int32_t *h_oow = nullptr;          // pinned: [0, 2) copy of the counters, [2, 4) the loose range
    int32_t loose_begin = 0, loose_end = 0;   // particles binned as loose (host copy)
    mpmw::ScratchBuf temp, temp2;
    double origin[3] = {0, 0, 0};      // grid origin the particles were binned on
    int64_t steps_since = 0;           // steps since the out-of-window counter was reset
    int64_t sorts = 0;
    int64_t resorts = 0;               // re-binnings forced by the deterministic P2G
};
