// vr_kernels.h - kernel argument blocks and launchers shared by the runtime.
#pragma once
#include "vr_device.h"

namespace vr {

struct HitPlanes {
    uint32_t* hitw;  // voxel (bits 0-23) | side << 24 | (D[side] > 0) << 26
    float* dist;
    uint32_t* seed;
};

// A plane of per-pixel reservoirs, as three structure-of-arrays records.
// r0 holds what a reuse pass tests before it takes a reservoir, so a
// neighbour or the previous frame's pixel that fails the tests costs one
// 16-byte read: {M (bits 0-23) | Nlight face << 24 | the pixel's hit side
// bits << 27 (side and the sign of D[side], bits 24-26 of the hit word) |
// kTraced | kVisible, hit distance, W, invPdf}.  r1 is the sample {pos.xyz,
// light index}, r2 the sums {wsum, pHat}.
struct ResPlanes {
    uint4* r0;
    uint4* r1;
    float2* r2;
};
// r0.x bits 30-31: the candidate pass traced the shadow ray of this
// reservoir's sample from this pixel (kTraced), and found it unoccluded
// (kVisible).  A later ray from the same surface point to the same sample
// is the same ray with the same answer.
constexpr uint32_t kTraced = 1u << 30, kVisible = 1u << 31;

// ray counts of a frame (FrameArgs::rays, when counting is on)
enum RayCount { kRaysPrimary = 0, kRaysCandidate = 1, kRaysShading = 2, kRaysMis = 3, kRayKinds = 4 };

// Everything one frame's kernels read, passed by value (kernarg segment).
struct FrameArgs {
    WorldView w;
    LightView L;
    Camera cam;
    HitPlanes cur;
    ResPlanes A;  // final reservoirs: previous frame's on entry, this frame's on exit
    ResPlanes B;  // candidates + temporal output
    unsigned long long* rays;  // [kRayKinds] rays traced, or nullptr (not counted)
    float4* frame;
    const float* emitLut;
    uint32_t W, H;
    uint32_t R0, restirtemporalframe, temporal, spatial;
    uint32_t nCand, maxTemporal, spatialRadius, spatialTaps;
    float4 Nl, Nr, Nt, Nb;
    uint32_t point, gfxexp, source, axisOnlyFace, doubleRadiance;
    int32_t shadowSteps;
    uint32_t tileOrder;  // VR_TILES_*
    uint32_t misSpatial; // spatial merge with MIS weights (VR_FLAG_MIS_SPATIAL)
    uint32_t waves;      // occupancy target of this stage's kernel (0: kDefaultWaves)
    uint32_t dda;        // traversal of this stage's shadow rays (kDda*)
    // learned source: candidate slot j of pixel i at [j * W*H + i] holds a
    // cluster and the bits of its mixture mass; inside the cluster the light
    // is drawn with the cluster's power alias table
    const uint2* candSlot;
    const uint32_t* clusterStart;  // [257] offsets into the cluster-ordered light list
    // bucket j of the list (cluster c's buckets are clusterStart[c]..): the
    // alias table's (prob bits, light j, light of alias j, alias j local
    // index) and the two lights' in-cluster shares pmf_i / P_c, so one pair
    // of loads yields the light and its pdf
    const uint4* clusterPick;
    const float2* clusterPickShare;
};

// Tile order (block id -> 16x16 tile).  Workgroups are dealt round-robin to
// the 8 XCDs in linear id order and each XCD has its own L2.
//   VR_TILES_LINEAR   block b runs tile b (row-major): the tiles in flight
//                     form one compact strip, neighbours on different XCDs
//   VR_TILES_XCD_ROWS block b runs tile (b % 8) * (n/8) + b/8: each XCD
//                     renders one contiguous band of rows (the tail beyond a
//                     multiple of 8 keeps its own index)
//   VR_TILES_XCD_COLS with 8 | tiles per row, XCD b % 8 renders column band
//                     b % 8 row by row, so the tiles in flight still form
//                     one strip and each XCD's part of it is contiguous;
//                     other widths fall back to the linear order
//   VR_TILES_XCD_BLOCKS the tiles taken in blocks of 4x4 (64x64 pixels, row
//                     by row within a block, blocks row by row, the edge
//                     blocks clipped); each XCD takes 16 consecutive tiles
//                     of that sequence at a time, XCD after XCD, so a
//                     pixel's spatial taps mostly fall in tiles of its own
//                     XCD while all XCDs sweep the image together (the tail
//                     beyond a multiple of 128 keeps its own position)
// Every order is a bijection on [0, n).
__device__ __forceinline__ void tile_of_block(uint32_t order, int& tx, int& ty)
{
    const uint32_t nbx = gridDim.x, nby = gridDim.y, nb = gridDim.x * gridDim.y;
    const uint32_t b = blockIdx.y * nbx + blockIdx.x;
    uint32_t t = b;
    if (order == 1u) {
        const uint32_t full = nb & ~7u;
        t = b < full ? (b & 7u) * (full >> 3) + (b >> 3) : b;
    } else if (order == 2u && (nbx & 7u) == 0u) {
        const uint32_t cw = nbx >> 3, i = b >> 3;
        t = (i / cw) * nbx + (b & 7u) * cw + i % cw;
    } else if (order == 3u) {
        // position p in the block sequence: XCD b % 8 takes run b / 128 * 8
        // + b % 8 of 16 positions, one position per round of 8 workgroups
        const uint32_t full = nb & ~127u;
        const uint32_t p = b < full ? (b & ~127u) + ((b & 7u) << 4) + ((b >> 3) & 15u) : b;
        // p -> tile: block rows of 4 tile rows (the last may be shorter),
        // blocks of 4 columns (the last may be narrower), row-major inside
        const uint32_t R = p / (nbx * 4u), q = p - R * nbx * 4u;
        const uint32_t h = min(4u, nby - 4u * R);
        const uint32_t C = q / (4u * h), in = q - C * 4u * h;
        const uint32_t w = min(4u, nbx - 4u * C);
        tx = (int)(4u * C + in % w);
        ty = (int)(4u * R + in / w);
        return;
    }
    tx = (int)(t % nbx);
    ty = (int)(t / nbx);
}

// traversal of a stage's shadow rays (FrameArgs::dda; the values of
// VR_DDA_SCHEDULED and VR_DDA_NESTED): the wave-scheduled DDA, or the
// nested one of the primary rays
constexpr uint32_t kDdaScheduled = 0u, kDdaNested = 1u;

// default occupancy targets (waves per SIMD) of the stage kernels: primary,
// candidates, spatial reuse + shading.  The primary pass mostly runs beside
// the previous frame's last stage, where fewer of its waves slow that stage
// less (c2: 0.24 ms at 5 against 0.31 ms at 8, the same when alone); the
// last stage then runs best at 8 (0.57 against 0.62 ms at 7, the same when
// alone).  The MIS merge's variant holds its tile's ray list in LDS and more
// live values, and runs best at 5.
constexpr uint32_t kDefaultWaves[3] = {5, 7, 8};
constexpr uint32_t kDefaultWavesMis = 5;
// the split MIS merge's merge/list and weights/shading kernels (the tracer
// runs at 8)
constexpr uint32_t kDefaultWavesMisSplit = 8;
hipError_t launch_stage(int stage, const FrameArgs& a, hipStream_t s);

// The MIS merge of one view in three launches: the merge, which stores the
// merged reservoirs and lists the visibility rays its weights need; a
// tracer over that list, one ray per thread; the weights and the shading.  The list
// holds up to W*H*spatialTaps rays (owner pixel << 5 | tap, neighbour
// pixel).  `count` is zero before the first launch; the last launch sets it
// to zero again for the next frame.
struct MisBufs {
    uint2* list;
    uint32_t* count;  // [1] rays listed
    uint32_t* occ;    // [W*H] bit k: tap k's ray found the merged sample occluded
    uint4* info;      // [W*H] taps whose ray is listed, combined taps, sel + 1, own M
    float2* terms;    // [tap][W*H] M * T and T of each combined tap
};
// traceBlocks: workgroups of the tracer's grid (256 rays each), enough for
// the list's capacity
hipError_t launch_mis_split(const FrameArgs& a, const MisBufs& m, uint32_t traceBlocks, hipStream_t s);
// Up to kChunkViews views rendered by one launch per stage: their
// FrameArgs travel in the kernel argument (8 x 448 B + the tile table,
// under the 4 KiB kernel-argument limit), so pointers in them are global
// pointers to the compiler as in the one-view kernels.
constexpr int kChunkViews = 8;
struct ViewChunk {
    FrameArgs v[kChunkViews];
    uint32_t start[kChunkViews + 1];  // first 16x16 tile of each view; start[n] = total
    uint32_t n;
};
static_assert(sizeof(ViewChunk) <= 4096, "kernel argument limit");

// one launch of stage `stage` over the tiles of a chunk's views: one
// workgroup per tile (blocks = 0) or a persistent grid of `blocks`
hipError_t launch_stage_views(int stage, const ViewChunk& c, uint32_t blocks, bool anyMis, hipStream_t s);
// the same launch as a graph kernel node (args[0] receives the argument
// pointer; c and args must stay valid while the node is added or updated)
hipError_t stage_views_node(int stage, const ViewChunk* c, uint32_t blocks, bool anyMis, void** args,
                            hipKernelNodeParams* p);
hipError_t launch_commit(const uint32_t* ids, const uint32_t* src, uint32_t* pool, uint32_t n, uint32_t cap,
                         hipStream_t s);
hipError_t launch_uber(const uint32_t* grid, uint32_t* uber, hipStream_t s);
hipError_t launch_trace_batch(const WorldView& w, const float* O, const float* D, uint32_t n, uint32_t kind,
                              int32_t steps, uint32_t* voxel, float* dist, uint32_t* side, hipStream_t s);
hipError_t launch_ground_truth(const FrameArgs& a, uint32_t spp, uint32_t frame0, double* accum, hipStream_t s);

}  // namespace vr
