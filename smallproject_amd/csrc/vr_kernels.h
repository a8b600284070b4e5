// vr_kernels.h - kernel argument blocks and launchers shared by the runtime.
#pragma once
#include "vr_device.h"

namespace vr {

struct HitPlanes {
    uint32_t* hitw;  // voxel (bits 0-23) | side << 24 | (D[side] > 0) << 26
    float* dist;
    uint32_t* seed;
};

struct ResPlanes {
    float4* a;    // wsum, W, pHat, invPdf
    float4* b;    // pos.xyz, lightIndex bits
    uint32_t* m;  // M | Nlight face << 24
};

// Everything one frame's kernels read, passed by value (kernarg segment).
struct FrameArgs {
    WorldView w;
    LightView L;
    Camera cam;
    HitPlanes cur;
    const uint32_t* prevHitw;
    const float* prevDist;
    ResPlanes A;  // final reservoirs: previous frame's on entry, this frame's on exit
    ResPlanes B;  // candidates + temporal output
    float4* frame;
    const float* emitLut;
    uint32_t W, H;
    uint32_t R0, restirtemporalframe, temporal, spatial;
    uint32_t nCand, maxTemporal, spatialRadius, spatialTaps;
    float4 Nl, Nr, Nt, Nb;
    uint32_t point, gfxexp, source, axisOnlyFace, doubleRadiance;
    int32_t shadowSteps;
};

hipError_t launch_stage(int stage, const FrameArgs& a, hipStream_t s);
hipError_t launch_commit(const uint32_t* ids, const uint32_t* src, uint32_t* pool, uint32_t n, uint32_t cap,
                         hipStream_t s);
hipError_t launch_uber(const uint32_t* grid, uint8_t* uber, hipStream_t s);
hipError_t launch_trace_batch(const WorldView& w, const float* O, const float* D, uint32_t n, uint32_t kind,
                              int32_t steps, uint32_t* voxel, float* dist, uint32_t* side, hipStream_t s);
hipError_t launch_ground_truth(const FrameArgs& a, uint32_t spp, uint32_t frame0, double* accum, hipStream_t s);

}  // namespace vr
