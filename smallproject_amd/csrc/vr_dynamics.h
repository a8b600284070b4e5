// vr_dynamics.h - launchers of the device-side scene motion (moving
// emitters, random-walk brick edits).
#pragma once
#include "vr_device.h"

namespace vr {

constexpr uint32_t kClaimBuckets = 1u << 16;  // target buckets of the light moves
constexpr uint32_t kNoTarget = 0xffffffffu;

// the light records (as the render kernels read them) and the motion state
struct LightMotionView {
    float4* la;              // (centre.xyz, size)
    float4* lb;              // (Le.rgb, packed position bits)
    float4* pos;             // sub-voxel position (xyz)
    float4* vel;             // voxels per frame (xyz)
    const uint32_t* voxel;   // voxel word written at the light's voxel
    uint32_t* target;        // packed target voxel of this frame or kNoTarget
    uint32_t* claim;         // [kClaimBuckets] lowest light index per target bucket
    uint32_t n;
};

hipError_t launch_edit_bricks(uint32_t* pool, uint32_t highWater, uint32_t frame, uint32_t seed, uint32_t permille,
                              uint32_t steps, hipStream_t s);
hipError_t launch_move_lights(const LightMotionView& m, const uint32_t* grid, uint32_t* pool, uint32_t cap,
                              hipStream_t s);

}  // namespace vr
