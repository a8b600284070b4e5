// vr_dynamics.hip - scene motion on the device: moving emitters and random-
// walk brick edits (the dynamic-scene benchmark configuration; no
// counterpart in the reference, whose worlds change only through host edits
// and World::Commit).
//
// Both run between frames on the context's stream and change the device
// copy of the world in place: brick payloads and the light records.  The top
// grid and the uber grid never change (no cell changes kind), so the DDA's
// acceleration data stays valid.
#include "vr_dynamics.h"

namespace vr {

namespace {

__device__ __forceinline__ uint32_t edit_hash(uint32_t id, uint32_t frame, uint32_t seed)
{
    return wang_hash((id * 0x9e3779b1u) ^ seed ^ (frame * 0x85ebca77u));
}

// Brick `id` (below the pool's high water) is edited in `frame` when
// edit_hash % 1000 < permille.  Its walker starts at voxel (s & 7, s >> 3 & 7,
// s >> 6 & 7) of s = wang_hash(h) with colour 0xF000 | (s >> 9 & 0xFFF)
// (opaque, no emission); each of `steps` steps draws r = xorshift(s), moves
// one voxel along axis r % 3 (up if r >> 8 & 1, wrapping inside the brick)
// and toggles the voxel it lands on between empty and the colour; emissive
// voxels are left alone.  One lane per brick.
__global__ __launch_bounds__(256) void k_edit_bricks(uint32_t* __restrict__ pool, uint32_t highWater, uint32_t frame,
                                                     uint32_t seed, uint32_t permille, uint32_t steps)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= highWater) return;
    const uint32_t h = edit_hash(id, frame, seed);
    if (h % 1000u >= permille) return;
    uint32_t s = wang_hash(h);
    uint32_t c[3] = {s & 7u, (s >> 3) & 7u, (s >> 6) & 7u};
    const uint32_t colour = 0xF000u | ((s >> 9) & 0xFFFu);
    uint32_t* b = pool + (size_t)id * 512u;
    for (uint32_t k = 0; k < steps; k++) {
        const uint32_t r = xorshift(s);
        const uint32_t axis = r % 3u;
        const uint32_t step = ((r >> 8) & 1u) ? 1u : 7u;
        if (axis == 0u) c[0] = (c[0] + step) & 7u;
        else if (axis == 1u) c[1] = (c[1] + step) & 7u;
        else c[2] = (c[2] + step) & 7u;
        uint32_t* v = b + c[0] + c[1] * 8u + c[2] * 64u;
        const uint32_t old = *v;
        if ((old >> 16) & 255u) continue;
        *v = old ? 0u : colour;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

// floor, clamped to the map [0, 1023] (NaN -> 0)
__device__ __forceinline__ uint32_t to_voxel(float f)
{
    const float fl = floorf(f);
    return !(fl > 0.0f) ? 0u : (fl >= 1023.0f ? 1023u : (uint32_t)fl);
}

// pool slot of voxel (x, y, z) when its top cell is a brick
__device__ __forceinline__ bool voxel_slot(const uint32_t* grid, uint32_t cap, uint32_t x, uint32_t y, uint32_t z,
                                           size_t& slot)
{
    const uint32_t g = grid[(x >> 3) + (z >> 3) * 128u + (y >> 3) * 16384u];
    if (!(g & 1u) || (g >> 1) >= cap) return false;
    slot = (size_t)(g >> 1) * 512u + (x & 7u) + (y & 7u) * 8u + (z & 7u) * 64u;
    return true;
}

__device__ __forceinline__ uint32_t claim_bucket(uint32_t packed) { return wang_hash(packed) & (kClaimBuckets - 1u); }

__device__ __forceinline__ void bounce(const LightMotionView& m, uint32_t i, uint32_t cx, uint32_t cy, uint32_t cz)
{
    const float4 v = m.vel[i];
    m.vel[i] = make_float4(-v.x, -v.y, -v.z, 0.0f);
    m.pos[i] = make_float4((float)cx + 0.5f, (float)cy + 0.5f, (float)cz + 0.5f, 0.0f);
}

// Light i of size 1 advances its position by its velocity.  Staying in its
// voxel just keeps the new position.  A new voxel that is empty and inside a
// brick cell becomes its target and the light claims the target's bucket
// (lowest index wins, so the outcome does not depend on timing); any other
// new voxel makes it bounce: velocity reversed, position back to its voxel
// centre.
__global__ __launch_bounds__(256) void k_light_claim(LightMotionView m, const uint32_t* __restrict__ grid,
                                                     const uint32_t* __restrict__ pool, uint32_t cap)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= m.n) return;
    m.target[i] = kNoTarget;
    if (m.la[i].w != 1.0f) return;
    const uint32_t cur = __float_as_uint(m.lb[i].w);
    const uint32_t cx = cur & 1023u, cz = (cur >> 10) & 1023u, cy = cur >> 20;
    const float4 p = m.pos[i], v = m.vel[i];
    const float nx = p.x + v.x, ny = p.y + v.y, nz = p.z + v.z;
    const uint32_t tx = to_voxel(nx), ty = to_voxel(ny), tz = to_voxel(nz);
    if (tx == cx && ty == cy && tz == cz) {
        m.pos[i] = make_float4(nx, ny, nz, 0.0f);
        return;
    }
    size_t slot;
    if (voxel_slot(grid, cap, tx, ty, tz, slot) && pool[slot] == 0u) {
        const uint32_t t = tx | (tz << 10) | (ty << 20);
        m.target[i] = t;
        atomicMin(&m.claim[claim_bucket(t)], i);
    } else {
        bounce(m, i, cx, cy, cz);
    }
}

// The claim winners move: voxel word written at the target, the old voxel
// cleared, the light record (centre, packed position) updated; the losers
// bounce.
__global__ __launch_bounds__(256) void k_light_move(LightMotionView m, const uint32_t* __restrict__ grid,
                                                    uint32_t* __restrict__ pool, uint32_t cap)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= m.n) return;
    const uint32_t t = m.target[i];
    if (t == kNoTarget) return;
    const uint32_t cur = __float_as_uint(m.lb[i].w);
    const uint32_t cx = cur & 1023u, cz = (cur >> 10) & 1023u, cy = cur >> 20;
    if (m.claim[claim_bucket(t)] != i) {
        bounce(m, i, cx, cy, cz);
        return;
    }
    const uint32_t tx = t & 1023u, tz = (t >> 10) & 1023u, ty = t >> 20;
    size_t st = 0, sc = 0;
    (void)voxel_slot(grid, cap, tx, ty, tz, st);  // the claim checked it
    if (voxel_slot(grid, cap, cx, cy, cz, sc)) pool[sc] = 0u;
    pool[st] = m.voxel[i];
    const float4 p = m.pos[i], v = m.vel[i];
    m.pos[i] = make_float4(p.x + v.x, p.y + v.y, p.z + v.z, 0.0f);
    m.la[i] = make_float4((float)tx + 0.5f, (float)ty + 0.5f, (float)tz + 0.5f, 1.0f);
    const float4 b = m.lb[i];
    m.lb[i] = make_float4(b.x, b.y, b.z, __uint_as_float(t));
}

}  // namespace

hipError_t launch_edit_bricks(uint32_t* pool, uint32_t highWater, uint32_t frame, uint32_t seed, uint32_t permille,
                              uint32_t steps, hipStream_t s)
{
    if (highWater == 0 || permille == 0 || steps == 0) return hipSuccess;
    hipLaunchKernelGGL(k_edit_bricks, dim3((highWater + 255) / 256), dim3(256), 0, s, pool, highWater, frame, seed,
                       permille, steps);
    return hipGetLastError();
}

hipError_t launch_move_lights(const LightMotionView& m, const uint32_t* grid, uint32_t* pool, uint32_t cap,
                              hipStream_t s)
{
    if (m.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(kClaimBuckets / 256), dim3(256), 0, s, m.claim, kClaimBuckets, kNoTarget);
    hipLaunchKernelGGL(k_light_claim, dim3((m.n + 255) / 256), dim3(256), 0, s, m, grid, pool, cap);
    hipLaunchKernelGGL(k_light_move, dim3((m.n + 255) / 256), dim3(256), 0, s, m, grid, pool, cap);
    return hipGetLastError();
}

}  // namespace vr
