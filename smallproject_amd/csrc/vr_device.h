// vr_device.h - device-side building blocks of the Voxel ReSTIR path (gfx950).
//
// Arithmetic contract (shared with every kernel of the library): binary32
// with correctly rounded + - * / sqrt, no FMA contraction (-ffp-contract=off),
// saturating float->int conversions.  Under this contract the kernels are a
// deterministic function of their inputs, so results are reproducible bit for
// bit on any IEEE machine that follows the same operation order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vr {

constexpr float kInvPi = 0.31830988618379067153777f;  // template/common.h:197
constexpr float kTwoPi = 6.28318530717958647692528f;  // template/common.h:199

// ---------------------------------------------------------------------------
// device views
// ---------------------------------------------------------------------------
struct WorldView {
    const uint32_t* __restrict__ grid;    // [128^3] cell: 0 empty, c<<1 solid, id<<1|1 brick
    const uint32_t* __restrict__ bricks;  // pool [id*512 + lx + ly*8 + lz*64]
    const uint32_t* __restrict__ uber;    // [32^3 bits] cell x + z*32 + y*1024: some of its 4^3 cells non-empty
};

// Light record prepared by the host from struct Light (common.h:134-139):
// a = (centre.xyz, size), b = (Le.rgb, position bits).  Le is
// ToFloatRGB(v) * EmitStrength(v) [* 1/pi with GFXEXP] as phat.cl:8-12
// computes it per evaluation.
struct LightView {
    const float4* __restrict__ a;
    const float4* __restrict__ b;
    // power source: bucket i of the alias table as (prob_i, alias_i bits,
    // pmf_i, pmf_alias_i), so one 16-byte load yields the light and its pdf;
    // null for the uniform source
    const float4* __restrict__ pick;
    uint32_t n;
};

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------
// Saturating conversions, truncating toward zero, NaN -> 0.  gfx950's
// v_cvt_u32_f32 / v_cvt_i32_f32 define exactly this for every input, so each
// is one instruction; written in C++ the range checks cost 4 selects (f2u)
// or a branch nest (f2i) per conversion.
__device__ __forceinline__ uint32_t f2u(float x)
{
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ int32_t f2i(float x)
{
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ uint32_t clampu(uint32_t v, uint32_t lo, uint32_t hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 normalize(f3 v)
{
    const float inv = 1.0f / sqrtf(dot(v, v));
    return f3{v.x * inv, v.y * inv, v.z * inv};
}

// ---------------------------------------------------------------------------
// RNG (cl/tools.cl:17-19)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wang_hash(uint32_t s)
{
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    return s ^ (s >> 15);
}

__device__ __forceinline__ uint32_t xorshift(uint32_t& s)
{
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

__device__ __forceinline__ float rnd(uint32_t& s) { return (float)xorshift(s) * 2.3283064365387e-10f; }

// ---------------------------------------------------------------------------
// voxel codec (cl/tools.cl:294-322)
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 to_rgb(uint32_t v)
{
    return f3{(float)((v >> 8) & 15u) * (1.0f / 15.0f), (float)((v >> 4) & 15u) * (1.0f / 15.0f),
              (float)(v & 15u) * (1.0f / 15.0f)};
}
__device__ __forceinline__ float alphaf(uint32_t v) { return (float)((v & 0xf000u) >> 12) / 15.0f; }
__device__ __forceinline__ bool is_emitter(uint32_t v) { return ((v >> 16) & 255u) != 0u; }

// VoxelNormal (cl/tools.cl:416-421)
__device__ __forceinline__ f3 voxel_normal(uint32_t side, f3 D)
{
    const float dx = side == 0 ? D.x : (side == 1 ? D.y : D.z);
    const float s = dx > 0 ? -1.0f : 1.0f;
    return f3{side == 0 ? s : 0.0f, side == 1 ? s : 0.0f, side == 2 ? s : 0.0f};
}

// face index of an axis normal: +x,+y,+z,-x,-y,-z = 0..5 (tools.cl:502-510)
__device__ __forceinline__ f3 face_normal(uint32_t f)
{
    if (f > 5u) return f3{0.0f, 0.0f, 0.0f};
    const float s = f < 3u ? 1.0f : -1.0f;
    const uint32_t a = f < 3u ? f : f - 3u;
    return f3{a == 0 ? s : 0.0f, a == 1 ? s : 0.0f, a == 2 ? s : 0.0f};
}

// ---------------------------------------------------------------------------
// sine/cosine: Cody-Waite reduction by pi/2 and Taylor polynomials, all in
// binary64 with a fixed operation order, rounded once to binary32.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sincos_det(float angle, float& s_out, float& c_out)
{
    const double x = (double)angle;
    const double q = floor(x * 0.63661977236758134308 + 0.5);
    const double r = (x - q * 1.5707963267341256) - q * 6.0771005065061922e-11;
    const double r2 = r * r;
    double sp = 1.0 / 121645100408832000.0;
    sp = sp * -r2 + 1.0 / 355687428096000.0;
    sp = sp * -r2 + 1.0 / 1307674368000.0;
    sp = sp * -r2 + 1.0 / 6227020800.0;
    sp = sp * -r2 + 1.0 / 39916800.0;
    sp = sp * -r2 + 1.0 / 362880.0;
    sp = sp * -r2 + 1.0 / 5040.0;
    sp = sp * -r2 + 1.0 / 120.0;
    sp = sp * -r2 + 1.0 / 6.0;
    sp = sp * -r2 + 1.0;
    const double sn = sp * r;
    double cp = 1.0 / 6402373705728000.0;
    cp = cp * -r2 + 1.0 / 20922789888000.0;
    cp = cp * -r2 + 1.0 / 87178291200.0;
    cp = cp * -r2 + 1.0 / 479001600.0;
    cp = cp * -r2 + 1.0 / 3628800.0;
    cp = cp * -r2 + 1.0 / 40320.0;
    cp = cp * -r2 + 1.0 / 720.0;
    cp = cp * -r2 + 1.0 / 24.0;
    cp = cp * -r2 + 1.0 / 2.0;
    cp = cp * -r2 + 1.0;
    const int quad = ((int)q) & 3;
    double so, co;
    if (quad == 0) so = sn, co = cp;
    else if (quad == 1) so = cp, co = -sn;
    else if (quad == 2) so = -sn, co = -cp;
    else so = -cp, co = sn;
    s_out = (float)so;
    c_out = (float)co;
}

// ---------------------------------------------------------------------------
// Three-level DDA over the 32^3 uber grid, 128^3 top grid and 8^3 bricks,
// cl/trace.cl:121-362.  Positions are packed x<<20 | y<<10 | z at each level
// so one add steps a coordinate and one mask test detects leaving a cell.
// ---------------------------------------------------------------------------
enum TraceKind { kTraceAll = 0, kTraceOpaque = 1, kTraceThrough = 2 };

__device__ __forceinline__ uint32_t uber_bit(const uint32_t* uber, uint32_t i) { return (uber[i >> 5] >> (i & 31u)) & 1u; }

template <int KIND>
__device__ __forceinline__ bool voxel_stops(uint32_t v, uint32_t through)
{
    if (KIND == kTraceAll) return v != 0u;
    if (KIND == kTraceOpaque) return (v & 0xf000u) == 0xf000u;
    return (v & 0x00ffffu) != (through & 0x00ffffu);
}

constexpr uint32_t kTopMask3 = ((1023u - 7u) << 20) + ((1023u - 7u) << 10) + (1023u - 7u);
constexpr uint32_t kUberMask3 = (1020u << 20) + (1020u << 10) + 1020u;

template <int KIND>
__device__ uint32_t trace(const WorldView& w, f3 A, f3 B, int steps, uint32_t through, float& dist,
                          uint32_t& side)
{
    // FixZeroDeltas (trace.cl:37-43)
    if (fabsf(B.x) < 1e-8f) B.x = B.x < 0 ? -1e-8f : 1e-8f;
    if (fabsf(B.y) < 1e-8f) B.y = B.y < 0 ? -1e-8f : 1e-8f;
    if (fabsf(B.z) < 1e-8f) B.z = B.z < 0 ? -1e-8f : 1e-8f;
    const f3 V = B;
    const f3 rV = f3{1.0f / V.x, 1.0f / V.y, 1.0f / V.z};
    const uint32_t ox = V.x > 0 ? 1u : 0u, oy = V.y > 0 ? 1u : 0u, oz = V.z > 0 ? 1u : 0u;
    const uint32_t dx = V.x > 0 ? (1u << 20) : (0xfffu << 20);
    const uint32_t dy = V.y > 0 ? (1u << 10) : 0xfffffc00u;
    const uint32_t dz = V.z > 0 ? 1u : 0xffffffffu;
    // cell crossing steps sign(V) * (1/V), which is |1/V| exactly (V != 0):
    // the absolute value is a free source modifier, so no registers hold it
    const f3 td = f3{fabsf(rV.x), fabsf(rV.y), fabsf(rV.z)};
    float to = 0.0f;
    uint32_t last = 0;
    if (A.x < 0 || A.y < 0 || A.z < 0 || A.x > 1024.0f || A.y > 1024.0f || A.z > 1024.0f) {
        // slab clip against [0,1024]^3 (trace.cl:303-317)
        const float tx1 = -A.x * rV.x, tx2 = (1024.0f - A.x) * rV.x;
        float tmin = fminf(tx1, tx2), tmax = fmaxf(tx1, tx2);
        const float ty1 = -A.y * rV.y, ty2 = (1024.0f - A.y) * rV.y;
        tmin = fmaxf(tmin, fminf(ty1, ty2)), tmax = fminf(tmax, fmaxf(ty1, ty2));
        const float tz1 = -A.z * rV.z, tz2 = (1024.0f - A.z) * rV.z;
        tmin = fmaxf(tmin, fminf(tz1, tz2)), tmax = fminf(tmax, fmaxf(tz1, tz2));
        if (tmax < tmin || tmax <= 0) return 0u;
        A.x += tmin * V.x, A.y += tmin * V.y, A.z += tmin * V.z;
        to = tmin;
        if (A.y < 0.01f || A.y > (1024.0f - 1.01f)) last = 1;
        if (A.z < 0.01f || A.z > (1024.0f - 1.01f)) last = 2;
    }
    uint32_t up = (min(f2u(A.x) >> 5, 31u) << 20) + (min(f2u(A.y) >> 5, 31u) << 10) + min(f2u(A.z) >> 5, 31u);
    float tmx = ((float)((up >> 20) + ox) - A.x * 0.03125f) * rV.x;
    float tmy = ((float)(((up >> 10) & 31u) + oy) - A.y * 0.03125f) * rV.y;
    float tmz = ((float)((up & 31u) + oz) - A.z * 0.03125f) * rV.z;
    float t = 0.0f;
    uint32_t o = uber_bit(w.uber, (up >> 20) + ((up & 31u) << 5) + (((up >> 10) & 31u) << 10));
    while (true) {
        if ((steps -= 4) <= 0) break;
        if (KIND == kTraceThrough ? (o != 0u || through != 0u) : (o != 0u)) {
            const float ux = tmx, uy = tmy, uz = tmz;
            t *= 4.0f;
            const uint32_t p4x = f2u(0.125f * A.x + V.x * t);
            const uint32_t p4y = f2u(0.125f * A.y + V.y * t);
            const uint32_t p4z = f2u(0.125f * A.z + V.z * t);
            const uint32_t bx4 = (up >> 20) << 2, by4 = ((up >> 10) & 31u) << 2, bz4 = (up & 31u) << 2;
            uint32_t tp = (clampu(p4x, bx4, bx4 + 3) << 20) + (clampu(p4y, by4, by4 + 3) << 10) +
                          clampu(p4z, bz4, bz4 + 3);
            const uint32_t tq = tp & kUberMask3;
            tmx = ((float)((tp >> 20) + ox) - A.x * 0.125f) * rV.x;
            tmy = ((float)(((tp >> 10) & 127u) + oy) - A.y * 0.125f) * rV.y;
            tmz = ((float)((tp & 127u) + oz) - A.z * 0.125f) * rV.z;
            o = w.grid[(tp >> 20) + ((tp & 127u) << 7) + (((tp >> 10) & 127u) << 14)];
            while (true) {
                if (!--steps) break;
                const uint32_t v = o >> 1;
                if ((o & 1u) == 0u) {
                    if (voxel_stops<KIND>(v, through)) {
                        dist = t * 8.0f + to, side = last;
                        return v;
                    }
                } else {
                    const float gx = tmx, gy = tmy, gz = tmz;
                    t *= 8.0f;
                    const uint32_t q4x = f2u(A.x + V.x * t);
                    const uint32_t q4y = f2u(A.y + V.y * t);
                    const uint32_t q4z = f2u(A.z + V.z * t);
                    const uint32_t cx8 = (tp >> 20) << 3, cy8 = ((tp >> 10) & 127u) << 3, cz8 = (tp & 127u) << 3;
                    uint32_t p = (clampu(q4x, cx8, cx8 + 7) << 20) + (clampu(q4y, cy8, cy8 + 7) << 10) +
                                 clampu(q4z, cz8, cz8 + 7);
                    tmx = ((float)((p >> 20) + ox) - A.x) * rV.x;
                    tmy = ((float)(((p >> 10) & 1023u) + oy) - A.y) * rV.y;
                    tmz = ((float)((p & 1023u) + oz) - A.z) * rV.z;
                    p &= 7u + (7u << 10) + (7u << 20);
                    const uint32_t* __restrict__ brick = w.bricks + (size_t)(o >> 1) * 512u;
                    // at most 22 steps cross a brick; the reference unrolls 24.
                    // Software-pipelined: the next voxel's position does not
                    // depend on the current voxel, so its load is issued
                    // before the current one is tested and two loads are in
                    // flight per lane.  The state on return and on leaving
                    // the loop is the plain loop's.
                    uint32_t vv = brick[(p >> 20) + ((p >> 7) & 56u) + ((p & 7u) << 6)];
                    for (int k = 0; k < 24; k++) {
                        const float tn = fminf(tmx, fminf(tmy, tmz));
                        float nx = tmx, ny = tmy, nz = tmz;
                        uint32_t pn = p, ln = 0;
                        if (tn == tmx) nx += td.x, pn += dx;
                        if (tn == tmy) ny += td.y, pn += dy, ln = 1;
                        if (tn == tmz) nz += td.z, pn += dz, ln = 2;
                        const bool more = !(pn & kTopMask3) && k + 1 < 24;
                        const uint32_t vn = more ? brick[(pn >> 20) + ((pn >> 7) & 56u) + ((pn & 7u) << 6)] : 0u;
                        if (voxel_stops<KIND>(vv, through)) {
                            dist = t + to, side = last;
                            return vv;
                        }
                        t = tn, last = ln, tmx = nx, tmy = ny, tmz = nz, p = pn;
                        if (!more) break;
                        vv = vn;
                    }
                    tmx = gx, tmy = gy, tmz = gz;
                }
                t = fminf(tmx, fminf(tmy, tmz)), last = 0;
                if (t == tmx) tmx += td.x, tp += dx;
                if (t == tmy) tmy += td.y, tp += dy, last = 1;
                if (t == tmz) tmz += td.z, tp += dz, last = 2;
                if ((tp & kUberMask3) != tq) break;
                o = w.grid[(tp >> 20) + ((tp & 127u) << 7) + (((tp >> 10) & 127u) << 14)];
            }
            tmx = ux, tmy = uy, tmz = uz;
        }
        t = fminf(tmx, fminf(tmy, tmz)), last = 0;
        if (t == tmx) tmx += td.x, up += dx;
        if (t == tmy) tmy += td.y, up += dy, last = 1;
        if (t == tmz) tmz += td.z, up += dz, last = 2;
        if (up & 0xfe0f83e0u) break;
        o = uber_bit(w.uber, (up >> 20) + ((up & 31u) << 5) + (((up >> 10) & 31u) << 10));
    }
    return 0u;
}

// voxel lx + 8 ly + 64 lz of a brick from a packed in-brick position
// (lx << 20 | ly << 10 | lz, each 0..7): the three fields land on disjoint
// bits, so a field extract and two shift-ors build the index
__device__ __forceinline__ uint32_t brick_voxel(uint32_t p)
{
    return (p >> 20) | (__builtin_amdgcn_ubfe(p, 10u, 3u) << 3) | ((p & 7u) << 6);
}

// TraceOpaqueRay (trace.cl:205-282) for rays that go their own ways, as
// shadow rays do.  trace() nests the brick walk inside the top-grid loop
// inside the uber loop, so a wave runs each brick walk while the lanes that
// are not inside a brick wait (on c2's shadow rays about 4 of 64 lanes take
// each brick step).  Here the same traversal is cut into units - one visit
// and step of an uber cell, of a top cell, or of a voxel - and each turn of
// the wave's loop runs the unit of one level, chosen from how many of the
// wave's lanes are at each (ballot counts, so the choice is a scalar
// branch); the other lanes keep their state.  Every lane performs the operations
// of trace() in the same order (each level keeps its own tm registers, which
// trace() saves and restores around a descent), so the results are the same
// bits.  `resume`: the lane's next unit at its level is the step that
// follows a child level's exit (trace() steps the parent right after the
// child loop ends).  While its steps advance a coordinate, a brick walk
// leaves its brick within 22 steps (7 per axis and the one that leaves), so
// trace()'s cap of 24 steps per brick never ends it; a ray whose steps do
// not advance (a NaN direction: a shadow ray gets one only when its sample
// point is the shading point itself) is ended by that cap and the step
// budget, and kCap keeps the cap (see trace_sched below).
template <bool kCap>
__device__ uint32_t trace_sched_walk(const WorldView& w, f3 A, f3 B, int steps, float& dist, uint32_t& side)
{
    if (fabsf(B.x) < 1e-8f) B.x = B.x < 0 ? -1e-8f : 1e-8f;
    if (fabsf(B.y) < 1e-8f) B.y = B.y < 0 ? -1e-8f : 1e-8f;
    if (fabsf(B.z) < 1e-8f) B.z = B.z < 0 ? -1e-8f : 1e-8f;
    const f3 V = B;
    const f3 rV = f3{1.0f / V.x, 1.0f / V.y, 1.0f / V.z};
    const uint32_t ox = V.x > 0 ? 1u : 0u, oy = V.y > 0 ? 1u : 0u, oz = V.z > 0 ? 1u : 0u;
    const uint32_t dx = V.x > 0 ? (1u << 20) : (0xfffu << 20);
    const uint32_t dy = V.y > 0 ? (1u << 10) : 0xfffffc00u;
    const uint32_t dz = V.z > 0 ? 1u : 0xffffffffu;
    const f3 td = f3{fabsf(rV.x), fabsf(rV.y), fabsf(rV.z)};
    float to = 0.0f;
    uint32_t last = 0;
    bool active = true;
    if (A.x < 0 || A.y < 0 || A.z < 0 || A.x > 1024.0f || A.y > 1024.0f || A.z > 1024.0f) {
        const float tx1 = -A.x * rV.x, tx2 = (1024.0f - A.x) * rV.x;
        float tmin = fminf(tx1, tx2), tmax = fmaxf(tx1, tx2);
        const float ty1 = -A.y * rV.y, ty2 = (1024.0f - A.y) * rV.y;
        tmin = fmaxf(tmin, fminf(ty1, ty2)), tmax = fminf(tmax, fmaxf(ty1, ty2));
        const float tz1 = -A.z * rV.z, tz2 = (1024.0f - A.z) * rV.z;
        tmin = fmaxf(tmin, fminf(tz1, tz2)), tmax = fminf(tmax, fmaxf(tz1, tz2));
        if (tmax < tmin || tmax <= 0) active = false;
        A.x += tmin * V.x, A.y += tmin * V.y, A.z += tmin * V.z;
        to = tmin;
        if (A.y < 0.01f || A.y > (1024.0f - 1.01f)) last = 1;
        if (A.z < 0.01f || A.z > (1024.0f - 1.01f)) last = 2;
    }
    uint32_t up = (min(f2u(A.x) >> 5, 31u) << 20) + (min(f2u(A.y) >> 5, 31u) << 10) + min(f2u(A.z) >> 5, 31u);
    // tm of the uber level (u*), of the top grid (g*) and of the brick (b*)
    float ux = ((float)((up >> 20) + ox) - A.x * 0.03125f) * rV.x;
    float uy = ((float)(((up >> 10) & 31u) + oy) - A.y * 0.03125f) * rV.y;
    float uz = ((float)((up & 31u) + oz) - A.z * 0.03125f) * rV.z;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f;
    float t = 0.0f;
    uint32_t tp = 0u, p = 0u, brick = 0u;
    // the lane's level: 0 uber, 1 top grid, 2 brick; 3 done without a hit,
    // 4 hit a solid top-grid cell, 5 hit a voxel.  `resume` (see above) for
    // the level it is at.  Inside a unit the lanes of that level update their
    // state with selects (the other lanes keep theirs), so the loop has
    // little control flow for the compiler to copy registers across.
    uint32_t lv = active ? 0u : 3u, resume = 0u;
    // the occupancy of the lane's current cell: uber bit, top-grid word or voxel
    uint32_t o = active ? uber_bit(w.uber, (up >> 20) + ((up & 31u) << 5) + (((up >> 10) & 31u) << 10)) : 0u;
    while (true) {
        const uint32_t n0 = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(lv == 0u));
        const uint32_t n1 = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(lv == 1u));
        const uint32_t n2 = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(lv == 2u));
        if ((n0 | n1 | n2) == 0u) break;
        // The level to run: bricks when at least half as many lanes are in
        // a brick as at either other level, then the top grid when at least
        // twice as many lanes are there as at the uber level, else the uber
        // level (of the weightings tried on c2, the fastest: brick units are
        // the cheapest, and running the uber level first gathers lanes into
        // the top grid's more costly units)
        if (2u * n2 >= n1 && 2u * n2 >= n0) {
            // voxels of the brick walks, for as long as that holds (an inner
            // loop: its state stays in place between turns)
            uint32_t m1 = n1, m2 = n2;
            do {
                const bool mine = lv == 2u;
                const bool stop = mine && (o & 0xf000u) == 0xf000u;
                const bool go = mine && !stop;
                const float tn = fminf(bx, fminf(by, bz));
                const bool sx = go && tn == bx, sy = go && tn == by, sz = go && tn == bz;
                bx = sx ? bx + td.x : bx;
                by = sy ? by + td.y : by;
                bz = sz ? bz + td.z : bz;
                p += (sx ? dx : 0u) + (sy ? dy : 0u) + (sz ? dz : 0u);
                last = go ? (sz ? 2u : (sy ? 1u : 0u)) : last;
                t = go ? tn : t;
                const bool out = go && ((p & kTopMask3) != 0u || (kCap && resume == 23u));
                lv = stop ? 5u : (out ? 1u : lv);
                resume = out ? 1u : (kCap && go ? resume + 1u : resume);
                if (go && !out) o = w.bricks[(size_t)brick * 512u + brick_voxel(p)];
                m1 = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(lv == 1u));
                m2 = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(lv == 2u));
            } while (2u * m2 >= m1 && 2u * m2 >= n0 && m2 != 0u);
        } else if (n1 >= 2u * n0) {
            // a cell of the top grid: visit (budget, solid cell, brick), then
            // step unless the lane stopped or went into a brick
            const bool mine = lv == 1u;
            const bool visit = mine && resume == 0u;
            steps = visit ? steps - 1 : steps;
            const bool spent = visit && steps == 0;
            const bool solid = visit && !spent && (o & 1u) == 0u;
            const bool stop = solid && ((o >> 1) & 0xf000u) == 0xf000u;
            const bool down = visit && !spent && (o & 1u) != 0u;
            if (down) {
                t *= 8.0f;
                const uint32_t q4x = f2u(A.x + V.x * t);
                const uint32_t q4y = f2u(A.y + V.y * t);
                const uint32_t q4z = f2u(A.z + V.z * t);
                const uint32_t cx8 = (tp >> 20) << 3, cy8 = ((tp >> 10) & 127u) << 3, cz8 = (tp & 127u) << 3;
                p = (clampu(q4x, cx8, cx8 + 7) << 20) + (clampu(q4y, cy8, cy8 + 7) << 10) + clampu(q4z, cz8, cz8 + 7);
                bx = ((float)((p >> 20) + ox) - A.x) * rV.x;
                by = ((float)(((p >> 10) & 1023u) + oy) - A.y) * rV.y;
                bz = ((float)((p & 1023u) + oz) - A.z) * rV.z;
                p &= 7u + (7u << 10) + (7u << 20);
                brick = o >> 1;
                o = w.bricks[(size_t)brick * 512u + brick_voxel(p)];
            }
            const bool go = mine && !spent && !stop && !down;
            const float tn = fminf(gx, fminf(gy, gz));
            const bool sx = go && tn == gx, sy = go && tn == gy, sz = go && tn == gz;
            gx = sx ? gx + td.x : gx;
            gy = sy ? gy + td.y : gy;
            gz = sz ? gz + td.z : gz;
            tp += (sx ? dx : 0u) + (sy ? dy : 0u) + (sz ? dz : 0u);
            last = go ? (sz ? 2u : (sy ? 1u : 0u)) : last;
            t = go ? tn : t;
            const bool out = go && (tp & kUberMask3) != (up << 2);
            lv = stop ? 4u : (down ? 2u : ((spent || out) ? 0u : lv));
            resume = (spent || out) ? 1u : (mine ? 0u : resume);
            if (go && !out) o = w.grid[(tp >> 20) + ((tp & 127u) << 7) + (((tp >> 10) & 127u) << 14)];
        } else {
            // a cell of the uber grid: visit (budget, occupied cell), then
            // step unless the lane stopped or went into the top grid
            const bool mine = lv == 0u;
            const bool visit = mine && resume == 0u;
            steps = visit ? steps - 4 : steps;
            const bool spent = visit && steps <= 0;
            const bool down = visit && !spent && o != 0u;
            if (down) {
                t *= 4.0f;
                const uint32_t p4x = f2u(0.125f * A.x + V.x * t);
                const uint32_t p4y = f2u(0.125f * A.y + V.y * t);
                const uint32_t p4z = f2u(0.125f * A.z + V.z * t);
                const uint32_t bx4 = (up >> 20) << 2, by4 = ((up >> 10) & 31u) << 2, bz4 = (up & 31u) << 2;
                tp = (clampu(p4x, bx4, bx4 + 3) << 20) + (clampu(p4y, by4, by4 + 3) << 10) + clampu(p4z, bz4, bz4 + 3);
                gx = ((float)((tp >> 20) + ox) - A.x * 0.125f) * rV.x;
                gy = ((float)(((tp >> 10) & 127u) + oy) - A.y * 0.125f) * rV.y;
                gz = ((float)((tp & 127u) + oz) - A.z * 0.125f) * rV.z;
                o = w.grid[(tp >> 20) + ((tp & 127u) << 7) + (((tp >> 10) & 127u) << 14)];
            }
            const bool go = mine && !spent && !down;
            const float tn = fminf(ux, fminf(uy, uz));
            const bool sx = go && tn == ux, sy = go && tn == uy, sz = go && tn == uz;
            ux = sx ? ux + td.x : ux;
            uy = sy ? uy + td.y : uy;
            uz = sz ? uz + td.z : uz;
            up += (sx ? dx : 0u) + (sy ? dy : 0u) + (sz ? dz : 0u);
            last = go ? (sz ? 2u : (sy ? 1u : 0u)) : last;
            t = go ? tn : t;
            const bool out = go && (up & 0xfe0f83e0u) != 0u;
            lv = (spent || out) ? 3u : (down ? 1u : lv);
            resume = mine ? 0u : resume;
            if (go && !out) o = uber_bit(w.uber, (up >> 20) + ((up & 31u) << 5) + (((up >> 10) & 31u) << 10));
        }
    }
    uint32_t hit = 0u;
    if (lv == 5u) hit = o, dist = t + to, side = last;
    if (lv == 4u) hit = o >> 1, dist = t * 8.0f + to, side = last;
    return hit;
}

// With every coordinate of the origin and the direction within +-2^20 (so
// finite: a NaN fails the comparisons), 2^-27 < |1/V| < 2^27 after
// FixZeroDeltas, and every tm, t and position product of the walk stays
// finite, so each step advances a coordinate.  A wave with a ray outside
// those bounds (a NaN direction; vr_trace_rays refuses non-finite rays but
// takes large ones) walks with trace()'s cap of 24 steps per brick, counted
// in `resume`; other waves do not pay for the count.
__device__ __forceinline__ uint32_t trace_sched(const WorldView& w, f3 A, f3 B, int steps, float& dist, uint32_t& side)
{
    constexpr float kBound = 1048576.0f;
    const bool bounded = fabsf(A.x) <= kBound && fabsf(A.y) <= kBound && fabsf(A.z) <= kBound &&
                         fabsf(B.x) <= kBound && fabsf(B.y) <= kBound && fabsf(B.z) <= kBound;
    if (__builtin_amdgcn_ballot_w64(!bounded) != 0u) return trace_sched_walk<true>(w, A, B, steps, dist, side);
    return trace_sched_walk<false>(w, A, B, steps, dist, side);
}

// ---------------------------------------------------------------------------
// camera ray (cl/tools.cl:102-127, TAA=0, PANINI=0)
// ---------------------------------------------------------------------------
struct Camera {
    float ox, oy;            // oneOverRes
    f3 E, p0, d10, d20;      // eye, p0, p1-p0, p2-p0
};

__device__ __forceinline__ f3 camera_ray(const Camera& c, int x, int y)
{
    const float u = (float)x * c.ox, v = (float)y * c.oy;
    const f3 P = f3{c.p0.x + c.d10.x * u + c.d20.x * v, c.p0.y + c.d10.y * u + c.d20.y * v,
                    c.p0.z + c.d10.z * u + c.d20.z * v};
    return normalize(sub(P, c.E));
}

// ---------------------------------------------------------------------------
// target function p-hat (cl/phat.cl:2-41)
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 phat_full(bool point, f3 pos, f3 Nl, f3 sp, f3 N, f3 brdf, f3 Le, float size)
{
    const f3 R = sub(pos, sp);
    const f3 L = normalize(R);
    const float NdotL = fmaxf(dot(N, L), 0.0f);
    const float d2 = fmaxf(1e-5f, dot(R, R));
    const float area = size * size;
    const float nldl = point ? 1.0f : fmaxf(dot(Nl, neg(L)), 0.0f);
    const float sa = (nldl * area) / d2;
    return f3{brdf.x * Le.x * NdotL * sa, brdf.y * Le.y * NdotL * sa, brdf.z * Le.z * NdotL * sa};
}

__device__ __forceinline__ float phat(bool point, f3 pos, f3 Nl, f3 sp, f3 N, f3 brdf, f3 Le, float size)
{
    const f3 v = phat_full(point, pos, Nl, sp, N, brdf, Le, size);
    const float r = sqrtf(dot(v, v));
    return (isnan(r) || isinf(r)) ? 0.0f : r;
}

// ---------------------------------------------------------------------------
// reservoir in registers (struct Reservoir, common.h:141-151)
// ---------------------------------------------------------------------------
struct Res {
    float wsum, W, pHat, invPdf;
    f3 pos;
    uint32_t light;
    uint32_t M;
    uint32_t face;  // Nlight as face index, 7 = zero vector
};

// true when the new sample takes the reservoir's place
__device__ __forceinline__ bool res_update(Res& r, float w, float pHat, uint32_t index, float r0, f3 pos,
                                           float invPdf, uint32_t face)
{
    // UpdateReservoir (reservoir.cl:3-15)
    r.M += 1u;
    r.wsum += w;
    const bool take = r0 < w / fmaxf(1e-5f, r.wsum);
    if (take) {
        r.light = index, r.pos = pos, r.invPdf = invPdf, r.face = face, r.pHat = pHat;
    }
    return take;
}

__device__ __forceinline__ void res_reweigh(Res& r, float newPHat, uint32_t M)
{
    // ReWeighSumOfWeights (reservoir.cl:22-28)
    r.wsum = r.W * (float)M * newPHat;
    r.M = M;
    r.pHat = newPHat;
}

// true when o's sample takes the reservoir's place
__device__ __forceinline__ bool res_combine(Res& r, const Res& o, float r0)
{
    // CombineReservoir (reservoir.cl:30-38)
    const uint32_t m1 = r.M;
    const bool take = res_update(r, o.wsum, o.pHat, o.light, r0, o.pos, o.invPdf, o.face);
    r.M = m1 + o.M;
    return take;
}

__device__ __forceinline__ void res_adjust(Res& r, float pHat)
{
    // AdjustWeight (reservoir.cl:40-43)
    r.W = r.wsum / fmaxf(1e-5f, pHat * (float)r.M);
}

}  // namespace vr
