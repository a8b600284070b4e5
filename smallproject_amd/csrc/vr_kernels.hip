// vr_kernels.hip - gfx950 kernels of the Voxel ReSTIR frame.
//
// One frame = three launches on the context stream, each one thread per
// pixel with 256-thread workgroups covering a 16x16 pixel tile (four 8x8
// wave64 tiles, so the rays of a wave are spatially coherent):
//   k_primary   renderAlbedo                cl/ris.cl:295-389
//   k_initial   perPixelInitialSampling     cl/ris.cl:20-124
//   k_resolve   perPixelSpatialResampling   cl/ris.cl:126-151 + spatialSampling.cl
//               then renderRIS / render_di_ris cl/ris.cl:155-293, 391-413
// Shading reads only its own pixel's final reservoir, which the spatial
// pass of the same thread has just written (no other pixel's spatial pass
// reads it), so the two run as one kernel.
// Per-pixel state lives in structure-of-arrays planes (12 B of hit data and
// 40 B per reservoir record: 16 + 16 + 8) so every pass streams coalesced
// 4/8/16-byte words.
#include "vr_kernels.h"

namespace vr {

__device__ __forceinline__ void tile_pixel(uint32_t order, int& x, int& y)
{
    int tx, ty;
    tile_of_block(order, tx, ty);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    x = tx * 16 + (wv & 1) * 8 + (lane & 7);
    y = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
}

__device__ __forceinline__ Res res_from(const uint4 r0, const uint4 r1, const float2 r2)
{
    Res o;
    o.M = r0.x & 0xffffffu;
    o.face = (r0.x >> 24) & 7u;
    o.W = __uint_as_float(r0.z), o.invPdf = __uint_as_float(r0.w);
    o.pos = f3{__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z)};
    o.light = r1.w;
    o.wsum = r2.x, o.pHat = r2.y;
    return o;
}

__device__ __forceinline__ Res load_res(const ResPlanes& r, size_t i) { return res_from(r.r0[i], r.r1[i], r.r2[i]); }

// A neighbour's or the previous frame's record (r0) and sample (r1), loaded
// together.  Its tests read the record word, then the distance, and only a
// record that passes needs the sample; left to itself the compiler loads
// each field inside the branch that uses it, three round trips to memory
// in a row.  The empty asm consumes all eight words where they are loaded,
// so both loads are issued back to back and one round trip serves the
// tests (spatial reuse: 0.565 -> 0.555 ms at 1080p).
__device__ __forceinline__ void load_record_sample(const ResPlanes& r, size_t i, uint4& r0, uint4& r1)
{
    uint4 s = r.r1[i];
    uint4 c = r.r0[i];
    asm volatile("" : "+v"(c.x), "+v"(c.y), "+v"(c.z), "+v"(c.w), "+v"(s.x), "+v"(s.y), "+v"(s.z), "+v"(s.w));
    r0 = c, r1 = s;
}

// the record of reservoir o at the pixel with hit word hitw and distance
// dist; shadow: kTraced / kVisible
__device__ __forceinline__ void store_res(const ResPlanes& r, size_t i, const Res& o, uint32_t hitw, float dist,
                                          uint32_t shadow)
{
    r.r0[i] = make_uint4((o.M & 0xffffffu) | ((o.face & 7u) << 24) | (((hitw >> 24) & 7u) << 27) | shadow,
                         __float_as_uint(dist), __float_as_uint(o.W), __float_as_uint(o.invPdf));
    r.r1[i] = make_uint4(__float_as_uint(o.pos.x), __float_as_uint(o.pos.y), __float_as_uint(o.pos.z), o.light);
    r.r2[i] = make_float2(o.wsum, o.pHat);
}

// a record word's hit side bits in hit-word position (bits 24-26)
__device__ __forceinline__ uint32_t rec_hit_bits(uint32_t w) { return (w >> 3) & (7u << 24); }

// Visibility of sample (light, pos) from a record's pixel as the candidate
// pass found it: 1 = visible, 0 = occluded, -1 = not traced (the record's
// sample is another one, or its ray was not traced).
__device__ __forceinline__ int known_visibility(uint32_t w0, const uint4 r1, uint32_t light, f3 pos)
{
    if (!(w0 & kTraced) || r1.w != light || r1.x != __float_as_uint(pos.x) || r1.y != __float_as_uint(pos.y) ||
        r1.z != __float_as_uint(pos.z))
        return -1;
    return (w0 & kVisible) ? 1 : 0;
}

// rays of kind k traced by this lane (only while a frame's rays are counted)
__device__ __forceinline__ void count_rays(unsigned long long* rays, int k, uint32_t n)
{
    if (rays && n) atomicAdd(&rays[k], (unsigned long long)n);
}

// GridCoordinatesFromHit + Uint3CubeEquals (cl/tools.cl:452-472) behind
// isLightOccluded (cl/ris.cl:1-17).  A shadow ray that finds no opaque voxel
// within its step budget counts as occluded.  Shadow rays of a wave head to
// different lights, so they take the wave-scheduled DDA: kPick picks the
// walk per wave (trace_sched: uncounted unless a ray is out of bounds), else
// the walk that counts its brick steps runs always (one copy of the loop:
// for the shading pass's rays a second copy costs more registers than the
// count costs their fewer rays; the MIS merge's kernel, with more registers
// per lane, picks).  kNested: the nested DDA instead
// (a stage's choice, kDdaNested: scenes whose rays stay coherent).
template <bool kPick, bool kNested>
__device__ __forceinline__ bool light_occluded(const WorldView& w, uint32_t lpos, uint32_t lsize, f3 spo, f3 L,
                                               int steps)
{
    float dist = 0.0f;
    uint32_t side = 0;
    const uint32_t v = kNested ? trace<kTraceOpaque>(w, spo, L, steps, 0u, dist, side)
                               : (kPick ? trace_sched(w, spo, L, steps, dist, side)
                                        : trace_sched_walk<true>(w, spo, L, steps, dist, side));
    if (v == 0u) return true;
    const f3 N = voxel_normal(side, L);
    const uint32_t hx = f2u((L.x * dist + spo.x) - N.x * 0.1f);
    const uint32_t hy = f2u((L.y * dist + spo.y) - N.y * 0.1f);
    const uint32_t hz = f2u((L.z * dist + spo.z) - N.z * 0.1f);
    const uint32_t ly = lpos >> 20, lz = (lpos >> 10) & 1023u, lx = lpos & 1023u;
    const bool inside = lx <= hx && lx + lsize > hx && ly <= hy && ly + lsize > hy && lz <= hz && lz + lsize > hz;
    return !inside;
}

// same-face test of temporal / spatial reuse; hit words carry side (bits
// 24-25) and the sign of D[side] (bit 26)
__device__ __forceinline__ bool same_face(uint32_t a, uint32_t b, bool axis_only)
{
    const uint32_t m = axis_only ? (3u << 24) : (7u << 24);
    return ((a ^ b) & m) == 0u;
}

// point of face i (normals +x,+y,+z,-x,-y,-z; tangent axis (a+1)%3,
// bitangent (a+2)%3 with the face's sign, tools.cl:502-518): centre + hs*N -
// hs*T - hs*B + r0*size*T + r1*size*B per axis, where the zero components of
// N, T, B leave the other terms unchanged
__device__ __forceinline__ f3 face_point(f3 c, float hs, float size, int a, float s, float r0, float r1)
{
    const float s0 = r0 * size, s1 = r1 * size;
    const float pa = (a == 0 ? c.x : (a == 1 ? c.y : c.z)) + hs * s;
    const float ct = a == 0 ? c.y : (a == 1 ? c.z : c.x);
    const float cb = a == 0 ? c.z : (a == 1 ? c.x : c.y);
    const float pt = (ct - hs * s) + s0 * s;
    const float pb = (cb - hs * s) + s1 * s;
    return a == 0 ? f3{pa, pt, pb} : (a == 1 ? f3{pb, pa, pt} : f3{pt, pb, pa});
}

// RandomPointOnVoxel (tools.cl:499-577): 12 draws for the six faces, one
// for the face CDF.  Face i = (axis a, sign s) has normal N = s*e_a, so
// dot(N, -L) = -s*L_a exactly, and its solid angle is non-zero only if
// s*R[a] < 0 (otherwise max(0, .) clamps a value <= 0, or NaN for R = 0).
// That holds for at most one face per axis (the +face iff fl(c+hs) < sp,
// the -face iff fl(c-hs) > sp), so three evaluations replace six; a zero's
// sign never reaches an output (it only enters sums and the 1/pdf that is
// zeroed when infinite).  The six faces still draw their two numbers each,
// in order, and the total and the CDF walk add the non-zero terms in face
// order (x + 0 = x exactly), so every output is the six-face computation's.
__device__ __forceinline__ void point_on_voxel(f3 c, float size, f3 sp, uint32_t& seed, f3& pos, float& invPdf,
                                               uint32_t& face)
{
    const float hs = 0.5f * size;
    const float area = size * size;
    // the face of each axis that can face the point is known before the
    // draws: the twelve numbers are drawn in face order and each axis keeps
    // only the pair of its front face (of its -face when neither faces it)
    bool plusA[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ca = a == 0 ? c.x : (a == 1 ? c.y : c.z);
        const float spa = a == 0 ? sp.x : (a == 1 ? sp.y : sp.z);
        plusA[a] = (ca + hs) - spa < 0.0f;
    }
    // the pair of each axis is picked as integers and converted once
    uint32_t w0[3], w1[3];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const uint32_t v0 = xorshift(seed), v1 = xorshift(seed);
        const int a = k % 3;
        if (plusA[a] == (k < 3)) w0[a] = v0, w1[a] = v1;
    }
    float q0[3], q1[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        q0[a] = (float)w0[a] * 2.3283064365387e-10f;
        q1[a] = (float)w1[a] * 2.3283064365387e-10f;
    }
    const float r2 = rnd(seed);
    float sol[3];
    int fidx[3];   // front face of each axis, or -1
    f3 fp0, fp1, fp2;  // the point drawn on that face (named: an array would be indexed at run time below)
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ca = a == 0 ? c.x : (a == 1 ? c.y : c.z);
        const float spa = a == 0 ? sp.x : (a == 1 ? sp.y : sp.z);
        const float Rp = (ca + hs) - spa, Rm = (ca - hs) - spa;
        const bool plus = plusA[a], minus = Rm > 0.0f;
        const float s = plus ? 1.0f : -1.0f;
        const int i = plus ? a : a + 3;
        const f3 fpa = face_point(c, hs, size, a, s, q0[a], q1[a]);
        if (a == 0) fp0 = fpa;
        else if (a == 1) fp1 = fpa;
        else fp2 = fpa;
        const f3 R = sub(fpa, sp);
        const float Ra = plus ? Rp : Rm;
        const float d = dot(R, R);
        const float inv = 1.0f / sqrtf(d);
        const float ndl = -s * (Ra * inv);
        const float d2 = fmaxf(1e-5f, d);
        sol[a] = (plus || minus) ? fmaxf(0.0f, ndl * area / d2) : 0.0f;
        fidx[a] = (plus || minus) ? i : -1;
    }
    // face order: +x +y +z -x -y -z
    float total = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) total += fidx[a] == a ? sol[a] : 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) total += fidx[a] == a + 3 ? sol[a] : 0.0f;
    // total is finite and >= 0.  The walk's pdf starts at 0/total: 0, or
    // NaN for total = 0; either way the weight below is 0, so 0 is kept.
    int index = 5;
    float pdf = 0.0f;
    if (total == 0.0f) {
        index = 0;  // 0/0: the walk stops at face 0 with pdf NaN
    } else if (r2 == 0.0f && fidx[0] != 0) {
        // cdf[0] = 0 >= r2 = 0: the walk stops at face 0 whatever its solid
        // angle (tools.cl:549-551); a +x face with none gives pdf 0, weight 0
        // (r2 is 0 only for a pixel stream of 0, since xorshift32 maps no
        // non-zero state to 0)
        index = 0;
    } else {
        // the two faces of an axis share its solid angle: one division each
        float q[3];
#pragma unroll
        for (int a = 0; a < 3; a++) q[a] = sol[a] / total;
        float cdf = 0.0f;
        bool done = false;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int a = k % 3;
            if (!done && fidx[a] == k) {
                const float p = q[a];
                cdf = cdf + p;
                if (cdf >= r2 || k == 5) index = k, pdf = p, done = cdf >= r2;
            }
        }
    }
    float wgt = 1.0f / pdf;
    if (isinf(wgt) || isnan(wgt)) wgt = 0.0f;
    invPdf = wgt;
    face = (uint32_t)index;
    // With a non-zero weight the chosen face is its axis's front face, whose
    // point the loop drew from the same two numbers.  With weight 0 the point
    // never reaches an output (p-hat is finite, times 0; the reservoir
    // cannot take a zero weight), so that axis's point serves as well.
    const int a = index % 3;
    pos.x = a == 0 ? fp0.x : (a == 1 ? fp1.x : fp2.x);
    pos.y = a == 0 ? fp0.y : (a == 1 ? fp1.y : fp2.y);
    pos.z = a == 0 ? fp0.z : (a == 1 ? fp1.z : fp2.z);
}

__device__ __forceinline__ uint32_t pick_light(const LightView& L, uint32_t source, uint32_t& seed, float& p)
{
    // GenerateLightCandidate + getRandomIndex (cl/tools.cl:423-426, 593-598)
    const uint32_t N = L.n;
    const float u = rnd(seed);
    int32_t i = f2i(u * (float)N);
    i = i < 0 ? 0 : (i > (int32_t)N - 1 ? (int32_t)N - 1 : i);
    if (source == 1u) {
        const float fr = u * (float)N - (float)i;
        const float4 e = L.pick[i];
        const bool keep = fr < e.x;
        p = keep ? e.z : e.w;  // pmf of the light returned
        return keep ? (uint32_t)i : __float_as_uint(e.y);
    }
    p = 1.0f / (float)N;
    return (uint32_t)i;
}

// ---------------------------------------------------------------------------
// k_primary: renderAlbedo (cl/ris.cl:295-389)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void primary_px(const FrameArgs& a, int x, int y)
{
    const size_t idx = (size_t)y * a.W + x;
    const uint32_t seed = wang_hash((uint32_t)x * 171u + (uint32_t)y * 1773u + a.R0);
    const f3 D = camera_ray(a.cam, x, y);
    uint32_t side = 0;
    float dist = 0.0f;
    uint32_t voxel = trace<kTraceAll>(a.w, a.cam.E, D, 999999, 0u, dist, side);
    float total = 0.0f;
    uint32_t enc;
    if (voxel != 0u && (voxel & 0xf000u) == 0xf000u) {
        // opaque first hit: the translucency loop is skipped
        // (remainingVisibility = 0); re-encode as FromFloatRGBA does
        const f3 c = to_rgb(voxel);
        total = dist;
        enc = (((voxel >> 16) & 255u) << 16) + (f2u(1.0f * 15.0f) << 12) + (f2u(c.x * 1.0f * 15.0f) << 8) +
              (f2u(c.y * 1.0f * 15.0f) << 4) + f2u(c.z * 1.0f * 15.0f);
    } else {
        float alpha = alphaf(voxel);
        f3 c = to_rgb(voxel);
        f3 out = f3{c.x * alpha, c.y * alpha, c.z * alpha};
        int iteration = 0;
        float remaining = 1.0f - alphaf(voxel);
        uint32_t prevSide = 0, prevVoxel = 0;
        while (prevVoxel != voxel && remaining > 0.0f) {
            prevVoxel = voxel;
            prevSide = side;
            total += dist + 0.01f;
            const f3 O = f3{a.cam.E.x + total * D.x, a.cam.E.y + total * D.y, a.cam.E.z + total * D.z};
            voxel = trace<kTraceThrough>(a.w, O, D, 999999, voxel, dist, side);
            count_rays(a.rays, kRaysPrimary, 1u);
            c = to_rgb(voxel);
            alpha = alphaf(voxel);
            out.x += remaining * c.x * alpha, out.y += remaining * c.y * alpha, out.z += remaining * c.z * alpha;
            remaining = remaining * (1 - alpha);
            if (++iteration == 25) remaining = 0.0f;
            if (dist <= 0.001f) break;
        }
        if (voxel != 0u) total += dist;
        else side = prevSide;
        enc = (((voxel >> 16) & 255u) << 16) + (f2u((1.0f - remaining) * 15.0f) << 12) + (f2u(out.x * 15.0f) << 8) +
              (f2u(out.y * 15.0f) << 4) + f2u(out.z * 15.0f);
    }
    const float ds = side == 0 ? D.x : (side == 1 ? D.y : D.z);
    a.cur.hitw[idx] = enc | (side << 24) | ((ds > 0 ? 1u : 0u) << 26);
    a.cur.dist[idx] = total;
    a.cur.seed[idx] = seed;
    count_rays(a.rays, kRaysPrimary, 1u);
}

// shading point set-up shared by the reuse passes
struct Shading {
    f3 D, N, brdf, sp;
};

// VoxelNormal (cl/tools.cl:416-421) from a hit word: side bits 24-25, the
// sign of D[side] in bit 26
__device__ __forceinline__ f3 hit_normal(uint32_t hitw)
{
    const uint32_t side = (hitw >> 24) & 3u;
    const float s = (hitw >> 26) & 1u ? -1.0f : 1.0f;
    return f3{side == 0 ? s : 0.0f, side == 1 ? s : 0.0f, side == 2 ? s : 0.0f};
}

// ToFloatRGB(voxel) * INVPI (cl/ris.cl:48)
__device__ __forceinline__ f3 hit_brdf(uint32_t hitw)
{
    const f3 c = to_rgb(hitw);
    return f3{c.x * kInvPi, c.y * kInvPi, c.z * kInvPi};
}

__device__ __forceinline__ Shading shading_at(const FrameArgs& a, int x, int y, uint32_t hitw, float dist)
{
    Shading s;
    s.D = camera_ray(a.cam, x, y);
    s.N = voxel_normal((hitw >> 24) & 3u, s.D);
    const f3 c = to_rgb(hitw);
    s.brdf = f3{c.x * kInvPi, c.y * kInvPi, c.z * kInvPi};
    s.sp = f3{s.D.x * dist + a.cam.E.x, s.D.y * dist + a.cam.E.y, s.D.z * dist + a.cam.E.z};
    return s;
}

// ---------------------------------------------------------------------------
// k_initial: candidates + visibility + temporal reuse (cl/ris.cl:20-124)
// ---------------------------------------------------------------------------
// LDS of the candidate pass (256 threads per workgroup, one slot each,
// planes of 16 bytes so that neighbouring threads hit different banks):
// what stays live through the shadow ray's traversal
struct ParkLds {
    float4 a[256], b[256], c[256];
    uint2 m[256];
};

__device__ __forceinline__ ParkLds& park_lds()
{
    __shared__ ParkLds park;
    return park;
}

__device__ __forceinline__ float4 vload(const float4& p)
{
    const volatile float4& v = p;
    return make_float4(v.x, v.y, v.z, v.w);
}

template <bool kNested>
__device__ __forceinline__ void initial_px(const FrameArgs& a, int x, int y)
{
    const size_t idx = (size_t)y * a.W + x;
    const uint32_t hitw = a.cur.hitw[idx];
    const float dist = a.cur.dist[idx];
    Res res;
    res.wsum = 0.0f, res.W = 0.0f, res.pHat = 0.0f, res.invPdf = 0.0f;
    res.pos = f3{0.0f, 0.0f, 0.0f}, res.light = 0u, res.M = 0u, res.face = 7u;
    uint32_t shadow = 0u;  // kTraced | kVisible of the B sample's ray
    uint32_t traced = 0u;  // a shadow ray was traced
    if ((hitw & 0xffffffu) != 0u) {
        uint32_t seed = a.cur.seed[idx];
        const Shading s = shading_at(a, x, y, hitw, dist);
        const f3 spo = f3{s.sp.x + 0.1f * s.N.x, s.sp.y + 0.1f * s.N.y, s.sp.z + 0.1f * s.N.z};
        const uint32_t nL = a.L.n;
        const uint32_t nC = nL < a.nCand ? nL : a.nCand;
        const size_t plane = (size_t)a.W * a.H;
        for (uint32_t i = 0; i < nC; i++) {
            float pr;
            uint32_t li;
            if (a.source == 2u) {
                // learned proposal: the slot's cluster and mixture mass; the
                // light inside the cluster by power (one draw for bucket and
                // alias, as pick_light), pdf = m_c * pmf_i / P_c
                const uint2 slot = a.candSlot[i * plane + idx];
                const uint32_t c = slot.x;
                const float mc = __uint_as_float(slot.y);
                const uint32_t c0 = c < 256u ? a.clusterStart[c] : 0u;
                const uint32_t cnt = c < 256u ? a.clusterStart[c + 1] - c0 : 0u;
                const float u = rnd(seed);
                if (cnt == 0u || !(mc > 0.0f)) {
                    // a target the proposal could not place: weight 0
                    res.M += 1u;
                    continue;
                }
                int32_t k = f2i(u * (float)cnt);
                k = k < 0 ? 0 : (k > (int32_t)cnt - 1 ? (int32_t)cnt - 1 : k);
                const float fr = u * (float)cnt - (float)k;
                const uint4 e = a.clusterPick[c0 + k];
                const float2 sh = a.clusterPickShare[c0 + k];
                const bool keep = fr < __uint_as_float(e.x);
                li = keep ? e.y : e.z;
                pr = mc * (keep ? sh.x : sh.y);
            } else {
                li = pick_light(a.L, a.source, seed, pr);
            }
            const float4 la = a.L.a[li], lb = a.L.b[li];
            f3 pos;
            float invPdf;
            uint32_t face;
            if (a.point) {
                pos = f3{la.x, la.y, la.z}, invPdf = 1.0f, face = 7u;
            } else {
                point_on_voxel(f3{la.x, la.y, la.z}, la.w, spo, seed, pos, invPdf, face);
            }
            const float ph = phat(a.point, pos, face_normal(face), s.sp, s.N, s.brdf, f3{lb.x, lb.y, lb.z}, la.w) * invPdf;
            res_update(res, ph / pr, ph, li, rnd(seed), pos, invPdf, face);
        }
        res_adjust(res, res.pHat);
        // The reservoir, the shading point and the stream wait in LDS while
        // the shadow ray traverses the grid, and are read back after it: the
        // traversal then has the registers to itself instead of spilling
        // them to scratch memory (HBM traffic).  The reads are volatile, so
        // they are not replaced by the values stored.
        ParkLds& pk = park_lds();
        const uint32_t t = threadIdx.x;
        pk.a[t] = make_float4(res.wsum, res.W, res.pHat, res.invPdf);
        pk.b[t] = make_float4(res.pos.x, res.pos.y, res.pos.z, __uint_as_float(res.light));
        pk.c[t] = make_float4(s.sp.x, s.sp.y, s.sp.z, __uint_as_float(seed));
        pk.m[t] = make_uint2(res.M, res.face);
        bool occ = false;
        if (res.wsum != 0.0f) {
            // with wsum == 0 the discard would leave W = wsum = 0 unchanged
            const float4 lb = a.L.b[res.light];
            const float4 la = a.L.a[res.light];
            const f3 L = normalize(sub(res.pos, spo));
            occ = light_occluded<true, kNested>(a.w, __float_as_uint(lb.w), (uint32_t)la.w, spo, L, a.shadowSteps);
            shadow = kTraced | (occ ? 0u : kVisible);
            traced = 1u;
        }
        const float4 pa = vload(pk.a[t]), pb = vload(pk.b[t]), pc = vload(pk.c[t]);
        const volatile uint2& pm = pk.m[t];
        res.wsum = occ ? 0.0f : pa.x, res.W = occ ? 0.0f : pa.y, res.pHat = pa.z, res.invPdf = pa.w;
        res.pos = f3{pb.x, pb.y, pb.z}, res.light = __float_as_uint(pb.w);
        res.M = pm.x, res.face = pm.y;
        seed = __float_as_uint(pc.w);
        if (a.restirtemporalframe > 0u && a.temporal) {
            // The normal and brdf again from the hit word (the same bits:
            // the word carries the side and the sign of D[side], and the
            // colour), the shading point from LDS: nothing of the shading
            // stays live through the shadow ray's traversal.  The empty asm
            // keeps the compiler from reusing the values computed before it.
            uint32_t hw = hitw;
            asm volatile("" : "+v"(hw));
            const Shading s2 = Shading{s.D, hit_normal(hw), hit_brdf(hw), f3{pc.x, pc.y, pc.z}};
            const Shading& s = s2;
            // ReprojectWorldPoint (cl/tools.cl:479-488)
            const float dl = dot(s.sp, f3{a.Nl.x, a.Nl.y, a.Nl.z}) - a.Nl.w;
            const float dr = dot(s.sp, f3{a.Nr.x, a.Nr.y, a.Nr.z}) - a.Nr.w;
            const float dt = dot(s.sp, f3{a.Nt.x, a.Nt.y, a.Nt.z}) - a.Nt.w;
            const float db = dot(s.sp, f3{a.Nb.x, a.Nb.y, a.Nb.z}) - a.Nb.w;
            const int px = f2i(roundf((float)a.W * (dl / (dl + dr))));
            const int py = f2i(roundf((float)a.H * (dt / (dt + db))));
            if (px >= 0 && py >= 0 && px < (int)a.W && py < (int)a.H) {
                // the previous frame's final reservoir at the reprojected
                // pixel: its record word and distance carry that pixel's
                // hit (the previous hit buffer's side bits and distance)
                const size_t pi = (size_t)py * a.W + px;
                const uint4 p0 = a.A.r0[pi];
                const float pd = __uint_as_float(p0.y);
                const uint32_t prevM = p0.x & 0xffffffu;
                const uint32_t cap = a.maxTemporal * nC;
                const uint32_t pm = cap < prevM ? cap : prevM;
                if (same_face(rec_hit_bits(p0.x), hitw, a.axisOnlyFace) && fabsf(dist - pd) <= 0.1f * dist && pm > 0u) {
                    const uint4 p1 = a.A.r1[pi];
                    Res prev = res_from(p0, p1, make_float2(0.0f, 0.0f));  // wsum and pHat are recomputed below
                    if (prev.light < nL) {
                        const float4 pa = a.L.a[prev.light], pb = a.L.b[prev.light];
                        const float pph = phat(a.point, prev.pos, face_normal(prev.face), s.sp, s.N, s.brdf,
                                               f3{pb.x, pb.y, pb.z}, pa.w) * prev.invPdf;
                        res_reweigh(prev, pph, pm);
                        if (res_combine(res, prev, rnd(seed))) shadow = 0u;  // the traced sample was replaced
                        const float4 ra = a.L.a[res.light], rb = a.L.b[res.light];
                        const float rph = phat(a.point, res.pos, face_normal(res.face), s.sp, s.N, s.brdf,
                                               f3{rb.x, rb.y, rb.z}, ra.w) * res.invPdf;
                        res_adjust(res, rph);
                    }
                }
            }
        }
        a.cur.seed[idx] = seed;
    }
    store_res(a.B, idx, res, hitw, dist, shadow);
    count_rays(a.rays, kRaysCandidate, traced);
}

// ---------------------------------------------------------------------------
// k_spatial: spatial reuse over a disc (cl/spatialSampling.cl:2-67)
// ---------------------------------------------------------------------------

// spatial tap i of a pixel: offset on the disc (cl/spatialSampling.cl:22-27)
__device__ __forceinline__ void tap_offset(uint32_t& seed, float rad, int& dx, int& dy)
{
    const float angle = rnd(seed) * kTwoPi;
    const float radius = rad * sqrtf(rnd(seed)) + 1.0f;
    float sn, cs;
    sincos_det(angle, sn, cs);
    dx = f2i(cs * radius), dy = f2i(sn * radius);
}

// The spatial merge of one pixel (cl/spatialSampling.cl:2-67).  `res` starts
// as the pixel's B reservoir; with a hit and spatial reuse on (`merged`)
// it is merged with the accepted taps and adjusted, `seed` is the stream
// after the taps, and the MIS merge's bookkeeping is kept: M0 (own M),
// seed0 (the stream before the taps), `mask` (combined taps) and `sel`
// (tap whose sample was taken last, -1 = own).  `shadow` is the own B
// record's kTraced | kVisible while its sample stays the merged one, else 0.
struct Merge {
    Res res;
    Shading s;
    uint32_t hitw, seed, seed0, M0, mask, shadow;
    float dist;
    int sel;
    bool merged;
};

template <bool kMis>
__device__ __forceinline__ Merge spatial_merge(const FrameArgs& a, int x, int y)
{
    Merge g;
    const size_t idx = (size_t)y * a.W + x;
    const uint4 o0 = a.B.r0[idx];
    g.res = res_from(o0, a.B.r1[idx], a.B.r2[idx]);
    g.shadow = o0.x & (kTraced | kVisible);
    g.hitw = a.cur.hitw[idx];
    g.dist = __uint_as_float(o0.y);  // the hit buffer's distance, as the candidate pass recorded it
    g.mask = 0u, g.sel = -1;
    g.merged = (g.hitw & 0xffffffu) != 0u && a.spatial;
    if (!g.merged) return g;
    Res& res = g.res;
    const float dist = g.dist;
    uint32_t seed = a.cur.seed[idx];
    g.seed0 = seed, g.M0 = res.M;
    g.s = shading_at(a, x, y, g.hitw, dist);
    const Shading& s = g.s;
    const float rad = (float)(a.spatialRadius - 1u);
    for (uint32_t i = 0; i < a.spatialTaps; i++) {
        int dx, dy;
        tap_offset(seed, rad, dx, dy);
        const int nx = x + dx, ny = y + dy;
        if ((nx == x && ny == y) || nx < 0 || ny < 0 || nx >= (int)a.W || ny >= (int)a.H) continue;
        const size_t ni = (size_t)ny * a.W + nx;
        // the neighbour's record word, distance, W and invPdf (16 bytes)
        // hold all its tests; its sample is needed only once they pass
        uint4 n0, n1;
        load_record_sample(a.B, ni, n0, n1);
        if ((n0.x & 0xffffffu) == 0u) continue;
        const float nd = __uint_as_float(n0.y);
        if (!same_face(rec_hit_bits(n0.x), g.hitw, a.axisOnlyFace) || !(fabsf(dist - nd) <= 0.1f * dist))
            continue;
        Res nr = res_from(n0, n1, make_float2(0.0f, 0.0f));  // wsum and pHat: the reweigh recomputes them
        if (nr.light >= a.L.n) continue;
        const float4 la = a.L.a[nr.light], lb = a.L.b[nr.light];
        const float ph = phat(a.point, nr.pos, face_normal(nr.face), s.sp, s.N, s.brdf, f3{lb.x, lb.y, lb.z}, la.w) *
                         nr.invPdf;
        res_reweigh(nr, ph, nr.M);
        const float r0 = rnd(seed);
        if (res_combine(res, nr, r0)) {
            g.sel = (int)i;
            g.shadow = 0u;
        }
        if (kMis) g.mask |= 1u << i;
    }
    res_adjust(res, res.pHat);
    g.seed = seed;
    return g;
}

// the reference's merge: merge, then store the final reservoir (and the
// stream) for shading
__device__ __forceinline__ void spatial_px(const FrameArgs& a, int x, int y)
{
    const Merge g = spatial_merge<false>(a, x, y);
    const size_t idx = (size_t)y * a.W + x;
    if (g.merged) a.cur.seed[idx] = g.seed;
    store_res(a.A, idx, g.res, g.hitw, g.dist, g.shadow);
}

// Unbiased contribution weight of the merged sample under generalized
// balance-heuristic MIS (weights evaluated for the selected sample only):
//     W = wsum * T_sel(y) / (pHat * sum_s M_s T_s(y))
// over the own reservoir (s = own, M0, T = phat without a ray: shading
// multiplies by the own visibility) and the combined neighbours in tap
// order (T = phat at the neighbour's surface point times its visibility of
// y, one shadow ray each unless the candidate pass traced that ray).  The
// neighbours are found again by replaying the tap draws from seed0: `mask`
// marks the combined taps, each of which drew one more number for its
// merge.  The common invPdf of y cancels.
//
// The visibility rays of a 16x16 tile are traced together: the taps are
// taken kMisChunk at a time; each pixel lists the rays its combined taps of
// the chunk need in LDS, all 256 threads trace the list, and each pixel then
// adds its chunk's terms in tap order.  The rays of a pixel no longer run
// one after another in its lane while other lanes idle.  Uniform control
// flow: every thread runs every chunk's barriers.
constexpr uint32_t kMisChunk = 5;

struct MisLds {
    float4 sample[256];               // the thread's merged sample: pos.xyz, light bits
    float term[kMisChunk][256];       // M_s * T_s(y) of tap k of the chunk, before its ray
    uint32_t occluded[256];           // bit k: tap k of the chunk found the sample occluded
    uint32_t task[kMisChunk * 256];   // owner thread | tap of the chunk << 8
    uint32_t taskPx[kMisChunk * 256]; // the tap's neighbour pixel
    uint32_t count[2];                // list length, per chunk parity
};

template <bool kNested>
__device__ __forceinline__ float mis_weight_tile(const FrameArgs& a, int x, int y, const Merge& g, bool in)
{
    __shared__ MisLds L;
    const uint32_t tid = threadIdx.x;
    const Res& res = g.res;
    const bool need = in && g.merged && res.wsum > 0.0f && res.light < a.L.n;
    float4 la = make_float4(0.0f, 0.0f, 0.0f, 0.0f), lb = la;
    f3 Nl = f3{0.0f, 0.0f, 0.0f};
    float den = 0.0f, psel = 0.0f;
    if (need) {
        la = a.L.a[res.light], lb = a.L.b[res.light];
        Nl = face_normal(res.face);
        const float p0 = phat(a.point, res.pos, Nl, g.s.sp, g.s.N, g.s.brdf, f3{lb.x, lb.y, lb.z}, la.w);
        den = (float)g.M0 * p0;
        psel = g.sel < 0 ? p0 : 0.0f;
    }
    L.sample[tid] = make_float4(res.pos.x, res.pos.y, res.pos.z, __uint_as_float(need ? res.light : 0u));
    L.occluded[tid] = 0u;
    if (tid < 2u) L.count[tid] = 0u;
    __syncthreads();
    const float rad = (float)(a.spatialRadius - 1u);
    uint32_t rs = g.seed0;
    uint32_t c = 0;
    for (uint32_t i0 = 0; i0 < a.spatialTaps; i0 += kMisChunk, c++) {
        uint32_t* count = &L.count[c & 1u];
        if (tid == 0u) L.count[(c + 1u) & 1u] = 0u;  // the next chunk's list (the last reader was chunk c - 1)
        uint32_t combined = 0u, traced = 0u;
        float pselChunk = 0.0f;  // T of tap `sel` if it is in this chunk, before its ray
#pragma unroll
        for (uint32_t k = 0; k < kMisChunk; k++) {
            const uint32_t i = i0 + k;
            if (!need || i >= a.spatialTaps || !(g.mask >> i)) continue;
            int dx, dy;
            tap_offset(rs, rad, dx, dy);
            if (!((g.mask >> i) & 1u)) continue;
            (void)rnd(rs);  // the merge's draw
            combined |= 1u << k;
            const int nx = x + dx, ny = y + dy;
            const size_t ni = (size_t)ny * a.W + nx;
            const uint32_t nh = a.cur.hitw[ni];
            const uint4 n0 = a.B.r0[ni];
            const Shading t = shading_at(a, nx, ny, nh, __uint_as_float(n0.y));
            float p = phat(a.point, res.pos, Nl, t.sp, t.N, t.brdf, f3{lb.x, lb.y, lb.z}, la.w);
            if (p > 0.0f) {
                const int known = known_visibility(n0.x, a.B.r1[ni], res.light, res.pos);
                if (known == 0) {
                    p = 0.0f;
                } else if (known < 0) {
                    const uint32_t slot = atomicAdd(count, 1u);
                    L.task[slot] = tid | (k << 8);
                    L.taskPx[slot] = (uint32_t)ni;
                    traced |= 1u << k;
                }
            }
            // the term as the one-lane loop adds it, (float)M * T: an occluded
            // tap adds (float)M * 0 = +0 instead
            L.term[k][tid] = (float)(n0.x & 0xffffffu) * p;
            if ((int)i == g.sel) pselChunk = p;
        }
        __syncthreads();
        const uint32_t n = *count;
        if (tid == 0u) count_rays(a.rays, kRaysMis, n);
        for (uint32_t t = tid; t < n; t += 256u) {
            const uint32_t tk = L.task[t], ni = L.taskPx[t], owner = tk & 255u;
            const float4 smp = L.sample[owner];
            const uint32_t light = __float_as_uint(smp.w);
            const int nx = (int)(ni % a.W), ny = (int)(ni / a.W);
            const float nd = __uint_as_float(a.B.r0[ni].y);
            const Shading sh = shading_at(a, nx, ny, a.cur.hitw[ni], nd);
            const f3 spo = f3{sh.sp.x + 0.1f * sh.N.x, sh.sp.y + 0.1f * sh.N.y, sh.sp.z + 0.1f * sh.N.z};
            const f3 Ld = normalize(sub(f3{smp.x, smp.y, smp.z}, spo));
            const float4 tla = a.L.a[light], tlb = a.L.b[light];
            if (light_occluded<true, kNested>(a.w, __float_as_uint(tlb.w), (uint32_t)tla.w, spo, Ld, a.shadowSteps))
                atomicOr(&L.occluded[owner], 1u << (tk >> 8));
        }
        __syncthreads();
        const uint32_t occ = L.occluded[tid];
        L.occluded[tid] = 0u;  // tracers of the next chunk write only after its first barrier
#pragma unroll
        for (uint32_t k = 0; k < kMisChunk; k++) {
            if (!((combined >> k) & 1u)) continue;
            const bool blocked = ((traced >> k) & 1u) && ((occ >> k) & 1u);
            den += blocked ? 0.0f : L.term[k][tid];
            if ((int)(i0 + k) == g.sel) psel = blocked ? 0.0f : pselChunk;
        }
    }
    if (!need) return 0.0f;
    const float d = res.pHat * den;
    const float W = (res.wsum * psel) / d;
    return (d > 0.0f && isfinite(W)) ? W : 0.0f;
}

// ---------------------------------------------------------------------------
// shading (cl/ris.cl:155-293 with sky dome and editor off)
// ---------------------------------------------------------------------------
// kDefer: a pixel whose sample needs a shadow ray that the candidate pass
// did not trace is left alone (no frame or reservoir write) and true is
// returned; shading it again with kDefer = false gives the one-pass result,
// because nothing it reads has changed.
// LDS of the shading pass (one slot per thread): the partial colour, the
// distance, the shading point and the pixel index wait here while the
// shadow ray traverses the grid, as the candidate pass parks its reservoir
struct ShadeParkLds {
    float4 c[256], p[256];
};

__device__ __forceinline__ ShadeParkLds& shade_park_lds()
{
    __shared__ ShadeParkLds park;
    return park;
}

// kPark: the call that traces (kDefer false) parks the live values in LDS
// across the traversal (the MIS merge's kernel has no LDS to spare for it)
template <bool kDefer = false, bool kPark = false, bool kNested = false>
__device__ __forceinline__ bool shade_px(const FrameArgs& a, int x, int y)
{
    constexpr bool park = kPark && !kDefer;
    const size_t idx = (size_t)y * a.W + x;
    const uint32_t hitw = a.cur.hitw[idx];
    const uint32_t voxel = hitw & 0xffffffu;
    const float dist = a.cur.dist[idx];
    f3 color = f3{0.0f, 0.0f, 0.0f};
    size_t oidx = idx;
    float odist = dist;
    uint32_t ovoxel = voxel;
    if (voxel != 0u) {
        const Shading s = shading_at(a, x, y, hitw, dist);
        if (is_emitter(voxel)) {
            const f3 c = to_rgb(voxel);
            const float es = a.emitLut[(voxel >> 16) & 255u];
            f3 e = f3{c.x * es, c.y * es, c.z * es};
            if (a.gfxexp) e = f3{e.x * kInvPi, e.y * kInvPi, e.z * kInvPi};
            color = f3{color.x + e.x, color.y + e.y, color.z + e.z};
        }
        ShadeParkLds& pk = shade_park_lds();
        const uint32_t t = threadIdx.x;
        if (park) {
            pk.c[t] = make_float4(color.x, color.y, color.z, dist);
            pk.p[t] = make_float4(s.sp.x, s.sp.y, s.sp.z, __uint_as_float((uint32_t)idx));
        }
        f3 add = f3{0.0f, 0.0f, 0.0f};
        if (a.gfxexp || !is_emitter(voxel)) {
            const uint4 r0 = a.A.r0[idx];
            const float W = __uint_as_float(r0.z);
            if (W > 0.0f && (r0.x & 0xffffffu) > 0u) {
                const uint4 r1 = a.A.r1[idx];
                const uint32_t light = r1.w;
                if (light < a.L.n) {
                    const f3 pos = f3{__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z)};
                    const f3 spo = f3{s.sp.x + 0.1f * s.N.x, s.sp.y + 0.1f * s.N.y, s.sp.z + 0.1f * s.N.z};
                    const float4 la = a.L.a[light], lb = a.L.b[light];
                    const f3 L = normalize(sub(pos, spo));
                    // the final record's shadow bits: the candidate pass traced
                    // this very ray (its sample survived both merges)
                    const int known = (r0.x & kTraced) ? ((r0.x & kVisible) ? 1 : 0) : -1;
                    if (kDefer && known < 0) return true;
                    const bool visible =
                        known >= 0 ? known == 1
                                   : !light_occluded<false, kNested>(a.w, __float_as_uint(lb.w), (uint32_t)la.w, spo, L, a.shadowSteps);
                    // what the traversal did not keep in registers: from LDS
                    // when parked, else the values themselves
                    const float4 pp = park ? vload(pk.p[t]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    const size_t pidx = park ? (size_t)__float_as_uint(pp.w) : idx;
                    if (visible) {
                        // the sample, its light and the surface again from
                        // memory and the hit word (cache hits, the same bits),
                        // so that little stays live through the traversal
                        uint32_t h = park ? 0u : hitw;
                        size_t i2 = pidx;
                        asm volatile("" : "+v"(h), "+v"(i2));
                        if (park) h = a.cur.hitw[i2];
                        const f3 sp = park ? f3{pp.x, pp.y, pp.z} : s.sp;
                        const uint4 q0 = a.A.r0[i2], q1 = a.A.r1[i2];
                        const float4 qa = a.L.a[q1.w], qb = a.L.b[q1.w];
                        const float Wq = __uint_as_float(q0.z), invPdf = __uint_as_float(q0.w);
                        const f3 posq = f3{__uint_as_float(q1.x), __uint_as_float(q1.y), __uint_as_float(q1.z)};
                        const f3 inc = phat_full(a.point, posq, face_normal((q0.x >> 24) & 7u), sp, hit_normal(h),
                                                 hit_brdf(h), f3{qb.x, qb.y, qb.z}, qa.w);
                        add = f3{inc.x * Wq * invPdf, inc.y * Wq * invPdf, inc.z * Wq * invPdf};
                    } else {
                        // discard: W = wsum = 0 (ris.cl:276-281)
                        reinterpret_cast<uint32_t*>(&a.A.r0[pidx])[2] = 0u;
                        reinterpret_cast<float*>(&a.A.r2[pidx])[0] = 0.0f;
                    }
                }
            }
        }
        if (park) {
            const float4 pc = vload(pk.c[t]);
            const float4 pp = vload(pk.p[t]);
            color = f3{pc.x, pc.y, pc.z}, odist = pc.w, oidx = (size_t)__float_as_uint(pp.w);
            ovoxel = a.cur.hitw[oidx] & 0xffffffu;
        }
        // color + 0 = color exactly (color is never -0)
        color = f3{color.x + add.x, color.y + add.y, color.z + add.z};
        if (a.doubleRadiance) {
            const float al = alphaf(ovoxel);
            color.x += al * color.x + (1 - al) * 0.0f;
            color.y += al * color.y + (1 - al) * 0.0f;
            color.z += al * color.z + (1 - al) * 0.0f;
        }
    }
    a.frame[oidx] = make_float4(color.x, color.y, color.z, odist);
    return false;
}

// ---------------------------------------------------------------------------
// stage kernels: one view (FrameArgs by value, 16x16 tile per block in the
// tile order of the stage), or many views in one persistent launch
// ---------------------------------------------------------------------------
// The uber grid's 4 KiB of bits, staged in LDS by every workgroup of the
// frame kernels: the DDA's outermost level then reads LDS.
__device__ __forceinline__ uint32_t* uber_lds()
{
    __shared__ uint32_t s[1024];
    return s;
}

__device__ __forceinline__ void stage_uber(const uint32_t* __restrict__ src)
{
    uint32_t* dst = uber_lds();
    for (uint32_t i = threadIdx.x; i < 1024u; i += 256u) dst[i] = src[i];
    __syncthreads();
}

// kTraces: the stage traces rays (and so reads the staged uber grid)
template <bool kTraces = true, class F>
__device__ __forceinline__ void one_view(const FrameArgs& a, F f)
{
    if (kTraces) stage_uber(a.w.uber);
    int x, y;
    tile_pixel(a.tileOrder, x, y);
    FrameArgs b = a;
    if (kTraces) b.w.uber = uber_lds();
    if (x < (int)a.W && y < (int)a.H) f(b, x, y);
}

// A chunk's views' 16x16 tiles are numbered view after view (start[v] =
// first tile of view v, start[n] = total, row-major inside a view); block b
// renders tiles b, b + gridDim.x, ...  The chunk is a kernel argument, so
// the view of a tile is found by a uniform scan over start and its
// FrameArgs (and the pointers in them) are read like any kernel argument.
// tile t of the chunk; `staged` is the view whose uber grid LDS holds
// (~0u: none yet).  Uniform control flow: every thread of the block runs
// the staging barriers.
// kAll: f runs on every thread of the tile (it checks the bounds itself and
// may use barriers); otherwise only on the pixels inside the view.
template <bool kAll = false, class F>
__device__ __forceinline__ void view_tile(const ViewChunk& c, uint32_t t, uint32_t& staged, F f)
{
    uint32_t v = 0;
    while (v + 1 < c.n && c.start[v + 1] <= t) v++;
    if (v != staged) {
        if (staged != ~0u) __syncthreads();  // the previous view's tile is done with LDS
        stage_uber(c.v[v].w.uber);
        staged = v;
    }
    FrameArgs a = c.v[v];
    a.w.uber = uber_lds();
    const uint32_t lt = t - c.start[v];
    const uint32_t tilesX = (a.W + 15u) >> 4;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int x = (int)(lt % tilesX) * 16 + (wv & 1) * 8 + (lane & 7);
    const int y = (int)(lt / tilesX) * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (kAll || (x < (int)a.W && y < (int)a.H)) f(a, x, y);
}

// kPersistent: gridDim.x < total, each block loops; otherwise one tile per block
template <bool kPersistent, bool kAll = false, class F>
__device__ __forceinline__ void many_views(const ViewChunk& c, F f)
{
    const uint32_t total = c.start[c.n];
    uint32_t staged = ~0u;
    if (!kPersistent) {
        if (blockIdx.x < total) view_tile<kAll>(c, blockIdx.x, staged, f);
        return;
    }
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) view_tile<kAll>(c, t, staged, f);
}

// Occupancy targets (waves per SIMD) of the ray-tracing stages: their DDA
// loops are bound by dependent-load latency, and more waves in flight hide
// more of it than the few registers the compiler then spills cost
// (measured on c2 at 1080p: candidates 1.83 -> 1.68 ms at 7 waves; the
// defaults of the other stages are explained where they are set).
#define VR_WAVES(n) __attribute__((amdgpu_waves_per_eu(n, 8)))

// One-view stage kernels, instantiated for occupancy targets of 5-8 waves
// per SIMD (FrameArgs::waves, picked per stage by vr_set_stage_waves or the
// autotuner; the defaults are the measured best above).
template <int W>
__global__ __launch_bounds__(256) VR_WAVES(W) void k_primary(FrameArgs a)
{
    one_view(a, [](const FrameArgs& v, int x, int y) { primary_px(v, x, y); });
}
template <int W, bool kNested>
__global__ __launch_bounds__(256) VR_WAVES(W) void k_initial(FrameArgs a)
{
    one_view(a, [](const FrameArgs& v, int x, int y) { initial_px<kNested>(v, x, y); });
}
// Spatial reuse, then shading of the tile's 256 pixels.  Only the pixels
// whose final sample needs a shadow ray the candidate pass did not trace
// (about one in seven on c2) trace one, and a wave would wait for the
// longest of its few: the tile's ray pixels are gathered in LDS (in pixel
// order) and shaded by the first threads, so they fill as few waves as
// possible and the other waves finish at once.  Uniform control flow: every
// thread reaches every barrier (the MIS weight's included).
// kPark: shading parks its live values in LDS across the traversal (not
// where a kernel also holds the MIS merge's LDS)
// Shading of a 16x16 tile: the pixels whose final sample needs a shadow ray
// that the candidate pass did not trace are gathered in LDS and shaded by
// the first threads, so their rays fill as few waves as possible.  Every
// thread of the workgroup calls it (barriers).
template <bool kPark, bool kNested>
__device__ __forceinline__ void shade_tile(const FrameArgs& a, int x, int y, bool in)
{
    __shared__ uint32_t rayPx[256];  // threads of the tile whose pixel needs a ray
    __shared__ uint32_t waveRays[4];
    const bool defer = in && shade_px<true, kPark, kNested>(a, x, y);
    const uint64_t m = __ballot(defer);
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    if (lane == 0u) waveRays[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t first = 0u, n = 0u;
    for (uint32_t k = 0; k < 4u; k++) {
        first += k < wv ? waveRays[k] : 0u;
        n += waveRays[k];
    }
    if (defer) rayPx[first + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = threadIdx.x;
    __syncthreads();
    if (threadIdx.x == 0u) count_rays(a.rays, kRaysShading, n);
    if (threadIdx.x < n) {
        // the queued thread's pixel: same tile, its own wave quadrant and lane
        const uint32_t q = rayPx[threadIdx.x], qw = q >> 6, ql = q & 63u;
        const int x0 = x - (int)((wv & 1u) * 8u + (lane & 7u)), y0 = y - (int)((wv >> 1) * 8u + (lane >> 3));
        shade_px<false, kPark, kNested>(a, x0 + (int)((qw & 1u) * 8u + (ql & 7u)), y0 + (int)((qw >> 1) * 8u + (ql >> 3)));
    }
}

template <bool kMis, bool kNested, bool kPark = !kMis>
__device__ __forceinline__ void resolve_tile(const FrameArgs& a, int x, int y)
{
    const bool in = x < (int)a.W && y < (int)a.H;
    if (kMis) {
        Merge g{};
        if (in) g = spatial_merge<true>(a, x, y);
        // The merged reservoir is stored before its visibility rays are
        // traced and its W written after them, so the fields the weight does
        // not need are not held in registers across the traversal.
        const size_t idx = (size_t)y * a.W + x;
        if (in) {
            if (g.merged) a.cur.seed[idx] = g.seed;
            store_res(a.A, idx, g.res, g.hitw, g.dist, g.shadow);
        }
        const float W = mis_weight_tile<kNested>(a, x, y, g, in);
        if (in && g.merged) a.A.r0[idx].z = __float_as_uint(W);
    } else if (in) {
        spatial_px(a, x, y);
    }
    shade_tile<kPark, kNested>(a, x, y, in);
}

template <bool kMis, int W, bool kNested>
__global__ __launch_bounds__(256) VR_WAVES(W) void k_resolve(FrameArgs a)
{
    stage_uber(a.w.uber);
    int x, y;
    tile_pixel(a.tileOrder, x, y);
    FrameArgs b = a;
    b.w.uber = uber_lds();
    resolve_tile<kMis, kNested>(b, x, y);
}

// ---------------------------------------------------------------------------
// The MIS merge in three launches (one view).  The weights' visibility rays
// are listed in device memory and traced one per thread by compact waves,
// instead of inside the merge kernel, whose registers and LDS allow 5 waves
// per SIMD.  Every operation is mis_weight_tile's, in its order.
// ---------------------------------------------------------------------------

// Combined tap i of a pixel's MIS weight for its merged sample (res, with
// the sample's face normal Nl and light records la, lb): the neighbour's
// target T (0 when the candidate pass found the sample occluded from there),
// whether a ray must decide its visibility, its pixel and its M.  `rs` is
// the replayed stream at tap i's offset draws.
struct MisTap {
    float p;
    bool ray;
    uint32_t ni, M;
};

__device__ __forceinline__ MisTap mis_tap(const FrameArgs& a, int x, int y, const Res& res, f3 Nl, float4 la,
                                          float4 lb, uint32_t& rs, float rad)
{
    int dx, dy;
    tap_offset(rs, rad, dx, dy);
    (void)rnd(rs);  // the merge's draw
    const int nx = x + dx, ny = y + dy;
    const size_t ni = (size_t)ny * a.W + nx;
    const uint32_t nh = a.cur.hitw[ni];
    const uint4 n0 = a.B.r0[ni];
    const Shading t = shading_at(a, nx, ny, nh, __uint_as_float(n0.y));
    MisTap o;
    o.p = phat(a.point, res.pos, Nl, t.sp, t.N, t.brdf, f3{lb.x, lb.y, lb.z}, la.w);
    o.ray = false;
    if (o.p > 0.0f) {
        const int known = known_visibility(n0.x, a.B.r1[ni], res.light, res.pos);
        if (known == 0) o.p = 0.0f;
        o.ray = known < 0;
    }
    o.ni = (uint32_t)ni;
    o.M = n0.x & 0xffffffu;
    return o;
}

// skip tap i's offset draws (a tap that was not combined)
__device__ __forceinline__ void mis_skip_tap(uint32_t& rs, float rad)
{
    int dx, dy;
    tap_offset(rs, rad, dx, dy);
}

template <int W>
__global__ __launch_bounds__(256) VR_WAVES(W) void k_mis_list(FrameArgs a, MisBufs m)
{
    int x, y;
    tile_pixel(a.tileOrder, x, y);
    const bool in = x < (int)a.W && y < (int)a.H;
    const size_t idx = in ? (size_t)y * a.W + x : 0u;
    uint32_t rays = 0u;  // taps whose visibility ray is listed
    Merge g{};
    const float rad = (float)(a.spatialRadius - 1u);
    float4 la = make_float4(0.0f, 0.0f, 0.0f, 0.0f), lb = la;
    f3 Nl = f3{0.0f, 0.0f, 0.0f};
    bool need = false;
    if (in) {
        g = spatial_merge<true>(a, x, y);
        if (g.merged) a.cur.seed[idx] = g.seed;
        store_res(a.A, idx, g.res, g.hitw, g.dist, g.shadow);  // W: k_mis_finish
        m.occ[idx] = 0u;
        need = g.merged && g.res.wsum > 0.0f && g.res.light < a.L.n;
        if (need) {
            la = a.L.a[g.res.light], lb = a.L.b[g.res.light];
            Nl = face_normal(g.res.face);
            const size_t npx = (size_t)a.W * a.H;
            uint32_t rs = g.seed0;
            for (uint32_t i = 0; i < a.spatialTaps && (g.mask >> i); i++) {
                if (!((g.mask >> i) & 1u)) {
                    mis_skip_tap(rs, rad);
                    continue;
                }
                const MisTap t = mis_tap(a, x, y, g.res, Nl, la, lb, rs, rad);
                // the term as the one-kernel merge adds it, (float)M * T
                m.terms[i * npx + idx] = make_float2((float)t.M * t.p, t.p);
                if (t.ray) rays |= 1u << i;
            }
        }
        m.info[idx] = make_uint4(rays, g.mask, (uint32_t)(g.sel + 1), g.M0);
    }
    // one atomic per wave: the lanes' rays go to consecutive slots
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n = (uint32_t)__popc(rays);
    uint32_t incl = n;
    for (uint32_t o = 1u; o < 64u; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    uint32_t base = 0u;
    if (lane == 63u && total != 0u) base = atomicAdd(m.count, total);
    base = __shfl(base, 63, 64);
    if (rays != 0u) {
        // replay the offsets of the listed taps
        uint32_t slot = base + incl - n;
        uint32_t rs = g.seed0;
        for (uint32_t i = 0; i < a.spatialTaps && (rays >> i); i++) {
            int dx, dy;
            tap_offset(rs, rad, dx, dy);
            if (!((g.mask >> i) & 1u)) continue;
            (void)rnd(rs);
            if ((rays >> i) & 1u)
                m.list[slot++] = make_uint2(((uint32_t)idx << 5) | i, (uint32_t)((size_t)(y + dy) * a.W + (x + dx)));
        }
    }
}

template <bool kNested>
__global__ __launch_bounds__(256) VR_WAVES(8) void k_mis_trace(FrameArgs a, MisBufs m)
{
    // one thread per listed ray: the grid covers the list's capacity and the
    // workgroups past the count leave at once (workgroup 0 counts the rays)
    const uint32_t n = *m.count;
    if (blockIdx.x * 256u >= n && blockIdx.x != 0u) return;
    stage_uber(a.w.uber);
    FrameArgs b = a;
    b.w.uber = uber_lds();
    if (blockIdx.x == 0u && threadIdx.x == 0u) count_rays(b.rays, kRaysMis, n);
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < n; t += gridDim.x * 256u) {
        const uint2 task = m.list[t];
        const uint32_t owner = task.x >> 5, k = task.x & 31u, ni = task.y;
        const uint4 smp = b.A.r1[owner];  // the merged sample, stored by k_mis_list
        const int nx = (int)(ni % b.W), ny = (int)(ni / b.W);
        const float nd = __uint_as_float(b.B.r0[ni].y);
        const Shading sh = shading_at(b, nx, ny, b.cur.hitw[ni], nd);
        const f3 spo = f3{sh.sp.x + 0.1f * sh.N.x, sh.sp.y + 0.1f * sh.N.y, sh.sp.z + 0.1f * sh.N.z};
        const f3 Ld = normalize(sub(f3{__uint_as_float(smp.x), __uint_as_float(smp.y), __uint_as_float(smp.z)}, spo));
        const float4 tla = b.L.a[smp.w], tlb = b.L.b[smp.w];
        if (light_occluded<true, kNested>(b.w, __float_as_uint(tlb.w), (uint32_t)tla.w, spo, Ld, b.shadowSteps))
            atomicOr(&m.occ[owner], 1u << k);
    }
}

template <int W, bool kNested>
__global__ __launch_bounds__(256) VR_WAVES(W) void k_mis_finish(FrameArgs a, MisBufs m)
{
    stage_uber(a.w.uber);
    int x, y;
    tile_pixel(a.tileOrder, x, y);
    FrameArgs b = a;
    b.w.uber = uber_lds();
    if (blockIdx.x == 0u && blockIdx.y == 0u && threadIdx.x == 0u) *m.count = 0u;  // the tracer has read it
    const bool in = x < (int)b.W && y < (int)b.H;
    if (in) {
        const size_t idx = (size_t)y * b.W + x;
        const uint32_t hitw = b.cur.hitw[idx];
        if ((hitw & 0xffffffu) != 0u && b.spatial) {
            // mis_weight_tile's weight of the merged reservoir
            const Res res = load_res(b.A, idx);
            const uint4 info = m.info[idx];
            const uint32_t rays = info.x, mask = info.y, occ = m.occ[idx];
            const int sel = (int)info.z - 1;
            float Wm = 0.0f;
            if (res.wsum > 0.0f && res.light < b.L.n) {
                const float dist = __uint_as_float(b.A.r0[idx].y);
                const Shading s = shading_at(b, x, y, hitw, dist);
                const float4 la = b.L.a[res.light], lb = b.L.b[res.light];
                const f3 Nl = face_normal(res.face);
                const float p0 = phat(b.point, res.pos, Nl, s.sp, s.N, s.brdf, f3{lb.x, lb.y, lb.z}, la.w);
                float den = (float)info.w * p0;
                float psel = sel < 0 ? p0 : 0.0f;
                // the combined taps' terms, stored by k_mis_list, in tap order
                const size_t npx = (size_t)b.W * b.H;
                for (uint32_t i = 0; i < b.spatialTaps && (mask >> i); i++) {
                    if (!((mask >> i) & 1u)) continue;
                    const float2 t = m.terms[i * npx + idx];
                    const bool blocked = ((rays >> i) & 1u) && ((occ >> i) & 1u);
                    den += blocked ? 0.0f : t.x;
                    if ((int)i == sel) psel = blocked ? 0.0f : t.y;
                }
                const float d = res.pHat * den;
                const float Wq = (res.wsum * psel) / d;
                Wm = (d > 0.0f && isfinite(Wq)) ? Wq : 0.0f;
            }
            b.A.r0[idx].z = __float_as_uint(Wm);
        }
    }
    shade_tile<false, kNested>(b, x, y, in);
}

template <bool kPersistent>
__global__ __launch_bounds__(256) VR_WAVES(8) void k_primary_views(ViewChunk c)
{
    many_views<kPersistent>(c, [](const FrameArgs& v, int x, int y) { primary_px(v, x, y); });
}
template <bool kPersistent>
__global__ __launch_bounds__(256) VR_WAVES(7) void k_initial_views(ViewChunk c)
{
    many_views<kPersistent>(c, [](const FrameArgs& v, int x, int y) { initial_px<false>(v, x, y); });
}
// kAnyMis: some view merges with MIS weights (the kernel then carries the
// MIS variant's registers).  Batch frames do not overlap one another, so the
// reference merge's kernel keeps the 7 waves it runs best at alone.
template <bool kPersistent, bool kAnyMis>
__global__ __launch_bounds__(256) VR_WAVES(kAnyMis ? kDefaultWavesMis : 7) void k_resolve_views(ViewChunk c)
{
    many_views<kPersistent, true>(c, [](const FrameArgs& v, int x, int y) {
        if (kAnyMis && v.misSpatial) resolve_tile<true, false>(v, x, y);
        else resolve_tile<false, false, !kAnyMis>(v, x, y);
    });
}

// ---------------------------------------------------------------------------
// support kernels
// ---------------------------------------------------------------------------

// commit (cl/kernels.cl:448-477): one 128-thread group per staged brick,
// 16 bytes per thread
__global__ __launch_bounds__(128) void k_commit(const uint32_t* __restrict__ ids, const uint4* __restrict__ src,
                                                uint4* __restrict__ pool, uint32_t n, uint32_t cap)
{
    const uint32_t b = blockIdx.x;
    if (b >= n) return;
    const uint32_t id = ids[b];
    if (id >= cap) return;
    pool[(size_t)id * 128u + threadIdx.x] = src[(size_t)b * 128u + threadIdx.x];
}

// updateUberGrid (cl/kernels.cl:498-513), stored as bits: cell
// i = x + z*32 + y*1024 is bit i & 31 of word i >> 5.  A wave covers 64
// consecutive cells; its ballot holds their two words.
__global__ __launch_bounds__(256) void k_uber(const uint32_t* __restrict__ grid, uint32_t* __restrict__ uber)
{
    const uint32_t task = blockIdx.x * 256u + threadIdx.x;  // 32768 tasks: the grid is exact
    const uint32_t x = task & 31u, z = (task >> 5) & 31u, y = task >> 10;
    uint32_t any = 0u;
    for (uint32_t b = 0; b < 4; b++)
        for (uint32_t c = 0; c < 4; c++) {
            const uint4 row = *reinterpret_cast<const uint4*>(grid + (x * 4) + (z * 4 + c) * 128u + (y * 4 + b) * 16384u);
            any |= row.x | row.y | row.z | row.w;
        }
    const uint64_t bits = __ballot(any != 0u);
    const uint32_t lane = threadIdx.x & 63u;
    if (lane == 0u) uber[task >> 5] = (uint32_t)bits;
    if (lane == 32u) uber[task >> 5] = (uint32_t)(bits >> 32);
}

// traceBatch (cl/kernels.cl:395-417) returning voxel, dist, side
__global__ __launch_bounds__(256) void k_trace_batch(WorldView w, const float* __restrict__ O, const float* __restrict__ D,
                                                     uint32_t n, uint32_t kind, int32_t steps, uint32_t* voxel,
                                                     float* dist, uint32_t* side)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const f3 A = f3{O[3 * i], O[3 * i + 1], O[3 * i + 2]};
    const f3 B = f3{D[3 * i], D[3 * i + 1], D[3 * i + 2]};
    float d = 0.0f;
    uint32_t s = 0;
    const uint32_t v = kind == 1u ? trace_sched(w, A, B, steps, d, s) : trace<kTraceAll>(w, A, B, steps, 0u, d, s);
    voxel[i] = v;
    dist[i] = v ? d : 0.0f;
    side[i] = v ? s : 0u;
}

// ground truth: render_di_gfxexp (cl/kernels.cl:177-241) with
// UniformSampleHemisphere (cl/tools.cl:31-52), binary64 accumulation
__global__ __launch_bounds__(256) void k_ground_truth(FrameArgs a, uint32_t spp, uint32_t frame0, double* accum)
{
    int x, y;
    tile_pixel(a.tileOrder, x, y);
    if (x >= (int)a.W || y >= (int)a.H) return;
    const size_t idx = (size_t)y * a.W + x;
    const f3 D = camera_ray(a.cam, x, y);
    float dist = 0.0f;
    uint32_t side = 0;
    const uint32_t voxel = trace<kTraceAll>(a.w, a.cam.E, D, 999999, 0u, dist, side);
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (voxel != 0u) {
        const f3 base = to_rgb(voxel);
        const float es = a.emitLut[(voxel >> 16) & 255u];
        const f3 contrib = f3{base.x * es * kInvPi, base.y * es * kInvPi, base.z * es * kInvPi};
        const f3 brdf = f3{kInvPi * base.x, kInvPi * base.y, kInvPi * base.z};
        const f3 I = f3{a.cam.E.x + D.x * dist, a.cam.E.y + D.y * dist, a.cam.E.z + D.z * dist};
        const f3 N = voxel_normal(side, D);
        f3 Nt;
        if (fabsf(N.x) > fabsf(N.y)) {
            const float l = sqrtf(N.x * N.x + N.z * N.z);
            Nt = f3{N.z / l, 0.0f / l, -N.x / l};
        } else {
            const float l = sqrtf(N.y * N.y + N.z * N.z);
            Nt = f3{0.0f / l, -N.z / l, N.y / l};
        }
        const f3 Nb = f3{N.y * Nt.z - N.z * Nt.y, N.z * Nt.x - N.x * Nt.z, N.x * Nt.y - N.y * Nt.x};
        const f3 O2 = f3{I.x + 0.1f * N.x, I.y + 0.1f * N.y, I.z + 0.1f * N.z};
        uint32_t r0state = a.R0;
        for (uint32_t f = 0; f < spp; f++) {
            const uint32_t R0 = xorshift(r0state);
            uint32_t seed = wang_hash((uint32_t)x * 171u + (uint32_t)y * 1773u + R0);
            const uint32_t frame = (frame0 + f) & 255u;
            const float inv16 = 1.0f / 16.0f;
            const float r0 = (float)(int)(frame & 15u) * inv16 + rnd(seed) * inv16;
            const float r1 = (float)(int)(frame >> 4) * inv16 + rnd(seed) * inv16;
            const float sinTheta = sqrtf(1 - r0 * r0);
            float sn, cs;
            sincos_det(kTwoPi * r1, sn, cs);
            const float xs = sinTheta * cs, zs = sinTheta * sn;
            const f3 R = f3{xs * Nb.x + r0 * N.x + zs * Nt.x, xs * Nb.y + r0 * N.y + zs * Nt.y,
                            xs * Nb.z + r0 * N.z + zs * Nt.z};
            const float NdotL = fmaxf(0.0f, dot(N, R));
            float d2 = 0.0f;
            uint32_t s2 = 0;
            const uint32_t v2 = trace<kTraceAll>(a.w, O2, R, 128, 0u, d2, s2);
            const f3 sl0 = to_rgb(v2);
            const float es2 = a.emitLut[(v2 >> 16) & 255u];
            const f3 sl = f3{sl0.x * es2 * kInvPi, sl0.y * es2 * kInvPi, sl0.z * es2 * kInvPi};
            acc0 += (double)((sl.x * kTwoPi * NdotL) * brdf.x + contrib.x);
            acc1 += (double)((sl.y * kTwoPi * NdotL) * brdf.y + contrib.y);
            acc2 += (double)((sl.z * kTwoPi * NdotL) * brdf.z + contrib.z);
        }
    }
    accum[idx * 3 + 0] += acc0;
    accum[idx * 3 + 1] += acc1;
    accum[idx * 3 + 2] += acc2;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 frame_grid(uint32_t W, uint32_t H) { return dim3((W + 15) / 16, (H + 15) / 16); }

template <int W>
static const void* stage_kernel_w(int stage, bool mis, bool nested)
{
    switch (stage) {
    case 0: return reinterpret_cast<const void*>(&k_primary<W>);
    case 1: return nested ? reinterpret_cast<const void*>(&k_initial<W, true>)
                          : reinterpret_cast<const void*>(&k_initial<W, false>);
    case 2:
        if (mis)
            return nested ? reinterpret_cast<const void*>(&k_resolve<true, W, true>)
                          : reinterpret_cast<const void*>(&k_resolve<true, W, false>);
        return nested ? reinterpret_cast<const void*>(&k_resolve<false, W, true>)
                      : reinterpret_cast<const void*>(&k_resolve<false, W, false>);
    default: return nullptr;
    }
}

// waves 0 = the stage's default target; dda: kDda* of the stage's shadow
// rays (the primary pass traces its coherent rays with the nested DDA
// either way)
static const void* stage_kernel(int stage, bool mis, uint32_t waves, uint32_t dda)
{
    if (waves == 0) waves = stage == 2 && mis ? kDefaultWavesMis : kDefaultWaves[stage < 0 || stage > 2 ? 0 : stage];
    if (dda > kDdaNested) return nullptr;
    const bool nested = dda == kDdaNested;
    switch (waves) {
    case 5: return stage_kernel_w<5>(stage, mis, nested);
    case 6: return stage_kernel_w<6>(stage, mis, nested);
    case 7: return stage_kernel_w<7>(stage, mis, nested);
    case 8: return stage_kernel_w<8>(stage, mis, nested);
    default: return nullptr;
    }
}

hipError_t launch_stage(int stage, const FrameArgs& a, hipStream_t s)
{
    const void* k = stage_kernel(stage, a.misSpatial != 0u, a.waves, a.dda);
    if (!k) return hipErrorInvalidValue;
    void* args[] = {const_cast<FrameArgs*>(&a)};
    return hipLaunchKernel(k, frame_grid(a.W, a.H), dim3(256), args, 0, s);
}

template <int W>
static const void* mis_kernel_w(int which, bool nested)
{
    if (which == 0) return reinterpret_cast<const void*>(&k_mis_list<W>);
    return nested ? reinterpret_cast<const void*>(&k_mis_finish<W, true>)
                  : reinterpret_cast<const void*>(&k_mis_finish<W, false>);
}

// which: 0 the merge and ray list, 1 the weights and shading
static const void* mis_kernel(int which, uint32_t waves, bool nested)
{
    switch (waves == 0 ? kDefaultWavesMisSplit : waves) {
    case 5: return mis_kernel_w<5>(which, nested);
    case 6: return mis_kernel_w<6>(which, nested);
    case 7: return mis_kernel_w<7>(which, nested);
    case 8: return mis_kernel_w<8>(which, nested);
    default: return nullptr;
    }
}

hipError_t launch_mis_split(const FrameArgs& a, const MisBufs& m, uint32_t traceBlocks, hipStream_t s)
{
    if (a.dda > kDdaNested || traceBlocks == 0u) return hipErrorInvalidValue;
    const bool nested = a.dda == kDdaNested;
    const void* k0 = mis_kernel(0, a.waves, nested);
    const void* k2 = mis_kernel(1, a.waves, nested);
    if (!k0 || !k2) return hipErrorInvalidValue;
    const void* k1 = nested ? reinterpret_cast<const void*>(&k_mis_trace<true>)
                            : reinterpret_cast<const void*>(&k_mis_trace<false>);
    void* args[] = {const_cast<FrameArgs*>(&a), const_cast<MisBufs*>(&m)};
    hipError_t e = hipLaunchKernel(k0, frame_grid(a.W, a.H), dim3(256), args, 0, s);
    if (e == hipSuccess) e = hipLaunchKernel(k1, dim3(traceBlocks), dim3(256), args, 0, s);
    if (e == hipSuccess) e = hipLaunchKernel(k2, frame_grid(a.W, a.H), dim3(256), args, 0, s);
    return e;
}

template <bool kP>
static const void* stage_views_kernel_t(int stage, bool anyMis)
{
    switch (stage) {
    case 0: return reinterpret_cast<const void*>(&k_primary_views<kP>);
    case 1: return reinterpret_cast<const void*>(&k_initial_views<kP>);
    case 2:
        return anyMis ? reinterpret_cast<const void*>(&k_resolve_views<kP, true>)
                      : reinterpret_cast<const void*>(&k_resolve_views<kP, false>);
    default: return nullptr;
    }
}

static const void* stage_views_kernel(int stage, bool anyMis, bool persistent)
{
    return persistent ? stage_views_kernel_t<true>(stage, anyMis) : stage_views_kernel_t<false>(stage, anyMis);
}

static uint32_t chunk_grid(const ViewChunk& c, uint32_t blocks)
{
    const uint32_t total = c.start[c.n];
    return blocks && blocks < total ? blocks : total;
}

hipError_t launch_stage_views(int stage, const ViewChunk& c, uint32_t blocks, bool anyMis, hipStream_t s)
{
    if (c.n == 0 || c.n > (uint32_t)kChunkViews) return hipErrorInvalidValue;
    const uint32_t g = chunk_grid(c, blocks);
    const void* k = stage_views_kernel(stage, anyMis, g < c.start[c.n]);
    if (!k) return hipErrorInvalidValue;
    if (g == 0) return hipSuccess;
    void* args[] = {const_cast<ViewChunk*>(&c)};
    return hipLaunchKernel(k, dim3(g), dim3(256), args, 0, s);
}

hipError_t stage_views_node(int stage, const ViewChunk* c, uint32_t blocks, bool anyMis, void** args,
                            hipKernelNodeParams* p)
{
    if (c->n == 0 || c->n > (uint32_t)kChunkViews) return hipErrorInvalidValue;
    const uint32_t g = chunk_grid(*c, blocks);
    const void* k = stage_views_kernel(stage, anyMis, g < c->start[c->n]);
    if (!k) return hipErrorInvalidValue;
    args[0] = const_cast<ViewChunk*>(c);
    *p = hipKernelNodeParams{};
    p->func = const_cast<void*>(k);
    p->gridDim = dim3(g);
    p->blockDim = dim3(256);
    p->sharedMemBytes = 0;
    p->kernelParams = args;
    p->extra = nullptr;
    return hipSuccess;
}

hipError_t launch_commit(const uint32_t* ids, const uint32_t* src, uint32_t* pool, uint32_t n, uint32_t cap,
                         hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_commit, dim3(n), dim3(128), 0, s, ids, reinterpret_cast<const uint4*>(src),
                       reinterpret_cast<uint4*>(pool), n, cap);
    return hipGetLastError();
}

hipError_t launch_uber(const uint32_t* grid, uint32_t* uber, hipStream_t s)
{
    hipLaunchKernelGGL(k_uber, dim3(128), dim3(256), 0, s, grid, uber);
    return hipGetLastError();
}

hipError_t launch_trace_batch(const WorldView& w, const float* O, const float* D, uint32_t n, uint32_t kind,
                              int32_t steps, uint32_t* voxel, float* dist, uint32_t* side, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_batch, dim3((n + 255) / 256), dim3(256), 0, s, w, O, D, n, kind, steps, voxel, dist,
                       side);
    return hipGetLastError();
}

hipError_t launch_ground_truth(const FrameArgs& a, uint32_t spp, uint32_t frame0, double* accum, hipStream_t s)
{
    hipLaunchKernelGGL(k_ground_truth, frame_grid(a.W, a.H), dim3(256), 0, s, a, spp, frame0, accum);
    return hipGetLastError();
}

}  // namespace vr
