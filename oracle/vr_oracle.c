/*
 * vr_oracle.c - CPU restatement of the reference Voxel ReSTIR hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  Build: gcc -O2 -ffp-contract=off -fopenmp -shared (oracle/build.py).
 *
 * The reference kernels are OpenCL C (the cl/ directory).  This file follows them
 * statement by statement in semantics; each function cites its source.  See
 * vr_oracle.h for the floating-point contract.
 */
#include "vr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INVPI 0.31830988618379067153777f  /* template/common.h:197 */
#define TWOPI 6.28318530717958647692528f  /* template/common.h:199 */

/* ------------------------------------------------------------------ */
/* scalar helpers                                                       */
/* ------------------------------------------------------------------ */

/* convert_uint / (uint) of a float: truncate toward zero, saturate, NaN -> 0 */
static inline uint32_t f2u(float x)
{
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}

/* convert_int / (int) of a float: truncate toward zero, saturate, NaN -> 0 */
static inline int32_t f2i(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

static inline uint32_t clampu(uint32_t v, uint32_t lo, uint32_t hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

static inline float dot3(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* OpenCL normalize(): v * (1 / sqrt(dot(v, v))) */
static inline void normalize3(const float v[3], float out[3])
{
    const float inv = 1.0f / sqrtf(dot3(v, v));
    out[0] = v[0] * inv, out[1] = v[1] * inv, out[2] = v[2] * inv;
}

/* ------------------------------------------------------------------ */
/* RNG: cl/tools.cl:17-19; host xorshift template/template.cpp:933-951  */
/* ------------------------------------------------------------------ */
uint32_t or_wang_hash(uint32_t s)
{
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}

uint32_t or_random_int(uint32_t* s)
{
    *s ^= *s << 13;
    *s ^= *s >> 17;
    *s ^= *s << 5;
    return *s;
}

float or_random_float(uint32_t* s)
{
    return (float)or_random_int(s) * 2.3283064365387e-10f;
}

/* ------------------------------------------------------------------ */
/* voxel codec: cl/tools.cl:294-322, template/common.h:203-211          */
/* ------------------------------------------------------------------ */
uint32_t or_get_alpha(uint32_t v) { return (v & 0xf000u) >> 12; }

float or_get_alphaf(uint32_t v)
{
    const uint32_t a = (v & 0xf000u) >> 12;
    return (float)a / 15.0f;
}

void or_to_float_rgb(uint32_t v, float out[3])
{
    out[0] = (float)((v >> 8) & 15u) * (1.0f / 15.0f);
    out[1] = (float)((v >> 4) & 15u) * (1.0f / 15.0f);
    out[2] = (float)(v & 15u) * (1.0f / 15.0f);
}

uint32_t or_from_float_rgba(const float rgba[4], uint32_t emit)
{
    return (emit << 16) + (f2u(rgba[3] * 15.0f) << 12) + (f2u(rgba[0] * 15.0f) << 8) +
           (f2u(rgba[1] * 15.0f) << 4) + f2u(rgba[2] * 15.0f);
}

static inline uint32_t is_emitter(uint32_t v) { return ((v >> 16) & 255u) > 0u; }

static inline float emit_strength(const or_config* c, uint32_t v)
{
    return c->emit_lut[(v >> 16) & 255u];
}

/* ------------------------------------------------------------------ */
/* Deterministic sin/cos.  The reference calls OpenCL cos()/sin() under  */
/* -cl-fast-relaxed-math (cl/spatialSampling.cl:24-26), whose bits are   */
/* implementation-defined.  Definition used here (and by the product):   */
/* Cody-Waite reduction by pi/2 in binary64, degree-18/19 Taylor         */
/* polynomials in binary64 (Horner, no FMA), rounded once to binary32.   */
/* ------------------------------------------------------------------ */
void or_sincos(float angle, float* s, float* c)
{
    const double x = (double)angle;
    const double q = floor(x * 0.63661977236758134308 + 0.5); /* 2/pi */
    const double r = (x - q * 1.5707963267341256) - q * 6.0771005065061922e-11;
    const double r2 = r * r;
    double sp = 1.0 / 121645100408832000.0;                 /* 1/19! */
    sp = sp * -r2 + 1.0 / 355687428096000.0;                /* 1/17! */
    sp = sp * -r2 + 1.0 / 1307674368000.0;                  /* 1/15! */
    sp = sp * -r2 + 1.0 / 6227020800.0;                     /* 1/13! */
    sp = sp * -r2 + 1.0 / 39916800.0;                       /* 1/11! */
    sp = sp * -r2 + 1.0 / 362880.0;                         /* 1/9!  */
    sp = sp * -r2 + 1.0 / 5040.0;                           /* 1/7!  */
    sp = sp * -r2 + 1.0 / 120.0;                            /* 1/5!  */
    sp = sp * -r2 + 1.0 / 6.0;                              /* 1/3!  */
    sp = sp * -r2 + 1.0;
    const double sn = sp * r;
    double cp = 1.0 / 6402373705728000.0;                   /* 1/18! */
    cp = cp * -r2 + 1.0 / 20922789888000.0;                 /* 1/16! */
    cp = cp * -r2 + 1.0 / 87178291200.0;                    /* 1/14! */
    cp = cp * -r2 + 1.0 / 479001600.0;                      /* 1/12! */
    cp = cp * -r2 + 1.0 / 3628800.0;                        /* 1/10! */
    cp = cp * -r2 + 1.0 / 40320.0;                          /* 1/8!  */
    cp = cp * -r2 + 1.0 / 720.0;                            /* 1/6!  */
    cp = cp * -r2 + 1.0 / 24.0;                             /* 1/4!  */
    cp = cp * -r2 + 1.0 / 2.0;                              /* 1/2!  */
    cp = cp * -r2 + 1.0;
    const double cs = cp;
    const int quad = ((int)q) & 3;
    double so, co;
    if (quad == 0) so = sn, co = cs;
    else if (quad == 1) so = cs, co = -sn;
    else if (quad == 2) so = -sn, co = -cs;
    else so = -cs, co = sn;
    *s = (float)so;
    *c = (float)co;
}

/* ------------------------------------------------------------------ */
/* camera: cl/tools.cl:102-127 (TAA=0, PANINI=0)                        */
/* ------------------------------------------------------------------ */
void or_camera_ray(const or_params* p, int x, int y, float D[3])
{
    const float u = (float)x * p->oneOverRes[0];
    const float v = (float)y * p->oneOverRes[1];
    float P[3], R[3];
    for (int i = 0; i < 3; i++) {
        P[i] = p->p0[i] + (p->p1[i] - p->p0[i]) * u + (p->p2[i] - p->p0[i]) * v;
        R[i] = P[i] - p->E[i];
    }
    normalize3(R, D);
}

/* VoxelNormal, cl/tools.cl:416-421 */
static inline void voxel_normal(uint32_t side, const float D[3], float N[3])
{
    N[0] = N[1] = N[2] = 0.0f;
    if (side > 2) return; /* unreachable: side is 0, 1 or 2 */
    N[side] = D[side] > 0 ? -1.0f : 1.0f;
}

/* ------------------------------------------------------------------ */
/* brick-grid DDA: cl/trace.cl:121-362                                  */
/*   kind ALL:     TraceRay        (285-362), hit iff v != 0             */
/*   kind OPAQUE:  TraceOpaqueRay  (205-282), hit iff GetAlpha(v) == 15  */
/*   kind THROUGH: TraceThrough    (121-200), hit iff (v & mask) !=     */
/*                 (through & mask); uber cells entered when o || through */
/* ------------------------------------------------------------------ */
#define TOPMASK3  (((1023u - 7u) << 20) + ((1023u - 7u) << 10) + (1023u - 7u))
#define UBERMASK3 ((1020u << 20) + (1020u << 10) + 1020u)

static inline int zero_test(int kind, uint32_t v, uint32_t through, uint32_t mask)
{
    if (kind == OR_TRACE_ALL) return v != 0;
    if (kind == OR_TRACE_OPAQUE) return ((v & 0xf000u) >> 12) == 0xfu;
    return (v & mask) != (through & mask);
}

uint32_t or_trace(const or_world_view* w, const float Ain[3], const float B[3], int steps,
                  int kind, uint32_t through, uint32_t mask, float* dist, uint32_t* side)
{
    /* FixZeroDeltas trace.cl:37-43 */
    float V[3] = {B[0], B[1], B[2]};
    for (int i = 0; i < 3; i++)
        if (fabsf(V[i]) < 1e-8f) V[i] = V[i] < 0 ? -1e-8f : 1e-8f;
    const float rV[3] = {1.0f / V[0], 1.0f / V[1], 1.0f / V[2]};
    /* 'bits' magic trace.cl:301 -> per-axis direction and plane offset */
    int dir[3], offs[3];
    for (int i = 0; i < 3; i++) dir[i] = V[i] > 0 ? 1 : -1, offs[i] = V[i] > 0 ? 1 : 0;
    const uint32_t dx = ((uint32_t)dir[0]) << 20, dy = ((uint32_t)dir[1]) << 10,
                   dz = (uint32_t)dir[2];
    float A[3] = {Ain[0], Ain[1], Ain[2]};
    float to = 0;
    uint32_t last = 0;
    if (A[0] < 0 || A[1] < 0 || A[2] < 0 || A[0] > 1024.0f || A[1] > 1024.0f || A[2] > 1024.0f) {
        /* slab test trace.cl:303-317 */
        const float tx1 = -A[0] * rV[0], tx2 = (1024.0f - A[0]) * rV[0];
        float tmin = fminf(tx1, tx2), tmax = fmaxf(tx1, tx2);
        const float ty1 = -A[1] * rV[1], ty2 = (1024.0f - A[1]) * rV[1];
        tmin = fmaxf(tmin, fminf(ty1, ty2)), tmax = fminf(tmax, fmaxf(ty1, ty2));
        const float tz1 = -A[2] * rV[2], tz2 = (1024.0f - A[2]) * rV[2];
        tmin = fmaxf(tmin, fminf(tz1, tz2)), tmax = fminf(tmax, fmaxf(tz1, tz2));
        if (tmax < tmin || tmax <= 0) return 0;
        A[0] += tmin * V[0], A[1] += tmin * V[1], A[2] += tmin * V[2];
        to = tmin;
        if (A[1] < 0.01f || A[1] > (1024.0f - 1.01f)) last = 1;
        if (A[2] < 0.01f || A[2] > (1024.0f - 1.01f)) last = 2;
    }
    uint32_t up = (clampu(f2u(A[0]) >> 5, 0u, 31u) << 20) + (clampu(f2u(A[1]) >> 5, 0u, 31u) << 10) +
                  clampu(f2u(A[2]) >> 5, 0u, 31u);
    float tm[3];
    tm[0] = ((float)((up >> 20) + offs[0]) - A[0] * 0.03125f) * rV[0];
    tm[1] = ((float)(((up >> 10) & 31u) + offs[1]) - A[1] * 0.03125f) * rV[1];
    tm[2] = ((float)((up & 31u) + offs[2]) - A[2] * 0.03125f) * rV[2];
    float t = 0;
    const float td[3] = {(float)dir[0] * rV[0], (float)dir[1] * rV[1], (float)dir[2] * rV[2]};
    uint32_t o = w->uber[(up >> 20) + ((up & 31u) << 5) + (((up >> 10) & 31u) << 10)];
    while (1) {
        if ((steps -= 4) <= 0) break;
        const int enter = (kind == OR_TRACE_THROUGH) ? (o != 0 || through != 0) : (o != 0);
        if (enter) {
            const float um[3] = {tm[0], tm[1], tm[2]}; /* uber state backup */
            t *= 4;
            const uint32_t p4x = f2u(0.125f * A[0] + V[0] * t);
            const uint32_t p4y = f2u(0.125f * A[1] + V[1] * t);
            const uint32_t p4z = f2u(0.125f * A[2] + V[2] * t);
            uint32_t tp = (clampu(p4x, up >> 18, (up >> 18) + 3) << 20) +
                          (clampu(p4y, (up >> 8) & 1023u, ((up >> 8) & 1023u) + 3) << 10) +
                          clampu(p4z, (up << 2) & 1023u, ((up << 2) & 1023u) + 3);
            const uint32_t tq = tp & UBERMASK3;
            tm[0] = ((float)((tp >> 20) + offs[0]) - A[0] * 0.125f) * rV[0];
            tm[1] = ((float)(((tp >> 10) & 127u) + offs[1]) - A[1] * 0.125f) * rV[1];
            tm[2] = ((float)((tp & 127u) + offs[2]) - A[2] * 0.125f) * rV[2];
            o = w->grid[(tp >> 20) + (tp & 127u) * 128u + ((tp >> 10) & 127u) * 16384u];
            while (1) {
                /* GRIDSTEP trace.cl:75-111 */
                if (!--steps) break;
                const uint32_t v = o >> 1;
                if ((zero_test(kind, v, through, mask) && (o & 1u) == 0) || (o & 1u) == 1) {
                    if ((o & 1u) == 0) {
                        *dist = t * 8.0f + to, *side = last;
                        return v;
                    }
                    const float gm[3] = {tm[0], tm[1], tm[2]}; /* top-grid state backup */
                    t *= 8;
                    const uint32_t q4x = f2u(A[0] + V[0] * t);
                    const uint32_t q4y = f2u(A[1] + V[1] * t);
                    const uint32_t q4z = f2u(A[2] + V[2] * t);
                    uint32_t p = (clampu(q4x, tp >> 17, (tp >> 17) + 7) << 20) +
                                 (clampu(q4y, (tp >> 7) & 1023u, ((tp >> 7) & 1023u) + 7) << 10) +
                                 clampu(q4z, (tp << 3) & 1023u, ((tp << 3) & 1023u) + 7);
                    tm[0] = ((float)((p >> 20) + offs[0]) - A[0]) * rV[0];
                    tm[1] = ((float)(((p >> 10) & 1023u) + offs[1]) - A[1]) * rV[1];
                    tm[2] = ((float)((p & 1023u) + offs[2]) - A[2]) * rV[2];
                    p &= 7u + (7u << 10) + (7u << 20);
                    const uint32_t base = (o - 1u) << 8; /* 'o = --o << 8' trace.cl:94 */
                    for (int k = 0; k < 24; k++) {
                        /* BRICKSTEP trace.cl:50-57 */
                        const uint32_t vv = w->bricks[(size_t)base + (p >> 20) + ((p >> 7) & 56u) + (p & 7u) * 64u];
                        if (zero_test(kind, vv, through, mask)) {
                            *dist = t + to, *side = last;
                            return vv;
                        }
                        t = fminf(tm[0], fminf(tm[1], tm[2])), last = 0;
                        if (t == tm[0]) tm[0] += td[0], p += dx;
                        if (t == tm[1]) tm[1] += td[1], p += dy, last = 1;
                        if (t == tm[2]) tm[2] += td[2], p += dz, last = 2;
                        if (p & TOPMASK3) break;
                    }
                    tm[0] = gm[0], tm[1] = gm[1], tm[2] = gm[2];
                }
                t = fminf(tm[0], fminf(tm[1], tm[2])), last = 0;
                if (t == tm[0]) tm[0] += td[0], tp += dx;
                if (t == tm[1]) tm[1] += td[1], tp += dy, last = 1;
                if (t == tm[2]) tm[2] += td[2], tp += dz, last = 2;
                if ((tp & UBERMASK3) - tq) break;
                o = w->grid[(tp >> 20) + (tp & 127u) * 128u + ((tp >> 10) & 127u) * 16384u];
            }
            tm[0] = um[0], tm[1] = um[1], tm[2] = um[2];
        }
        t = fminf(tm[0], fminf(tm[1], tm[2])), last = 0;
        if (t == tm[0]) tm[0] += td[0], up += dx;
        if (t == tm[1]) tm[1] += td[1], up += dy, last = 1;
        if (t == tm[2]) tm[2] += td[2], up += dz, last = 2;
        if (up & 0xfe0f83e0u) break;
        o = w->uber[(up >> 20) + ((up & 31u) << 5) + (((up >> 10) & 31u) << 10)];
    }
    return 0u;
}

void or_trace_batch(const or_world_view* w, const float* O, const float* D, uint32_t n,
                    int steps, int kind, uint32_t* voxel, float* dist, uint32_t* side)
{
    for (uint32_t i = 0; i < n; i++) {
        float d = 0.0f;
        uint32_t s = 0;
        voxel[i] = or_trace(w, O + 3 * i, D + 3 * i, steps, kind, 0, 0, &d, &s);
        dist[i] = d, side[i] = s;
    }
}

/* ------------------------------------------------------------------ */
/* light sampling: cl/tools.cl:423-434, 452-472, 499-621                */
/* ------------------------------------------------------------------ */

/* indexToCoordinates, cl/tools.cl:428-434 */
static inline void index_to_coords(uint32_t i, uint32_t c[3])
{
    c[1] = i / (1024u * 1024u);
    c[2] = (i / 1024u) % 1024u;
    c[0] = i % 1024u;
}

/* RandomPointOnVoxel, cl/tools.cl:499-577 */
void or_random_point_on_voxel(const float center[3], const float sp[3], int isize,
                              uint32_t* seed, float* invPdf, float Nl[3], float pos[3])
{
    static const float normals[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                        {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    static const int tix[6] = {1, 2, 0, 4, 5, 3};
    static const int bix[6] = {2, 0, 1, 5, 3, 4};
    const float size = (float)isize;
    float solid[6], pts[6][3];
    float total = 0;
    for (int i = 0; i < 6; i++) {
        const float* N = normals[i];
        const float* T = normals[tix[i]];
        const float* Bn = normals[bix[i]];
        const float r0 = or_random_float(seed);
        const float r1 = or_random_float(seed);
        float R[3], L[3];
        for (int k = 0; k < 3; k++) {
            pts[i][k] = center[k] + (0.5f * size * N[k]) - (0.5f * size * T[k]) - (0.5f * size * Bn[k]) +
                        (r0 * size * T[k]) + (r1 * size * Bn[k]);
            R[k] = pts[i][k] - sp[k];
        }
        normalize3(R, L);
        const float d2 = fmaxf(1e-5f, dot3(R, R));
        const float negL[3] = {-L[0], -L[1], -L[2]};
        const float ndl = dot3(N, negL);
        const float area = size * size;
        const float sa = ndl * area / d2;
        solid[i] = fmaxf(0.0f, sa);
        total += fmaxf(0.0f, sa);
    }
    float pdf[6], cdf[6];
    const float r2 = or_random_float(seed);
    uint32_t index = 0;
    pdf[0] = solid[0] / total;
    cdf[0] = pdf[0];
    if (cdf[0] < r2) {
        for (uint32_t i = 1; i < 6; i++) {
            pdf[i] = solid[i] / total;
            cdf[i] = cdf[i - 1] + pdf[i];
            index = i;
            if (cdf[i] >= r2) break;
        }
    }
    float weight = 1.0f / pdf[index];
    if (isinf(weight) || isnan(weight)) weight = 0.0f;
    *invPdf = weight;
    Nl[0] = normals[index][0], Nl[1] = normals[index][1], Nl[2] = normals[index][2];
    pos[0] = pts[index][0], pos[1] = pts[index][1], pos[2] = pts[index][2];
}

/* PointOnVoxelLight, cl/tools.cl:600-621 */
static void point_on_voxel_light(const or_config* c, const or_light* light, const float sp[3],
                                 uint32_t* seed, float pos[3], float* invPdf, float Nl[3])
{
    uint32_t coords[3];
    index_to_coords(light->position, coords);
    const float hs = 0.5f * (float)light->size;
    const float center[3] = {(float)coords[0] + hs, (float)coords[1] + hs, (float)coords[2] + hs};
    if (c->point_lights) {
        *invPdf = 1.0f;
        pos[0] = center[0], pos[1] = center[1], pos[2] = center[2];
        Nl[0] = Nl[1] = Nl[2] = 0.0f;
    } else {
        or_random_point_on_voxel(center, sp, (int)light->size, seed, invPdf, Nl, pos);
    }
}

/* evaluatePHatFull, cl/phat.cl:2-29 */
void or_eval_phat_full(const or_config* c, const float pos[3], const float Nl[3],
                       const float sp[3], const float N[3], const float brdf[3],
                       uint32_t lvoxel, uint32_t lsize, float out[3])
{
    const float size = (float)lsize;
    float em[3];
    or_to_float_rgb(lvoxel, em);
    const float es = emit_strength(c, lvoxel);
    em[0] = em[0] * es, em[1] = em[1] * es, em[2] = em[2] * es;
    if (c->gfxexp) em[0] = em[0] * INVPI, em[1] = em[1] * INVPI, em[2] = em[2] * INVPI;
    float R[3], L[3];
    R[0] = pos[0] - sp[0], R[1] = pos[1] - sp[1], R[2] = pos[2] - sp[2];
    normalize3(R, L);
    const float NdotL = fmaxf(dot3(N, L), 0.0f);
    const float d2 = fmaxf(1e-5f, dot3(R, R));
    const float area = size * size;
    float nldl = 1.0f;
    if (!c->point_lights) {
        const float negL[3] = {-L[0], -L[1], -L[2]};
        nldl = fmaxf(dot3(Nl, negL), 0.0f);
    }
    const float sa = (nldl * area) / d2;
    for (int k = 0; k < 3; k++) out[k] = brdf[k] * em[k] * NdotL * sa;
}

/* evaluatePHat, cl/phat.cl:32-41 */
float or_eval_phat(const or_config* c, const float pos[3], const float Nl[3],
                   const float sp[3], const float N[3], const float brdf[3],
                   uint32_t lvoxel, uint32_t lsize)
{
    float v[3];
    or_eval_phat_full(c, pos, Nl, sp, N, brdf, lvoxel, lsize, v);
    const float r = sqrtf(dot3(v, v));
    if (isnan(r) || isinf(r)) return 0.0f;
    return r;
}

/* ------------------------------------------------------------------ */
/* reservoir algebra: cl/reservoir.cl                                   */
/* ------------------------------------------------------------------ */
void or_reservoir_update(or_reservoir* r, float w, float pHat, uint32_t index, float r0,
                         const float pos[3], float invPdf, const float Nl[3])
{
    /* UpdateReservoir reservoir.cl:3-15 */
    r->M += 1;
    r->wsum += w;
    if (r0 < w / fmaxf(1e-5f, r->wsum)) {
        r->lightIndex = index;
        r->pos.x = pos[0], r->pos.y = pos[1], r->pos.z = pos[2];
        r->invPdf = invPdf;
        r->Nlight.x = Nl[0], r->Nlight.y = Nl[1], r->Nlight.z = Nl[2];
        r->pHat = pHat;
    }
}

void or_reservoir_reweigh(or_reservoir* r, float newPHat, uint32_t M)
{
    /* ReWeighSumOfWeights reservoir.cl:22-28 */
    r->wsum = r->W * (float)M * newPHat;
    r->M = M;
    r->pHat = newPHat;
}

void or_reservoir_combine(or_reservoir* r, const or_reservoir* that, float r0)
{
    /* CombineReservoir reservoir.cl:30-38 */
    const uint32_t m1 = r->M, m2 = that->M;
    const float pos[3] = {that->pos.x, that->pos.y, that->pos.z};
    const float nl[3] = {that->Nlight.x, that->Nlight.y, that->Nlight.z};
    or_reservoir_update(r, that->wsum, that->pHat, that->lightIndex, r0, pos, that->invPdf, nl);
    r->M = m1 + m2;
}

void or_reservoir_adjust(or_reservoir* r, float pHat)
{
    /* AdjustWeight reservoir.cl:40-43 */
    r->W = r->wsum / fmaxf(1e-5f, pHat * (float)r->M);
}

/* ReprojectWorldPoint, cl/tools.cl:479-488 */
void or_reproject(const or_params* p, uint32_t W, uint32_t H, const float sp[3], int out[2])
{
    const float dl = dot3(sp, p->Nleft) - p->Nleft[3];
    const float dr = dot3(sp, p->Nright) - p->Nright[3];
    const float dt = dot3(sp, p->Ntop) - p->Ntop[3];
    const float db = dot3(sp, p->Nbottom) - p->Nbottom[3];
    const float u = (float)W * (dl / (dl + dr));
    const float v = (float)H * (dt / (dt + db));
    out[0] = f2i(roundf(u));
    out[1] = f2i(roundf(v));
}

/* ------------------------------------------------------------------ */
/* source pdf                                                           */
/* ------------------------------------------------------------------ */

/* power-proportional alias table (Vose).  Power of light i is
 * lightContribution (cl/tools.cl:579-591) reduced to the sum of its rgb. */
int or_build_power_table(const or_config* c, const or_light* lights, uint32_t n,
                         float* prob, uint32_t* alias, float* pmf)
{
    if (n == 0) return 0;
    double* pw = (double*)malloc(sizeof(double) * n);
    uint32_t* small = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* large = (uint32_t*)malloc(sizeof(uint32_t) * n);
    double total = 0;
    for (uint32_t i = 0; i < n; i++) {
        float rgb[3];
        or_to_float_rgb(lights[i].voxel, rgb);
        const double es = (double)emit_strength(c, lights[i].voxel);
        const double s = (double)lights[i].size;
        const double geo = c->point_lights ? 4.0 * 3.14159265358979323846 : 6.0 * s * s * 3.14159265358979323846;
        pw[i] = ((double)rgb[0] + (double)rgb[1] + (double)rgb[2]) * es * geo;
        total += pw[i];
    }
    if (!(total > 0)) {
        for (uint32_t i = 0; i < n; i++) pw[i] = 1.0;
        total = (double)n;
    }
    double* sc = (double*)malloc(sizeof(double) * n);
    uint32_t ns = 0, nl = 0;
    for (uint32_t i = 0; i < n; i++) {
        pmf[i] = (float)(pw[i] / total);
        sc[i] = pw[i] / total * (double)n;
        if (sc[i] < 1.0) small[ns++] = i;
        else large[nl++] = i;
        alias[i] = i;
    }
    while (ns > 0 && nl > 0) {
        const uint32_t s = small[--ns], l = large[--nl];
        prob[s] = (float)sc[s];
        alias[s] = l;
        sc[l] = (sc[l] + sc[s]) - 1.0;
        if (sc[l] < 1.0) small[ns++] = l;
        else large[nl++] = l;
    }
    while (nl > 0) { const uint32_t l = large[--nl]; prob[l] = 1.0f; alias[l] = l; }
    while (ns > 0) { const uint32_t s = small[--ns]; prob[s] = 1.0f; alias[s] = s; }
    free(pw), free(small), free(large), free(sc);
    return 0;
}

/* GenerateLightCandidate, cl/tools.cl:593-598 (+ getRandomIndex 423-426).
 * Power source: the same single uniform draw picks a bucket and, from its
 * fractional part, the bucket's own light or its alias. */
static inline uint32_t gen_candidate(const or_config* c, const or_alias* src, uint32_t N,
                                     uint32_t* seed, float* p)
{
    const float u = or_random_float(seed);
    int32_t i = f2i(u * (float)N);
    if (i < 0) i = 0;
    if (i > (int32_t)N - 1) i = (int32_t)N - 1;
    if (c->source_pdf == OR_SOURCE_POWER && src && src->prob) {
        const float fr = u * (float)N - (float)i;
        const uint32_t k = fr < src->prob[i] ? (uint32_t)i : src->alias[i];
        *p = src->pmf[k];
        return k;
    }
    *p = 1.0f / (float)N;
    return (uint32_t)i;
}

/* GridCoordinatesFromHit + Uint3CubeEquals, cl/tools.cl:452-472;
 * isLightOccluded cl/ris.cl:1-17.  A ray that hits nothing (step budget
 * exhausted) leaves dist/side unset in the reference; here it is occluded. */
static int light_occluded(const or_config* c, const or_world_view* w, const or_light* light,
                          const float spo[3], const float L[3])
{
    uint32_t ec[3];
    index_to_coords(light->position, ec);
    float dist = 0;
    uint32_t side = 0;
    const int steps = (c->quirks & OR_QUIRK_SHADOW_STEP_CAP) ? 128 : 999999;
    const uint32_t v = or_trace(w, spo, L, steps, OR_TRACE_OPAQUE, 0, 0, &dist, &side);
    if (v == 0) return 1;
    float N[3], hc[3];
    voxel_normal(side, L, N);
    for (int k = 0; k < 3; k++) hc[k] = (L[k] * dist + spo[k]) - N[k] * 0.1f;
    const uint32_t h[3] = {f2u(hc[0]), f2u(hc[1]), f2u(hc[2])};
    const uint32_t s = light->size;
    const int inside = ec[0] <= h[0] && ec[0] + s > h[0] && ec[1] <= h[1] && ec[1] + s > h[1] &&
                       ec[2] <= h[2] && ec[2] + s > h[2];
    return !inside;
}

/* same-face test of temporal and spatial reuse */
static inline int same_face(const or_config* c, const or_ray* a, const or_ray* b)
{
    if (a->side != b->side) return 0;
    if (c->quirks & OR_QUIRK_AXIS_ONLY_FACE) return 1;
    const float da = a->side == 0 ? a->D.x : (a->side == 1 ? a->D.y : a->D.z);
    const float db = b->side == 0 ? b->D.x : (b->side == 1 ? b->D.y : b->D.z);
    return (da > 0) == (db > 0);
}

/* ------------------------------------------------------------------ */
/* stages                                                               */
/* ------------------------------------------------------------------ */

/* renderAlbedo, cl/ris.cl:295-389 */
void or_stage_albedo(const or_config* c, const or_params* p, const or_world_view* w,
                     or_ray* hits, int y0, int y1)
{
    const uint32_t W = c->width;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < (int)W; x++) {
            const uint32_t seed = or_wang_hash((uint32_t)x * 171u + (uint32_t)y * 1773u + p->R0);
            float D[3];
            or_camera_ray(p, x, y, D);
            uint32_t side = 0;
            float dist = 0;
            float origin[3] = {p->E[0], p->E[1], p->E[2]};
            uint32_t voxel = or_trace(w, origin, D, 999999, OR_TRACE_ALL, 0, 0, &dist, &side);
            float alpha = or_get_alphaf(voxel);
            float col[3], out[3];
            or_to_float_rgb(voxel, col);
            out[0] = col[0] * alpha, out[1] = col[1] * alpha, out[2] = col[2] * alpha;
            float total = 0.f;
            int iteration = 0;
            float remaining = 1.0f - or_get_alphaf(voxel);
            uint32_t prevSide = 0, prevVoxel = 0;
            while (prevVoxel != voxel && remaining > 0.f) {
                prevVoxel = voxel;
                prevSide = side;
                total += dist + 0.01f;
                for (int k = 0; k < 3; k++) origin[k] = p->E[k] + total * D[k];
                voxel = or_trace(w, origin, D, 999999, OR_TRACE_THROUGH, voxel, 0x00ffffu, &dist, &side);
                or_to_float_rgb(voxel, col);
                alpha = or_get_alphaf(voxel);
                for (int k = 0; k < 3; k++) out[k] += remaining * col[k] * alpha;
                remaining = remaining * (1 - alpha);
                if (++iteration == 25) remaining = 0.f;
                if (dist <= 0.001f) break;
            }
            if (voxel != 0) total += dist;
            else side = prevSide;
            const float rgba[4] = {out[0], out[1], out[2], 1.0f - remaining};
            /* the emit byte is carried over unchanged (EmitStrength of the
             * reference equals the byte; with an intensity LUT it is not) */
            voxel = or_from_float_rgba(rgba, (voxel >> 16) & 255u);
            or_ray* h = &hits[(size_t)y * W + x];
            h->voxel = voxel, h->side = side, h->seed = seed, h->dist = total;
            h->D.x = D[0], h->D.y = D[1], h->D.z = D[2];
        }
}

/* perPixelInitialSampling, cl/ris.cl:20-124 */
void or_stage_initial(const or_config* c, const or_params* p, const or_world_view* w,
                      const or_light* lights, const or_alias* src, const or_ray* prevHits,
                      or_ray* hits, or_reservoir* out, const or_reservoir* prevRes,
                      int y0, int y1)
{
    const uint32_t W = c->width, H = c->height;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < (int)W; x++) {
            or_ray* ray = &hits[(size_t)y * W + x];
            const float dist = ray->dist;
            const uint32_t voxel = ray->voxel;
            or_reservoir res;
            memset(&res, 0, sizeof(res));
            if (voxel != 0) {
                uint32_t* seed = &ray->seed;
                const float D[3] = {ray->D.x, ray->D.y, ray->D.z};
                const uint32_t side = ray->side;
                float N[3], brdf[3], sp[3], spo[3];
                voxel_normal(side, D, N);
                or_to_float_rgb(voxel, brdf);
                for (int k = 0; k < 3; k++) {
                    brdf[k] = brdf[k] * INVPI;
                    sp[k] = D[k] * dist + p->E[k];
                    spo[k] = sp[k] + 0.1f * N[k];
                }
                const uint32_t nL = p->numberOfLights;
                const uint32_t nC = nL < p->numberOfCandidates ? nL : p->numberOfCandidates;
                for (uint32_t i = 0; i < nC; i++) {
                    float pr;
                    uint32_t li;
                    if (c->source_pdf == OR_SOURCE_SLOTS) {
                        const size_t slot = (size_t)i * W * H + (size_t)y * W + x;
                        const uint32_t cl = c->cand_idx[slot];
                        const float mc = c->cand_pdf[slot];
                        const uint32_t c0 = cl < 256u ? c->cl_start[cl] : 0u;
                        const uint32_t cnt = cl < 256u ? c->cl_start[cl + 1] - c0 : 0u;
                        const float u = or_random_float(seed);
                        if (cnt == 0 || !(mc > 0.0f)) {
                            res.M += 1;
                            continue;
                        }
                        int32_t k = f2i(u * (float)cnt);
                        if (k < 0) k = 0;
                        if (k > (int32_t)cnt - 1) k = (int32_t)cnt - 1;
                        const float fr = u * (float)cnt - (float)k;
                        const uint32_t kk = fr < c->cl_prob[c0 + k] ? (uint32_t)k : c->cl_alias[c0 + k];
                        li = c->cl_lights[c0 + kk];
                        pr = mc * c->cl_share[li];
                    } else {
                        li = gen_candidate(c, src, nL, seed, &pr);
                    }
                    const or_light* light = &lights[li];
                    float pos[3], invPdf, Nl[3];
                    point_on_voxel_light(c, light, spo, seed, pos, &invPdf, Nl);
                    const float pHat = or_eval_phat(c, pos, Nl, sp, N, brdf, light->voxel, light->size) * invPdf;
                    or_reservoir_update(&res, pHat / pr, pHat, li, or_random_float(seed), pos, invPdf, Nl);
                }
                or_reservoir_adjust(&res, res.pHat);
                if (nL > 0 && res.wsum != 0.0f) {
                    /* with wsum == 0 (hence W == 0) the discard below changes nothing */
                    const or_light* light = &lights[res.lightIndex];
                    float R[3], L[3];
                    R[0] = res.pos.x - spo[0], R[1] = res.pos.y - spo[1], R[2] = res.pos.z - spo[2];
                    normalize3(R, L);
                    if (light_occluded(c, w, light, spo, L)) res.W = 0, res.wsum = 0;
                }
                if (p->restirtemporalframe > 0 && p->temporal) {
                    int pxy[2];
                    or_reproject(p, W, H, sp, pxy);
                    if (pxy[0] >= 0 && pxy[1] >= 0 && pxy[0] < (int)W && pxy[1] < (int)H) {
                        const size_t pi = (size_t)pxy[1] * W + pxy[0];
                        const or_ray* pr = &prevHits[pi];
                        if (same_face(c, pr, ray) && fabsf(dist - pr->dist) <= 0.1f * dist) {
                            or_reservoir prev = prevRes[pi];
                            const uint32_t cap = p->numberOfMaxTemporalImportance * nC;
                            const uint32_t pm = cap < prev.M ? cap : prev.M;
                            if (pm > 0 && prev.lightIndex < nL) {
                                const or_light* pl = &lights[prev.lightIndex];
                                const float ppos[3] = {prev.pos.x, prev.pos.y, prev.pos.z};
                                const float pnl[3] = {prev.Nlight.x, prev.Nlight.y, prev.Nlight.z};
                                const float pph = or_eval_phat(c, ppos, pnl, sp, N, brdf, pl->voxel, pl->size) * prev.invPdf;
                                or_reservoir_reweigh(&prev, pph, pm);
                                or_reservoir_combine(&res, &prev, or_random_float(seed));
                                const or_light* l2 = &lights[res.lightIndex];
                                const float rpos[3] = {res.pos.x, res.pos.y, res.pos.z};
                                const float rnl[3] = {res.Nlight.x, res.Nlight.y, res.Nlight.z};
                                const float ph = or_eval_phat(c, rpos, rnl, sp, N, brdf, l2->voxel, l2->size) * res.invPdf;
                                or_reservoir_adjust(&res, ph);
                            }
                        }
                    }
                }
            }
            out[(size_t)y * W + x] = res;
        }
}

/* Unbiased contribution weight of the spatially merged sample y under
 * generalized balance-heuristic MIS (Bitterli et al. 2020, "Spatiotemporal
 * reservoir resampling", sec. 4: weights evaluated for the selected sample
 * only).  Input s of the merge (own reservoir, then each combined neighbour
 * in tap order) has target T_s(y) = phat_s(y) * V_s(y): the target of its
 * pixel times visibility from that pixel (one shadow ray per neighbour);
 * the own term omits the ray because the shading pass multiplies by the
 * own visibility.  With resampling weights phat(y_s) W_s M_s the estimate
 * is unbiased when
 *     W = wsum * T_sel(y) / (pHat * sum_s M_s T_s(y)),
 * pHat being the merged reservoir's own target of y (phat * invPdf; the
 * common invPdf cancels between T_sel and the sum).  W = 0 when the
 * denominator is 0 or not finite. */
static float mis_weight(const or_config* c, const or_params* p, const or_world_view* w, const or_light* lights,
                        const or_ray* hits, const or_reservoir* in, const or_reservoir* res, const float sp[3],
                        const float N[3], const float brdf[3], uint32_t M0, const size_t* acc, int nacc, int sel)
{
    if (!(res->wsum > 0.0f) || res->lightIndex >= p->numberOfLights) return 0.0f;
    const or_light* light = &lights[res->lightIndex];
    const float pos[3] = {res->pos.x, res->pos.y, res->pos.z};
    const float nl[3] = {res->Nlight.x, res->Nlight.y, res->Nlight.z};
    const float p0 = or_eval_phat(c, pos, nl, sp, N, brdf, light->voxel, light->size);
    float den = (float)M0 * p0;
    float psel = sel < 0 ? p0 : 0.0f;
    for (int j = 0; j < nacc; j++) {
        const or_ray* nr = &hits[acc[j]];
        const float nD[3] = {nr->D.x, nr->D.y, nr->D.z};
        float nN[3], nbrdf[3], nsp[3], nspo[3];
        voxel_normal(nr->side, nD, nN);
        or_to_float_rgb(nr->voxel, nbrdf);
        for (int k = 0; k < 3; k++) {
            nbrdf[k] = nbrdf[k] * INVPI;
            nsp[k] = nD[k] * nr->dist + p->E[k];
            nspo[k] = nsp[k] + 0.1f * nN[k];
        }
        float pj = or_eval_phat(c, pos, nl, nsp, nN, nbrdf, light->voxel, light->size);
        if (pj > 0.0f) {
            float R[3], L[3];
            R[0] = pos[0] - nspo[0], R[1] = pos[1] - nspo[1], R[2] = pos[2] - nspo[2];
            normalize3(R, L);
            if (light_occluded(c, w, light, nspo, L)) pj = 0.0f;
        }
        den += (float)in[acc[j]].M * pj;
        if (j == sel) psel = pj;
    }
    const float d = res->pHat * den;
    const float W = (res->wsum * psel) / d;
    return (d > 0.0f && isfinite(W)) ? W : 0.0f;
}

/* perPixelSpatialResampling + _perPixelSpatialResampling,
 * cl/ris.cl:126-151, cl/spatialSampling.cl:2-67 */
void or_stage_spatial(const or_config* c, const or_params* p, const or_world_view* w, const or_light* lights,
                      or_ray* hits, or_reservoir* out, const or_reservoir* in, int y0, int y1)
{
    const uint32_t W = c->width, H = c->height;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < (int)W; x++) {
            const size_t idx = (size_t)y * W + x;
            or_reservoir res = in[idx];
            or_ray* ray = &hits[idx];
            if (ray->voxel != 0 && p->spatial) {
                uint32_t* seed = &ray->seed;
                const float dist = ray->dist;
                const float D[3] = {ray->D.x, ray->D.y, ray->D.z};
                float N[3], brdf[3], sp[3];
                voxel_normal(ray->side, D, N);
                or_to_float_rgb(ray->voxel, brdf);
                for (int k = 0; k < 3; k++) brdf[k] = brdf[k] * INVPI, sp[k] = D[k] * dist + p->E[k];
                /* MIS-weighted merge: the combined neighbours (pixel index)
                 * and which input the selected sample came from (-1 = own) */
                const uint32_t M0 = res.M;
                size_t acc[OR_MIS_MAX_TAPS];
                int nacc = 0, sel = -1;
                for (int i = 0; i < (int)p->spatialTaps; i++) {
                    const float angle = or_random_float(seed) * TWOPI;
                    const float radius = (float)(p->spatialRadius - 1u) * sqrtf(or_random_float(seed)) + 1.0f;
                    float sn, cs;
                    or_sincos(angle, &sn, &cs);
                    const int xo = f2i(cs * radius), yo = f2i(sn * radius);
                    const int nx = x + xo, ny = y + yo;
                    if ((nx == x && ny == y) || nx < 0 || ny < 0 || nx >= (int)W || ny >= (int)H) continue;
                    const size_t ni = (size_t)ny * W + nx;
                    or_reservoir nr = in[ni];
                    if (nr.M > 0) {
                        const or_ray* nray = &hits[ni];
                        if (same_face(c, nray, ray) && fabsf(dist - nray->dist) <= 0.1f * dist &&
                            nr.lightIndex < p->numberOfLights) {
                            const or_light* nl = &lights[nr.lightIndex];
                            const float npos[3] = {nr.pos.x, nr.pos.y, nr.pos.z};
                            const float nnl[3] = {nr.Nlight.x, nr.Nlight.y, nr.Nlight.z};
                            const float ph = or_eval_phat(c, npos, nnl, sp, N, brdf, nl->voxel, nl->size) * nr.invPdf;
                            or_reservoir_reweigh(&nr, ph, nr.M);
                            const float r0 = or_random_float(seed);
                            /* the test of UpdateReservoir: does this input take over? */
                            const int taken = r0 < nr.wsum / fmaxf(1e-5f, res.wsum + nr.wsum);
                            or_reservoir_combine(&res, &nr, r0);
                            if (c->spatial_mis && nacc < OR_MIS_MAX_TAPS) {
                                if (taken) sel = nacc;
                                acc[nacc++] = ni;
                            }
                        }
                    }
                }
                or_reservoir_adjust(&res, res.pHat);
                if (c->spatial_mis) res.W = mis_weight(c, p, w, lights, hits, in, &res, sp, N, brdf, M0, acc, nacc, sel);
            }
            out[idx] = res;
        }
}

/* renderRIS + render_di_ris, cl/ris.cl:155-293, 391-413 (editor and sky
 * dome off: skyDomeSampling=0 for every configuration) */
void or_stage_shade(const or_config* c, const or_params* p, const or_world_view* w,
                    const or_light* lights, const or_ray* hits, float* frame,
                    or_reservoir* out, or_reservoir* in, int y0, int y1)
{
    const uint32_t W = c->width;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < (int)W; x++) {
            const size_t idx = (size_t)y * W + x;
            const or_ray* hit = &hits[idx];
            const float dist = hit->dist;
            const float D[3] = {hit->D.x, hit->D.y, hit->D.z};
            const uint32_t voxel = hit->voxel;
            or_reservoir res = in[idx];
            float color[3] = {0, 0, 0};
            if (voxel != 0) {
                float N[3];
                voxel_normal(hit->side, D, N);
                if (is_emitter(voxel)) {
                    float ec[3];
                    or_to_float_rgb(voxel, ec);
                    const float es = emit_strength(c, voxel);
                    for (int k = 0; k < 3; k++) {
                        ec[k] = ec[k] * es;
                        if (c->gfxexp) ec[k] = ec[k] * INVPI;
                        color[k] += ec[k];
                    }
                }
                if ((c->gfxexp || !is_emitter(voxel)) && res.W > 0 && res.M > 0 &&
                    res.lightIndex < p->numberOfLights) {
                    float sp[3], spo[3];
                    for (int k = 0; k < 3; k++) sp[k] = D[k] * dist + p->E[k], spo[k] = sp[k] + 0.1f * N[k];
                    const or_light* light = &lights[res.lightIndex];
                    float R[3], L[3];
                    R[0] = res.pos.x - spo[0], R[1] = res.pos.y - spo[1], R[2] = res.pos.z - spo[2];
                    normalize3(R, L);
                    if (!light_occluded(c, w, light, spo, L)) {
                        float brdf[3], inc[3];
                        or_to_float_rgb(voxel, brdf);
                        for (int k = 0; k < 3; k++) brdf[k] = brdf[k] * INVPI;
                        const float pos[3] = {res.pos.x, res.pos.y, res.pos.z};
                        const float nl[3] = {res.Nlight.x, res.Nlight.y, res.Nlight.z};
                        or_eval_phat_full(c, pos, nl, sp, N, brdf, light->voxel, light->size, inc);
                        for (int k = 0; k < 3; k++) color[k] += inc[k] * res.W * res.invPdf;
                    } else {
                        res.W = 0, res.wsum = 0;
                        in[idx] = res;
                    }
                }
                if (c->quirks & OR_QUIRK_DOUBLE_RADIANCE) {
                    /* color += alphaF*color + (1-alphaF)*sky, sky = 0 */
                    const float a = or_get_alphaf(voxel);
                    for (int k = 0; k < 3; k++) color[k] += a * color[k] + (1 - a) * 0.0f;
                }
            }
            frame[idx * 4 + 0] = color[0], frame[idx * 4 + 1] = color[1];
            frame[idx * 4 + 2] = color[2], frame[idx * 4 + 3] = dist;
            out[idx] = in[idx];
        }
}

/* ------------------------------------------------------------------ */
/* frame driver: world.cpp:2127-2340                                    */
/* ------------------------------------------------------------------ */
or_state* or_state_create(uint32_t W, uint32_t H)
{
    or_state* s = (or_state*)calloc(1, sizeof(or_state));
    const size_t n = (size_t)W * H;
    s->W = W, s->H = H;
    for (int i = 0; i < 2; i++) {
        s->hits[i] = (or_ray*)calloc(n, sizeof(or_ray));
        s->res[i] = (or_reservoir*)calloc(n, sizeof(or_reservoir));
    }
    s->frame = (float*)calloc(n * 4, sizeof(float));
    s->initial = (or_reservoir*)calloc(n, sizeof(or_reservoir));
    s->hitCur = 0, s->hitPrev = 1;
    s->resOut = 0, s->resIn = 1;
    return s;
}

void or_state_destroy(or_state* s)
{
    if (!s) return;
    for (int i = 0; i < 2; i++) free(s->hits[i]), free(s->res[i]);
    free(s->frame);
    free(s->initial);
    free(s);
}

#define SWAPI(a, b) do { int t_ = a; a = b; b = t_; } while (0)

void or_render_frame(or_state* s, const or_config* c, const or_params* p,
                     const or_world_view* w, const or_light* lights, const or_alias* src,
                     int nthreads)
{
    const int H = (int)c->height;
    /* buffer rotation world.cpp:2209-2223 */
    const int iOut = s->resOut, iIn = s->resIn;
    SWAPI(s->resOut, s->resIn);
    const int sOut = s->resOut, sIn = s->resIn;
    SWAPI(s->resOut, s->resIn);
    const int hOut = s->resOut, hIn = s->resIn;
    SWAPI(s->resOut, s->resIn);
    SWAPI(s->hitCur, s->hitPrev);
    s->initialOut = iOut, s->spatialOut = sOut, s->shadeIn = hIn;
    or_ray* cur = s->hits[s->hitCur];
    const or_ray* prev = s->hits[s->hitPrev];
    if (nthreads < 1) nthreads = 1;
    const int rows = 4;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int y = 0; y < H; y += rows) or_stage_albedo(c, p, w, cur, y, y + rows < H ? y + rows : H);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int y = 0; y < H; y += rows)
        or_stage_initial(c, p, w, lights, src, prev, cur, s->res[iOut], s->res[iIn], y, y + rows < H ? y + rows : H);
    memcpy(s->initial, s->res[iOut], (size_t)s->W * s->H * sizeof(or_reservoir));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int y = 0; y < H; y += rows) or_stage_spatial(c, p, w, lights, cur, s->res[sOut], s->res[sIn], y, y + rows < H ? y + rows : H);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int y = 0; y < H; y += rows)
        or_stage_shade(c, p, w, lights, cur, s->frame, s->res[hOut], s->res[hIn], y, y + rows < H ? y + rows : H);
}

/* ------------------------------------------------------------------ */
/* ground truth: render_di_gfxexp cl/kernels.cl:177-241 with            */
/* UniformSampleHemisphere cl/tools.cl:31-52, accumulated in binary64.  */
/* Sample f uses R0 = xorshift32^(f+1)(p->R0) and frame index           */
/* (frame0 + f) & 255 for the 16x16 strata.                             */
/* ------------------------------------------------------------------ */
void or_ground_truth(const or_config* c, const or_params* p, const or_world_view* w,
                     uint32_t spp, uint32_t frame0, double* accum, int y0, int y1)
{
    const uint32_t W = c->width;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < (int)W; x++) {
            float D[3];
            or_camera_ray(p, x, y, D);
            float dist = 0;
            uint32_t side = 0;
            const uint32_t voxel = or_trace(w, p->E, D, 999999, OR_TRACE_ALL, 0, 0, &dist, &side);
            double acc[3] = {0, 0, 0};
            if (voxel != 0) {
                float base[3], contrib[3], brdf[3], N[3], I[3];
                or_to_float_rgb(voxel, base);
                const float es = emit_strength(c, voxel);
                for (int k = 0; k < 3; k++) {
                    contrib[k] = base[k] * es * INVPI;
                    brdf[k] = INVPI * base[k];
                    I[k] = p->E[k] + D[k] * dist;
                }
                voxel_normal(side, D, N);
                /* tangent frame, tools.cl:33-38 */
                float Nt[3], Nb[3];
                if (fabsf(N[0]) > fabsf(N[1])) {
                    const float l = sqrtf(N[0] * N[0] + N[2] * N[2]);
                    Nt[0] = N[2] / l, Nt[1] = 0.0f / l, Nt[2] = -N[0] / l;
                } else {
                    const float l = sqrtf(N[1] * N[1] + N[2] * N[2]);
                    Nt[0] = 0.0f / l, Nt[1] = -N[2] / l, Nt[2] = N[1] / l;
                }
                Nb[0] = N[1] * Nt[2] - N[2] * Nt[1];
                Nb[1] = N[2] * Nt[0] - N[0] * Nt[2];
                Nb[2] = N[0] * Nt[1] - N[1] * Nt[0];
                float O2[3];
                for (int k = 0; k < 3; k++) O2[k] = I[k] + 0.1f * N[k];
                uint32_t r0state = p->R0;
                for (uint32_t f = 0; f < spp; f++) {
                    const uint32_t R0 = or_random_int(&r0state);
                    uint32_t seed = or_wang_hash((uint32_t)x * 171u + (uint32_t)y * 1773u + R0);
                    const uint32_t frame = (frame0 + f) & 255u;
                    const float INV16 = 1.0f / 16.0f;
                    const float r0 = (float)(int)(frame & 15u) * INV16 + or_random_float(&seed) * INV16;
                    const float r1 = (float)(int)(frame >> 4) * INV16 + or_random_float(&seed) * INV16;
                    const float sinTheta = sqrtf(1 - r0 * r0);
                    const float phi = TWOPI * r1;
                    float sn, cs;
                    or_sincos(phi, &sn, &cs);
                    const float xs = sinTheta * cs, zs = sinTheta * sn;
                    float R[3];
                    for (int k = 0; k < 3; k++) R[k] = xs * Nb[k] + r0 * N[k] + zs * Nt[k];
                    const float NdotL = fmaxf(0.0f, dot3(N, R));
                    float d2 = 0;
                    uint32_t s2 = 0;
                    const uint32_t v2 = or_trace(w, O2, R, 128, OR_TRACE_ALL, 0, 0, &d2, &s2);
                    float sl[3];
                    or_to_float_rgb(v2, sl);
                    const float es2 = emit_strength(c, v2);
                    for (int k = 0; k < 3; k++) {
                        sl[k] = sl[k] * es2 * INVPI;
                        const float inc = (sl[k] * TWOPI * NdotL) * brdf[k] + contrib[k];
                        acc[k] += (double)inc;
                    }
                }
            }
            double* a = &accum[((size_t)y * W + x) * 3];
            a[0] += acc[0], a[1] += acc[1], a[2] += acc[2];
        }
}

/* ------------------------------------------------------------------ */
/* world build semantics                                                */
/* template/world.h:293-606, template/world.cpp:47-94, 228-246,         */
/* 438-563, 655-683, 726-745 (THREADSAFEWORLD=1, world.h:6)             */
/* ------------------------------------------------------------------ */
struct or_world {
    uint32_t cap;              /* BRICKCOUNT (common.h:186), power of two */
    uint32_t* grid;            /* [128^3] */
    uint32_t* bricks;          /* [cap * 512] */
    uint32_t* zeroes;          /* [cap] */
    uint32_t* trash;           /* [cap] circular list of free brick ids */
    uint32_t* modified;        /* [cap / 32] dirty bits */
    uint32_t trashHead, trashTail;
    uint32_t highWater;        /* 1 + largest brick id handed out */
    or_light* lights;
    uint32_t numLights, lightCap;
};

or_world* or_world_create(uint32_t cap)
{
    or_world* w = (or_world*)calloc(1, sizeof(or_world));
    w->cap = cap;
    w->grid = (uint32_t*)calloc(OR_GRIDSIZE, 4);
    w->bricks = (uint32_t*)calloc((size_t)cap * OR_BRICKSIZE, 4);
    w->zeroes = (uint32_t*)calloc(cap, 4);
    w->trash = (uint32_t*)calloc(cap, 4);
    w->modified = (uint32_t*)calloc(cap / 32 + 1, 4);
    or_world_clear(w);
    return w;
}

void or_world_destroy(or_world* w)
{
    if (!w) return;
    free(w->grid), free(w->bricks), free(w->zeroes), free(w->trash), free(w->modified), free(w->lights);
    free(w);
}

void or_world_clear(or_world* w)
{
    /* World::Clear world.cpp:726-745 */
    memset(w->grid, 0, (size_t)OR_GRIDSIZE * 4);
    for (uint32_t i = 0; i < w->cap; i++) w->trash[(i * 31u) & (w->cap - 1)] = i;
    w->trashHead = w->cap, w->trashTail = 0;
    if (w->highWater) memset(w->bricks, 0, (size_t)w->highWater * OR_BRICKSIZE * 4);
    memset(w->zeroes, 0, (size_t)w->cap * 4);
    memset(w->modified, 0, (size_t)(w->cap / 32 + 1) * 4);
    w->highWater = 0;
    w->numLights = 0;
}

static inline void mark(or_world* w, uint32_t idx) { w->modified[idx >> 5] |= 1u << (idx & 31); }

static inline uint32_t new_brick(or_world* w)
{
    /* NewBrick world.h:585-595: InterlockedAdd(&trashTail, 31) - 31 */
    const uint32_t item = w->trashTail;
    w->trashTail += 31;
    const uint32_t id = w->trash[item & (w->cap - 1)];
    if (id + 1 > w->highWater) w->highWater = id + 1;
    return id;
}

static inline void free_brick(or_world* w, uint32_t idx)
{
    /* FreeBrick world.h:596-606 */
    const uint32_t item = w->trashHead;
    w->trashHead += 31;
    w->trash[item & (w->cap - 1)] = idx;
}

uint32_t or_world_get(const or_world* w, uint32_t x, uint32_t y, uint32_t z)
{
    /* World::Get world.h:293-306 */
    const uint32_t bx = (x / 8) & 127, by = (y / 8) & 127, bz = (z / 8) & 127;
    const uint32_t cv = w->grid[bx + bz * 128 + by * 16384];
    if ((cv & 1) == 0) return cv >> 1;
    const uint32_t lx = x & 7, ly = y & 7, lz = z & 7;
    return w->bricks[(size_t)(cv >> 1) * 512 + lx + ly * 8 + lz * 64];
}

/* SetupLightBuffer(ls, numberOfLights) append, world.cpp:563-612 */
static void lights_append(or_world* w, uint32_t pos, uint32_t voxel, uint32_t size)
{
    if (w->numLights == w->lightCap) {
        w->lightCap = w->lightCap ? w->lightCap * 2 : 64;
        w->lights = (or_light*)realloc(w->lights, sizeof(or_light) * w->lightCap);
    }
    or_light* l = &w->lights[w->numLights++];
    l->position = pos, l->voxel = voxel, l->size = size;
}

/* World::AddLight world.cpp:228-246 */
static void add_light(or_world* w, uint32_t x, uint32_t y, uint32_t z, uint32_t size, uint32_t c)
{
    lights_append(w, x + z * 1024u + y * 1024u * 1024u, c, size);
}

/* World::RemoveLight world.cpp:506-561: keeps every light at another
 * position and sets the count to max(0, n - 1), whatever was removed */
static void remove_light(or_world* w, uint32_t x, uint32_t y, uint32_t z)
{
    if (w->numLights < 1) return;
    const uint32_t newCount = w->numLights - 1;
    const uint32_t target = x + z * 1024u + y * 1024u * 1024u;
    uint32_t k = 0;
    for (uint32_t i = 0; i < w->numLights; i++)
        if (w->lights[i].position != target) w->lights[k++] = w->lights[i];
    w->numLights = newCount;
}

void or_world_add_brick(or_world* w, uint32_t bx, uint32_t by, uint32_t bz, uint32_t v)
{
    /* World::AddBrick world.h:308-352 */
    if (bx >= 128 || by >= 128 || bz >= 128) return;
    const uint32_t ci = bx + bz * 128 + by * 16384;
    const uint32_t cv = w->grid[ci], off = cv >> 1;
    if ((cv & 1) != 0) {
        free_brick(w, off);
        for (uint32_t ly = 0; ly < 8; ly++)
            for (uint32_t lz = 0; lz < 8; lz++)
                for (uint32_t lx = 0; lx < 8; lx++) {
                    const uint32_t x = lx + bx * 8, y = ly + by * 8, z = lz + bz * 8;
                    if (is_emitter(or_world_get(w, x, y, z))) remove_light(w, x, y, z);
                }
    } else if (is_emitter(cv >> 1)) {
        remove_light(w, bx * 8, by * 8, bz * 8);
    }
    if (is_emitter(v)) add_light(w, bx * 8, by * 8, bz * 8, 8, v);
    mark(w, 0);
    w->grid[ci] = v << 1;
}

void or_world_remove_brick(or_world* w, uint32_t bx, uint32_t by, uint32_t bz)
{
    /* World::RemoveBrick world.h:354-390 */
    if (bx >= 128 || by >= 128 || bz >= 128) return;
    const uint32_t ci = bx + bz * 128 + by * 16384;
    const uint32_t cv = w->grid[ci], off = cv >> 1;
    w->grid[ci] = 0;
    if ((cv & 1) != 0) {
        w->zeroes[off] = 512;
        free_brick(w, off);
        /* the cell is already empty here, so Get() returns 0 for every voxel
         * and no light is removed (world.h:365-377 reads after clearing) */
        for (uint32_t ly = 0; ly < 8; ly++)
            for (uint32_t lz = 0; lz < 8; lz++)
                for (uint32_t lx = 0; lx < 8; lx++) {
                    const uint32_t x = lx + bx * 8, y = ly + by * 8, z = lz + bz * 8;
                    if (is_emitter(or_world_get(w, x, y, z))) remove_light(w, x, y, z);
                }
    } else if (is_emitter(cv >> 1)) {
        remove_light(w, bx * 8, by * 8, bz * 8);
    }
    mark(w, 0);
}

void or_world_set(or_world* w, uint32_t x, uint32_t y, uint32_t z, uint32_t v)
{
    /* World::Set world.h:456-513 */
    const uint32_t bx = x / 8, by = y / 8, bz = z / 8;
    if (bx >= 128 || by >= 128 || bz >= 128) return;
    const uint32_t ci = bx + bz * 128 + by * 16384;
    uint32_t cv = w->grid[ci];
    const uint32_t color = cv >> 1;
    uint32_t off = cv >> 1;
    if ((cv & 1) == 0) {
        if (v == color) return;
        /* SplitSolidBrick world.h:392-445 */
        off = new_brick(w);
        uint32_t* b = &w->bricks[(size_t)off * 512];
        for (int i = 0; i < 512; i++) b[i] = color;
        w->zeroes[off] = (color == 0) * 512u;
        w->grid[ci] = (off << 1) | 1u;
    }
    const uint32_t lx = x & 7, ly = y & 7, lz = z & 7;
    const size_t vi = (size_t)off * 512 + lx + ly * 8 + lz * 64;
    const uint32_t orig = w->bricks[vi];
    if (is_emitter(orig)) remove_light(w, x, y, z);
    if (is_emitter(v)) add_light(w, x, y, z, 1, v);
    const int change = (orig != 0 && v == 0) - (orig == 0 && v != 0);
    w->zeroes[off] += (uint32_t)change;
    if (w->zeroes[off] < 512) {
        w->bricks[vi] = v;
        mark(w, off);
    } else {
        w->grid[ci] = 0;
        free_brick(w, off);
    }
}

uint32_t or_world_optimize(or_world* w)
{
    /* World::OptimizeBricks world.cpp:655-683 */
    uint32_t replaced = 0;
    for (uint32_t i = 0; i < OR_GRIDSIZE; i++) {
        const uint32_t value = w->grid[i];
        if (!(value & 1)) continue;
        const uint32_t* b = &w->bricks[(size_t)(value >> 1) * 512];
        const uint32_t first = b[0];
        int solid = 1;
        for (int j = 1; j < 512; j++)
            if (b[j] != first) { solid = 0; break; }
        if (solid) {
            w->grid[i] = (first << 1) | 0u;
            free_brick(w, value >> 1);
            replaced++;
        }
    }
    return replaced;
}

uint32_t or_world_find_lights(const or_world* w, or_light* out, uint32_t cap)
{
    /* World::FindLightsInWord world.cpp:438-504 */
    uint32_t n = 0;
    for (uint32_t by = 0; by < 128; by++)
        for (uint32_t bz = 0; bz < 128; bz++)
            for (uint32_t bx = 0; bx < 128; bx++) {
                const uint32_t g = w->grid[bx + bz * 128 + by * 16384];
                if ((g & 1) == 0) {
                    const uint32_t vox = g >> 1;
                    if (is_emitter(vox)) {
                        if (n < cap) out[n].position = bx * 8 + bz * 8 * 1024 + by * 8 * 1024 * 1024, out[n].voxel = vox, out[n].size = 8;
                        n++;
                    }
                } else {
                    for (uint32_t ly = 0; ly < 8; ly++)
                        for (uint32_t lz = 0; lz < 8; lz++)
                            for (uint32_t lx = 0; lx < 8; lx++) {
                                const uint32_t vox = w->bricks[(size_t)(g >> 1) * 512 + lx + ly * 8 + lz * 64];
                                if (is_emitter(vox)) {
                                    const uint32_t x = lx + bx * 8, y = ly + by * 8, z = lz + bz * 8;
                                    if (n < cap) out[n].position = x + z * 1024 + y * 1024 * 1024, out[n].voxel = vox, out[n].size = 1;
                                    n++;
                                }
                            }
                }
            }
    return n;
}

uint32_t or_world_num_lights(const or_world* w) { return w->numLights; }
const or_light* or_world_lights(const or_world* w) { return w->lights; }
const uint32_t* or_world_grid(const or_world* w) { return w->grid; }
const uint32_t* or_world_bricks(const or_world* w) { return w->bricks; }
const uint32_t* or_world_zeroes(const or_world* w) { return w->zeroes; }
uint32_t or_world_brick_high_water(const or_world* w) { return w->highWater; }

uint32_t or_world_dirty(const or_world* w, uint32_t* ids, uint32_t cap)
{
    uint32_t n = 0;
    for (uint32_t j = 0; j < w->cap / 32; j++)
        if (w->modified[j])
            for (uint32_t k = 0; k < 32; k++)
                if (w->modified[j] & (1u << k)) {
                    if (n < cap) ids[n] = j * 32 + k;
                    n++;
                }
    return n;
}

/* updateUberGrid, cl/kernels.cl:498-513 */
void or_build_uber(const uint32_t* grid, uint8_t* uber)
{
    for (uint32_t y = 0; y < 32; y++)
        for (uint32_t z = 0; z < 32; z++)
            for (uint32_t x = 0; x < 32; x++) {
                int empty = 1;
                for (int a = 0; a < 4 && empty; a++)
                    for (int b = 0; b < 4 && empty; b++)
                        for (int c = 0; c < 4; c++) {
                            const uint32_t gx = x * 4 + a, gy = y * 4 + b, gz = z * 4 + c;
                            if (grid[gx + gz * 128 + gy * 16384]) { empty = 0; break; }
                        }
                uber[x + z * 32 + y * 1024] = empty ? 0 : 1;
            }
}

/* ------------------------------------------------------------------ */
/* scene motion: the product's vr_scene_step (moving emitters and       */
/* random-walk brick edits), restated sequentially.  Brick edits come   */
/* first; light moves see the edited bricks.                             */
/* ------------------------------------------------------------------ */
static uint32_t or_to_voxel(float f)
{
    const float fl = floorf(f);
    return !(fl > 0.0f) ? 0u : (fl >= 1023.0f ? 1023u : (uint32_t)fl);
}

static int or_voxel_slot(const uint32_t* grid, uint32_t cap, uint32_t x, uint32_t y, uint32_t z, size_t* slot)
{
    const uint32_t g = grid[(x >> 3) + (z >> 3) * 128u + (y >> 3) * 16384u];
    if (!(g & 1u) || (g >> 1) >= cap) return 0;
    *slot = (size_t)(g >> 1) * 512u + (x & 7u) + (y & 7u) * 8u + (z & 7u) * 64u;
    return 1;
}

static void or_bounce(float* pos, float* vel, uint32_t i, uint32_t cx, uint32_t cy, uint32_t cz)
{
    vel[4 * i] = -vel[4 * i], vel[4 * i + 1] = -vel[4 * i + 1], vel[4 * i + 2] = -vel[4 * i + 2];
    pos[4 * i] = (float)cx + 0.5f, pos[4 * i + 1] = (float)cy + 0.5f, pos[4 * i + 2] = (float)cz + 0.5f;
}

void or_scene_step(const uint32_t* grid, uint32_t* bricks, uint32_t cap, or_light* lights, float* pos, float* vel,
                   uint32_t n, uint32_t frame, uint32_t seed, uint32_t permille, uint32_t steps,
                   uint32_t high_water)
{
    /* edits: brick id < high_water with hash % 1000 < permille; a walker of
     * `steps` xorshift steps toggling non-emissive voxels */
    for (uint32_t id = 0; steps > 0 && id < high_water && id < cap; id++) {
        const uint32_t h = or_wang_hash((id * 0x9e3779b1u) ^ seed ^ (frame * 0x85ebca77u));
        if (h % 1000u >= permille) continue;
        uint32_t st = or_wang_hash(h);
        uint32_t c[3] = {st & 7u, (st >> 3) & 7u, (st >> 6) & 7u};
        const uint32_t colour = 0xF000u | ((st >> 9) & 0xFFFu);
        uint32_t* b = bricks + (size_t)id * 512u;
        for (uint32_t k = 0; k < steps; k++) {
            const uint32_t r = or_random_int(&st);
            const uint32_t step = ((r >> 8) & 1u) ? 1u : 7u;
            c[r % 3u] = (c[r % 3u] + step) & 7u;
            uint32_t* v = b + c[0] + c[1] * 8u + c[2] * 64u;
            if ((*v >> 16) & 255u) continue;
            *v = *v ? 0u : colour;
        }
    }
    /* lights: targets and claims (lowest index per bucket), then moves */
    uint32_t* target = (uint32_t*)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    uint32_t* claim = (uint32_t*)malloc((size_t)65536 * sizeof(uint32_t));
    if (!target || !claim) {
        free(target), free(claim);
        return;
    }
    for (uint32_t b = 0; b < 65536u; b++) claim[b] = 0xffffffffu;
    for (uint32_t i = 0; i < n; i++) {
        target[i] = 0xffffffffu;
        if (lights[i].size != 1u) continue;
        const uint32_t cur = lights[i].position;
        const uint32_t cx = cur & 1023u, cz = (cur >> 10) & 1023u, cy = cur >> 20;
        const float nx = pos[4 * i] + vel[4 * i], ny = pos[4 * i + 1] + vel[4 * i + 1];
        const float nz = pos[4 * i + 2] + vel[4 * i + 2];
        const uint32_t tx = or_to_voxel(nx), ty = or_to_voxel(ny), tz = or_to_voxel(nz);
        if (tx == cx && ty == cy && tz == cz) {
            pos[4 * i] = nx, pos[4 * i + 1] = ny, pos[4 * i + 2] = nz;
            continue;
        }
        size_t slot = 0;
        if (or_voxel_slot(grid, cap, tx, ty, tz, &slot) && bricks[slot] == 0u) {
            const uint32_t t = tx | (tz << 10) | (ty << 20);
            const uint32_t bk = or_wang_hash(t) & 65535u;
            target[i] = t;
            if (i < claim[bk]) claim[bk] = i;
        } else {
            or_bounce(pos, vel, i, cx, cy, cz);
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t t = target[i];
        if (t == 0xffffffffu) continue;
        const uint32_t cur = lights[i].position;
        const uint32_t cx = cur & 1023u, cz = (cur >> 10) & 1023u, cy = cur >> 20;
        if (claim[or_wang_hash(t) & 65535u] != i) {
            or_bounce(pos, vel, i, cx, cy, cz);
            continue;
        }
        const uint32_t tx = t & 1023u, tz = (t >> 10) & 1023u, ty = t >> 20;
        size_t st = 0, sc = 0;
        (void)or_voxel_slot(grid, cap, tx, ty, tz, &st);
        if (or_voxel_slot(grid, cap, cx, cy, cz, &sc)) bricks[sc] = 0u;
        bricks[st] = lights[i].voxel;
        pos[4 * i] += vel[4 * i], pos[4 * i + 1] += vel[4 * i + 1], pos[4 * i + 2] += vel[4 * i + 2];
        lights[i].position = t;
    }
    free(target);
    free(claim);
}

