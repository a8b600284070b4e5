/*
 * vr_oracle.h - CPU restatement of the reference Voxel ReSTIR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product (smallproject_amd/) never links, loads or
 * calls it, and the product fails loudly when its own HIP library is missing.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Floating-point semantics: IEEE-754 binary32
 * with round-to-nearest, correctly rounded + - * / sqrt, no FMA contraction
 * (built with -ffp-contract=off), float->uint conversions truncate toward zero
 * and saturate (NaN -> 0), float->int conversions truncate and saturate.  The
 * reference itself compiles with -cl-fast-relaxed-math, whose float bits are
 * implementation-defined; this restatement pins one exact definition.
 *
 * Each function is pinned by the tests named in its comment or in
 * tests/test_oracle_*.py (known answers, closed forms, brute force).
 */
#ifndef VR_ORACLE_H
#define VR_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* template/common.h:16-21, 176-189 */
#define OR_MAPDIM     1024u
#define OR_BRICKDIM   8u
#define OR_GRIDDIM    128u
#define OR_UBERDIM    32u
#define OR_GRIDSIZE   (128u * 128u * 128u)
#define OR_UBERSIZE   (32u * 32u * 32u)
#define OR_BRICKSIZE  512u
#define OR_MAXCOMMITS 8192u

/* trace kinds: TraceRay / TraceOpaqueRay / TraceThrough (cl/trace.cl:113-119) */
#define OR_TRACE_ALL     0
#define OR_TRACE_OPAQUE  1
#define OR_TRACE_THROUGH 2

/* parity quirks (reference behaviour reproduced when the bit is set) */
#define OR_QUIRK_DOUBLE_RADIANCE 1u /* cl/ris.cl:254-255 color += alphaF*color          */
#define OR_QUIRK_AXIS_ONLY_FACE  2u /* cl/ris.cl:93, cl/spatialSampling.cl:46 side==side */
#define OR_QUIRK_SHADOW_STEP_CAP 4u /* cl/ris.cl:10 TraceOpaqueRay(..., GRIDWIDTH)       */
#define OR_QUIRKS_REFERENCE      7u

/* source pdf for candidate generation */
#define OR_SOURCE_UNIFORM 0u  /* cl/tools.cl:593-598, p = 1/N */
#define OR_SOURCE_POWER   1u  /* alias table over lightContribution, cl/tools.cl:579-591 */
#define OR_SOURCE_SLOTS   2u  /* per-pixel candidate lists given by the caller (learned proposal) */

typedef struct or_f3 { float x, y, z; } or_f3;

/* struct CLRay, template/common.h:125-132 (float3 stored unpadded) */
typedef struct or_ray {
    uint32_t voxel;
    float dist;
    uint32_t side;
    uint32_t seed;
    or_f3 D;
} or_ray;

/* struct Light, template/common.h:134-139 */
typedef struct or_light {
    uint32_t position;
    uint32_t voxel;
    uint32_t size;
} or_light;

/* struct Reservoir, template/common.h:141-151 */
typedef struct or_reservoir {
    float wsum;        /* sumOfWeights */
    uint32_t M;        /* streamLength */
    uint32_t lightIndex;
    float W;           /* adjustedWeight */
    or_f3 pos;         /* positionOnVoxel */
    float invPdf;      /* invPositionProbability */
    or_f3 Nlight;
    float pHat;
} or_reservoir;

/* struct RenderParams, template/common.h:100-118: the fields the ReSTIR path reads.
 * Field order and types match vr_render_params in include/vr_api.h. */
typedef struct or_params {
    float oneOverRes[2];
    float E[3], p0[3], p1[3], p2[3];
    uint32_t R0, frame, framecount, restirtemporalframe;
    float Nleft[4], Nright[4], Ntop[4], Nbottom[4];
    int32_t accumulate, spatial, temporal, skyDomeSampling;
    uint32_t numberOfLights, numberOfCandidates, numberOfMaxTemporalImportance;
    uint32_t spatialRadius, spatialTaps;
} or_params;

/* compile-time switches of the reference turned into run-time fields */
typedef struct or_config {
    uint32_t width, height;      /* SCRWIDTH, SCRHEIGHT (common.h:2-3) */
    uint32_t point_lights;       /* VOXELSAREPOINTLIGHTS (common.h:70) */
    uint32_t gfxexp;             /* GFXEXP (common.h:68) */
    uint32_t quirks;             /* OR_QUIRK_* */
    uint32_t source_pdf;         /* OR_SOURCE_* */
    float emit_lut[256];         /* EmitStrength(v) = emit_lut[(v >> 16) & 255] */
    /* OR_SOURCE_SLOTS (learned proposal): candidate i of pixel p is a draw
     * from cluster c = cand_idx[i*W*H + p] with cluster mass m =
     * cand_pdf[i*W*H + p].  One random number u picks the light inside the
     * cluster with its alias table (bucket k = (int)(u*cnt), kept when the
     * fraction u*cnt - k < cl_prob[c0+k], else cl_alias[c0+k]); the source
     * pdf of light i is m * cl_share[i].  An empty cluster or m <= 0 is an
     * empty slot (M += 1, u consumed, no update). */
    const uint32_t* cand_idx;
    const float* cand_pdf;
    const uint32_t* cl_start;   /* [257] offsets into cl_lights */
    const uint32_t* cl_lights;  /* light ids grouped by cluster */
    const float* cl_prob;       /* alias tables aligned with cl_lights */
    const uint32_t* cl_alias;   /* cluster-local indices */
    const float* cl_share;      /* by light id */
    /* spatial merge with generalized balance-heuristic MIS weights and one
     * visibility ray per combined neighbour (unbiased); at most
     * OR_MIS_MAX_TAPS combined neighbours enter the weights */
    uint32_t spatial_mis;
} or_config;
#define OR_MIS_MAX_TAPS 32

/* read-only view of a device world: top grid, brick pool, uber grid */
typedef struct or_world_view {
    const uint32_t* grid;    /* [128^3], index bx + bz*128 + by*128^2 (world.h:299) */
    const uint32_t* bricks;  /* pool, brick id * 512 + lx + ly*8 + lz*64 (world.h:305) */
    const uint8_t* uber;     /* [32^3], index ux + uz*32 + uy*1024 (kernels.cl:498-513) */
} or_world_view;

/* alias table for the power-proportional source (built by or_build_power_table) */
typedef struct or_alias {
    const float* prob;     /* [N] acceptance probability of the bucket */
    const uint32_t* alias; /* [N] alias index */
    const float* pmf;      /* [N] probability of selecting light i */
} or_alias;

/* ---- RNG, cl/tools.cl:17-19 and template/template.cpp:933-951 ---- */
uint32_t or_wang_hash(uint32_t s);
uint32_t or_random_int(uint32_t* s);
float or_random_float(uint32_t* s);

/* ---- voxel codec, cl/tools.cl:294-322 and common.h:203-211 ---- */
uint32_t or_get_alpha(uint32_t v);
float or_get_alphaf(uint32_t v);
void or_to_float_rgb(uint32_t v, float out[3]);
uint32_t or_from_float_rgba(const float rgba[4], uint32_t emit);

/* ---- deterministic sine/cosine used by the spatial taps ---- */
void or_sincos(float angle, float* s, float* c);

/* ---- camera ray, cl/tools.cl:102-127 ---- */
void or_camera_ray(const or_params* p, int x, int y, float D[3]);

/* ---- brick-grid DDA, cl/trace.cl:121-362 ---- */
uint32_t or_trace(const or_world_view* w, const float A[3], const float B[3], int steps,
                  int kind, uint32_t through, uint32_t mask, float* dist, uint32_t* side);
void or_trace_batch(const or_world_view* w, const float* O, const float* D, uint32_t n,
                    int steps, int kind, uint32_t* voxel, float* dist, uint32_t* side);

/* ---- light sampling helpers, cl/tools.cl:499-621, cl/phat.cl ---- */
void or_random_point_on_voxel(const float center[3], const float sp[3], int size,
                              uint32_t* seed, float* invPdf, float Nl[3], float pos[3]);
float or_eval_phat(const or_config* c, const float pos[3], const float Nl[3],
                   const float sp[3], const float N[3], const float brdf[3],
                   uint32_t lvoxel, uint32_t lsize);
void or_eval_phat_full(const or_config* c, const float pos[3], const float Nl[3],
                       const float sp[3], const float N[3], const float brdf[3],
                       uint32_t lvoxel, uint32_t lsize, float out[3]);
void or_reservoir_update(or_reservoir* r, float w, float pHat, uint32_t index, float r0,
                         const float pos[3], float invPdf, const float Nl[3]);
void or_reservoir_combine(or_reservoir* r, const or_reservoir* that, float r0);
void or_reservoir_reweigh(or_reservoir* r, float newPHat, uint32_t M);
void or_reservoir_adjust(or_reservoir* r, float pHat);
void or_reproject(const or_params* p, uint32_t W, uint32_t H, const float sp[3], int out[2]);
int or_build_power_table(const or_config* c, const or_light* lights, uint32_t n,
                         float* prob, uint32_t* alias, float* pmf);

/* ---- per-pixel stages, cl/ris.cl ---- */
void or_stage_albedo(const or_config* c, const or_params* p, const or_world_view* w,
                     or_ray* hits, int y0, int y1);
void or_stage_initial(const or_config* c, const or_params* p, const or_world_view* w,
                      const or_light* lights, const or_alias* src, const or_ray* prevHits,
                      or_ray* hits, or_reservoir* out, const or_reservoir* prevRes,
                      int y0, int y1);
void or_stage_spatial(const or_config* c, const or_params* p, const or_world_view* w, const or_light* lights,
                      or_ray* hits, or_reservoir* out, const or_reservoir* in, int y0, int y1);
void or_stage_shade(const or_config* c, const or_params* p, const or_world_view* w,
                    const or_light* lights, const or_ray* hits, float* frame,
                    or_reservoir* out, or_reservoir* in, int y0, int y1);

/* ---- whole-frame driver, world.cpp:2127-2340 buffer rotation ---- */
typedef struct or_state {
    uint32_t W, H;
    or_ray* hits[2];
    or_reservoir* res[2];
    float* frame;           /* W*H*4: rgb + dist (ris.cl:292,411) */
    int hitCur, hitPrev;    /* world.cpp:2221-2223 statics */
    int resOut, resIn;      /* world.cpp:2209-2219 statics */
    int initialOut, spatialOut, shadeIn;  /* buffer indices used by the last frame */
    or_reservoir* initial;  /* copy of the last frame's candidates+temporal output
                               (the rotation overwrites that buffer with the
                               shading stage's copy, ris.cl:412) */
} or_state;

or_state* or_state_create(uint32_t W, uint32_t H);
void or_state_destroy(or_state* s);
void or_render_frame(or_state* s, const or_config* c, const or_params* p,
                     const or_world_view* w, const or_light* lights, const or_alias* src,
                     int nthreads);

/* ---- ground truth, cl/kernels.cl:177-241 (render_di_gfxexp) in FP64 accumulation ---- */
void or_ground_truth(const or_config* c, const or_params* p, const or_world_view* w,
                     uint32_t spp, uint32_t frame0, double* accum, int y0, int y1);

/* ---- scene motion (the product's vr_scene_step; no reference counterpart):
 * brick edits, then light moves, on the oracle's copy of the world.
 * pos/vel hold 4 floats per light (xyz used); lights of size 1 move. ---- */
void or_scene_step(const uint32_t* grid, uint32_t* bricks, uint32_t cap, or_light* lights, float* pos, float* vel,
                   uint32_t n, uint32_t frame, uint32_t seed, uint32_t permille, uint32_t steps,
                   uint32_t high_water);

/* ---- world build semantics, template/world.h:293-606, world.cpp:47-683 ---- */
typedef struct or_world or_world;
or_world* or_world_create(uint32_t brick_capacity);
void or_world_destroy(or_world* w);
void or_world_clear(or_world* w);
uint32_t or_world_get(const or_world* w, uint32_t x, uint32_t y, uint32_t z);
void or_world_set(or_world* w, uint32_t x, uint32_t y, uint32_t z, uint32_t v);
void or_world_add_brick(or_world* w, uint32_t bx, uint32_t by, uint32_t bz, uint32_t v);
void or_world_remove_brick(or_world* w, uint32_t bx, uint32_t by, uint32_t bz);
uint32_t or_world_optimize(or_world* w);
uint32_t or_world_find_lights(const or_world* w, or_light* out, uint32_t cap);
uint32_t or_world_num_lights(const or_world* w);
const or_light* or_world_lights(const or_world* w);
const uint32_t* or_world_grid(const or_world* w);
const uint32_t* or_world_bricks(const or_world* w);
const uint32_t* or_world_zeroes(const or_world* w);
uint32_t or_world_brick_high_water(const or_world* w);
uint32_t or_world_dirty(const or_world* w, uint32_t* ids, uint32_t cap);
void or_build_uber(const uint32_t* grid, uint8_t* uber);

#ifdef __cplusplus
}
#endif
#endif
