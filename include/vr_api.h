/*
 * vr_api.h - C ABI of the MI355X Voxel ReSTIR render path.
 *
 * This library replaces what sits below World::Render()/World::Commit() in
 * the reference: the OpenCL Kernel/Buffer runtime (template/template.cpp:
 * 1266-1754) and the ReSTIR kernels of cl/ris.cl, cl/trace.cl, cl/tools.cl,
 * cl/phat.cl, cl/reservoir.cl, cl/spatialSampling.cl and the commit/ubergrid
 * kernels of cl/kernels.cl.  Above it, the C++ World class (world.h beside
 * this header) keeps the reference's host API and calls only these functions.
 *
 * Conventions: every function returns VR_OK (0) or a VR_ERR_* code and never
 * throws; vr_last_error() describes the last failure of a context.  Pointers
 * are host pointers unless a name says otherwise.  A context is not
 * thread-safe: one host thread drives it (as the reference's single render
 * thread does).  Rendering is asynchronous on the context's HIP stream (two
 * while it is the only context: see vr_stage_times); read-backs and vr_sync()
 * block.
 */
#ifndef VR_API_H
#define VR_API_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* world dimensions, template/common.h:16-21, 176-193 */
#define VR_MAPDIM     1024u
#define VR_BRICKDIM   8u
#define VR_GRIDDIM    128u
#define VR_UBERDIM    32u
#define VR_GRIDSIZE   (128u * 128u * 128u)
#define VR_BRICKSIZE  512u
#define VR_BRICKCOUNT (1u << 20)
#define VR_MAXCOMMITS 8192u

/* status codes */
#define VR_OK          0
#define VR_ERR_ARG     1   /* bad argument (null pointer, size mismatch, range) */
#define VR_ERR_HIP     2   /* a HIP runtime call failed */
#define VR_ERR_STATE   3   /* call not valid in the current state */
#define VR_ERR_NOMEM   4   /* device or host allocation failed */
#define VR_ERR_UNSUPPORTED 5

/* light modes: VOXELSAREPOINTLIGHTS (template/common.h:70) */
#define VR_LIGHTS_AREA  0u
#define VR_LIGHTS_POINT 1u

/* candidate source pdf */
#define VR_SOURCE_UNIFORM 0u  /* cl/tools.cl:593-598: p = 1/N (reference) */
#define VR_SOURCE_POWER   1u  /* alias table over lightContribution (cl/tools.cl:579-591) */
#define VR_SOURCE_LEARNED 2u  /* MLP proposal over light clusters (no reference counterpart) */

/* reference quirks reproduced when set (VR_QUIRKS_REFERENCE = all three) */
#define VR_QUIRK_DOUBLE_RADIANCE 1u /* cl/ris.cl:254-255                      */
#define VR_QUIRK_AXIS_ONLY_FACE  2u /* cl/ris.cl:93, cl/spatialSampling.cl:46 */
#define VR_QUIRK_SHADOW_STEP_CAP 4u /* cl/ris.cl:10 (128-step shadow rays)    */
#define VR_QUIRKS_REFERENCE      7u

/* trace kinds for vr_trace_rays: TraceRay / TraceOpaqueRay (cl/trace.cl) */
#define VR_TRACE_ALL    0u
#define VR_TRACE_OPAQUE 1u

/* stages, in pipeline order (used by vr_stage_times) */
#define VR_STAGE_PRIMARY 0  /* renderAlbedo             cl/ris.cl:295-389 */
#define VR_STAGE_INITIAL 1  /* perPixelInitialSampling  cl/ris.cl:20-124  */
#define VR_STAGE_RESOLVE 2  /* perPixelSpatialResampling cl/ris.cl:126-151, then
                               renderRIS cl/ris.cl:391-413, in one kernel */
#define VR_NUM_STAGES    3

/* order in which workgroups take the 16x16 pixel tiles, per stage
 * (vr_config.flags, or vr_set_tile_order); results do not depend on it */
#define VR_TILES_LINEAR   0u  /* row-major */
#define VR_TILES_XCD_ROWS 1u  /* each of the 8 XCDs takes one band of tile rows */
#define VR_TILES_XCD_COLS 2u  /* each XCD takes one band of tile columns, row by row */
#define VR_TILES_XCD_BLOCKS 3u /* each XCD takes whole 4x4-tile blocks, dealt round-robin */

/* struct Light, template/common.h:134-139 */
typedef struct vr_light {
    uint32_t position; /* x + z*1024 + y*1024*1024 of the light's min corner */
    uint32_t voxel;    /* voxel word 0x00MMARGB */
    uint32_t size;     /* edge length in voxels (1 voxel light, 8 brick light) */
} vr_light;

/* struct RenderParams, template/common.h:100-118: the fields the ReSTIR
 * path reads, same names and meaning.  float3 fields are packed (no
 * OpenCL 16-byte padding).  Produced per frame by World::Render(). */
typedef struct vr_render_params {
    float oneOverRes[2];
    float E[3], p0[3], p1[3], p2[3];
    uint32_t R0, frame, framecount, restirtemporalframe;
    float Nleft[4], Nright[4], Ntop[4], Nbottom[4];
    int32_t accumulate, spatial, temporal, skyDomeSampling;
    uint32_t numberOfLights, numberOfCandidates, numberOfMaxTemporalImportance;
    uint32_t spatialRadius, spatialTaps;
} vr_render_params;

/* compile-time switches of the reference (template/common.h) as a
 * per-context configuration */
/* vr_config.flags: spatial reuse merges with generalized balance-heuristic
 * MIS weights instead of the reference's 1/M normalisation (unbiased; no
 * counterpart in the reference).  The merged sample and the random stream
 * are the reference's; only W changes, to
 *     W = wsum * T_sel(y) / (pHat * sum_s M_s T_s(y))
 * over the own reservoir and every combined neighbour s, T_s(y) being the
 * target of s's pixel times its visibility of y (one shadow ray per
 * combined neighbour; the own term has no ray).  At most VR_MIS_MAX_TAPS
 * spatial taps.  The first frame with the merge allocates the visibility
 * rays' list and the taps' terms (16 bytes per pixel and tap, plus 20 bytes
 * per pixel), as does a later frame with more taps. */
#define VR_FLAG_MIS_SPATIAL (1u << 8)
#define VR_MIS_MAX_TAPS 32

typedef struct vr_config {
    uint32_t width, height;   /* SCRWIDTH, SCRHEIGHT */
    uint32_t light_mode;      /* VR_LIGHTS_* */
    uint32_t gfxexp;          /* GFXEXP: emission scaled by 1/pi, emitters also lit */
    uint32_t quirks;          /* VR_QUIRK_* */
    uint32_t source_pdf;      /* VR_SOURCE_* */
    uint32_t brick_capacity;  /* device brick pool size in bricks (reference: VR_BRICKCOUNT) */
    uint32_t flags;           /* bits 2s..2s+1: tile order VR_TILES_* of stage s
                                 (s < VR_NUM_STAGES); VR_FLAG_MIS_SPATIAL; other
                                 bits zero */
    float emit_lut[256];      /* EmitStrength(v) = emit_lut[(v >> 16) & 255]; the
                                 reference's is emit_lut[m] = m (common.h:208-211) */
} vr_config;

typedef struct vr_ctx vr_ctx;

/* Fill *cfg with the reference defaults for a width x height view:
 * area lights, GFXEXP=1, all quirks, uniform source, 2^20 bricks,
 * identity emission LUT. */
void vr_default_config(vr_config* cfg, uint32_t width, uint32_t height);

/* Create a render context on the current HIP device: allocates the top grid
 * (8 MiB), uber grid, brick pool (brick_capacity * 2 KiB), two hit buffers
 * (12 B/px each), two reservoir buffers (36 B/px each) and the frame (16 B/px).
 * Replaces World::World's buffer set-up (template/world.cpp:47-193). */
int vr_create(const vr_config* cfg, vr_ctx** out);
int vr_destroy(vr_ctx* ctx);
const char* vr_last_error(const vr_ctx* ctx);

/* World::Commit device side (template/world.cpp:2357-2420, cl/kernels.cl:
 * 448-513): copy the full top grid (VR_GRIDSIZE u32, index bx+bz*128+by*16384),
 * scatter n changed bricks (ids[n], payload[n*512]) into the pool, rebuild
 * the uber grid.  n may exceed VR_MAXCOMMITS (uploaded in several batches). */
int vr_commit(vr_ctx* ctx, const uint32_t* grid, const uint32_t* brick_ids,
              const uint32_t* brick_payload, uint32_t n);

/* Upload the light list (World::SetupLightBuffer, template/world.cpp:563-612). */
int vr_set_lights(vr_ctx* ctx, const vr_light* lights, uint32_t n);

/* World::Render device side (template/world.cpp:2127-2340): swap the hit
 * buffers, then primary hit -> candidates + visibility + temporal reuse ->
 * spatial reuse -> shading.  Asynchronous. */
int vr_render_frame(vr_ctx* ctx, const vr_render_params* params);

/* ---- scene motion on the device (no counterpart in the reference, whose
 * worlds change through host edits and World::Commit) ----
 * These change the device copy of the world only: brick payloads and light
 * records.  The host World keeps its own copy (vr_read_lights and
 * vr_read_bricks return the device's).  Grid cells never change kind, so the
 * top grid and uber grid stay as committed. */
/* velocity[3n] in voxels per frame for the n lights of the current list
 * (|component| <= 64); only lights of size 1 move.  Each light's sub-voxel
 * position restarts at its voxel centre.  vr_set_lights drops the motion. */
int vr_set_light_motion(vr_ctx* ctx, const float* velocity, uint32_t n);
/* One frame of motion, queued on the context's stream:
 *  - brick edits: brick id < high_water is edited when
 *    wang_hash(id*0x9e3779b1 ^ seed ^ frame*0x85ebca77) % 1000 < edit_permille;
 *    a walker from a hashed voxel takes walk_steps one-voxel xorshift steps
 *    inside the brick, toggling each voxel it lands on between empty and an
 *    opaque hashed colour (emissive voxels are never touched);
 *  - lights: position += velocity.  A light whose voxel changes moves when the
 *    new voxel is empty, lies in a brick cell, and no lower-indexed light
 *    targets the same claim bucket (hash of the voxel, 65536 buckets): its
 *    old voxel is cleared and its voxel word written at the new one.  Any
 *    other change of voxel bounces it (velocity reversed, position reset to
 *    its voxel centre).  Deterministic: it does not depend on thread timing. */
int vr_scene_step(vr_ctx* ctx, uint32_t frame, uint32_t seed, uint32_t edit_permille, uint32_t walk_steps,
                  uint32_t high_water);
/* the device's light list (positions after motion), n = light count */
int vr_read_lights(vr_ctx* ctx, vr_light* out, uint32_t n);
/* payloads of n bricks (out[n*512]) from the device pool */
int vr_read_bricks(vr_ctx* ctx, const uint32_t* ids, uint32_t n, uint32_t* out);

/* ---- batches of views (no counterpart in the reference) ----
 * A batch renders the frames of up to VR_MAX_VIEWS contexts (views of
 * different resolutions and worlds on one device) with one launch per stage
 * per chunk of up to 8 views, over all their tiles.  Each view keeps its
 * own world, lights, buffers and
 * configuration; its frames are the ones vr_render_frame would render.  The
 * batch runs on its own stream, after the work queued so far on each view's
 * stream, and each view's later work (read-backs, commits) waits for the
 * batch frame.  Destroy a batch before its views. */
#define VR_MAX_VIEWS 64
typedef struct vr_batch vr_batch;
int vr_batch_create(vr_ctx* const* views, uint32_t n, vr_batch** out);
int vr_batch_destroy(vr_batch* b);
const char* vr_batch_last_error(const vr_batch* b);
/* workgroups per stage launch: 0 = one per 16x16 tile, else a persistent
 * grid of `blocks` workgroups striding over the tiles */
int vr_batch_set_blocks(vr_batch* b, uint32_t blocks);
/* views per chunk (1-8, default 8).  Each chunk of views is one chain of
 * launches (its learned views' proposals, then the four stages, one launch
 * each); chains are independent and may run concurrently. */
int vr_batch_set_chunk_views(vr_batch* b, uint32_t views_per_chunk);
/* replay each frame's launches from a hipGraph (one branch per chain),
 * built once per in-flight frame slot and updated with each frame's
 * arguments; rebuilt when the launch shape changes */
int vr_batch_set_graph(vr_batch* b, int enable);
/* params[n]: one frame of every view (each view's own render params) */
int vr_batch_render(vr_batch* b, const vr_render_params* params);
int vr_batch_sync(vr_batch* b);
/* GPU time of whole batch frames since vr_batch_profile_begin */
int vr_batch_profile_begin(vr_batch* b);
int vr_batch_profile_read(vr_batch* b, double* ms_sum, uint32_t* frames);

/* Block until all work of the context has finished. */
int vr_sync(vr_ctx* ctx);

/* Tile order VR_TILES_* of stage VR_STAGE_* (-1: every stage) for the
 * following frames (a performance setting with no counterpart in the
 * reference; outputs are identical for every order). */
int vr_set_tile_order(vr_ctx* ctx, int32_t stage, uint32_t order);

/* Occupancy target (waves per SIMD, 5-8; 0 = the stage's default) of stage
 * VR_STAGE_* (-1: every stage) for the following frames: a performance
 * setting with no counterpart in the reference (outputs are identical).
 * More waves hide more of the DDA's load latency; fewer spill less. */
int vr_set_stage_waves(vr_ctx* ctx, int32_t stage, uint32_t waves);

/* Traversal of the shadow rays of stage VR_STAGE_INITIAL or
 * VR_STAGE_RESOLVE (-1: both; the primary stage's rays always take the
 * nested DDA) for the following frames: VR_DDA_SCHEDULED (the default) runs
 * each turn of a wave's loop at the DDA level most of its lanes are at,
 * VR_DDA_NESTED walks the uber, top-grid and brick levels as nested loops,
 * which is faster where a wave's rays stay coherent.  A performance setting
 * with no counterpart in the reference (outputs are identical). */
#define VR_DDA_SCHEDULED 0u
#define VR_DDA_NESTED    1u
int vr_set_stage_dda(vr_ctx* ctx, int32_t stage, uint32_t mode);

/* Read-backs (blocking), n = width*height.
 * hits (current frame): hitword = voxel | side<<24 | (D[side] > 0)<<26,
 * dist, seed.  Any output pointer may be NULL. */
int vr_read_hits(vr_ctx* ctx, uint32_t* hitword, float* dist, uint32_t* seed);
/* which = 0: the final reservoirs of the last frame (after spatial reuse and
 * shading's discard), 1: the output of candidate sampling + temporal reuse.
 * a = (wsum, W, pHat, invPdf) per pixel, b = (pos.xyz, lightIndex bits),
 * m = M | Nlight face << 24 (face 0..5 = +x,+y,+z,-x,-y,-z; 7 = none). */
int vr_read_reservoirs(vr_ctx* ctx, int which, float* a4, float* b4, uint32_t* m);
/* frame: rgb radiance and hit distance per pixel (cl/ris.cl:292, 411) */
int vr_read_frame(vr_ctx* ctx, float* rgba);

/* Per-stage device time of the last rendered frame, milliseconds, measured
 * with HIP events around each stage (template/world.cpp:2439-2471).  When
 * the context is the only one in the process and nothing changed the world
 * since the previous frame's candidate pass, the primary pass runs on a
 * second stream beside that frame's last stage; its time then spans its run
 * next to that stage. */
int vr_stage_times(vr_ctx* ctx, float ms[VR_NUM_STAGES]);

/* Rays traced by the following frames (a diagnostic with no counterpart in
 * the reference, which traces 3 per pixel: primary, candidate and shading
 * shadow rays).  vr_set_ray_counting(ctx, 1) waits for the queued frames,
 * zeroes the counters and counts from the next frame on, with one atomic
 * add per traced ray; 0 stops.
 * vr_read_ray_counts waits for the queued frames and returns the sums:
 * [0] primary rays (with renderAlbedo's translucency continuations),
 * [1] the candidate pass's shadow rays, [2] shading's (those the candidate
 * pass had not traced), [3] the MIS merge's neighbour visibility rays. */
#define VR_RAY_KINDS 4
int vr_set_ray_counting(vr_ctx* ctx, int enable);
int vr_read_ray_counts(vr_ctx* ctx, uint64_t counts[VR_RAY_KINDS]);

/* Stage profiling over many frames (the reference keeps 100-frame averages,
 * template/world.cpp:2342-2353): vr_profile_begin starts summing every
 * frame's per-stage event times; vr_profile_read waits for the profiled
 * frames and returns the sums (ms) and the number of frames. */
int vr_profile_begin(vr_ctx* ctx);
int vr_profile_read(vr_ctx* ctx, double sum_ms[VR_NUM_STAGES], uint32_t* frames);

/* Trace n rays through the committed world (World::TraceBatch,
 * template/world.cpp:2051-2068, kernel traceBatch cl/kernels.cl:395-417).
 * origins/dirs: n*3 floats; kind VR_TRACE_*; steps: DDA step budget.
 * Outputs voxel, dist, side per ray (voxel 0 = miss, dist/side then 0). */
int vr_trace_rays(vr_ctx* ctx, const float* origins, const float* dirs, uint32_t n,
                  uint32_t kind, int32_t steps, uint32_t* voxel, float* dist, uint32_t* side);

/* vr_trace_rays without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The rays and hits are staged in a fixed device scratch of
 * VR_NOALLOC_MAX_RAYS rays (9 MiB per device, held until the process ends);
 * larger n is rejected with VR_ERR_ARG.  Same arguments, checks and results
 * as vr_trace_rays.  Every context on a device shares that device's
 * scratch: calls that overlap on one device (from several threads, or on
 * several contexts' streams) race and must not be made. */
#define VR_NOALLOC_MAX_RAYS (1u << 18)
int vr_trace_rays_noalloc(vr_ctx* ctx, const float* origins, const float* dirs, uint32_t n,
                          uint32_t kind, int32_t steps, uint32_t* voxel, float* dist, uint32_t* side);

/* ---- learned proposal (VR_SOURCE_LEARNED; no counterpart in the reference) ----
 * A 9-64-64-256 MLP maps a shading point (position / 1024, one-hot voxel
 * face) to a softmax q over light clusters.  A candidate is a cluster drawn
 * from the defensive mixture m = (1 - eps) q + eps P (P_c = the cluster's
 * share of the total light power), then a light drawn inside the cluster in
 * proportion to its power; its source pdf m_c * pmf_i / P_c enters the RIS
 * weight p-hat/p (eps = 1 is exactly the power source).  The n candidates
 * of a pixel (n = min(lights, numberOfCandidates)) take the targets
 * (j + u_j)/n of m's CDF, u_j uniform: stratified, one per stratum.  BF16
 * MFMA with FP32 accumulation for inference and training;
 * FP32 master weights and Adam. */
#define VR_MLP_PARAMS 21440u  /* W1[64][9] b1[64] W2[64][64] b2[64] W3[256][64] b3[256] */

/* cluster of every uploaded light (0..nclusters-1, nclusters <= 256); must
 * be called again after vr_set_lights */
int vr_set_light_clusters(vr_ctx* ctx, const uint32_t* cluster_of_light, uint32_t nlights, uint32_t nclusters);
/* eps of the defensive mixture, 0..1 (default 0.25) */
int vr_set_proposal_mix(vr_ctx* ctx, float eps);
/* how the light inside a drawn cluster is chosen: in proportion to its
 * power (VR_CLUSTER_DRAW_POWER, the default; source pdf m_c * pmf_i / P_c)
 * or uniformly (VR_CLUSTER_DRAW_UNIFORM; source pdf m_c / n_c, the
 * many-light configuration's "uniform within the cluster").  Takes effect
 * with the next vr_set_light_clusters, which must follow before a learned
 * frame. */
#define VR_CLUSTER_DRAW_POWER   0u
#define VR_CLUSTER_DRAW_UNIFORM 1u
int vr_set_cluster_draw(vr_ctx* ctx, uint32_t mode);
/* the cluster tables (any pointer may be NULL): start[257] offsets into
 * lights[n] (light ids grouped by cluster), per-cluster power alias tables
 * prob[n] / alias[n] (cluster-local indices) aligned with lights, share[n]
 * = pmf_i / P_c by light id, power[256] = P_c */
int vr_read_cluster_tables(vr_ctx* ctx, uint32_t* start, uint32_t* lights, float* prob, uint32_t* alias,
                           float* share, float* power);
/* The network starts from a fixed He-uniform initialisation (seed 1, biases
 * 0, output layer scaled by 1/8 so the first proposal is near-uniform over
 * clusters).  FP32 master parameters in the flat layout above; setting them
 * resets Adam */
int vr_mlp_set_params(vr_ctx* ctx, const float* params, uint32_t n);
int vr_mlp_get_params(vr_ctx* ctx, float* params, uint32_t n);
/* cluster probabilities q[B][256] for inputs x[B][9] */
int vr_mlp_probs(vr_ctx* ctx, const float* x, uint32_t B, float* q);
/* gradient (VR_MLP_PARAMS floats) of (1/B) sum_b w_b * -log q(t_b | x_b) and the loss */
int vr_mlp_grad(vr_ctx* ctx, const float* x, const uint32_t* t, const float* w, uint32_t B, float* grad,
                double* loss);
/* online training: gather nsamples (x, cluster of the final reservoir's
 * light, weight 1 if the reservoir holds a visible sample) from the last
 * frame rendered by `from` (NULL: by ctx itself; otherwise a context on the
 * same device with the same light list, e.g. one running the plain
 * source), at hashed random pixels, then one Adam step (beta 0.9/0.999,
 * rate lr) on each of `steps` equal minibatches of them; *loss (optional)
 * receives the last minibatch's loss before its update */
int vr_mlp_train(vr_ctx* ctx, const vr_ctx* from, uint32_t nsamples, float lr, uint32_t steps, double* loss);
/* one Adam step on an explicit batch (x[B][9], cluster targets t[B],
 * weights w[B]); *loss (optional) receives the batch loss before the step */
int vr_mlp_step(vr_ctx* ctx, const float* x, const uint32_t* t, const float* w, uint32_t B, float lr, double* loss);
/* the last frame's candidate slots: cluster[32][W*H] (0xffffffff = empty)
 * and mixture mass[32][W*H] */
int vr_read_candidates(vr_ctx* ctx, uint32_t* idx, float* pdf);

/* FP64 converged reference image: render_di_gfxexp (cl/kernels.cl:177-241)
 * accumulated over spp samples in binary64 (the reference's ACCUMULATOR=1
 * path, cl/kernels.cl:304-326, keeps an FP32 running mean instead).
 * Sample f uses R0 = xorshift32^(f+1)(params->R0) and strata frame
 * (frame0 + f) & 255.  accum: width*height*3 doubles, added to (host). */
int vr_ground_truth(vr_ctx* ctx, const vr_render_params* params, uint32_t spp,
                    uint32_t frame0, double* accum);

#ifdef __cplusplus
}
#endif
#endif
