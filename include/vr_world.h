/*
 * vr_world.h - C ABI over the host World (the reference's WrldTmpl8 World
 * class, template/world.h:164-758, and worldapi, template/worldapi.h:1-101),
 * plus the seeded scene generators and camera paths of the benchmark
 * configurations.
 *
 * A vrw_world owns the authoritative host copy of the brick grid, brick pool
 * and light list and a lazily created vr_ctx (include/vr_api.h) that mirrors
 * them on the device.  Editing calls (vrw_set, vrw_add_brick, ...) touch host
 * memory only and work without a GPU; vrw_commit/vrw_force_sync/vrw_render
 * create the device context on first use.
 */
#ifndef VR_WORLD_H
#define VR_WORLD_H
#include <stddef.h>
#include <stdint.h>

#include "vr_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vrw_world vrw_world;

/* World::World (template/world.cpp:47-193) + InitReSTIR (195-208) */
int vrw_create(const vr_config* cfg, vrw_world** out);
int vrw_destroy(vrw_world* w);
const char* vrw_last_error(const vrw_world* w);
/* the device context (created on first call); NULL if no GPU is usable */
vr_ctx* vrw_context(vrw_world* w);

/* ClearWorld / Plot / Read (template/worldapi.h:5-16) and World::Set/Get/
 * AddBrick/RemoveBrick/OptimizeBricks (template/world.h:293-513,
 * template/world.cpp:655-683) */
void vrw_clear(vrw_world* w);
uint32_t vrw_get(const vrw_world* w, uint32_t x, uint32_t y, uint32_t z);
void vrw_set(vrw_world* w, uint32_t x, uint32_t y, uint32_t z, uint32_t v);
/* n voxels at xyz[3n] with values v[n], applied in order */
void vrw_set_many(vrw_world* w, const uint32_t* xyz, const uint32_t* v, size_t n);
void vrw_add_brick(vrw_world* w, uint32_t bx, uint32_t by, uint32_t bz, uint32_t v);
void vrw_remove_brick(vrw_world* w, uint32_t bx, uint32_t by, uint32_t bz);
void vrw_optimize(vrw_world* w);
/* Box (template/template.cpp:85-88): x1..x2-1 etc. */
void vrw_box(vrw_world* w, int x1, int y1, int z1, int x2, int y2, int z2, uint32_t c);
/* World::Sphere (template/world.cpp:932-945): corners (u,v,w) within r of
 * (x,y,z), in binary32 */
void vrw_sphere(vrw_world* w, float x, float y, float z, float r, uint32_t c);
/* World::HDisc (template/world.cpp:949-961): the disc of radius r in the
 * plane (uint)y */
void vrw_hdisc(vrw_world* w, float x, float y, float z, float r, uint32_t c);
/* FillWorld / World::Fill (template/worldapi.h:7, template/world.cpp:747-757):
 * every top cell a solid cell of colour c, every brick free */
void vrw_fill(vrw_world* w, uint32_t c);
/* Copy (template/worldapi.h:21-24, template/template.cpp:92-99): the box
 * (x1,y1,z1)..(x2,y2,z2), inclusive, to (dx,dy,dz) */
void vrw_copy(vrw_world* w, int x1, int y1, int z1, int x2, int y2, int z2, int dx, int dy, int dz);
/* World::Print / World::PrintZ (template/world.h:219-220, world.cpp:965-1001):
 * text in the glyph sheet's 6x10 font along +x (Print) or +z (PrintZ).  The
 * sheet is a PNG read on first use from the path vrw_set_font names
 * ("assets/font.png" by default, as in World::World, world.cpp:192).  Print
 * returns VR_ERR_ARG (and draws nothing) when the sheet cannot be read. */
int vrw_set_font(vrw_world* w, const char* png_path);
int vrw_print(vrw_world* w, const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c);
int vrw_printz(vrw_world* w, const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c);

/* lights (template/world.cpp:228-612) */
uint32_t vrw_find_lights(const vrw_world* w, vr_light* out, uint32_t cap);
uint32_t vrw_num_lights(const vrw_world* w);
uint32_t vrw_get_lights(const vrw_world* w, vr_light* out, uint32_t cap);
void vrw_setup_light_buffer(vrw_world* w, const vr_light* ls, uint32_t n, int pos);
void vrw_add_light(vrw_world* w, int x, int y, int z, uint32_t size, uint32_t c);
void vrw_remove_light(vrw_world* w, int x, int y, int z);

/* host data (template/world.h:207-211) */
const uint32_t* vrw_grid(const vrw_world* w);
const uint32_t* vrw_bricks(const vrw_world* w);
const uint32_t* vrw_zeroes(const vrw_world* w);
uint32_t vrw_brick_high_water(const vrw_world* w);
uint32_t vrw_dirty_bricks(const vrw_world* w, uint32_t* ids, uint32_t cap);

/* camera: LookAt (template/template.cpp:204-215) and the focal distance of
 * the pinhole (2.2 in template/world.cpp:2155-2157; 1/tan(fov/2) in general) */
void vrw_look_at(vrw_world* w, const float pos[3], const float target[3]);
void vrw_set_focal(vrw_world* w, float focal);
/* the World's RenderParams (mutable, like World::GetRenderParams()) */
vr_render_params* vrw_params(vrw_world* w);
/* R0 of the next frame (otherwise the World's xorshift32, seed 0x12345678) */
void vrw_set_next_r0(vrw_world* w, uint32_t r0);

/* per frame (template/world.cpp:2127-2483) */
int vrw_render(vrw_world* w);
/* Render() without the launch: vrw_prepare_frame fills vrw_params() for the
 * next frame (camera, reprojection planes, R0, light count), vrw_finish_frame
 * advances the frame counters */
void vrw_prepare_frame(vrw_world* w);
void vrw_finish_frame(vrw_world* w);
/* Render() up to the launch, for a batch of views (vr_batch_render with
 * vrw_params() of each world, then vrw_finish_frame): creates the device
 * context, uploads the light list if it changed, prepares the frame */
int vrw_begin_frame(vrw_world* w);
int vrw_commit(vrw_world* w);
int vrw_force_sync(vrw_world* w);
int vrw_stage_times(vrw_world* w, float ms[VR_NUM_STAGES]);

/* ---- scene files: the reference's gzip .vx format ---- */
/* LoadSprite + StampSpriteTo (template/world.cpp:1005-1038, 1547-1563): stamp
 * frame 0 of a .vx file at (x,y,z).  16-bit MRGB4444 payloads become
 * 0x00MMARGB with A=F and the emission nibble as emission byte; 8-bit RGB332
 * payloads are expanded to 4-bit channels.  stats (optional, 4 u32): entries,
 * non-zero entries, payload size, frames. */
int vrw_load_vx(vrw_world* w, const char* path, int x, int y, int z, uint32_t* stats);
/* CreateSprite over a box + SaveSprite (world.cpp:1307-1325, 1347-1386),
 * 32-bit payload */
int vrw_save_vx(const vrw_world* w, const char* path, int x, int y, int z, int sx, int sy, int sz);
/* the world editor's world file (WorldEditor::LoadWorld, worldeditor.cpp:
 * 1347-1452): clear, write the listed top-grid cells and bricks, find the
 * lights, optimize bricks (the device copy follows on vrw_commit or
 * vrw_force_sync).  stats (optional, 3 u32): grid entries, bricks, lights */
int vrw_load_nbt(vrw_world* w, const char* path, uint32_t* stats);
/* WorldEditor::SaveWorld (worldeditor.cpp:1454-1600): non-empty cells in
 * index order, brick ids renumbered from 0 */
int vrw_save_nbt(const vrw_world* w, const char* path);

/* ---- image metrics and raster dumps (mse.cpp:33-65, mygame.cpp:518-549) ---- */
/* out[0] = MSE, out[1] = mean absolute error over the rgb channels of two
 * float4 rasters: per-pixel differences in binary32, sums in binary64 */
int vrw_image_error(const float* gt, const float* img, uint32_t width, uint32_t height, double out[2]);
/* .dat raster: u32 width, u32 height, float4[width*height] */
int vrw_write_dat(const char* path, uint32_t width, uint32_t height, const float* rgba);
/* reads the size; pixels too when rgba is non-null and holds capacity_px */
int vrw_read_dat(const char* path, uint32_t* width, uint32_t* height, float* rgba, size_t capacity_px);

/* ---- seeded scene generators of the benchmark configurations ---- */
/* value-noise heightfield over [0,size)^2: h = scale*(40 + 24*n(x/scale, z/scale))
 * with n a 4-octave value noise normalised to [-1,1] */
float vrw_terrain_height(float x, float z, uint32_t scale, uint32_t seed);
int vrw_scene_terrain(vrw_world* w, uint32_t size, uint32_t scale, uint32_t seed);
/* n axis-aligned boxes, edges uniform in [emin, emax], resting in the terrain */
int vrw_scene_boxes(vrw_world* w, uint32_t n, uint32_t emin, uint32_t emax, uint32_t size, uint32_t scale,
                    uint32_t terrain_seed, uint32_t seed);
/* n emissive voxels at uniform (x,z), hmin..hmax voxels above the terrain,
 * emit byte uniform in [mmin, mmax] */
int vrw_scene_lights_above(vrw_world* w, uint32_t n, uint32_t size, uint32_t scale, uint32_t terrain_seed,
                           uint32_t seed, uint32_t hmin, uint32_t hmax, uint32_t mmin, uint32_t mmax);
/* n emissive voxels replacing the top voxel of uniformly drawn columns */
int vrw_scene_lights_surface(vrw_world* w, uint32_t n, uint32_t size, uint32_t seed, uint32_t mmin, uint32_t mmax);
/* flat floor (height floor_h, a multiple of 8) over [0,size)^2 plus n boxes
 * with edges uniform in [emin, emax] standing on it */
int vrw_scene_floor_boxes(vrw_world* w, uint32_t n, uint32_t emin, uint32_t emax, uint32_t size, uint32_t floor_h,
                          uint32_t seed);
/* n emissive voxels in nclusters Gaussian clusters (std. dev. sigma) with
 * centres uniform over [0,size)^2 x [ymin,ymax), in empty voxels only */
int vrw_scene_light_clusters(vrw_world* w, uint32_t n, uint32_t nclusters, float sigma, uint32_t size, uint32_t ymin,
                             uint32_t ymax, uint32_t seed, uint32_t mmin, uint32_t mmax);
/* k-means (k-means++ seeding, iters Lloyd steps) of the light list's
 * positions; out[i] = cluster of light i (for vr_set_light_clusters) */
int vrw_cluster_lights(const vrw_world* w, uint32_t k, uint32_t iters, uint32_t seed, uint32_t* out);
/* emission LUT: lut[0] = 0, lut[m] = lo * (hi/lo)^((m-1)/254) for m >= 1 */
void vrw_log_lut(float lo, float hi, float lut[256]);
/* camera path: kind 0 = straight line at 2 voxels/frame toward the centre,
 * kind 1 = Catmull-Rom spline through seeded control points above the terrain
 * (above y = 0 when scale is 0, for flat-floor scenes) */
int vrw_camera_path(uint32_t kind, uint32_t seed, uint32_t frame, uint32_t nframes, uint32_t size, uint32_t scale,
                    uint32_t terrain_seed, float pos[3], float target[3]);

#ifdef __cplusplus
}
#endif
#endif
