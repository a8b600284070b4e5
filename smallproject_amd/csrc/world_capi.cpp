// world_capi.cpp - C ABI over vrhost::World (include/vr_world.h).
#include <cstring>
#include <new>

#include "vr_world.h"
#include "world.h"
#include "worldapi.h"

using vrhost::World;

static inline World* W(vrw_world* w) { return reinterpret_cast<World*>(w); }
static inline const World* W(const vrw_world* w) { return reinterpret_cast<const World*>(w); }

extern "C" {

int vrw_create(const vr_config* cfg, vrw_world** out)
{
    if (!cfg || !out) return VR_ERR_ARG;
    *out = nullptr;
    try {
        *out = reinterpret_cast<vrw_world*>(new World(*cfg));
    } catch (const std::bad_alloc&) {
        return VR_ERR_NOMEM;
    } catch (...) {
        return VR_ERR_ARG;
    }
    return VR_OK;
}

int vrw_destroy(vrw_world* w)
{
    if (!w) return VR_ERR_ARG;
    delete W(w);
    return VR_OK;
}

const char* vrw_last_error(const vrw_world* w) { return w ? W(w)->LastError().c_str() : "null world"; }
vr_ctx* vrw_context(vrw_world* w) { return w ? W(w)->Context() : nullptr; }

void vrw_clear(vrw_world* w) { W(w)->Clear(); }
uint32_t vrw_get(const vrw_world* w, uint32_t x, uint32_t y, uint32_t z) { return W(w)->Get(x, y, z); }
void vrw_set(vrw_world* w, uint32_t x, uint32_t y, uint32_t z, uint32_t v) { W(w)->Set(x, y, z, v); }
void vrw_set_many(vrw_world* w, const uint32_t* xyz, const uint32_t* v, size_t n)
{
    World* p = W(w);
    for (size_t i = 0; i < n; i++) p->Set(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], v[i]);
}
void vrw_add_brick(vrw_world* w, uint32_t bx, uint32_t by, uint32_t bz, uint32_t v) { W(w)->AddBrick(bx, by, bz, v); }
void vrw_remove_brick(vrw_world* w, uint32_t bx, uint32_t by, uint32_t bz) { W(w)->RemoveBrick(bx, by, bz); }
void vrw_optimize(vrw_world* w) { W(w)->OptimizeBricks(); }
void vrw_box(vrw_world* w, int x1, int y1, int z1, int x2, int y2, int z2, uint32_t c)
{
    // Box, template/template.cpp:121-124
    for (int y = y1; y < y2; y++)
        for (int z = z1; z < z2; z++)
            for (int x = x1; x < x2; x++) W(w)->Set((uint32_t)x, (uint32_t)y, (uint32_t)z, c);
}
void vrw_sphere(vrw_world* w, float x, float y, float z, float r, uint32_t c) { W(w)->Sphere(x, y, z, r, c); }
void vrw_hdisc(vrw_world* w, float x, float y, float z, float r, uint32_t c) { W(w)->HDisc(x, y, z, r, c); }
void vrw_fill(vrw_world* w, uint32_t c) { W(w)->Fill(c); }
void vrw_copy(vrw_world* w, int x1, int y1, int z1, int x2, int y2, int z2, int dx, int dy, int dz)
{
    vrhost::World* prev = vrhost::GetWorld();
    vrhost::SetWorld(W(w));
    vrhost::Copy(vrhost::int3{x1, y1, z1}, vrhost::int3{x2, y2, z2}, vrhost::int3{dx, dy, dz});
    vrhost::SetWorld(prev);
}
int vrw_set_font(vrw_world* w, const char* png_path)
{
    if (!w || !png_path) return VR_ERR_ARG;
    return W(w)->SetFont(png_path) ? VR_OK : VR_ERR_ARG;
}
static int print_text(vrw_world* w, const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c, bool alongZ)
{
    if (!w || !text) return VR_ERR_ARG;
    if (!W(w)->SetFontIfMissing()) return VR_ERR_ARG;
    if (alongZ) W(w)->PrintZ(text, x, y, z, c);
    else W(w)->Print(text, x, y, z, c);
    return VR_OK;
}
int vrw_print(vrw_world* w, const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c)
{
    return print_text(w, text, x, y, z, c, false);
}
int vrw_printz(vrw_world* w, const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c)
{
    return print_text(w, text, x, y, z, c, true);
}

uint32_t vrw_find_lights(const vrw_world* w, vr_light* out, uint32_t cap)
{
    std::vector<vr_light> ls;
    W(w)->FindLightsInWord(ls);
    for (size_t i = 0; i < ls.size() && i < cap; i++) out[i] = ls[i];
    return (uint32_t)ls.size();
}
uint32_t vrw_num_lights(const vrw_world* w) { return (uint32_t)W(w)->GetLights().size(); }
uint32_t vrw_get_lights(const vrw_world* w, vr_light* out, uint32_t cap)
{
    const auto& ls = W(w)->GetLights();
    for (size_t i = 0; i < ls.size() && i < cap; i++) out[i] = ls[i];
    return (uint32_t)ls.size();
}
void vrw_setup_light_buffer(vrw_world* w, const vr_light* ls, uint32_t n, int pos)
{
    std::vector<vr_light> v(ls, ls + n);
    W(w)->SetupLightBuffer(v, pos);
}
void vrw_add_light(vrw_world* w, int x, int y, int z, uint32_t size, uint32_t c)
{
    W(w)->AddLight(vrhost::int3{x, y, z}, size, c);
}
void vrw_remove_light(vrw_world* w, int x, int y, int z) { W(w)->RemoveLight(vrhost::int3{x, y, z}); }

const uint32_t* vrw_grid(const vrw_world* w) { return W(w)->GetGrid(); }
const uint32_t* vrw_bricks(const vrw_world* w) { return W(w)->GetBrick(); }
const uint32_t* vrw_zeroes(const vrw_world* w) { return W(w)->GetZeroes(); }
uint32_t vrw_brick_high_water(const vrw_world* w) { return W(w)->BrickHighWater(); }
uint32_t vrw_dirty_bricks(const vrw_world* w, uint32_t* ids, uint32_t cap)
{
    std::vector<uint32_t> v;
    const uint32_t n = W(w)->DirtyBricks(&v);
    for (uint32_t i = 0; i < n && i < cap; i++) ids[i] = v[i];
    return n;
}

void vrw_look_at(vrw_world* w, const float pos[3], const float target[3])
{
    W(w)->SetCameraMatrix(vrhost::mat4::LookAt(vrhost::make_float3(pos[0], pos[1], pos[2]),
                                               vrhost::make_float3(target[0], target[1], target[2])));
}
void vrw_set_focal(vrw_world* w, float focal) { W(w)->SetFocalDistance(focal); }
vr_render_params* vrw_params(vrw_world* w) { return &W(w)->GetRenderParams(); }
void vrw_set_next_r0(vrw_world* w, uint32_t r0) { W(w)->SetNextR0(r0); }

int vrw_render(vrw_world* w)
{
    W(w)->Render();
    return W(w)->LastStatus();
}
void vrw_prepare_frame(vrw_world* w) { W(w)->PrepareFrame(); }
int vrw_begin_frame(vrw_world* w)
{
    W(w)->BeginFrame();
    return W(w)->LastStatus();
}
void vrw_finish_frame(vrw_world* w) { W(w)->FinishFrame(); }
int vrw_commit(vrw_world* w)
{
    W(w)->Commit();
    return W(w)->LastStatus();
}
int vrw_force_sync(vrw_world* w)
{
    W(w)->ForceSyncAllBricks();
    return W(w)->LastStatus();
}
int vrw_stage_times(vrw_world* w, float ms[VR_NUM_STAGES])
{
    vr_ctx* c = W(w)->Context();
    if (!c) return VR_ERR_STATE;
    return vr_stage_times(c, ms);
}

}  // extern "C"
