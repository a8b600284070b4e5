// render_driver.cpp - a C++ caller of the C ABI alone (include/vr_world.h,
// include/vr_api.h), shaped like the reference's main loop
// (template/template.cpp:519-530: world->Render() before the game's Tick,
// world->Commit() after): build a scene through the World API, render
// frames along a camera path, read the frame back.
//
// Usage: render_driver WIDTH HEIGHT FRAMES
// Prints one line: "frames F width W height H bits <FNV-1a of the frame's
// float bits> lights N".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vr_world.h"

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s WIDTH HEIGHT FRAMES\n", argv[0]);
        return 2;
    }
    const uint32_t W = (uint32_t)std::atoi(argv[1]), H = (uint32_t)std::atoi(argv[2]);
    const uint32_t frames = (uint32_t)std::atoi(argv[3]);
    vr_config cfg;
    vr_default_config(&cfg, W, H);
    cfg.light_mode = VR_LIGHTS_POINT;
    cfg.brick_capacity = 1u << 14;
    vrw_world* w = nullptr;
    if (vrw_create(&cfg, &w) != VR_OK) {
        std::fprintf(stderr, "vrw_create failed\n");
        return 1;
    }
    // scene: 128^3 heightfield with 64 point lights above it (the seeded
    // generators of the point-light configuration)
    vrw_scene_terrain(w, 128, 1, 1);
    vrw_optimize(w);
    vrw_scene_lights_above(w, 64, 128, 1, 1, 2, 2, 20, 1, 10);
    vrw_set_focal(w, 1.7320508f);  // 60 degree field of view
    // the scene goes to the device once before the loop (the reference
    // renders its initial device state on the first frame instead; here
    // vr_render_frame refuses to render before any commit)
    if (vrw_commit(w) != VR_OK) {
        std::fprintf(stderr, "commit: %s\n", vrw_last_error(w));
        return 1;
    }
    vr_render_params* p = vrw_params(w);
    p->numberOfCandidates = 32, p->spatialTaps = 5, p->spatialRadius = 30, p->numberOfMaxTemporalImportance = 20;
    for (uint32_t f = 0; f < frames; f++) {
        float pos[3], tgt[3];
        vrw_camera_path(0, 3, f, 8, 128, 1, 1, pos, tgt);
        vrw_look_at(w, pos, tgt);
        if (vrw_render(w) != VR_OK || vrw_commit(w) != VR_OK) {
            std::fprintf(stderr, "frame %u: %s\n", f, vrw_last_error(w));
            return 1;
        }
    }
    std::vector<float> rgba((size_t)W * H * 4);
    vr_ctx* ctx = vrw_context(w);
    if (!ctx || vr_read_frame(ctx, rgba.data()) != VR_OK) {
        std::fprintf(stderr, "read-back failed\n");
        return 1;
    }
    uint64_t h = 1469598103934665603ull;
    const uint8_t* b = reinterpret_cast<const uint8_t*>(rgba.data());
    for (size_t i = 0; i < rgba.size() * 4; i++) h = (h ^ b[i]) * 1099511628211ull;
    std::printf("frames %u width %u height %u bits %016llx lights %u\n", frames, W, H, (unsigned long long)h,
                vrw_num_lights(w));
    vrw_destroy(w);
    return 0;
}
