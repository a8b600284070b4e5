// world.h - host-side brick-grid World, the reference's WrldTmpl8 API
// (template/world.h:164-758) re-implemented over the vr_* C ABI
// (vr_api.h), for C++ callers such as the reference's scene code.
//
// The public surface keeps the reference's names and meaning: Set/Get/
// AddBrick/RemoveBrick/OptimizeBricks/FindLightsInWord/SetupLightBuffer/
// AddLight/RemoveLight, Sphere/HDisc/Print/PrintZ/Fill, the camera matrix
// accessors, GetRenderParams(), Render() and Commit().  The host keeps the
// authoritative grid, brick pool, zero counts, dirty bits and light list;
// Commit() mirrors them to the device.  worldapi.h adds the reference's
// free functions (Plot, Read, Box, Copy, ...) over a current World.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "vr_api.h"

namespace vrhost {

struct float3 {
    float x, y, z;
};
struct int3 {
    int x, y, z;
};

inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }
inline float3 operator+(float3 a, float3 b) { return float3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator-(float3 a, float3 b) { return float3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float3 operator*(float3 a, float s) { return float3{a.x * s, a.y * s, a.z * s}; }
inline float3 operator*(float s, float3 a) { return float3{a.x * s, a.y * s, a.z * s}; }
inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float3 cross(float3 a, float3 b)
{
    return float3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
float3 normalize(float3 v);

// row-major 4x4 matrix; columns 0-2 hold the camera axes, 3 the position
// (template/precomp.h:825-835)
struct mat4 {
    float cell[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float& operator[](int i) { return cell[i]; }
    float operator[](int i) const { return cell[i]; }
    static mat4 LookAt(float3 P, float3 T);
};
float3 TransformPosition(float3 a, const mat4& M);

typedef vr_light Light;
typedef vr_render_params RenderParams;

class World {
public:
    explicit World(const vr_config& cfg);
    ~World();
    World(const World&) = delete;
    World& operator=(const World&) = delete;

    // initialization (template/world.h:170-180)
    void Clear();
    void OptimizeBricks();
    void ForceSyncAllBricks();

    // camera (template/world.h:182-187)
    void SetCameraMatrix(const mat4& m) { camMat = m; }
    float3 GetCameraViewDir() const { return make_float3(camMat[2], camMat[6], camMat[10]); }
    float3 GetCameraPos() const { return make_float3(camMat[3], camMat[7], camMat[11]); }
    void SetCameraPos(float3 P) { camMat[3] = P.x, camMat[7] = P.y, camMat[11] = P.z; }
    mat4& GetCameraMatrix() { return camMat; }
    RenderParams& GetRenderParams() { return params; }

    // per frame (template/world.cpp:2127-2483)
    void Render();
    // Render() up to the launch: device context, light upload when the list
    // changed, the next frame's params (for a batch of views)
    bool BeginFrame();
    void Commit();
    // the two halves of Render() around the launch, for drivers that render
    // through another path: camera/reprojection/R0 set-up, frame counters
    void PrepareFrame();
    void FinishFrame();

    // per-stage timing of the last frame, ms (template/world.h:222-227).
    // Spatial reuse and shading run as one kernel: GetSpatialTime() covers
    // both and GetShadingTime() is 0.
    float GetAlbedoTime() const { return stageMs[VR_STAGE_PRIMARY]; }
    float GetCandidateTime() const { return stageMs[VR_STAGE_INITIAL]; }
    float GetSpatialTime() const { return stageMs[VR_STAGE_RESOLVE]; }
    float GetShadingTime() const { return 0.0f; }
    float GetRenderTime() const;

    // low-level voxel access (template/world.h:293-513)
    uint32_t Get(uint32_t x, uint32_t y, uint32_t z) const;
    void Set(uint32_t x, uint32_t y, uint32_t z, uint32_t v);
    void AddBrick(uint32_t bx, uint32_t by, uint32_t bz, uint32_t v);
    void RemoveBrick(uint32_t bx, uint32_t by, uint32_t bz);
    // World::Fill (template/world.cpp:747-757): every top cell becomes a
    // solid cell of colour c, every brick is free again, no dirty marks
    void Fill(uint32_t c);

    // high-level voxel access (template/world.h:217-220, world.cpp:932-1000)
    void Sphere(float x, float y, float z, float r, uint32_t c);
    void HDisc(float x, float y, float z, float r, uint32_t c);
    // text in the glyph sheet's 6x10 font: Print along +x in the plane z,
    // PrintZ along +z in the plane x; glyph pixels that are black in the
    // sheet are set to c.  The sheet is read on first use from the path
    // SetFont names (the reference loads "assets/font.png" in World::World,
    // template/world.cpp:192); without one, Print sets LastError and draws
    // nothing.
    void Print(const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c);
    void PrintZ(const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c);
    bool SetFont(const char* pngPath);
    // the glyph sheet is (or could now be) read: false if it cannot be
    bool SetFontIfMissing() { return EnsureFont(); }

    // lights (template/world.h:245-262)
    void FindLightsInWord(std::vector<Light>& ls) const;
    void SetupLightBuffer()
    {
        std::vector<Light> none;
        SetupLightBuffer(none);
    }
    void SetupLightBuffer(const std::vector<Light>& ls, int pos = 0);
    void AddLight(int3 pos, uint32_t size, uint32_t c);
    void RemoveLight(int3 pos);
    const std::vector<Light>& GetLights() const { return lights; }

    // data access (template/world.h:207-211)
    const uint32_t* GetGrid() const { return grid.data(); }
    const uint32_t* GetBrick() const { return bricks; }
    const uint32_t* GetZeroes() const { return zeroes.data(); }
    uint32_t BrickHighWater() const { return highWater; }
    uint32_t BrickCapacity() const { return cap; }
    uint32_t DirtyBricks(std::vector<uint32_t>* ids) const;

    // raw state writes of WorldEditor::LoadWorld (worldeditor.cpp:1376-1426):
    // a top-grid cell, and a brick's payload and zero count at a pool
    // offset (marked dirty, with one ring allocation per loaded brick as the
    // reference makes); false if the index or offset is out of range
    bool SetGridCell(uint32_t index, uint32_t value);
    bool LoadBrick(uint32_t offset, const uint32_t* payload, uint32_t zeroCount);

    // extensions beyond the reference
    void SetFocalDistance(float f) { focal = f; }  // 2.2 in world.cpp:2155-2157
    float GetFocalDistance() const { return focal; }
    void SetNextR0(uint32_t r0) { r0Override = true, r0Next = r0; }
    uint32_t RngState() const { return rng; }
    void SetRngState(uint32_t s) { rng = s; }
    vr_ctx* Context() { return EnsureContext() ? ctx : nullptr; }
    int LastStatus() const { return status; }
    const std::string& LastError() const { return error; }
    const vr_config& Config() const { return cfg; }

private:
    uint32_t NewBrick();
    void FreeBrick(uint32_t idx);
    void Mark(uint32_t idx) { modified[idx >> 5] |= 1u << (idx & 31); }
    uint32_t SplitSolidBrick(uint32_t color, uint32_t cellIndex);
    bool EnsureFont();
    void PrintGlyphs(const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c, bool alongZ);
    int Check(int rc, const char* what);
    bool EnsureContext();

    vr_config cfg;
    vr_ctx* ctx = nullptr;
    uint32_t cap;
    std::vector<uint32_t> grid;   // [128^3]
    uint32_t* bricks = nullptr;   // [cap * 512], zero-filled on demand
    std::vector<uint32_t> zeroes; // [cap]
    std::vector<uint32_t> trash;  // [cap]
    std::vector<uint32_t> modified;
    uint32_t trashHead, trashTail, highWater = 0;
    bool gridDirty = true, lightsDirty = true, firstFrame = true, firstRender = true;
    std::vector<Light> lights;
    mat4 camMat;
    RenderParams params;
    float focal = 2.2f;
    uint32_t rng = 0x12345678u;  // RandomUInt state, template/template.cpp:933-941
    bool r0Override = false;
    uint32_t r0Next = 0;
    float stageMs[VR_NUM_STAGES] = {};
    int status = VR_OK;
    std::string error;
    std::string fontPath = "assets/font.png";
    std::vector<uint32_t> font;  // glyph sheet, 0xRRGGBB per pixel
    uint32_t fontW = 0, fontH = 0;
};

// PNG reader for the glyph sheet (8-bit grey, RGB or RGBA, not interlaced):
// pixels as 0xRRGGBB, row-major from the top, as the reference's
// Surface::LoadImage stores them (template/template.cpp:1776-1798).
bool ReadPngRgb(const char* path, uint32_t& w, uint32_t& h, std::vector<uint32_t>& px, std::string& err);

}  // namespace vrhost
