// world.cpp - host World over the vr_* C ABI.  Brick allocation, zero
// counts, dirty bits and light bookkeeping follow template/world.h:293-606
// and template/world.cpp:47-94, 228-246, 438-612, 655-683, 726-745 with
// THREADSAFEWORLD=1 (world.h:6): brick ids come from a circular free list
// walked with stride 31, so a fresh world hands out ids 0, 1, 2, ...
#include "world.h"

#include <sys/mman.h>

#include "worldapi.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace vrhost {

float3 normalize(float3 v)
{
    const float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}

mat4 mat4::LookAt(float3 P, float3 T)
{
    // template/precomp.h:825-835
    const float3 z = normalize(T - P);
    const float3 x = normalize(cross(z, make_float3(0, 1, 0)));
    const float3 y = cross(x, z);
    mat4 M;
    M[3] = P.x, M[7] = P.y, M[11] = P.z;
    M[0] = x.x, M[4] = x.y, M[8] = x.z;
    M[1] = y.x, M[5] = y.y, M[9] = y.z;
    M[2] = z.x, M[6] = z.y, M[10] = z.z;
    return M;
}

float3 TransformPosition(float3 a, const mat4& M)
{
    // make_float3(make_float4(a, 1) * M), template/template.cpp:1052-1062
    return make_float3(M[0] * a.x + M[1] * a.y + M[2] * a.z + M[3], M[4] * a.x + M[5] * a.y + M[6] * a.z + M[7],
                       M[8] * a.x + M[9] * a.y + M[10] * a.z + M[11]);
}

static inline bool IsEmitter(uint32_t v) { return ((v >> 16) & 255u) > 0u; }

World::World(const vr_config& c) : cfg(c), cap(c.brick_capacity)
{
    if (cap == 0 || (cap & (cap - 1)) != 0) throw std::invalid_argument("brick_capacity must be a power of two");
    grid.assign(VR_GRIDSIZE, 0u);
    // the pool is reserved, not touched: pages are committed as bricks are written
    const size_t bytes = (size_t)cap * VR_BRICKSIZE * 4;
    void* p = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (p == MAP_FAILED) throw std::bad_alloc();
    bricks = static_cast<uint32_t*>(p);
    zeroes.assign(cap, 0u);
    trash.assign(cap, 0u);
    modified.assign(cap / 32 + 1, 0u);
    std::memset(&params, 0, sizeof(params));
    // World::InitReSTIR (template/world.cpp:195-208); the sky dome is off
    // because the sky HDR and blue-noise tables are not part of this path
    params.spatial = 1, params.temporal = 1;
    params.spatialTaps = 5, params.spatialRadius = 30;
    params.numberOfCandidates = 32, params.numberOfMaxTemporalImportance = 20;
    params.skyDomeSampling = 0;
    Clear();
}

bool World::EnsureContext()
{
    // the device context is created on first use, so the host-side world
    // (building, editing, light bookkeeping) works without a GPU
    if (ctx) return true;
    const int rc = vr_create(&cfg, &ctx);
    if (rc != VR_OK) {
        ctx = nullptr;
        status = rc;
        error = "vr_create failed with status " + std::to_string(rc);
        return false;
    }
    return true;
}

World::~World()
{
    if (ctx) vr_destroy(ctx);
    if (bricks) munmap(bricks, (size_t)cap * VR_BRICKSIZE * 4);
}

int World::Check(int rc, const char* what)
{
    status = rc;
    if (rc != VR_OK) error = std::string(what) + ": " + (ctx ? vr_last_error(ctx) : "no context");
    return rc;
}

void World::Clear()
{
    // World::Clear, template/world.cpp:726-745
    std::fill(grid.begin(), grid.end(), 0u);
    for (uint32_t i = 0; i < cap; i++) trash[(i * 31u) & (cap - 1)] = i;
    trashHead = cap, trashTail = 0;
    if (highWater) std::memset(bricks, 0, (size_t)highWater * VR_BRICKSIZE * 4);
    std::fill(zeroes.begin(), zeroes.end(), 0u);
    std::fill(modified.begin(), modified.end(), 0u);
    highWater = 0;
    lights.clear();
    params.numberOfLights = 0;
    gridDirty = true, lightsDirty = true;
}

uint32_t World::NewBrick()
{
    // NewBrick, template/world.h:585-595 (InterlockedAdd(&trashTail, 31) - 31)
    const uint32_t item = trashTail;
    trashTail += 31;
    const uint32_t id = trash[item & (cap - 1)];
    if (id + 1 > highWater) highWater = id + 1;
    return id;
}

bool World::SetGridCell(uint32_t index, uint32_t value)
{
    if (index >= VR_GRIDSIZE) return false;
    grid[index] = value;
    gridDirty = true;
    return true;
}

bool World::LoadBrick(uint32_t offset, const uint32_t* payload, uint32_t zeroCount)
{
    if (offset >= cap) return false;
    std::memcpy(bricks + (size_t)offset * VR_BRICKSIZE, payload, VR_BRICKSIZE * 4);
    zeroes[offset] = zeroCount;
    Mark(offset);
    (void)NewBrick();  // keeps the free ring in step with the loaded bricks
    if (offset + 1 > highWater) highWater = offset + 1;
    return true;
}

void World::FreeBrick(uint32_t idx)
{
    // FreeBrick, template/world.h:596-606
    const uint32_t item = trashHead;
    trashHead += 31;
    trash[item & (cap - 1)] = idx;
}

uint32_t World::Get(uint32_t x, uint32_t y, uint32_t z) const
{
    const uint32_t bx = (x / 8) & 127u, by = (y / 8) & 127u, bz = (z / 8) & 127u;
    const uint32_t cv = grid[bx + bz * 128u + by * 16384u];
    if ((cv & 1u) == 0u) return cv >> 1;
    return bricks[(size_t)(cv >> 1) * 512u + (x & 7u) + (y & 7u) * 8u + (z & 7u) * 64u];
}

uint32_t World::SplitSolidBrick(uint32_t color, uint32_t ci)
{
    // template/world.h:392-445
    const uint32_t off = NewBrick();
    uint32_t* b = bricks + (size_t)off * 512u;
    for (int i = 0; i < 512; i++) b[i] = color;
    zeroes[off] = (color == 0u) * 512u;
    grid[ci] = (off << 1) | 1u;
    gridDirty = true;
    return off;
}

void World::Set(uint32_t x, uint32_t y, uint32_t z, uint32_t v)
{
    // World::Set, template/world.h:456-513
    const uint32_t bx = x / 8, by = y / 8, bz = z / 8;
    if (bx >= 128u || by >= 128u || bz >= 128u) return;
    const uint32_t ci = bx + bz * 128u + by * 16384u;
    const uint32_t cv = grid[ci];
    uint32_t off = cv >> 1;
    if ((cv & 1u) == 0u) {
        if (v == (cv >> 1)) return;
        off = SplitSolidBrick(cv >> 1, ci);
    }
    const size_t vi = (size_t)off * 512u + (x & 7u) + (y & 7u) * 8u + (z & 7u) * 64u;
    const uint32_t orig = bricks[vi];
    if (IsEmitter(orig)) RemoveLight(int3{(int)x, (int)y, (int)z});
    if (IsEmitter(v)) AddLight(int3{(int)x, (int)y, (int)z}, 1, v);
    const int change = (orig != 0u && v == 0u) - (orig == 0u && v != 0u);
    zeroes[off] += (uint32_t)change;
    if (zeroes[off] < 512u) {
        bricks[vi] = v;
        Mark(off);
    } else {
        grid[ci] = 0;
        gridDirty = true;
        FreeBrick(off);
    }
}

void World::AddBrick(uint32_t bx, uint32_t by, uint32_t bz, uint32_t v)
{
    // World::AddBrick, template/world.h:308-352
    if (bx >= 128u || by >= 128u || bz >= 128u) return;
    const uint32_t ci = bx + bz * 128u + by * 16384u;
    const uint32_t cv = grid[ci];
    if ((cv & 1u) != 0u) {
        FreeBrick(cv >> 1);
        for (uint32_t ly = 0; ly < 8; ly++)
            for (uint32_t lz = 0; lz < 8; lz++)
                for (uint32_t lx = 0; lx < 8; lx++) {
                    const uint32_t x = lx + bx * 8, y = ly + by * 8, z = lz + bz * 8;
                    if (IsEmitter(Get(x, y, z))) RemoveLight(int3{(int)x, (int)y, (int)z});
                }
    } else if (IsEmitter(cv >> 1)) {
        RemoveLight(int3{(int)(bx * 8), (int)(by * 8), (int)(bz * 8)});
    }
    if (IsEmitter(v)) AddLight(int3{(int)(bx * 8), (int)(by * 8), (int)(bz * 8)}, 8, v);
    Mark(0);
    grid[ci] = v << 1;
    gridDirty = true;
}

void World::RemoveBrick(uint32_t bx, uint32_t by, uint32_t bz)
{
    // World::RemoveBrick, template/world.h:354-390.  The cell is emptied
    // before its voxels are inspected, so Get() sees zeros there and no
    // voxel light is removed - the reference's order, kept as is.
    if (bx >= 128u || by >= 128u || bz >= 128u) return;
    const uint32_t ci = bx + bz * 128u + by * 16384u;
    const uint32_t cv = grid[ci];
    grid[ci] = 0;
    gridDirty = true;
    if ((cv & 1u) != 0u) {
        zeroes[cv >> 1] = 512u;
        FreeBrick(cv >> 1);
        for (uint32_t ly = 0; ly < 8; ly++)
            for (uint32_t lz = 0; lz < 8; lz++)
                for (uint32_t lx = 0; lx < 8; lx++) {
                    const uint32_t x = lx + bx * 8, y = ly + by * 8, z = lz + bz * 8;
                    if (IsEmitter(Get(x, y, z))) RemoveLight(int3{(int)x, (int)y, (int)z});
                }
    } else if (IsEmitter(cv >> 1)) {
        RemoveLight(int3{(int)(bx * 8), (int)(by * 8), (int)(bz * 8)});
    }
    Mark(0);
}

// the corner range [lo, hi) of a primitive along one axis: (uint)max(0, p - r)
// to (uint)min(dim - 1, p + r + 1), as World::Sphere and World::HDisc take it
// (template/world.cpp:934-936, 951-952)
static inline void span(float p, float r, uint32_t& lo, uint32_t& hi)
{
    lo = (uint32_t)std::fmax(0.0f, p - r);
    hi = (uint32_t)std::fmin((float)VR_MAPDIM - 1.0f, p + r + 1.0f);
}

void World::Sphere(float x, float y, float z, float r, uint32_t c)
{
    // World::Sphere, template/world.cpp:932-945: voxel corners (u, v, w)
    // with |(u, v, w) - (x, y, z)|^2 < r^2, in binary32
    uint32_t x1, x2, y1, y2, z1, z2;
    span(x, r, x1, x2), span(y, r, y1, y2), span(z, r, z1, z2);
    if (x1 >= VR_MAPDIM || y1 >= VR_MAPDIM || z1 >= VR_MAPDIM || x2 >= VR_MAPDIM || y2 >= VR_MAPDIM ||
        z2 >= VR_MAPDIM)
        return;
    const float r2 = r * r;
    for (uint32_t u = x1; u < x2; u++)
        for (uint32_t v = y1; v < y2; v++)
            for (uint32_t w = z1; w < z2; w++) {
                const float dx = (float)u - x, dy = (float)v - y, dz = (float)w - z;
                if (dx * dx + dy * dy + dz * dz < r2) Set(u, v, w, c);
            }
}

void World::HDisc(float x, float y, float z, float r, uint32_t c)
{
    // World::HDisc, template/world.cpp:949-961: the horizontal disc in the
    // plane (uint)y
    uint32_t x1, x2, z1, z2;
    span(x, r, x1, x2), span(z, r, z1, z2);
    if (x1 >= VR_MAPDIM || !(y < (float)VR_MAPDIM) || z1 >= VR_MAPDIM || x2 >= VR_MAPDIM || z2 >= VR_MAPDIM) return;
    const float r2 = r * r;
    for (uint32_t u = x1; u < x2; u++)
        for (uint32_t w = z1; w < z2; w++) {
            const float dx = (float)u - x, dz = (float)w - z;
            if (dx * dx + dz * dz < r2) Set(u, (uint32_t)y, w, c);
        }
}

void World::Fill(uint32_t c)
{
    // World::Fill, template/world.cpp:747-757 (the lights are untouched)
    std::fill(grid.begin(), grid.end(), c << 1);
    for (uint32_t i = 0; i < cap; i++) trash[(i * 31u) & (cap - 1)] = i;
    trashHead = cap, trashTail = 0;
    std::fill(modified.begin(), modified.end(), 0u);
    gridDirty = true;
}

bool World::SetFont(const char* pngPath)
{
    if (!pngPath) return false;
    fontPath = pngPath;
    font.clear();
    return EnsureFont();
}

bool World::EnsureFont()
{
    if (!font.empty()) return true;
    std::string err;
    if (!ReadPngRgb(fontPath.c_str(), fontW, fontH, font, err)) {
        font.clear();
        status = VR_ERR_ARG;
        error = "font: " + err;
        return false;
    }
    return true;
}

void World::PrintGlyphs(const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c, bool alongZ)
{
    // World::Print / World::PrintZ, template/world.cpp:965-1001: glyph t of
    // the table below is the 6x10 block at column t % 13, row t / 13 of the
    // sheet; a character not in the table draws glyph 0, one at or below
    // ' ' (or above 127) draws nothing but still takes its 6 voxels.
    static const char table[] =
        "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz1234567890+-=/\\*:;()[]{}<>!?.,'\"&#$_%^|";
    if (!text || !EnsureFont()) return;
    int glyph[256] = {};
    for (int i = 0; table[i]; i++) glyph[(unsigned char)table[i]] = i;
    const size_t n = std::strlen(text);
    for (size_t i = 0; i < n; i++) {
        const signed char ch = (signed char)text[i];
        if (ch <= 32) continue;
        const int t = glyph[(unsigned char)ch], cx = t % 13, cy = t / 13;
        for (uint32_t v = 0; v < 10; v++)
            for (uint32_t u = 0; u < 6; u++) {
                const uint32_t sx = (uint32_t)cx * 6 + u, sy = (uint32_t)cy * 10 + v;
                if (sx >= fontW || sy >= fontH || (font[sx + sy * fontW] & 0xffffffu) != 0u) continue;
                const uint32_t along = u + (uint32_t)i * 6;
                if (alongZ) Set(x, y + (9 - v), z + along, c);
                else Set(x + along, y + (9 - v), z, c);
            }
    }
}

void World::Print(const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c)
{
    PrintGlyphs(text, x, y, z, c, false);
}

void World::PrintZ(const char* text, uint32_t x, uint32_t y, uint32_t z, uint32_t c)
{
    PrintGlyphs(text, x, y, z, c, true);
}

void World::OptimizeBricks()
{
    // World::OptimizeBricks, template/world.cpp:655-683
    for (uint32_t i = 0; i < VR_GRIDSIZE; i++) {
        const uint32_t value = grid[i];
        if (!(value & 1u)) continue;
        const uint32_t* b = bricks + (size_t)(value >> 1) * 512u;
        const uint32_t first = b[0];
        bool solid = true;
        for (int j = 1; j < 512; j++)
            if (b[j] != first) {
                solid = false;
                break;
            }
        if (solid) {
            grid[i] = (first << 1) | 0u;
            FreeBrick(value >> 1);
            gridDirty = true;
        }
    }
}

void World::FindLightsInWord(std::vector<Light>& ls) const
{
    // template/world.cpp:438-504
    ls.clear();
    for (uint32_t by = 0; by < 128; by++)
        for (uint32_t bz = 0; bz < 128; bz++)
            for (uint32_t bx = 0; bx < 128; bx++) {
                const uint32_t g = grid[bx + bz * 128u + by * 16384u];
                if ((g & 1u) == 0u) {
                    if (IsEmitter(g >> 1))
                        ls.push_back(Light{bx * 8 + bz * 8 * 1024 + by * 8 * 1024 * 1024, g >> 1, 8});
                } else {
                    const uint32_t* b = bricks + (size_t)(g >> 1) * 512u;
                    for (uint32_t ly = 0; ly < 8; ly++)
                        for (uint32_t lz = 0; lz < 8; lz++)
                            for (uint32_t lx = 0; lx < 8; lx++) {
                                const uint32_t v = b[lx + ly * 8 + lz * 64];
                                if (IsEmitter(v)) {
                                    const uint32_t x = lx + bx * 8, y = ly + by * 8, z = lz + bz * 8;
                                    ls.push_back(Light{x + z * 1024 + y * 1024 * 1024, v, 1});
                                }
                            }
                }
            }
}

void World::SetupLightBuffer(const std::vector<Light>& ls, int pos)
{
    // template/world.cpp:563-612: writes ls at pos (0 by default).  The
    // count becomes max(count, pos + size); the reference adds ls.size() to
    // the count whatever pos is, which for pos < count reads past its old
    // buffer.
    const size_t at = pos < 0 ? lights.size() : (size_t)pos;
    if (lights.size() < at + ls.size()) lights.resize(at + ls.size());
    for (size_t i = 0; i < ls.size(); i++) lights[at + i] = ls[i];
    params.numberOfLights = (uint32_t)lights.size();
    lightsDirty = true;
}

void World::AddLight(int3 pos, uint32_t size, uint32_t c)
{
    // template/world.cpp:228-246
    const uint32_t index = (uint32_t)pos.x + (uint32_t)pos.z * 1024u + (uint32_t)pos.y * 1024u * 1024u;
    lights.push_back(Light{index, c, size});
    params.numberOfLights = (uint32_t)lights.size();
    params.restirtemporalframe = 0;
    lightsDirty = true;
}

void World::RemoveLight(int3 pos)
{
    // template/world.cpp:506-561: every light at pos is dropped and the
    // count becomes max(0, n - 1), however many matched
    if (lights.empty()) return;
    const size_t newCount = lights.size() - 1;
    const uint32_t target = (uint32_t)pos.x + (uint32_t)pos.z * 1024u + (uint32_t)pos.y * 1024u * 1024u;
    size_t k = 0;
    for (size_t i = 0; i < lights.size(); i++)
        if (lights[i].position != target) lights[k++] = lights[i];
    lights.resize(newCount);
    params.numberOfLights = (uint32_t)newCount;
    params.restirtemporalframe = 0;
    lightsDirty = true;
}

uint32_t World::DirtyBricks(std::vector<uint32_t>* ids) const
{
    uint32_t n = 0;
    for (uint32_t j = 0; j < cap / 32; j++)
        if (modified[j])
            for (uint32_t k = 0; k < 32; k++)
                if (modified[j] & (1u << k)) {
                    if (ids) ids->push_back(j * 32 + k);
                    n++;
                }
    return n;
}

void World::ForceSyncAllBricks()
{
    // template/world.cpp:619-633: the whole pool in use goes to the device
    if (!EnsureContext()) return;
    std::vector<uint32_t> ids(highWater);
    for (uint32_t i = 0; i < highWater; i++) ids[i] = i;
    if (Check(vr_commit(ctx, grid.data(), ids.data(), bricks, highWater), "ForceSyncAllBricks") != VR_OK) return;
    std::fill(modified.begin(), modified.end(), 0u);
    gridDirty = false, firstFrame = false;
    if (lightsDirty && Check(vr_set_lights(ctx, lights.data(), (uint32_t)lights.size()), "SetupLightBuffer") == VR_OK)
        lightsDirty = false;
}

void World::Commit()
{
    // World::Commit, template/world.cpp:2357-2483: gather at most
    // MAXCOMMITS dirty bricks (the rest waits for the next frame), send the
    // top grid with them, then collect the timing of the rendered frame
    if (!EnsureContext()) return;
    std::vector<uint32_t> ids;
    std::vector<uint32_t> payload;
    uint32_t tasks = 0, words = 0;
    for (uint32_t j = 0; j < cap / 32; j++) {
        if (!modified[j]) continue;
        for (uint32_t k = 0; k < 32; k++) {
            if (!(modified[j] & (1u << k))) continue;
            const uint32_t i = j * 32 + k;
            ids.push_back(i);
            payload.insert(payload.end(), bricks + (size_t)i * 512u, bricks + (size_t)i * 512u + 512u);
            tasks++;
        }
        words = j + 1;
        if (tasks + 32 >= VR_MAXCOMMITS) break;
    }
    if (tasks > 0 || firstFrame || gridDirty) {
        // the dirty bits of the sent bricks are cleared only once the upload
        // succeeded: after a failure they are sent again by the next Commit
        if (Check(vr_commit(ctx, grid.data(), ids.data(), payload.data(), tasks), "Commit") != VR_OK) return;
        std::fill(modified.begin(), modified.begin() + words, 0u);
        firstFrame = false, gridDirty = false;
    }
    if (lightsDirty) {
        if (Check(vr_set_lights(ctx, lights.data(), (uint32_t)lights.size()), "SetupLightBuffer") != VR_OK) return;
        lightsDirty = false;
    }
    float ms[VR_NUM_STAGES];
    if (vr_stage_times(ctx, ms) == VR_OK)
        for (int s = 0; s < VR_NUM_STAGES; s++) stageMs[s] = ms[s];
}

float World::GetRenderTime() const
{
    float t = 0;
    for (int s = 0; s < VR_NUM_STAGES; s++) t += stageMs[s];
    return t;
}

void World::PrepareFrame()
{
    // the parameter half of World::Render, template/world.cpp:2145-2191
    const float3 prevE = make_float3(params.E[0], params.E[1], params.E[2]);
    float3 pE = prevE;
    float3 pP0 = make_float3(params.p0[0], params.p0[1], params.p0[2]);
    float3 pP1 = make_float3(params.p1[0], params.p1[1], params.p1[2]);
    float3 pP2 = make_float3(params.p2[0], params.p2[1], params.p2[2]);
    const mat4 M = camMat;
    const float aspect = (float)cfg.width / (float)cfg.height;
    const float3 E = TransformPosition(make_float3(0, 0, 0), M);
    const float3 p0 = TransformPosition(make_float3(aspect, 1, focal), M);
    const float3 p1 = TransformPosition(make_float3(-aspect, 1, focal), M);
    const float3 p2 = TransformPosition(make_float3(aspect, -1, focal), M);
    if (firstRender) {
        pE = E, pP0 = p0, pP1 = p1, pP2 = p2;
        firstRender = false;
    }
    const float3 pP3 = pP1 + (pP2 - pP0);
    params.oneOverRes[0] = 1.0f / (float)cfg.width, params.oneOverRes[1] = 1.0f / (float)cfg.height;
    params.E[0] = E.x, params.E[1] = E.y, params.E[2] = E.z;
    params.p0[0] = p0.x, params.p0[1] = p0.y, params.p0[2] = p0.z;
    params.p1[0] = p1.x, params.p1[1] = p1.y, params.p1[2] = p1.z;
    params.p2[0] = p2.x, params.p2[1] = p2.y, params.p2[2] = p2.z;
    auto plane = [&](float* out, float3 a, float3 b) {
        const float3 n = normalize(cross(a - pE, b - pE));
        out[0] = n.x, out[1] = n.y, out[2] = n.z, out[3] = dot(pE, n);
    };
    plane(params.Nleft, pP0, pP2);
    plane(params.Nright, pP3, pP1);
    plane(params.Ntop, pP0, pP1);
    plane(params.Nbottom, pP3, pP2);
    if (r0Override) {
        params.R0 = r0Next;
        r0Override = false;
    } else {
        rng ^= rng << 13, rng ^= rng >> 17, rng ^= rng << 5;
        params.R0 = rng;
    }
    params.numberOfLights = (uint32_t)lights.size();
}

void World::FinishFrame()
{
    // frame counters, template/world.cpp:2335-2337
    params.frame = (params.frame + 1) & 255u;
    params.framecount = params.framecount + 1;
    params.restirtemporalframe = params.restirtemporalframe + 1;
}

bool World::BeginFrame()
{
    if (!EnsureContext()) return false;
    if (lightsDirty) {
        if (Check(vr_set_lights(ctx, lights.data(), (uint32_t)lights.size()), "SetupLightBuffer") != VR_OK)
            return false;
        lightsDirty = false;
    }
    PrepareFrame();
    return true;
}

void World::Render()
{
    // World::Render, template/world.cpp:2127-2340
    if (!BeginFrame()) return;
    if (Check(vr_render_frame(ctx, &params), "Render") != VR_OK) return;
    FinishFrame();
}

// ---------------------------------------------------------------------------
// the free-function world API (worldapi.h) over the current World
// ---------------------------------------------------------------------------
static World* g_world = nullptr;

World* GetWorld() { return g_world; }
void SetWorld(World* w) { g_world = w; }

void Box(const int x1, const int y1, const int z1, const int x2, const int y2, const int z2, const uint32_t c)
{
    // Box, template/template.cpp:85-88
    for (int y = y1; y < y2; y++)
        for (int z = z1; z < z2; z++)
            for (int x = x1; x < x2; x++) Plot((uint32_t)x, (uint32_t)y, (uint32_t)z, c);
}

void Copy(const int3 s1, const int3 s2, const int3 D)
{
    // Copy, template/template.cpp:92-99
    const int3 e = int3{s2.x - s1.x, s2.y - s1.y, s2.z - s1.z};
    for (int z = 0; z <= e.z; z++)
        for (int y = 0; y <= e.y; y++)
            for (int x = 0; x <= e.x; x++)
                Plot(int3{D.x + x, D.y + y, D.z + z}, Read(int3{s1.x + x, s1.y + y, s1.z + z}));
}

}  // namespace vrhost
