// vxio.cpp - the reference's .vx scene format (gzip):
//   i32 payloadSize, i32 frames, then per frame
//   i32 size[3], u32 n, u32 pos[n] (x | y<<10 | z<<20), PAYLOAD val[n]
// (SpriteManager::LoadSprite/SaveSprite, template/world.cpp:1005-1038,
// 1307-1325; World::CreateSprite 1347-1386; World::StampSpriteTo 1547-1563).
#include <zlib.h>

#include <cstring>
#include <vector>

#include "vr_world.h"
#include "world.h"

namespace {

// payload conversions to the 32-bit word 0x00MMARGB
uint32_t from16(uint32_t v)
{
    // VOXEL16 is MRGB4444 with M = emitter strength (template/common.h:27-31)
    return (((v >> 12) & 15u) << 16) | 0xf000u | (v & 0xfffu);
}

uint32_t from8(uint32_t v)
{
    // VOXEL8 is RGB332 (template/common.h:22-26)
    const uint32_t r3 = (v >> 5) & 7u, g3 = (v >> 2) & 7u, b2 = v & 3u;
    return 0xf000u | (((r3 * 15u + 3u) / 7u) << 8) | (((g3 * 15u + 3u) / 7u) << 4) | (b2 * 5u);
}

bool readn(gzFile f, void* dst, size_t bytes)
{
    uint8_t* p = static_cast<uint8_t*>(dst);
    while (bytes > 0) {
        const unsigned chunk = bytes > (1u << 30) ? (1u << 30) : (unsigned)bytes;
        const int got = gzread(f, p, chunk);
        if (got <= 0) return false;
        p += got, bytes -= (size_t)got;
    }
    return true;
}

}  // namespace

extern "C" {

int vrw_load_vx(vrw_world* wp, const char* path, int x, int y, int z, uint32_t* stats)
{
    if (!wp || !path) return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    gzFile f = gzopen(path, "rb");
    if (!f) return VR_ERR_ARG;
    int32_t pls = 0, frames = 0;
    int rc = VR_OK;
    if (!readn(f, &pls, 4) || !readn(f, &frames, 4) || frames < 1 || !(pls == 1 || pls == 2 || pls == 4)) {
        gzclose(f);
        return VR_ERR_ARG;  // the reference aborts on a payload mismatch (world.cpp:1015)
    }
    int32_t size[3];
    uint32_t n = 0;
    if (!readn(f, size, 12) || !readn(f, &n, 4) || n > (1u << 30)) {
        gzclose(f);
        return VR_ERR_ARG;
    }
    std::vector<uint32_t> pos(n), val(n);
    if (!readn(f, pos.data(), (size_t)n * 4)) rc = VR_ERR_ARG;
    if (rc == VR_OK) {
        if (pls == 4) {
            if (!readn(f, val.data(), (size_t)n * 4)) rc = VR_ERR_ARG;
        } else {
            std::vector<uint8_t> raw((size_t)n * pls);
            if (!readn(f, raw.data(), raw.size())) rc = VR_ERR_ARG;
            for (uint32_t i = 0; rc == VR_OK && i < n; i++)
                val[i] = pls == 2 ? ((uint32_t)raw[2 * i] | ((uint32_t)raw[2 * i + 1] << 8)) : raw[i];
        }
    }
    gzclose(f);
    if (rc != VR_OK) return rc;
    // StampSpriteTo of frame 0: Set(pos + offset, val | 0xF000) for every
    // draw-list entry, including the zero entries CreateSprite leaves at the
    // front of the list (world.cpp:1356), which set the offset voxel to 0xF000
    uint32_t nonzero = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t v = pos[i];
        const uint32_t vx = v & 1023u, vy = (v >> 10) & 1023u, vz = (v >> 20) & 1023u;
        const uint32_t word = (pls == 4 ? val[i] : (pls == 2 ? from16(val[i]) : from8(val[i]))) | 0xf000u;
        nonzero += val[i] != 0;
        w.Set(vx + (uint32_t)x, vy + (uint32_t)y, vz + (uint32_t)z, word);
    }
    if (stats) stats[0] = n, stats[1] = nonzero, stats[2] = (uint32_t)pls, stats[3] = (uint32_t)frames;
    return VR_OK;
}

int vrw_save_vx(const vrw_world* wp, const char* path, int x, int y, int z, int sx, int sy, int sz)
{
    // World::CreateSprite over the box + SaveSprite, 32-bit payload
    if (!wp || !path || sx <= 0 || sy <= 0 || sz <= 0 || sx > 1024 || sy > 1024 || sz > 1024) return VR_ERR_ARG;
    const vrhost::World& w = *reinterpret_cast<const vrhost::World*>(wp);
    const size_t vpf = (size_t)sx * sy * sz;
    std::vector<uint32_t> pos(vpf / 4, 0u), val(vpf / 4, 0u);  // the reference's zero-filled estimate
    for (int xx = 0; xx < sx; xx++)
        for (int yy = 0; yy < sy; yy++)
            for (int zz = 0; zz < sz; zz++) {
                const uint32_t v = w.Get((uint32_t)(xx + x), (uint32_t)(yy + y), (uint32_t)(zz + z));
                if (v) pos.push_back((uint32_t)xx + ((uint32_t)yy << 10) + ((uint32_t)zz << 20)), val.push_back(v);
            }
    gzFile f = gzopen(path, "wb");
    if (!f) return VR_ERR_ARG;
    const int32_t hdr[2] = {4, 1};
    const int32_t size[3] = {sx, sy, sz};
    const uint32_t n = (uint32_t)pos.size();
    bool ok = gzwrite(f, hdr, 8) == 8 && gzwrite(f, size, 12) == 12 && gzwrite(f, &n, 4) == 4;
    ok = ok && gzwrite(f, pos.data(), n * 4u) == (int)(n * 4u) && gzwrite(f, val.data(), n * 4u) == (int)(n * 4u);
    gzclose(f);
    return ok ? VR_OK : VR_ERR_ARG;
}

}  // extern "C"
