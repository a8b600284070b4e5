// nbtio.cpp - the world editor's world file (worlddata.nbt): load and save
// (WorldEditor::LoadWorld / SaveWorld, worldeditor.cpp:1347-1600, tag
// encoding worldeditor.cpp:13-172).
//
// The file is a little-endian NBT variant: a tag is a type byte, for named
// tags an i16 name length and the name, then the payload (TAG_Int: 4 bytes;
// TAG_Byte_Array: i32 length + bytes; TAG_List: element type byte, i32
// count, then the elements as type byte + payload without names;
// TAG_Compound: named tags up to TAG_End).  The top level holds
//   TAG_List "grid list":  compounds {TAG_Int "grid index", TAG_Int "grid value"}
//   TAG_List "brick list": compounds {TAG_Int "brick buffer offset",
//                           TAG_Byte_Array "brick value" (512 u32),
//                           TAG_Int "brick zeroes"}
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "vr_world.h"
#include "world.h"

namespace {

enum { TAG_End = 0, TAG_Int = 3, TAG_Byte_Array = 7, TAG_List = 9, TAG_Compound = 10 };

struct Tag {
    int type = TAG_End;
    std::string name;
    std::vector<uint8_t> payload;
    std::vector<Tag> tags;
    uint32_t u32() const
    {
        return payload.size() < 4 ? 0u
                                  : (uint32_t)payload[0] | (uint32_t)payload[1] << 8 | (uint32_t)payload[2] << 16 |
                                        (uint32_t)payload[3] << 24;
    }
};

struct Reader {
    const std::vector<uint8_t>& d;
    size_t pos = 0;
    bool ok = true;
    bool take(void* out, size_t n)
    {
        if (!ok || pos + n > d.size()) return ok = false;
        std::memcpy(out, d.data() + pos, n);
        pos += n;
        return true;
    }
    int32_t i32()
    {
        int32_t v = 0;
        take(&v, 4);
        return v;
    }
    // one tag; depth bounds nesting of malformed files
    bool tag(Tag& t, bool named, int depth)
    {
        if (depth > 8) return ok = false;
        uint8_t ty = 0;
        if (!take(&ty, 1)) return false;
        t.type = ty;
        if (ty == TAG_End) return true;
        if (named) {
            int16_t len = 0;
            if (!take(&len, 2) || len < 0) return ok = false;
            t.name.resize((size_t)len);
            if (len && !take(&t.name[0], (size_t)len)) return false;
        }
        switch (ty) {
        case TAG_Int:
            t.payload.resize(4);
            return take(t.payload.data(), 4);
        case TAG_Byte_Array: {
            const int32_t n = i32();
            if (!ok || n < 0 || pos + (size_t)n > d.size()) return ok = false;
            t.payload.assign(d.begin() + pos, d.begin() + pos + n);
            pos += (size_t)n;
            return true;
        }
        case TAG_List: {
            uint8_t et = 0;
            if (!take(&et, 1)) return false;
            t.payload.push_back(et);
            const int32_t n = i32();
            if (!ok || n < 0 || (size_t)n > d.size()) return ok = false;
            t.tags.resize((size_t)n);
            for (auto& e : t.tags)
                if (!tag(e, false, depth + 1)) return false;
            return true;
        }
        case TAG_Compound:
            while (true) {
                t.tags.emplace_back();
                if (!tag(t.tags.back(), true, depth + 1)) return false;
                if (t.tags.back().type == TAG_End) return true;
            }
        default:
            return ok = false;  // the world file uses no other tag types
        }
    }
};

void put(std::vector<uint8_t>& o, const void* p, size_t n)
{
    const uint8_t* b = static_cast<const uint8_t*>(p);
    o.insert(o.end(), b, b + n);
}

void put_name(std::vector<uint8_t>& o, uint8_t type, const char* name)
{
    o.push_back(type);
    const int16_t len = (int16_t)std::strlen(name);
    put(o, &len, 2);
    put(o, name, (size_t)len);
}

void put_int(std::vector<uint8_t>& o, const char* name, uint32_t v)
{
    put_name(o, TAG_Int, name);
    put(o, &v, 4);
}

}  // namespace

extern "C" {

int vrw_load_nbt(vrw_world* wp, const char* path, uint32_t* stats)
{
    if (!wp || !path) return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    std::vector<uint8_t> data;
    {
        FILE* f = std::fopen(path, "rb");
        if (!f) return VR_ERR_ARG;
        uint8_t buf[65536];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + n);
        std::fclose(f);
    }
    Reader r{data};
    std::vector<Tag> top;
    while (r.pos < data.size()) {
        top.emplace_back();
        if (!r.tag(top.back(), true, 0)) return VR_ERR_ARG;
    }
    // validate before touching the world
    std::vector<std::pair<uint32_t, uint32_t>> cells;
    struct Brick {
        uint32_t offset, zeroes;
        const std::vector<uint8_t>* value;
    };
    std::vector<Brick> bricks;
    for (const Tag& t : top) {
        if (t.type != TAG_List) continue;
        for (const Tag& c : t.tags) {
            if (c.type != TAG_Compound) return VR_ERR_ARG;
            if (t.name == "grid list") {
                uint32_t index = 0, value = 0;
                for (const Tag& e : c.tags) {
                    if (e.type != TAG_Int) continue;
                    if (e.name == "grid index") index = e.u32();
                    if (e.name == "grid value") value = e.u32();
                }
                if (index >= VR_GRIDSIZE) return VR_ERR_ARG;
                cells.emplace_back(index, value);
            } else if (t.name == "brick list") {
                Brick b{0, 0, nullptr};
                for (const Tag& e : c.tags) {
                    if (e.type == TAG_Int && e.name == "brick buffer offset") b.offset = e.u32();
                    if (e.type == TAG_Int && e.name == "brick zeroes") b.zeroes = e.u32();
                    if (e.type == TAG_Byte_Array && e.name == "brick value") b.value = &e.payload;
                }
                if (!b.value || b.value->size() != VR_BRICKSIZE * 4 || b.offset >= w.BrickCapacity())
                    return VR_ERR_ARG;
                bricks.push_back(b);
            }
        }
    }
    w.Clear();
    for (const auto& c : cells) w.SetGridCell(c.first, c.second);
    std::vector<uint32_t> payload(VR_BRICKSIZE);
    for (const Brick& b : bricks) {
        std::memcpy(payload.data(), b.value->data(), VR_BRICKSIZE * 4);
        w.LoadBrick(b.offset, payload.data(), b.zeroes);
    }
    // lights of the loaded world, then uniform bricks become solid cells
    // (worldeditor.cpp:1439-1451); the device copy follows on Commit or
    // ForceSyncAllBricks
    std::vector<vrhost::Light> ls;
    w.FindLightsInWord(ls);
    w.SetupLightBuffer(ls);
    w.OptimizeBricks();
    if (stats) stats[0] = (uint32_t)cells.size(), stats[1] = (uint32_t)bricks.size(), stats[2] = (uint32_t)ls.size();
    return VR_OK;
}

int vrw_save_nbt(const vrw_world* wp, const char* path)
{
    if (!wp || !path) return VR_ERR_ARG;
    const vrhost::World& w = *reinterpret_cast<const vrhost::World*>(wp);
    const uint32_t* grid = w.GetGrid();
    const uint32_t* brick = w.GetBrick();
    const uint32_t* zeroes = w.GetZeroes();
    // non-empty cells in index order; brick ids renumbered 0.. in order of
    // first use, bricks written in ascending old id (worldeditor.cpp:1511-1600)
    std::map<uint32_t, uint32_t> oldToNew;
    std::set<uint32_t> oldIds;
    std::vector<std::pair<uint32_t, uint32_t>> cells;
    for (uint32_t i = 0; i < VR_GRIDSIZE; i++) {
        const uint32_t g = grid[i];
        if (g == 0) continue;
        if (g & 1u) {
            const uint32_t old = g >> 1;
            if (old >= w.BrickCapacity()) return VR_ERR_STATE;
            auto it = oldToNew.find(old);
            if (it == oldToNew.end()) it = oldToNew.emplace(old, (uint32_t)oldToNew.size()).first;
            oldIds.insert(old);
            cells.emplace_back(i, (it->second << 1) | 1u);
        } else {
            cells.emplace_back(i, g);
        }
    }
    std::vector<uint8_t> o;
    put_name(o, TAG_List, "grid list");
    o.push_back(TAG_Compound);
    int32_t n = (int32_t)cells.size();
    put(o, &n, 4);
    for (const auto& c : cells) {
        o.push_back(TAG_Compound);
        put_int(o, "grid index", c.first);
        put_int(o, "grid value", c.second);
        o.push_back(TAG_End);
    }
    put_name(o, TAG_List, "brick list");
    o.push_back(TAG_Compound);
    n = (int32_t)oldIds.size();
    put(o, &n, 4);
    for (const uint32_t old : oldIds) {
        o.push_back(TAG_Compound);
        put_int(o, "brick buffer offset", oldToNew[old]);
        put_name(o, TAG_Byte_Array, "brick value");
        const int32_t bytes = VR_BRICKSIZE * 4;
        put(o, &bytes, 4);
        put(o, brick + (size_t)old * VR_BRICKSIZE, VR_BRICKSIZE * 4);
        put_int(o, "brick zeroes", zeroes[old]);
        o.push_back(TAG_End);
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return VR_ERR_ARG;
    const bool good = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    return (std::fclose(f) == 0 && good) ? VR_OK : VR_ERR_ARG;
}

}  // extern "C"
