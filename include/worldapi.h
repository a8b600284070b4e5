// worldapi.h - the reference's free-function world API (template/worldapi.h,
// implemented in template/template.cpp:50-130) over a current World, so that
// scene code written against WrldTmpl8 (e.g. letterscene.cpp) builds against
// this library.  SetWorld() names the World the functions act on (the
// reference's is the one its main() creates).  Sprites, tiles, particles,
// scrolling and the editor are not part of the render path and are not
// provided; a .vx sprite can be stamped with vrw_load_vx (vr_world.h).
#pragma once
#include <cstdint>

#include "world.h"

namespace vrhost {

World* GetWorld();
void SetWorld(World* w);

// template/template.cpp:52-63
inline void ClearWorld() { GetWorld()->Clear(); }
inline void FillWorld(const uint32_t c) { GetWorld()->Fill(c); }
inline void Plot(const uint32_t x, const uint32_t y, const uint32_t z, const uint32_t c) { GetWorld()->Set(x, y, z, c); }
inline void Plot(const int3 pos, const uint32_t c) { GetWorld()->Set(pos.x, pos.y, pos.z, c); }
inline uint32_t Read(const int x, const int y, const int z) { return GetWorld()->Get(x, y, z); }
inline uint32_t Read(const int3 pos) { return GetWorld()->Get(pos.x, pos.y, pos.z); }

// template/template.cpp:76-130
inline void Sphere(const float x, const float y, const float z, const float r, const uint32_t c)
{
    GetWorld()->Sphere(x, y, z, r, c);
}
inline void Sphere(const float3 pos, const float r, const uint32_t c) { GetWorld()->Sphere(pos.x, pos.y, pos.z, r, c); }
// x1..x2-1 by y1..y2-1 by z1..z2-1
void Box(const int x1, const int y1, const int z1, const int x2, const int y2, const int z2, const uint32_t c);
inline void Box(const int3 pos1, const int3 pos2, const uint32_t c) { Box(pos1.x, pos1.y, pos1.z, pos2.x, pos2.y, pos2.z, c); }
// the box s1..s2 (inclusive) read voxel by voxel and plotted at D onward, in
// z, y, x order
void Copy(const int3 s1, const int3 s2, const int3 D);
inline void Copy(const int3 s1, const int3 s2, const int x, const int y, const int z) { Copy(s1, s2, int3{x, y, z}); }
inline void Copy(const int x1, const int y1, const int z1, const int x2, const int y2, const int z2, const int3 D)
{
    Copy(int3{x1, y1, z1}, int3{x2, y2, z2}, D);
}
inline void HDisc(const float x, const float y, const float z, const float r, const uint32_t c)
{
    GetWorld()->HDisc(x, y, z, r, c);
}
inline void HDisc(const float3 pos, const float r, const uint32_t c) { GetWorld()->HDisc(pos.x, pos.y, pos.z, r, c); }
inline void Print(const char* text, const uint32_t x, const uint32_t y, const uint32_t z, const uint32_t c)
{
    GetWorld()->Print(text, x, y, z, c);
}
inline void Print(const char* text, const int3 pos, const uint32_t c) { GetWorld()->Print(text, pos.x, pos.y, pos.z, c); }

// LookAt (template/template.cpp:204-215)
inline void LookAt(const float3 pos, const float3 target) { GetWorld()->SetCameraMatrix(mat4::LookAt(pos, target)); }

}  // namespace vrhost
