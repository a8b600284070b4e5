// scenes.cpp - seeded scene generators and camera paths for the benchmark
// configurations.  They play the role of the reference's scene .cpp files
// (mountainscene.cpp, scatterscene.cpp, ...): callers of the World API that
// build a world voxel by voxel, with emission in the 0x00MM0000 byte.
#include <algorithm>
#include <cmath>

#include "vr_world.h"
#include "world.h"

namespace {

uint32_t hash32(uint32_t s)
{
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    return s ^ (s >> 15);
}

// xorshift32 stream seeded through a hash so that small seeds differ
struct Rng {
    uint32_t s;
    explicit Rng(uint32_t seed) : s(hash32(seed * 2654435761u + 0x9e3779b9u) | 1u) {}
    uint32_t u32()
    {
        s ^= s << 13, s ^= s >> 17, s ^= s << 5;
        return s;
    }
    double uniform() { return (double)u32() * (1.0 / 4294967296.0); }
    uint32_t range(uint32_t lo, uint32_t hi) { return lo + (uint32_t)(uniform() * (double)(hi - lo + 1)); }
};

double lattice(int i, int j, uint32_t seed)
{
    const uint32_t h = hash32((uint32_t)i * 73856093u ^ hash32((uint32_t)j * 19349663u ^ seed * 83492791u));
    return (double)(h >> 8) * (2.0 / 16777216.0) - 1.0;
}

// value noise with cosine interpolation between lattice values
// (the interpolation of template/template.cpp:975-979)
double vnoise(double u, double v, uint32_t seed)
{
    const double fi = std::floor(u), fj = std::floor(v);
    const int i = (int)fi, j = (int)fj;
    const double fu = (1.0 - std::cos((u - fi) * 3.14159265358979323846)) * 0.5;
    const double fv = (1.0 - std::cos((v - fj) * 3.14159265358979323846)) * 0.5;
    const double a = lattice(i, j, seed), b = lattice(i + 1, j, seed);
    const double c = lattice(i, j + 1, seed), d = lattice(i + 1, j + 1, seed);
    const double ab = a * (1 - fu) + b * fu, cd = c * (1 - fu) + d * fu;
    return ab * (1 - fv) + cd * fv;
}

// RGB332 byte -> 4-bit channels with A=F (0x0000FRGB)
uint32_t rgb332_to_4444(uint32_t c)
{
    const uint32_t r3 = (c >> 5) & 7u, g3 = (c >> 2) & 7u, b2 = c & 3u;
    const uint32_t r4 = (r3 * 15u + 3u) / 7u, g4 = (g3 * 15u + 3u) / 7u, b4 = b2 * 5u;
    return 0xf000u | (r4 << 8) | (g4 << 4) | b4;
}

uint32_t terrain_color(uint32_t y, float h, uint32_t scale)
{
    const float top = h - 2.0f * (float)scale;
    if ((float)y < top) return rgb332_to_4444(0x8du);       // dirt (4,3,1)
    const float rel = h / (float)scale;                       // 16 .. 64
    if (rel < 34.0f) return rgb332_to_4444(0x55u);            // grass (2,5,1)
    if (rel < 52.0f) return rgb332_to_4444(0x92u);            // rock (4,4,2)
    return rgb332_to_4444(0xffu);                             // snow
}

}  // namespace

extern "C" {

float vrw_terrain_height(float x, float z, uint32_t scale, uint32_t seed)
{
    const double s = (double)(scale ? scale : 1u);
    const double u = (double)x / s, v = (double)z / s;
    double n = 0, amp = 1, freq = 1.0 / 32.0, norm = 0;
    for (int o = 0; o < 4; o++) {
        n += amp * vnoise(u * freq, v * freq, seed + (uint32_t)o * 1013u);
        norm += amp;
        amp *= 0.5, freq *= 2.0;
    }
    return (float)(s * (40.0 + 24.0 * (n / norm)));
}

int vrw_scene_terrain(vrw_world* wp, uint32_t size, uint32_t scale, uint32_t seed)
{
    if (!wp || size == 0 || size > 1024) return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    std::vector<float> hs((size_t)size * size);
    for (uint32_t z = 0; z < size; z++)
        for (uint32_t x = 0; x < size; x++) hs[(size_t)z * size + x] = vrw_terrain_height((float)x, (float)z, scale, seed);
    const uint32_t dirt = rgb332_to_4444(0x8du);
    for (uint32_t bz = 0; bz * 8 < size; bz++)
        for (uint32_t bx = 0; bx * 8 < size; bx++) {
            // cells wholly below the coloured top layers of all their columns
            // are uniform dirt: one solid cell each (World::AddBrick)
            float hmin = 1e30f;
            for (uint32_t z = bz * 8; z < std::min(size, bz * 8 + 8); z++)
                for (uint32_t x = bx * 8; x < std::min(size, bx * 8 + 8); x++) hmin = std::min(hmin, hs[(size_t)z * size + x]);
            const float solidTop = hmin - 2.0f * (float)scale;
            uint32_t by = 0;
            const bool full = (bx * 8 + 8 <= size) && (bz * 8 + 8 <= size);
            while (full && (float)(by * 8 + 8) <= solidTop && by < 128) w.AddBrick(bx, by++, bz, dirt);
            for (uint32_t z = bz * 8; z < std::min(size, bz * 8 + 8); z++)
                for (uint32_t x = bx * 8; x < std::min(size, bx * 8 + 8); x++) {
                    const float h = hs[(size_t)z * size + x];
                    for (uint32_t y = by * 8; (float)y < h && y < 1024; y++) w.Set(x, y, z, terrain_color(y, h, scale));
                }
        }
    return VR_OK;
}

int vrw_scene_boxes(vrw_world* wp, uint32_t n, uint32_t emin, uint32_t emax, uint32_t size, uint32_t scale,
                    uint32_t terrain_seed, uint32_t seed)
{
    if (!wp || emin == 0 || emax < emin || size <= emax) return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    Rng rng(seed);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t ex = rng.range(emin, emax), ey = rng.range(emin, emax), ez = rng.range(emin, emax);
        const uint32_t x = rng.range(0, size - ex - 1), z = rng.range(0, size - ez - 1);
        const float h = scale ? vrw_terrain_height((float)(x + ex / 2), (float)(z + ez / 2), scale, terrain_seed) : 0.0f;
        const uint32_t sink = rng.range(0, ey / 2);
        const int y = std::max(0, (int)h - (int)sink);
        uint32_t c = rgb332_to_4444(rng.range(1, 255));
        for (uint32_t yy = (uint32_t)y; yy < (uint32_t)y + ey && yy < 1024; yy++)
            for (uint32_t zz = z; zz < z + ez; zz++)
                for (uint32_t xx = x; xx < x + ex; xx++) w.Set(xx, yy, zz, c);
    }
    return VR_OK;
}

int vrw_scene_lights_above(vrw_world* wp, uint32_t n, uint32_t size, uint32_t scale, uint32_t terrain_seed,
                           uint32_t seed, uint32_t hmin, uint32_t hmax, uint32_t mmin, uint32_t mmax)
{
    if (!wp || size == 0 || hmax < hmin || mmin == 0 || mmax < mmin || mmax > 255) return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    Rng rng(seed);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t x = rng.range(0, size - 1), z = rng.range(0, size - 1);
        const float h = vrw_terrain_height((float)x, (float)z, scale, terrain_seed);
        const uint32_t y = (uint32_t)std::ceil(h) + rng.range(hmin, hmax);
        const uint32_t c = rgb332_to_4444(rng.range(1, 255)), m = rng.range(mmin, mmax);
        w.Set(x, y, z, (m << 16) | c);
    }
    return VR_OK;
}

int vrw_scene_lights_surface(vrw_world* wp, uint32_t n, uint32_t size, uint32_t seed, uint32_t mmin, uint32_t mmax)
{
    if (!wp || size == 0 || mmin == 0 || mmax < mmin || mmax > 255) return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    Rng rng(seed);
    uint32_t placed = 0, tries = 0;
    while (placed < n && tries < n * 64u) {
        tries++;
        const uint32_t x = rng.range(0, size - 1), z = rng.range(0, size - 1);
        int y = 1023;
        while (y >= 0 && w.Get(x, (uint32_t)y, z) == 0u) y--;
        if (y < 0 || ((w.Get(x, (uint32_t)y, z) >> 16) & 255u)) continue;
        const uint32_t c = rgb332_to_4444(rng.range(1, 255)), m = rng.range(mmin, mmax);
        w.Set(x, (uint32_t)y, z, (m << 16) | c);
        placed++;
    }
    return placed == n ? VR_OK : VR_ERR_ARG;
}

int vrw_scene_floor_boxes(vrw_world* wp, uint32_t n, uint32_t emin, uint32_t emax, uint32_t size, uint32_t floor_h,
                          uint32_t seed)
{
    // a flat floor of height floor_h over [0,size)^2 (solid cells) and n
    // boxes with edges uniform in [emin, emax] standing on it
    if (!wp || emin == 0 || emax < emin || size <= emax || size > 1024 || floor_h == 0 || floor_h % 8 != 0)
        return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    const uint32_t floorc = rgb332_to_4444(0x6du);
    for (uint32_t bz = 0; bz < size / 8; bz++)
        for (uint32_t bx = 0; bx < size / 8; bx++)
            for (uint32_t by = 0; by < floor_h / 8; by++) w.AddBrick(bx, by, bz, floorc);
    Rng rng(seed);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t ex = rng.range(emin, emax), ey = rng.range(emin, emax), ez = rng.range(emin, emax);
        const uint32_t x = rng.range(0, size - ex - 1), z = rng.range(0, size - ez - 1);
        const uint32_t c = rgb332_to_4444(rng.range(1, 255));
        for (uint32_t yy = floor_h; yy < floor_h + ey; yy++)
            for (uint32_t zz = z; zz < z + ez; zz++)
                for (uint32_t xx = x; xx < x + ex; xx++) w.Set(xx, yy, zz, c);
    }
    return VR_OK;
}

int vrw_scene_light_clusters(vrw_world* wp, uint32_t n, uint32_t nclusters, float sigma, uint32_t size,
                             uint32_t ymin, uint32_t ymax, uint32_t seed, uint32_t mmin, uint32_t mmax)
{
    // n emissive voxels in nclusters Gaussian clusters (centres uniform over
    // [0,size)^2 x [ymin, ymax]), placed in empty voxels only
    if (!wp || nclusters == 0 || !(sigma > 0.0f) || size == 0 || ymax <= ymin || mmin == 0 || mmax < mmin ||
        mmax > 255)
        return VR_ERR_ARG;
    vrhost::World& w = *reinterpret_cast<vrhost::World*>(wp);
    Rng rng(seed);
    std::vector<double> cx(nclusters), cy(nclusters), cz(nclusters);
    for (uint32_t k = 0; k < nclusters; k++)
        cx[k] = rng.uniform() * size, cz[k] = rng.uniform() * size, cy[k] = ymin + rng.uniform() * (ymax - ymin);
    auto gauss = [&]() {
        // Box-Muller
        const double u1 = std::max(rng.uniform(), 1e-12), u2 = rng.uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    uint32_t placed = 0, tries = 0;
    while (placed < n && tries < n * 64u) {
        tries++;
        const uint32_t k = placed % nclusters;
        const double x = cx[k] + sigma * gauss(), y = cy[k] + sigma * gauss(), z = cz[k] + sigma * gauss();
        if (x < 0 || z < 0 || x >= size || z >= size || y < ymin || y >= 1024) continue;
        const uint32_t ix = (uint32_t)x, iy = (uint32_t)y, iz = (uint32_t)z;
        if (w.Get(ix, iy, iz) != 0u) continue;
        const uint32_t c = rgb332_to_4444(rng.range(1, 255)), m = rng.range(mmin, mmax);
        w.Set(ix, iy, iz, (m << 16) | c);
        placed++;
    }
    return placed == n ? VR_OK : VR_ERR_ARG;
}

int vrw_cluster_lights(const vrw_world* wp, uint32_t k, uint32_t iters, uint32_t seed, uint32_t* out)
{
    // k-means over light positions: k-means++ seeding, then Lloyd iterations
    if (!wp || !out || k == 0) return VR_ERR_ARG;
    const vrhost::World& w = *reinterpret_cast<const vrhost::World*>(wp);
    const auto& L = w.GetLights();
    const size_t n = L.size();
    if (n == 0) return VR_OK;
    std::vector<double> px(n), py(n), pz(n);
    for (size_t i = 0; i < n; i++) {
        const uint32_t p = L[i].position;
        px[i] = (double)(p % 1024u), pz[i] = (double)((p / 1024u) % 1024u), py[i] = (double)(p / (1024u * 1024u));
    }
    k = (uint32_t)std::min<size_t>(k, n);
    Rng rng(seed);
    std::vector<double> cx, cy, cz, d2(n, 1e300);
    size_t first = (size_t)(rng.uniform() * n) % n;
    cx.push_back(px[first]), cy.push_back(py[first]), cz.push_back(pz[first]);
    while (cx.size() < k) {
        double total = 0;
        for (size_t i = 0; i < n; i++) {
            const double dx = px[i] - cx.back(), dy = py[i] - cy.back(), dz = pz[i] - cz.back();
            d2[i] = std::min(d2[i], dx * dx + dy * dy + dz * dz);
            total += d2[i];
        }
        double r = rng.uniform() * total;
        size_t pick = n - 1;
        for (size_t i = 0; i < n; i++) {
            r -= d2[i];
            if (r <= 0) {
                pick = i;
                break;
            }
        }
        cx.push_back(px[pick]), cy.push_back(py[pick]), cz.push_back(pz[pick]);
    }
    std::vector<double> sx(k), sy(k), sz(k);
    std::vector<uint32_t> cnt(k);
    for (uint32_t it = 0; it <= iters; it++) {
        for (size_t i = 0; i < n; i++) {
            double best = 1e300;
            uint32_t bk = 0;
            for (uint32_t c = 0; c < k; c++) {
                const double dx = px[i] - cx[c], dy = py[i] - cy[c], dz = pz[i] - cz[c];
                const double d = dx * dx + dy * dy + dz * dz;
                if (d < best) best = d, bk = c;
            }
            out[i] = bk;
        }
        if (it == iters) break;
        std::fill(sx.begin(), sx.end(), 0.0), std::fill(sy.begin(), sy.end(), 0.0), std::fill(sz.begin(), sz.end(), 0.0);
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (size_t i = 0; i < n; i++) sx[out[i]] += px[i], sy[out[i]] += py[i], sz[out[i]] += pz[i], cnt[out[i]]++;
        for (uint32_t c = 0; c < k; c++)
            if (cnt[c]) cx[c] = sx[c] / cnt[c], cy[c] = sy[c] / cnt[c], cz[c] = sz[c] / cnt[c];
    }
    return VR_OK;
}

void vrw_log_lut(float lo, float hi, float lut[256])
{
    lut[0] = 0.0f;
    for (int m = 1; m < 256; m++) lut[m] = (float)((double)lo * std::pow((double)hi / (double)lo, (m - 1) / 254.0));
}

int vrw_camera_path(uint32_t kind, uint32_t seed, uint32_t frame, uint32_t nframes, uint32_t size, uint32_t scale,
                    uint32_t terrain_seed, float pos[3], float target[3])
{
    if (!pos || !target || size == 0) return VR_ERR_ARG;
    Rng rng(seed);
    const double half = (double)size * 0.5;
    if (kind == 0) {
        // straight line from a seeded corner region toward the centre,
        // 2 voxels per frame, looking ahead and down
        const double x0 = 0.1 * size + rng.uniform() * 0.15 * size, z0 = 0.1 * size + rng.uniform() * 0.15 * size;
        double dx = half - x0, dz = half - z0;
        const double l = std::sqrt(dx * dx + dz * dz);
        dx /= l, dz /= l;
        const double y = (64.0 + 26.0) * (scale ? scale : 1u);
        const double px = x0 + 2.0 * frame * dx, pz = z0 + 2.0 * frame * dz;
        pos[0] = (float)px, pos[1] = (float)y, pos[2] = (float)pz;
        target[0] = (float)(px + 0.5 * size * dx), target[1] = (float)(y - 0.4 * size), target[2] = (float)(pz + 0.5 * size * dz);
        return VR_OK;
    }
    if (kind == 1) {
        // Catmull-Rom spline through 8 seeded control points above the terrain
        double P[8][3];
        for (int k = 0; k < 8; k++) {
            const double x = 0.2 * size + rng.uniform() * 0.6 * size, z = 0.2 * size + rng.uniform() * 0.6 * size;
            const double h = scale ? vrw_terrain_height((float)x, (float)z, scale, terrain_seed) : 0.0;
            P[k][0] = x, P[k][1] = h + (0.08 + 0.12 * rng.uniform()) * size, P[k][2] = z;
        }
        auto eval = [&](double t, double out[3]) {
            t = std::min(std::max(t, 0.0), 4.999999);
            const int i = (int)t;
            const double u = t - i, u2 = u * u, u3 = u2 * u;
            for (int c = 0; c < 3; c++) {
                const double p0 = P[i][c], p1 = P[i + 1][c], p2 = P[i + 2][c], p3 = P[i + 3][c];
                out[c] = 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u2 + (-p0 + 3 * p1 - 3 * p2 + p3) * u3);
            }
        };
        const double t = nframes > 1 ? 5.0 * frame / (double)(nframes - 1) : 0.0;
        double a[3], b[3];
        eval(t, a);
        eval(t + 0.25, b);
        if (t + 0.25 > 4.999999) {  // past the end: keep looking along the last tangent
            double c0[3];
            eval(t - 0.25, c0);
            for (int c = 0; c < 3; c++) b[c] = a[c] + (a[c] - c0[c]);
        }
        pos[0] = (float)a[0], pos[1] = (float)a[1], pos[2] = (float)a[2];
        target[0] = (float)b[0], target[1] = (float)(b[1] - 0.06 * size), target[2] = (float)b[2];
        return VR_OK;
    }
    return VR_ERR_ARG;
}

}  // extern "C"
