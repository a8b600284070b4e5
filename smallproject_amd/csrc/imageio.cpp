// imageio.cpp - the reference's image metrics and raster dump format, and
// the PNG reader behind World::Print's glyph sheet.
//   .dat: u32 width, u32 height, float4 pixels[width*height] row-major
//         (saveScreenBuffer mygame.cpp:518-549, Image::Load mse.cpp:619-662)
//   MSE / "RMAE" (a mean absolute error): per-pixel float differences of the
//         rgb channels, summed in binary64, divided by 3N (mse.cpp:33-65)
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vr_world.h"
#include "world.h"

extern "C" {

int vrw_image_error(const float* gt, const float* img, uint32_t width, uint32_t height, double out[2])
{
    if (!gt || !img || !out || width == 0 || height == 0) return VR_ERR_ARG;
    const size_t n = (size_t)width * height;
    double se = 0.0, ae = 0.0;
    for (size_t i = 0; i < n; i++) {
        const float dx = gt[4 * i + 0] - img[4 * i + 0];
        const float dy = gt[4 * i + 1] - img[4 * i + 1];
        const float dz = gt[4 * i + 2] - img[4 * i + 2];
        se += (double)(dx * dx + dy * dy + dz * dz);
        ae += (double)(std::fabs(dx) + std::fabs(dy) + std::fabs(dz));
    }
    out[0] = se / (double)(n * 3);
    out[1] = ae / (double)(n * 3);
    return VR_OK;
}

int vrw_write_dat(const char* path, uint32_t width, uint32_t height, const float* rgba)
{
    if (!path || !rgba || width == 0 || height == 0) return VR_ERR_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return VR_ERR_ARG;
    const size_t n = (size_t)width * height;
    bool ok = std::fwrite(&width, 4, 1, f) == 1 && std::fwrite(&height, 4, 1, f) == 1 &&
              std::fwrite(rgba, 16, n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? VR_OK : VR_ERR_ARG;
}

int vrw_read_dat(const char* path, uint32_t* width, uint32_t* height, float* rgba, size_t capacity_px)
{
    if (!path || !width || !height) return VR_ERR_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return VR_ERR_ARG;
    uint32_t wh[2];
    if (std::fread(wh, 4, 2, f) != 2 || wh[0] == 0 || wh[1] == 0 || wh[0] > 65536 || wh[1] > 65536) {
        std::fclose(f);
        return VR_ERR_ARG;
    }
    *width = wh[0], *height = wh[1];
    int rc = VR_OK;
    if (rgba) {
        const size_t n = (size_t)wh[0] * wh[1];
        if (n > capacity_px || std::fread(rgba, 16, n, f) != n) rc = VR_ERR_ARG;
    }
    std::fclose(f);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// PNG reader for the World's glyph sheet (the reference reads it with
// stb_image in Surface::LoadImage, template/template.cpp:1776-1798): 8-bit
// grey, RGB or RGBA, not interlaced; chunks per the PNG specification
// (ISO/IEC 15948), the IDAT stream inflated with zlib and each row
// unfiltered (None, Sub, Up, Average, Paeth).
// ---------------------------------------------------------------------------
namespace vrhost {

static uint32_t be32(const unsigned char* p) { return ((uint32_t)p[0] << 24) | (p[1] << 16) | (p[2] << 8) | p[3]; }

bool ReadPngRgb(const char* path, uint32_t& w, uint32_t& h, std::vector<uint32_t>& px, std::string& err)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return false;
    }
    std::vector<unsigned char> d;
    unsigned char buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) d.insert(d.end(), buf, buf + got);
    std::fclose(f);
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (d.size() < 8 || std::memcmp(d.data(), sig, 8) != 0) {
        err = "not a PNG file";
        return false;
    }
    std::vector<unsigned char> idat;
    int depth = 0, type = -1, interlace = 0;
    w = h = 0;
    for (size_t i = 8; i + 12 <= d.size();) {
        const uint32_t len = be32(&d[i]);
        if (i + 12 + (size_t)len > d.size()) break;
        const unsigned char* t = &d[i + 4];
        const unsigned char* body = &d[i + 8];
        if (!std::memcmp(t, "IHDR", 4) && len >= 13) {
            w = be32(body), h = be32(body + 4), depth = body[8], type = body[9], interlace = body[12];
        } else if (!std::memcmp(t, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(t, "IEND", 4)) {
            break;
        }
        i += 12 + (size_t)len;
    }
    const int n = type == 0 ? 1 : type == 2 ? 3 : type == 6 ? 4 : 0;
    if (depth != 8 || n == 0 || interlace != 0 || w == 0 || h == 0 || w > 65536 || h > 65536) {
        err = "unsupported PNG layout (8-bit grey, RGB or RGBA, not interlaced)";
        return false;
    }
    const size_t stride = (size_t)w * n;
    std::vector<unsigned char> raw((stride + 1) * h);
    uLongf rawLen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rawLen, idat.data(), (uLong)idat.size()) != Z_OK || rawLen != raw.size()) {
        err = "corrupt PNG image data";
        return false;
    }
    std::vector<unsigned char> img(stride * h);
    for (uint32_t y = 0; y < h; y++) {
        const unsigned char filter = raw[y * (stride + 1)];
        const unsigned char* src = &raw[y * (stride + 1) + 1];
        unsigned char* row = &img[y * stride];
        const unsigned char* up = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t k = 0; k < stride; k++) {
            const int a = k >= (size_t)n ? row[k - n] : 0, b = up ? up[k] : 0,
                      c = (up && k >= (size_t)n) ? up[k - n] : 0;
            int pred = 0;
            switch (filter) {
            case 0: pred = 0; break;
            case 1: pred = a; break;
            case 2: pred = b; break;
            case 3: pred = (a + b) / 2; break;
            case 4: {
                const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: err = "bad PNG row filter"; return false;
            }
            row[k] = (unsigned char)(src[k] + pred);
        }
    }
    px.resize((size_t)w * h);
    for (size_t i = 0; i < px.size(); i++) {
        const unsigned char* p = &img[i * n];
        px[i] = n == 1 ? (uint32_t)p[0] * 0x010101u : ((uint32_t)p[0] << 16) | ((uint32_t)p[1] << 8) | p[2];
    }
    return true;
}

}  // namespace vrhost
