// imageio.cpp - the reference's image metrics and raster dump format.
//   .dat: u32 width, u32 height, float4 pixels[width*height] row-major
//         (saveScreenBuffer mygame.cpp:518-549, Image::Load mse.cpp:619-662)
//   MSE / "RMAE" (a mean absolute error): per-pixel float differences of the
//         rgb channels, summed in binary64, divided by 3N (mse.cpp:33-65)
#include <cmath>
#include <cstdio>
#include <cstring>

#include "vr_world.h"

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
