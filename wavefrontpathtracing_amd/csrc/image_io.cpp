// Image files for the renderer's outputs (host only): the linear energy or
// accumulator as PFM (three 32-bit floats per pixel, bottom row first, the
// byte order flagged by a negative scale for little endian), and the
// tone-mapped RGBA8 output as binary PPM (alpha dropped).  The reference
// presents its output to a swapchain (renderer.cpp:1017-1048); files are how
// this library hands images to a caller without a display.
#include "wf_host.h"

#include <cstdio>
#include <vector>

extern "C" WF_API wf_status wf_write_pfm(const char* path, const float* rgba, uint32_t width, uint32_t height)
{
    if (!path || !rgba || width == 0 || height == 0) return WF_ERROR_INVALID_ARGUMENT;
    FILE* f = std::fopen(path, "wb");
    if (!f) return WF_ERROR_INVALID_ARGUMENT;
    std::fprintf(f, "PF\n%u %u\n-1.0\n", width, height);
    std::vector<float> row(3 * (size_t)width);
    bool ok = true;
    for (uint32_t y = height; y-- > 0 && ok;) {        // PFM rows run bottom to top
        const float* src = rgba + 4 * (size_t)y * width;
        for (uint32_t x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) row[3 * (size_t)x + c] = src[4 * (size_t)x + c];
        ok = std::fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? WF_OK : WF_ERROR_DEVICE;
}

extern "C" WF_API wf_status wf_write_ppm(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height)
{
    if (!path || !rgba8 || width == 0 || height == 0) return WF_ERROR_INVALID_ARGUMENT;
    FILE* f = std::fopen(path, "wb");
    if (!f) return WF_ERROR_INVALID_ARGUMENT;
    std::fprintf(f, "P6\n%u %u\n255\n", width, height);
    std::vector<uint8_t> row(3 * (size_t)width);
    bool ok = true;
    for (uint32_t y = 0; y < height && ok; ++y) {
        const uint8_t* src = rgba8 + 4 * (size_t)y * width;
        for (uint32_t x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) row[3 * (size_t)x + c] = src[4 * (size_t)x + c];
        ok = std::fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? WF_OK : WF_ERROR_DEVICE;
}
