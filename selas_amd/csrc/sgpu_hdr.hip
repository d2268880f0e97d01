// Framebuffers to Radiance .hdr files: the replacement of FrameBuffer_Save
// (Source/Core/TextureLib/Framebuffer.cpp:43-59), which writes each layer with stb_image_write's
// HDR writer (StbImage.cpp:64-78). The file is the Radiance picture format: the header
// "#?RADIANCE", "FORMAT=32-bit_rle_rgbe", a blank line and the resolution "-Y height +X width",
// then one scanline per row, top row first. A pixel is Ward's shared-exponent RGBE: with m the
// largest component, frexp(m) = f 2^e, the bytes are floor(c f 256 / m) per component and e + 128
// (all zero for m < 1e-32). Scanlines of 8 to 32767 pixels are run-length coded per component
// (the marker 2 2 width_hi width_lo, then for each of the four byte planes: runs of 3 to 127 equal
// bytes as 128 + n, byte; literal stretches of up to 128 bytes as n, bytes); other widths are
// written flat. The RGBE conversion runs on the device, 4 bytes per pixel come back instead of 12.
#include "sgpu_internal.h"

#include <cstdio>
#include <string>
#include <vector>

namespace {

__global__ void k_rgbe(const float* __restrict rgb, size_t n, uint32_t* __restrict out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
    const float m = fmaxf(r, fmaxf(g, b));
    uint32_t word = 0;
    if (m >= 1e-32f) {
        int e;
        const float scale = frexpf(m, &e) * 256.0f / m;
        const uint32_t cr = (uint32_t)(unsigned char)(int)(r * scale);
        const uint32_t cg = (uint32_t)(unsigned char)(int)(g * scale);
        const uint32_t cb = (uint32_t)(unsigned char)(int)(b * scale);
        word = cr | (cg << 8) | (cb << 16) | ((uint32_t)(unsigned char)(e + 128) << 24);
    }
    out[i] = word;
}

// One byte plane of a scanline, run-length coded.
void rle_plane(const uint8_t* p, uint32_t w, std::vector<uint8_t>& out)
{
    uint32_t x = 0;
    while (x < w) {
        // the next run of at least 3 equal bytes
        uint32_t run_start = x;
        uint32_t run_len = 0;
        while (run_start < w) {
            run_len = 1;
            while (run_start + run_len < w && run_len < 127 && p[run_start + run_len] == p[run_start]) ++run_len;
            if (run_len >= 3) break;
            run_start += run_len;
            run_len = 0;
        }
        // the literal bytes before it
        while (x < run_start) {
            const uint32_t n = std::min<uint32_t>(128, run_start - x);
            out.push_back((uint8_t)n);
            out.insert(out.end(), p + x, p + x + n);
            x += n;
        }
        if (run_len >= 3) {
            out.push_back((uint8_t)(128 + run_len));
            out.push_back(p[run_start]);
            x = run_start + run_len;
        }
    }
}

}  // namespace

int sgpu_framebuffer_save_hdr(const char* path_prefix, const float* framebuffer, uint32_t width, uint32_t height,
                              uint32_t layers, void* stream)
{
    if (!path_prefix || !framebuffer) return sgpu_fail(SGPU_E_INVALID, "sgpu_framebuffer_save_hdr: NULL argument");
    if (width == 0 || height == 0 || layers == 0 || (uint64_t)width * height > (1ull << 28))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_framebuffer_save_hdr: empty or too large framebuffer");
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)width * height;
    uint32_t* d = nullptr;
    SGPU_HIP_TRY(hipMallocAsync((void**)&d, n * sizeof(uint32_t), st));
    std::vector<uint32_t> px(n);
    std::vector<uint8_t> line, planes(4 * (size_t)width);
    int rc = SGPU_OK;
    for (uint32_t layer = 0; layer < layers && rc == SGPU_OK; ++layer) {
        hipLaunchKernelGGL(k_rgbe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, framebuffer + 3 * n * layer, n, d);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(px.data(), d, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            rc = sgpu_fail(SGPU_E_HIP, std::string("sgpu_framebuffer_save_hdr: ") + hipGetErrorString(e));
            break;
        }
        const std::string path = std::string(path_prefix) + "_" + std::to_string(layer) + ".hdr";
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) {
            rc = sgpu_fail(SGPU_E_INVALID, "sgpu_framebuffer_save_hdr: cannot open " + path);
            break;
        }
        std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", height, width);
        const bool rle = width >= 8 && width < 32768;
        for (uint32_t y = 0; y < height; ++y) {
            const uint32_t* row = px.data() + (size_t)y * width;
            line.clear();
            if (rle) {
                line.push_back(2);
                line.push_back(2);
                line.push_back((uint8_t)(width >> 8));
                line.push_back((uint8_t)(width & 255));
                for (uint32_t c = 0; c < 4; ++c) {
                    for (uint32_t x = 0; x < width; ++x) planes[(size_t)c * width + x] = (uint8_t)(row[x] >> (8 * c));
                    rle_plane(planes.data() + (size_t)c * width, width, line);
                }
            } else {
                for (uint32_t x = 0; x < width; ++x)
                    for (uint32_t c = 0; c < 4; ++c) line.push_back((uint8_t)(row[x] >> (8 * c)));
            }
            if (std::fwrite(line.data(), 1, line.size(), f) != line.size()) rc = SGPU_E_INVALID;
        }
        if (std::fclose(f) != 0 || rc != SGPU_OK) rc = sgpu_fail(SGPU_E_INVALID, "sgpu_framebuffer_save_hdr: write failed");
    }
    (void)hipFreeAsync(d, st);
    return rc;
}
