// Frame output for the reference's renderers (newTest.cpp:23-37 output()):
// one file per step, Particle_data%04d.bin, holding every particle as a raw
// float4 (x, y, z, w) with w = 1.  The particle renderer reads x, y, z and
// takes w as the particle's density (read_particles.cpp:55-70), so a caller
// may pass a per-particle scalar (e.g. smoke density) for w instead.
// Positions come from device memory and go to the file through a pinned
// host buffer in chunks of 2^20 particles (16 MiB), so host memory stays
// bounded at any particle count.
#include <cstdio>
#include <string>

#include "pppm_internal.hpp"

#define FAIL(code, msg)                                  \
    do {                                                 \
        pppm_set_last_error((msg), __FILE__, __LINE__);  \
        return (code);                                   \
    } while (0)

namespace {

constexpr uint32_t kFrameChunk = 1u << 20;

// pinned staging and the open file, released on every exit path
struct FrameIo {
    float4* pos = nullptr;
    float* w = nullptr;
    FILE* f = nullptr;
    ~FrameIo() {
        if (f) std::fclose(f);
        std::lock_guard<std::mutex> lk(pppm::runtime_mutex());
        if (pos) (void)hipHostFree(pos);
        if (w) (void)hipHostFree(w);
    }
};

}  // namespace

extern "C" pppm_status pppm_write_frame(const char* dir, int32_t frame, const float* pos4, const float* w,
                                        uint32_t n, void* hip_stream) {
    if (!dir) FAIL(PPPM_ERR_INVALID, "null directory");
    if (frame < 0) FAIL(PPPM_ERR_INVALID, "frame must be >= 0");
    if (n > 0 && !pos4) FAIL(PPPM_ERR_INVALID, "null positions");
    char name[48];
    std::snprintf(name, sizeof name, "Particle_data%04d.bin", (int)frame);
    std::string path = dir;
    if (!path.empty() && path.back() != '/') path += '/';
    path += name;
    hipStream_t st = (hipStream_t)hip_stream;
    FrameIo io;
    const uint32_t cap = n < kFrameChunk ? n : kFrameChunk;
    if (cap > 0) {
        std::lock_guard<std::mutex> lk(pppm::runtime_mutex());
        PPPM_HIP_TRY(hipHostMalloc((void**)&io.pos, (size_t)cap * sizeof(float4), hipHostMallocDefault));
        if (w) PPPM_HIP_TRY(hipHostMalloc((void**)&io.w, (size_t)cap * sizeof(float), hipHostMallocDefault));
    }
    io.f = std::fopen(path.c_str(), "wb");
    if (!io.f) FAIL(PPPM_ERR_INVALID, "cannot open the frame file for writing");
    for (uint32_t s0 = 0; s0 < n; s0 += cap) {
        const uint32_t c = n - s0 < cap ? n - s0 : cap;
        PPPM_HIP_TRY(hipMemcpyAsync(io.pos, pos4 + 4 * (size_t)s0, (size_t)c * sizeof(float4),
                                    hipMemcpyDeviceToHost, st));
        if (w) PPPM_HIP_TRY(hipMemcpyAsync(io.w, w + s0, (size_t)c * sizeof(float), hipMemcpyDeviceToHost, st));
        PPPM_HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < c; i++) io.pos[i].w = w ? io.w[i] : 1.0f;
        if (std::fwrite(io.pos, sizeof(float4), c, io.f) != c) FAIL(PPPM_ERR_INVALID, "frame write failed");
    }
    const int rc = std::fclose(io.f);
    io.f = nullptr;
    if (rc != 0) FAIL(PPPM_ERR_INVALID, "frame close failed");
    return PPPM_OK;
}
