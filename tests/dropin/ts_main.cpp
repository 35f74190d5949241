// Drives the reference's time-step free functions (include/pppm_particle_mesh.h)
// through the reference's own order of one step
// (VortexSimulationSystem::solveTimeStep, hybridFastSummation/hybridFastSummation/
// VortexSimulationSystemImp.cpp:90-125, without the baroclinic force, the
// potential flow and shedding): convertVortToSeg (ComputePosb), the end
// velocities accumulated with AddURotation, applyFlow (move_particle twice,
// ComputePosa), convertSegToVort (ComputeVortex) and computeStretching
// (VectorAdd, the centres hashed with the reference's extern "C" names,
// ParticleGaussianSmooth with nu = 0.1, VectorAdd).  The end velocities come
// from the input; tests/test_gpu_timestep.py checks every stage with the
// oracle.
//
// usage: ts_main <in.bin> <out.bin>
//   in.bin : uint32 N, uint32 g, double origin[3], double h (hash of the new
//            centres), double seg_h, double dt, float4 cent[N], double w[3][N],
//            double ua[3][N], double ub[3][N]
//   out.bin: float4 enda0[N], float4 endb0[N], double kappa[N], float4 enda[N],
//            float4 endb[N], float4 cent[N], double dw[3][N], double w[3][N]
#include <cstdio>
#include <vector>

#include "pppm_dropin.hpp"
#include "pppm_particle_mesh.h"
#include "pppm_ref.h"

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }
static bool wr(FILE* f, const void* p, size_t n) { return fwrite(p, 1, n, f) == n; }

template <typename T>
static T* dalloc(size_t n) {
    void* p = nullptr;
    AllocateMemory((n ? n : 1) * sizeof(T), &p);
    return (T*)p;
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) return 2;
    uint32_t N = 0, g = 0;
    double o[3], h = 0, seg_h = 0, dt = 0;
    if (!rd(fi, &N, 4) || !rd(fi, &g, 4) || !rd(fi, o, 24) || !rd(fi, &h, 8) || !rd(fi, &seg_h, 8) ||
        !rd(fi, &dt, 8))
        return 2;
    std::vector<float> cent(4 * (size_t)N);
    std::vector<double> w(3 * (size_t)N), ua(3 * (size_t)N), ub(3 * (size_t)N);
    if (!rd(fi, cent.data(), 16 * (size_t)N) || !rd(fi, w.data(), 24 * (size_t)N) ||
        !rd(fi, ua.data(), 24 * (size_t)N) || !rd(fi, ub.data(), 24 * (size_t)N))
        return 2;
    std::fclose(fi);
    const size_t G = (size_t)g * g * g;

    float4 *d_cent = dalloc<float4>(N), *d_a = dalloc<float4>(N), *d_b = dalloc<float4>(N);
    float4* d_scent = dalloc<float4>(N);
    double *d_w[3], *d_dw[3], *d_sdw[3], *d_ua[3], *d_ub[3];
    double* d_kappa = dalloc<double>(N);
    double4 *d_Ua = dalloc<double4>(N), *d_Ub = dalloc<double4>(N);
    MemcpyHstToDev(d_cent, cent.data(), 16 * (size_t)N);
    // the ends start as copies of the centres (the reference allocates them
    // alongside; ComputePosb writes x, y, z only)
    MemcpyHstToDev(d_a, cent.data(), 16 * (size_t)N);
    MemcpyHstToDev(d_b, cent.data(), 16 * (size_t)N);
    for (int c = 0; c < 3; c++) {
        d_w[c] = dalloc<double>(N);
        d_dw[c] = dalloc<double>(N);
        d_sdw[c] = dalloc<double>(N);
        d_ua[c] = dalloc<double>(N);
        d_ub[c] = dalloc<double>(N);
        MemcpyHstToDev(d_w[c], w.data() + (size_t)c * N, 8 * (size_t)N);
        MemcpyHstToDev(d_ua[c], ua.data() + (size_t)c * N, 8 * (size_t)N);
        MemcpyHstToDev(d_ub[c], ub.data() + (size_t)c * N, 8 * (size_t)N);
    }
    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    bool ok = true;
    std::vector<char> hbuf(32 * (size_t)(N ? N : 1));
    auto dump = [&](const void* d, size_t bytes) {
        MemcpyDevToHst(d, hbuf.data(), bytes);
        ok = ok && wr(fo, hbuf.data(), bytes);
    };

    // convertVortToSeg
    ComputePosb(d_cent, d_a, d_b, d_w[0], d_w[1], d_w[2], d_kappa, seg_h, N);
    ThreadSync();
    dump(d_a, 16 * (size_t)N);
    dump(d_b, 16 * (size_t)N);
    dump(d_kappa, 8 * (size_t)N);
    // computeURotation's accumulation into zeroed double4 velocities
    mg_zerofy(d_Ua, 32 * (size_t)N);
    mg_zerofy(d_Ub, 32 * (size_t)N);
    AddURotation(d_Ua, d_ua[0], d_ua[1], d_ua[2], N);
    AddURotation(d_Ub, d_ub[0], d_ub[1], d_ub[2], N);
    // applyFlow
    move_particle(d_a, dt, d_Ua, N);
    move_particle(d_b, dt, d_Ub, N);
    ComputePosa(d_a, d_b, d_cent, N);
    // convertSegToVort
    for (int c = 0; c < 3; c++) mg_zerofy(d_dw[c], 8 * (size_t)N);
    ComputeVortex(d_a, d_b, d_w[0], d_w[1], d_w[2], d_dw[0], d_dw[1], d_dw[2], d_kappa, N);
    ThreadSync();
    dump(d_a, 16 * (size_t)N);
    dump(d_b, 16 * (size_t)N);
    dump(d_cent, 16 * (size_t)N);
    for (int c = 0; c < 3; c++) dump(d_dw[c], 8 * (size_t)N);
    // computeStretching
    for (int c = 0; c < 3; c++) VectorAdd(d_w[c], d_dw[c], d_dw[c], 1, 1, N);
    uint32_t *d_key = dalloc<uint32_t>(N), *d_idx = dalloc<uint32_t>(N);
    uint32_t *d_start = dalloc<uint32_t>(G), *d_end = dalloc<uint32_t>(G);
    HashParams hp{};
    hp.gridSize[0] = hp.gridSize[1] = hp.gridSize[2] = g;
    hp.numCells = (uint32_t)G;
    for (int c = 0; c < 3; c++) {
        hp.worldOrigin[c] = (float)o[c];
        hp.cellSize[c] = (float)h;
    }
    setParameters(&hp);
    calcHash(d_key, d_idx, (float*)d_cent, (int)N);
    sortParticles(d_key, d_idx, N);
    FindCellStart(d_start, d_end, d_key, d_idx, N, (uint32_t)G);
    reorder(d_idx, d_cent, 4, d_scent, 1, N);
    for (int c = 0; c < 3; c++) reorder(d_idx, d_dw[c], 1, d_sdw[c], 2, N);
    ParticleGaussianSmooth(d_start, d_end, d_cent, d_scent, 0.1, dt, d_dw[0], d_dw[1], d_dw[2], d_sdw[0], d_sdw[1],
                           d_sdw[2], (double)(float)h, make_uint3(g, g, g), N, make_double4(o[0], o[1], o[2], 0.0));
    for (int c = 0; c < 3; c++) VectorAdd(d_w[c], d_dw[c], d_w[c], 0, 1, N);
    ThreadSync();
    for (int c = 0; c < 3; c++) dump(d_w[c], 8 * (size_t)N);
    std::fclose(fo);
    const char* err = pppm_last_error();
    if (err && *err) std::fprintf(stderr, "last error: %s\n", err);
    return ok ? 0 : 1;
}
