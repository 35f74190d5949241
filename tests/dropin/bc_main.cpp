// Drives the reference's baroclinic free functions (include/pppm_particle_mesh.h)
// in the order of VortexSimulationSystem::addBaroclinicForce
// (hybridFastSummation/hybridFastSummation/VortexSimulationSystemImp.cpp:369-465):
// the smoke hashed on the density grid with the reference's extern "C" hash
// names, AssignDensity, GradDensity, ParticleGetBaroclinic.  The grid
// (origin, h, dim) comes from the input; tests/test_gpu_baroclinic.py checks
// every stage with the oracle.
//
// usage: bc_main <in.bin> <out.bin>
//   in.bin : uint32 nheat, uint32 n, uint32 dim, double origin[3], double h,
//            double dt_beta, float4 heat[nheat], float rho[nheat],
//            float4 vort[n], double w[3][n]
//   out.bin: float rho_grid[dim^3], float grad[dim^3][3], double w[3][n]
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
    uint32_t nh = 0, n = 0, dim = 0;
    double o[3], h = 0, dtb = 0;
    if (!rd(fi, &nh, 4) || !rd(fi, &n, 4) || !rd(fi, &dim, 4) || !rd(fi, o, 24) || !rd(fi, &h, 8) || !rd(fi, &dtb, 8))
        return 2;
    std::vector<float> heat(4 * (size_t)nh), rho(nh), vp(4 * (size_t)n);
    std::vector<double> w(3 * (size_t)n);
    if (!rd(fi, heat.data(), 16 * (size_t)nh) || !rd(fi, rho.data(), 4 * (size_t)nh) ||
        !rd(fi, vp.data(), 16 * (size_t)n) || !rd(fi, w.data(), 24 * (size_t)n))
        return 2;
    std::fclose(fi);
    const size_t G = (size_t)dim * dim * dim;
    float *d_heat = dalloc<float>(4 * (size_t)nh), *d_sheat = dalloc<float>(4 * (size_t)nh);
    float *d_rho = dalloc<float>(nh), *d_srho = dalloc<float>(nh);
    float* d_vp = dalloc<float>(4 * (size_t)n);
    double* d_w[3];
    uint32_t *d_key = dalloc<uint32_t>(nh), *d_idx = dalloc<uint32_t>(nh);
    uint32_t *d_start = dalloc<uint32_t>(G), *d_end = dalloc<uint32_t>(G);
    float *d_grid = dalloc<float>(G), *d_grad = dalloc<float>(3 * G);
    MemcpyHstToDev(d_heat, heat.data(), 16 * (size_t)nh);
    MemcpyHstToDev(d_rho, rho.data(), 4 * (size_t)nh);
    MemcpyHstToDev(d_vp, vp.data(), 16 * (size_t)n);
    for (int c = 0; c < 3; c++) {
        d_w[c] = dalloc<double>(n);
        MemcpyHstToDev(d_w[c], w.data() + (size_t)c * n, 8 * (size_t)n);
    }
    // the smoke's hash on the density grid (VortexSimulationSystemImp.cpp:405-412)
    HashParams hp{};
    hp.gridSize[0] = hp.gridSize[1] = hp.gridSize[2] = dim;
    hp.numCells = (uint32_t)G;
    for (int c = 0; c < 3; c++) {
        hp.worldOrigin[c] = (float)o[c];
        hp.cellSize[c] = (float)h;
    }
    setParameters(&hp);
    calcHash(d_key, d_idx, d_heat, (int)nh);
    sortParticles(d_key, d_idx, nh);
    FindCellStart(d_start, d_end, d_key, d_idx, nh, (uint32_t)G);
    reorder(d_idx, d_heat, 4, d_sheat, 1, nh);
    reorder(d_idx, d_rho, 1, d_srho, 1, nh);
    const uint3 gd = make_uint3(dim, dim, dim);
    const double4 org = make_double4(o[0], o[1], o[2], 0.0);
    AssignDensity(d_start, d_end, (float4*)d_sheat, d_srho, d_srho, h, d_grid, gd, gd, nh, org);
    GradDensity(d_grid, (float3*)d_grad, h, dim, dim, dim);
    ParticleGetBaroclinic((float4*)d_vp, d_w[0], d_w[1], d_w[2], (float3*)d_grad, dtb, n, dim, dim, dim,
                          1.0 / (double)(float)h, org);
    ThreadSync();
    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    bool ok = true;
    std::vector<char> hb(12 * G > 8 * (size_t)n ? 12 * G : 8 * (size_t)n);
    auto dump = [&](const void* d, size_t bytes) {
        MemcpyDevToHst(d, hb.data(), bytes);
        ok = ok && wr(fo, hb.data(), bytes);
    };
    dump(d_grid, 4 * G);
    dump(d_grad, 12 * G);
    for (int c = 0; c < 3; c++) dump(d_w[c], 8 * (size_t)n);
    std::fclose(fo);
    const char* err = pppm_last_error();
    if (err && *err) std::fprintf(stderr, "last error: %s\n", err);
    return ok ? 0 : 1;
}
