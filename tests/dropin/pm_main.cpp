// Drives the reference's particle-mesh free functions (include/pppm_particle_mesh.h)
// through the reference's own orchestration of one velocity evaluation
// (BiotSavartSolver::sortVort, m_ParticleToMesh, m_SolvePoisson, m_ComputeCurl,
// computeFarFieldBuffer and evaluateVelocity in
// hybridFastSummation/hybridFastSummation/BiotSavartSolverImp.cpp:410-816), with
// the hash through the reference's extern "C" names and the Poisson solves
// through MultiGridSolver3DPeriod_double.  Writes every stage for
// tests/test_gpu_particle_mesh.py to compare with the oracle.
//
// usage: pm_main <in.bin> <out.bin>
//   in.bin : uint32 N, uint32 g, int32 K, double origin[3], double h,
//            float4 pos[N], double w[3][N]
//   out.bin: double wgrid[3][G], double U[3][G], double far[3][G],
//            double u[3][N] (caller order), double m2p[N] (MeshToParticle of U_x)
#include <cstdio>
#include <cstring>
#include <vector>

#include "pppm_dropin.hpp"
#include "pppm_particle_mesh.h"
#include "pppm_ref.h"

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }
static bool wr(FILE* f, const void* p, size_t n) { return fwrite(p, 1, n, f) == n; }

template <typename T>
static T* dalloc(size_t n) {
    void* p = nullptr;
    AllocateMemory(n * sizeof(T), &p);
    return (T*)p;
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) return 2;
    uint32_t N = 0, g = 0;
    int32_t K = 0;
    double o[3], h = 0;
    if (!rd(fi, &N, 4) || !rd(fi, &g, 4) || !rd(fi, &K, 4) || !rd(fi, o, 24) || !rd(fi, &h, 8)) return 2;
    std::vector<float> pos(4 * (size_t)N);
    std::vector<double> w(3 * (size_t)N);
    if (!rd(fi, pos.data(), 16 * (size_t)N) || !rd(fi, w.data(), 24 * (size_t)N)) return 2;
    std::fclose(fi);
    const size_t G = (size_t)g * g * g;

    float* d_pos = dalloc<float>(4 * (size_t)N);
    float* d_spos = dalloc<float>(4 * (size_t)N);
    double* d_w[3];
    double* d_sw[3];
    uint32_t *d_key = dalloc<uint32_t>(N), *d_idx = dalloc<uint32_t>(N);
    uint32_t *d_start = dalloc<uint32_t>(G), *d_end = dalloc<uint32_t>(G);
    MemcpyHstToDev(d_pos, pos.data(), 16 * (size_t)N);
    for (int c = 0; c < 3; c++) {
        d_w[c] = dalloc<double>(N);
        d_sw[c] = dalloc<double>(N);
        MemcpyHstToDev(d_w[c], w.data() + (size_t)c * N, 8 * (size_t)N);
    }

    // hash (BiotSavartSolverImp.cpp:733-753 sortVort)
    HashParams hp{};
    hp.gridSize[0] = hp.gridSize[1] = hp.gridSize[2] = g;
    hp.numCells = (uint32_t)G;
    for (int c = 0; c < 3; c++) {
        hp.worldOrigin[c] = (float)o[c];
        hp.cellSize[c] = (float)h;
    }
    setParameters(&hp);
    calcHash(d_key, d_idx, d_pos, (int)N);
    sortParticles(d_key, d_idx, N);
    FindCellStart(d_start, d_end, d_key, d_idx, N, (uint32_t)G);
    reorder(d_idx, d_pos, 4, d_spos, 1, N);
    for (int c = 0; c < 3; c++) reorder(d_idx, d_w[c], 1, d_sw[c], 2, N);

    const double hh = (double)(float)h;
    const uint3 gd = make_uint3(g, g, g);
    const double4 org = make_double4(o[0], o[1], o[2], 0.0);
    // P2M + RHS (m_ParticleToMesh, BiotSavartSolverImp.cpp:517-602)
    double *wg[3], *rhs[3], *psi[3], *U[3], *far[3], *Ucopy[3];
    for (int c = 0; c < 3; c++) {
        wg[c] = dalloc<double>(G);
        rhs[c] = dalloc<double>(G);
        psi[c] = dalloc<double>(G);
        U[c] = dalloc<double>(G);
        far[c] = dalloc<double>(G);
        Ucopy[c] = dalloc<double>(G);
        ParticleToMesh(d_start, d_end, (float4*)d_spos, d_sw[c], hh, wg[c], gd, gd, N, org);
        MemcpyDevToDev(rhs[c], wg[c], 8 * G);
        ComputeRHS(rhs[c], hh * hh, -1.0, (uint32_t)G);
    }
    // Poisson solves (m_SolvePoisson, BiotSavartSolverImp.cpp:603-620)
    gf::MultiGridSolver3DPeriod_double mg;
    mg.m_InitialSystem((int)g, (int)g, (int)g);
    for (int c = 0; c < 3; c++) {
        GpuArrayd x, b;
        x.alloc(G);
        b.alloc(G);
        MemcpyDevToDev(b.getDevicePtr(), rhs[c], 8 * G);
        x.memset(0);
        double res = 0;
        mg.m_FullMultiGrid(&x, &b, 1e-10, res);
        MemcpyDevToDev(psi[c], x.getDevicePtr(), 8 * G);
        x.free();
        b.free();
    }
    // curl (m_ComputeCurl) and far-field buffer (computeFarFieldBuffer)
    getCurl(psi[0], psi[1], psi[2], U[0], U[1], U[2], g, g, g, hh);
    for (int c = 0; c < 3; c++) MemcpyDevToDev(Ucopy[c], U[c], 8 * G);
    BiotSavartComputeFarField(d_start, d_end, (float4*)d_spos, wg[0], wg[1], wg[2], psi[0], psi[1], psi[2], U[0], U[1],
                              U[2], far[0], far[1], far[2], hh, gd, gd, K, N, org);
    // evaluation at the vortices (evaluateVelocity, BiotSavartSolverImp.cpp:410-516)
    double *us[3], *uo[3];
    for (int c = 0; c < 3; c++) {
        us[c] = dalloc<double>(N);
        uo[c] = dalloc<double>(N);
    }
    BiotSavartInterpolateFarField((float4*)d_spos, far[0], far[1], far[2], us[0], us[1], us[2], hh, g, g, g, N, org);
    BiotSavartPPCorrScaleMN(d_start, d_end, (float4*)d_spos, (float4*)d_spos, 0, (float4*)d_spos, d_sw[0], d_sw[1],
                            d_sw[2], us[0], us[1], us[2], nullptr, nullptr, nullptr, hh, gd, gd, K, N, N, org);
    for (int c = 0; c < 3; c++) deorder(d_idx, us[c], 1, uo[c], 2, N);
    double* m2p = dalloc<double>(N);
    MeshToParticle((float4*)d_pos, Ucopy[0], m2p, hh, gd, gd, N, org);
    ThreadSync();

    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    std::vector<double> hbuf(G > N ? G : N);
    bool ok = true;
    auto dump = [&](double* d, size_t n) {
        MemcpyDevToHst(d, hbuf.data(), 8 * n);
        ok = ok && wr(fo, hbuf.data(), 8 * n);
    };
    for (int c = 0; c < 3; c++) dump(wg[c], G);
    for (int c = 0; c < 3; c++) dump(Ucopy[c], G);
    for (int c = 0; c < 3; c++) dump(far[c], G);
    for (int c = 0; c < 3; c++) dump(uo[c], N);
    dump(m2p, N);
    std::fclose(fo);
    const char* err = pppm_last_error();
    if (err && *err) std::fprintf(stderr, "last error: %s\n", err);
    return ok ? 0 : 4;
}
