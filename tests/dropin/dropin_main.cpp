// Drives the drop-in classes of include/pppm_dropin.hpp the way the
// reference's caller does (VortexSimulationSystem::computeURotation,
// hybridFastSummation/hybridFastSummation/VortexSimulationSystemImp.cpp:142-228)
// and writes the results for tests/test_gpu_dropin.py to compare with the
// oracle.
//
// usage: dropin_main <in.bin> <out.bin>
//   in.bin : uint32 N, uint32 g, int32 K, float4 cent[N], float4 endb[N],
//            double w[3][N], then uint32 mg_n and double b[mg_n^3]
//   out.bin: double u[3][N] (velocity at endb, segment mode with cent),
//            uint32 hash_start[g^3], uint32 hash_end[g^3], uint32 index[N]
//            (ParticleManager on cent with the default 64^3/h=1/64 grid is
//            replaced by g^3, h=1/g), double x[mg_n^3] (FMG solve of b),
//            double residual
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pppm_dropin.hpp"

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }
static bool wr(FILE* f, const void* p, size_t n) { return fwrite(p, 1, n, f) == n; }

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) return 2;
    uint32_t N = 0, g = 0, mgn = 0;
    int32_t K = 0;
    if (!rd(fi, &N, 4) || !rd(fi, &g, 4) || !rd(fi, &K, 4)) return 2;
    GpuArrayf4 cent, endb;
    cent.alloc(N);
    endb.alloc(N);
    GpuArrayd w[3];
    GpuArrayd* wp[3] = {&w[0], &w[1], &w[2]};
    for (auto& a : w) a.alloc(N);
    if (!rd(fi, cent.getHostPtr(), 16 * (size_t)N) || !rd(fi, endb.getHostPtr(), 16 * (size_t)N)) return 2;
    for (auto& a : w)
        if (!rd(fi, a.getHostPtr(), 8 * (size_t)N)) return 2;
    if (!rd(fi, &mgn, 4)) return 2;
    const size_t M3 = (size_t)mgn * mgn * mgn;
    GpuArrayd x, b;
    x.alloc(M3);
    b.alloc(M3);
    if (!rd(fi, b.getHostPtr(), 8 * M3)) return 2;
    std::fclose(fi);
    cent.copy(GpuArrayf4::HOST_TO_DEVICE);
    endb.copy(GpuArrayf4::HOST_TO_DEVICE);
    for (auto& a : w) a.copy(GpuArrayd::HOST_TO_DEVICE);
    b.copy(GpuArrayd::HOST_TO_DEVICE);

    // --- BiotSavartSolver as computeURotation uses it
    BiotSavartSolver solver;
    if (!solver.initializeSolver(g, g, g, false, K, N, N)) return 3;
    if (!solver.setVortParameter(N, &cent, wp)) return 3;
    solver.computeFarFieldBuffer();
    if (!solver.setEvalParameter(N, &endb)) return 3;
    if (!solver.evaluateVelocity(&cent, 1)) {
        std::fprintf(stderr, "evaluateVelocity failed: %s\n", pppm_last_error());
        return 3;
    }
    for (int c = 0; c < 3; c++) solver.getU(c)->copy(GpuArrayd::DEVICE_TO_HOST);
    solver.shutdown();

    // --- ParticleManager on the vortex centres
    ParticleManager pm;
    pm.initSpatialHash(N, g, g, g);
    pm.doSpatialHash(cent.getDevicePtr(), N);
    pppm_synchronize();
    std::vector<uint32_t> st((size_t)g * g * g), en((size_t)g * g * g), ix(N);
    pppm_memcpy_dtoh(st.data(), pm.getStartTable(), 4 * st.size());
    pppm_memcpy_dtoh(en.data(), pm.getEndTable(), 4 * en.size());
    pppm_memcpy_dtoh(ix.data(), pm.getIndexTable(), 4 * ix.size());

    // --- MultiGridSolver3DPeriod_double (reference FMG schedule)
    MultiGridSolver3DPeriod_double mg;
    mg.m_InitialSystem(mgn, mgn, mgn);
    x.memset(0);
    double res = 0;
    mg.m_FullMultiGrid(&x, &b, 1e-10, res);
    x.copy(GpuArrayd::DEVICE_TO_HOST);

    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    bool ok = true;
    for (int c = 0; c < 3; c++) ok = ok && wr(fo, solver.getU(c)->getHostPtr(), 8 * (size_t)N);
    ok = ok && wr(fo, st.data(), 4 * st.size()) && wr(fo, en.data(), 4 * en.size()) && wr(fo, ix.data(), 4 * ix.size());
    ok = ok && wr(fo, x.getHostPtr(), 8 * M3) && wr(fo, &res, 8);
    std::fclose(fo);
    return ok ? 0 : 4;
}
