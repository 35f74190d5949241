// Drives the potential-flow gradient path through the reference's free
// functions (include/pppm_particle_mesh.h, include/pppm_ref.h) in the order
// of PotentialFieldSolver::computeFarFieldBuffer and evaluateGradient
// (hybridFastSummation/hybridFastSummation/PotentialFieldSolver.cpp:409-739,
// with the large evaluation, whose points are hashed and sorted): domain (2 x
// the masses' extent), mass hash, ParticleToMesh, ComputeRHS, applyDirichlet with the
// masses' total and centre, three FMG passes, ComputeGradient,
// PotentialComputeFarField, PotentialInterpolateFarField, Potential_PPCorrMN,
// PotentialComputeGradForOutParticle; then the scalar potential
// (PotentialComputeScalarFarField, PotentialInterpolateFarFieldScalar,
// Potential_PPCorrMNScalar) on the same phi.  tests/test_gpu_potential.py checks
// every stage with the oracle.
//
// usage: pot_main <in.bin> <out.bin>
//   in.bin : uint32 n, uint32 m, uint32 g, int32 K, float4 mass_pos[n],
//            double mass[n], float4 eval[m]
//   out.bin: double rho[G], rhs[G], phi[G], grad[G][3], far[G][3], pf[m][3],
//            phi_far[G], phi_points[m]
#include <cmath>
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
    uint32_t n = 0, m = 0, g = 0;
    int32_t K = 0;
    if (!rd(fi, &n, 4) || !rd(fi, &m, 4) || !rd(fi, &g, 4) || !rd(fi, &K, 4)) return 2;
    std::vector<float> mp(4 * (size_t)n), ev(4 * (size_t)m);
    std::vector<double> mass(n);
    if (!rd(fi, mp.data(), 16 * (size_t)n) || !rd(fi, mass.data(), 8 * (size_t)n) || !rd(fi, ev.data(), 16 * (size_t)m))
        return 2;
    std::fclose(fi);
    const size_t G = (size_t)g * g * g;
    // setDomain (PotentialFieldSolver.cpp:409-439)
    double mn[3], mx[3];
    for (int c = 0; c < 3; c++) mn[c] = mx[c] = mp[c];
    for (uint32_t i = 1; i < n; i++)
        for (int c = 0; c < 3; c++) {
            double v = mp[4 * (size_t)i + c];
            if (v < mn[c]) mn[c] = v; else if (v > mx[c]) mx[c] = v;
        }
    double L = mx[0] - mn[0];
    if (mx[1] - mn[1] > L) L = mx[1] - mn[1];
    if (mx[2] - mn[2] > L) L = mx[2] - mn[2];
    L = 2.0 * L;
    double o[3];
    for (int c = 0; c < 3; c++) o[c] = (mn[c] + 0.5 * (mx[c] - mn[c])) - 0.5 * L;
    const float hf = (float)(L / (double)g);
    // mass hash (m_ParticleToMesh, PotentialFieldSolver.cpp:526-541)
    float *d_mp = dalloc<float>(4 * (size_t)n), *d_smp = dalloc<float>(4 * (size_t)n);
    float* d_ev = dalloc<float>(4 * (size_t)m);
    double *d_mass = dalloc<double>(n), *d_smass = dalloc<double>(n);
    uint32_t *d_key = dalloc<uint32_t>(n), *d_idx = dalloc<uint32_t>(n);
    uint32_t *d_start = dalloc<uint32_t>(G), *d_end = dalloc<uint32_t>(G);
    MemcpyHstToDev(d_mp, mp.data(), 16 * (size_t)n);
    MemcpyHstToDev(d_mass, mass.data(), 8 * (size_t)n);
    MemcpyHstToDev(d_ev, ev.data(), 16 * (size_t)m);
    HashParams hp{};
    hp.gridSize[0] = hp.gridSize[1] = hp.gridSize[2] = g;
    hp.numCells = (uint32_t)G;
    for (int c = 0; c < 3; c++) {
        hp.worldOrigin[c] = (float)o[c];
        hp.cellSize[c] = hf;
    }
    setParameters(&hp);
    calcHash(d_key, d_idx, d_mp, (int)n);
    sortParticles(d_key, d_idx, n);
    FindCellStart(d_start, d_end, d_key, d_idx, n, (uint32_t)G);
    reorder(d_idx, d_mp, 4, d_smp, 1, n);
    reorder(d_idx, d_mass, 1, d_smass, 2, n);
    const uint3 gd = make_uint3(g, g, g);
    const double4 org = make_double4(o[0], o[1], o[2], 0.0);
    double *rho = dalloc<double>(G), *rhs = dalloc<double>(G), *grad = dalloc<double>(3 * G), *far = dalloc<double>(3 * G);
    double* pf = dalloc<double>(3 * (size_t)m);
    ParticleToMesh(d_start, d_end, (float4*)d_smp, d_smass, (double)hf, rho, gd, gd, n, org);
    MemcpyDevToDev(rhs, rho, 8 * G);
    ComputeRHS(rhs, (double)(hf * hf), -1.0, (uint32_t)G);
    // total mass and its |m|-weighted centre over the sorted masses (host)
    std::vector<float> hsp(4 * (size_t)n);
    std::vector<double> hsm(n);
    MemcpyDevToHst(d_smp, hsp.data(), 16 * (size_t)n);
    MemcpyDevToHst(d_smass, hsm.data(), 8 * (size_t)n);
    double tw = 0, total = 0, cx = 0, cy = 0, cz = 0;
    for (uint32_t i = 0; i < n; i++) {
        tw += std::fabs(hsm[i]);
        total += hsm[i];
    }
    for (uint32_t i = 0; i < n; i++) {
        cx += hsp[4 * (size_t)i] * std::fabs(hsm[i]);
        cy += hsp[4 * (size_t)i + 1] * std::fabs(hsm[i]);
        cz += hsp[4 * (size_t)i + 2] * std::fabs(hsm[i]);
    }
    cx /= tw;
    cy /= tw;
    cz /= tw;
    applyDirichlet(rhs, make_double4(cx, cy, cz, 0), total, org, (double)hf, g, g, g);
    // m_SolvePoisson: three FMG passes on phi
    gf::MultiGridSolver3DPeriod_double mgs;
    mgs.m_InitialSystem((int)g, (int)g, (int)g);
    GpuArrayd x, b;
    x.alloc(G);
    b.alloc(G);
    MemcpyDevToDev(b.getDevicePtr(), rhs, 8 * G);
    x.memset(0);
    for (int q = 0; q < 3; q++) {
        double res = 0;
        mgs.m_FullMultiGrid(&x, &b, 1e-7, res);
    }
    double* phi = x.getDevicePtr();
    ComputeGradient(phi, (double3*)grad, (double)hf, 1.0, g, g, g);
    ThreadSync();
    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    bool ok = true;
    std::vector<double> hb(3 * (G > m ? G : m));
    auto dump = [&](const double* d, size_t cnt) {
        MemcpyDevToHst(d, hb.data(), 8 * cnt);
        ok = ok && wr(fo, hb.data(), 8 * cnt);
    };
    dump(rho, G);
    dump(rhs, G);
    dump(phi, G);
    dump(grad, 3 * G);
    // computeFarFieldBuffer's far field (phi and grad are overwritten)
    double* phi2 = dalloc<double>(G);
    MemcpyDevToDev(phi2, phi, 8 * G);
    PotentialComputeFarField(d_start, d_end, (float4*)d_mp, d_smass, rho, phi2, (double3*)grad, (double3*)far,
                             (double)hf, 1.0, 1.0, gd, gd, K, m, org);
    // evaluateGradient(true) (PotentialFieldSolver.cpp:441-485): the points
    // hashed on the mass grid and sorted, interpolation and near field in
    // that order, deordered into the caller's
    uint32_t *d_ekey = dalloc<uint32_t>(m), *d_eidx = dalloc<uint32_t>(m);
    float* d_sev = dalloc<float>(4 * (size_t)m);
    double* pfs = dalloc<double>(3 * (size_t)m);
    calcHash(d_ekey, d_eidx, d_ev, (int)m);
    sortParticles(d_ekey, d_eidx, m);
    reorder(d_eidx, d_ev, 4, d_sev, 1, m);
    mg_zerofy(pfs, 24 * (size_t)m);
    PotentialInterpolateFarField((float4*)d_sev, (double3*)far, (double3*)pfs, (double)hf, g, g, g, m, org);
    Potential_PPCorrMN(d_start, d_end, (float4*)d_sev, (float4*)d_smp, d_smass, (double3*)pfs, 1.0 / (double)hf,
                       (double)hf, 1.0, 1.0, gd, gd, make_uint2(g * g, g), make_uint2(g * g, g), K, m, n, org);
    deorder(d_eidx, pfs, 3, pf, 2, m);
    const float3 lo = make_float3((float)o[0], (float)o[1], (float)o[2]);
    const float3 hi = make_float3((float)((double)lo.x + L), (float)((double)lo.y + L), (float)((double)lo.z + L));
    PotentialComputeGradForOutParticle((float4*)d_ev, total, make_double3(cx, cy, cz), lo, hi, 1.0, 1.0,
                                       (double3*)pf, m);
    ThreadSync();
    dump(far, 3 * G);
    dump(pf, 3 * (size_t)m);
    // computeFarFieldPotentialBuffer + evaluateScalarPotential
    // (PotentialFieldSolver.cpp:695-739) on the same phi
    double* ps = dalloc<double>(m);
    mg_zerofy(ps, 8 * (size_t)m);
    PotentialComputeScalarFarField(d_smass, rho, phi, (double)hf, 1.0, 1.0, gd, gd, K);
    PotentialInterpolateFarFieldScalar((float4*)d_ev, phi, ps, (double)hf, g, g, g, m, org);
    Potential_PPCorrMNScalar(d_start, d_end, (float4*)d_ev, (float4*)d_smp, d_smass, ps, 1.0 / (double)hf,
                             (double)hf, 1.0, 1.0, gd, gd, make_uint2(g * g, g), make_uint2(g * g, g), K, m, n, org);
    ThreadSync();
    dump(phi, G);
    dump(ps, m);
    std::fclose(fo);
    const char* err = pppm_last_error();
    if (err && *err) std::fprintf(stderr, "last error: %s\n", err);
    return ok ? 0 : 1;
}
