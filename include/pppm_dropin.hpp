/*
 * pppm_dropin.hpp -- the reference's C++ classes with their names and
 * signatures, implemented header-only on the C ABI of libpppm_hip.so.
 *
 *   gf::gf_GpuArray<T>               multigrid/multigrid/gfGpuArray.h:25-391
 *   ParticleManager                  particle_hash/particle_hash/SpaceHashClass.h:10-84
 *   gf::MultiGrid<T>,
 *   gf::MultiGridSolver3DPeriod(_double)  multigrid/multigrid/MultiGrid.h:12-209
 *   BiotSavartSolver                 hybridFastSummation/hybridFastSummation/BiotSavartSolver.h:16-89
 *
 * Build a user of these classes with any C++ compiler:
 *   g++ -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -I<repo>/include app.cpp \
 *       -L<repo>/pppm_vortex_bouyant_flow_sample_amd/_build -lpppm_hip
 *
 * Differences from the reference (deliberate): methods return false on a
 * library error instead of exiting; gf_GpuArray has no OpenGL buffer (vbo
 * arguments must be false); memset writes zero whatever the value argument,
 * as the reference's does (gfGpuArray.h:371-380); BiotSavartSolver frees what
 * it allocates.
 */
#ifndef PPPM_DROPIN_HPP
#define PPPM_DROPIN_HPP

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pppm.h"

typedef unsigned int uint;

namespace gf {

/* gfGpuArray.h:25-89: device buffer with a host mirror */
template <class T>
class gf_GpuArray {
   public:
    enum Direction { HOST_TO_DEVICE, DEVICE_TO_HOST, DEVICE_TO_DEVICE };
    gf_GpuArray() : m_size(0), m_dptr(nullptr), m_hptr(nullptr) {}
    ~gf_GpuArray() { free(); }
    gf_GpuArray(const gf_GpuArray&) = delete;
    gf_GpuArray& operator=(const gf_GpuArray&) = delete;

    void alloc(size_t size, bool vbo = false, bool doubleBuffer = false, bool useElementArray = false) {
        (void)doubleBuffer;
        (void)useElementArray;
        free();
        if (vbo) {
            std::fprintf(stderr, "gf_GpuArray: OpenGL buffers are not supported\n");
            return;
        }
        m_size = size;
        void* p = nullptr;
        if (pppm_malloc(&p, sizeof(T) * (size ? size : 1)) != PPPM_OK) {
            m_size = 0;
            return;
        }
        m_dptr = (T*)p;
        m_hptr = new T[size ? size : 1]();
    }
    void allocPitch(size_t size, bool vbo = false, bool doubleBuffer = false, bool useElementArray = false) {
        alloc(size, vbo, doubleBuffer, useElementArray);
    }
    void free() {
        if (m_dptr) pppm_free(m_dptr);
        delete[] m_hptr;
        m_dptr = nullptr;
        m_hptr = nullptr;
        m_size = 0;
    }
    void swap() {}
    void map() {}
    void unmap() {}
    void copy(Direction dir, size_t start = 0, size_t count = 0) {
        if (count == 0) count = m_size;
        if (!m_dptr || count == 0) return;
        if (dir == HOST_TO_DEVICE)
            pppm_memcpy_htod(m_dptr + start, m_hptr + start, sizeof(T) * count);
        else if (dir == DEVICE_TO_HOST)
            pppm_memcpy_dtoh(m_hptr + start, m_dptr + start, sizeof(T) * count);
    }
    void copyPitch(Direction dir, size_t start = 0, size_t count = 0) { copy(dir, start, count); }
    /* gfGpuArray.h:371-380: writes zero whatever `value` is */
    void memset(T value, size_t start = 0, size_t count = 0) {
        (void)value;
        (void)start;
        if (count == 0) count = m_size;
        if (m_dptr && count) pppm_memset(m_dptr, 0, sizeof(T) * count);
    }
    T* getDevicePtr() { return m_dptr; }
    T* getDeviceWritePtr() { return m_dptr; }
    T* getHostPtr() { return m_hptr; }
    size_t getSize() const { return m_size; }
    size_t typeSize() const { return sizeof(T); }

   private:
    size_t m_size;
    T* m_dptr;
    T* m_hptr;
};

typedef gf_GpuArray<float> GpuArrayf;
typedef gf_GpuArray<int> GpuArrayi;
typedef gf_GpuArray<char> GpuArrayc;
typedef gf_GpuArray<double> GpuArrayd;
typedef gf_GpuArray<float2> GpuArrayf2;
typedef gf_GpuArray<double2> GpuArrayd2;
typedef gf_GpuArray<double3> GpuArrayd3;
typedef gf_GpuArray<float3> GpuArrayf3;
typedef gf_GpuArray<double4> GpuArrayd4;
typedef gf_GpuArray<float4> GpuArrayf4;

/* MultiGrid.h:12-28 */
struct uniform_grid_descriptor_3D {
    int gridx, gridy, gridz, system_size;
    uniform_grid_descriptor_3D() : gridx(0), gridy(0), gridz(0), system_size(0) {}
    uniform_grid_descriptor_3D(int x, int y, int z) : gridx(x), gridy(y), gridz(z), system_size(x * y * z) {}
    void set_grid_information(int x, int y, int z) {
        gridx = x;
        gridy = y;
        gridz = z;
        system_size = x * y * z;
    }
};

/* MultiGrid.h:29-139: level hierarchy; the device work lives in pppm_mg */
template <class T>
class MultiGrid {
   public:
    MultiGrid() : m_bInitialized(false), m_max_level(0), m_mg(nullptr) {}
    ~MultiGrid() { m_FinalMemoryEachLevel(); }
    void m_InitialSystem(int gridx, int gridy, int gridz) {
        m_FinalMemoryEachLevel();
        m_system_descriptor.set_grid_information(gridx, gridy, gridz);
        if (pppm_mg_create(gridx, gridy, gridz, sizeof(T) == 4 ? PPPM_F32 : PPPM_F64, &m_mg) != PPPM_OK) return;
        int L = 0;
        int dims[16][3];
        pppm_mg_levels(m_mg, &L, dims, 16);
        m_max_level = L;
        systemk.resize(L);
        for (int l = 0; l < L; l++) systemk[l].set_grid_information(dims[l][0], dims[l][1], dims[l][2]);
        m_bInitialized = true;
    }
    void m_FinalMemoryEachLevel() {
        if (m_mg) pppm_mg_destroy(m_mg);
        m_mg = nullptr;
        m_bInitialized = false;
    }
    bool m_bInitialized;
    uniform_grid_descriptor_3D m_system_descriptor;
    int m_max_level;
    std::vector<uniform_grid_descriptor_3D> systemk;

   protected:
    /* one FMG pass + one V(4,4) SOR(1.43f) cycle, MultiGrid.cpp:202-250 */
    bool fmg(T* x, const T* b, T& residual) {
        double rel = 0;
        int cyc = 0;
        pppm_status s = pppm_mg_solve(m_mg, x, b, PPPM_MG_REF_FMG, 4, 4, (double)1.43f, 0.0, 1, &cyc, &rel, nullptr);
        residual = (T)rel;
        return s == PPPM_OK;
    }
    bool vcycle(T* x, const T* b, T& residual) {
        double rel = 0;
        int cyc = 0;
        pppm_status s = pppm_mg_solve(m_mg, x, b, PPPM_MG_REF_VCYCLE, 4, 4, (double)1.43f, 0.0, 1, &cyc, &rel, nullptr);
        residual = (T)rel;
        return s == PPPM_OK;
    }
    pppm_mg* m_mg;
};

/* MultiGrid.h:167-184 */
class MultiGridSolver3DPeriod : public MultiGrid<float> {
   public:
    /* V-cycles run from the finest level (the reference's level argument
     * selects its internal coarse arrays, which are private here) */
    void m_Vcycle(GpuArrayf* x, GpuArrayf* b, float tol, float& residual, int level) {
        (void)tol;
        (void)level;
        vcycle(x->getDevicePtr(), b->getDevicePtr(), residual);
    }
    void m_FullMultiGrid(GpuArrayf* x, GpuArrayf* b, float tol, float& residual) {
        (void)tol;
        fmg(x->getDevicePtr(), b->getDevicePtr(), residual);
    }
};

/* MultiGrid.h:188-209 */
class MultiGridSolver3DPeriod_double : public MultiGrid<double> {
   public:
    void m_Vcycle(GpuArrayd* x, GpuArrayd* b, double tol, double& residual, int level) {
        (void)tol;
        (void)level;
        vcycle(x->getDevicePtr(), b->getDevicePtr(), residual);
    }
    void m_FullMultiGrid(GpuArrayd* x, GpuArrayd* b, double tol, double& residual) {
        (void)tol;
        fmg(x->getDevicePtr(), b->getDevicePtr(), residual);
    }
};

}  // namespace gf

/* SpaceHashClass.h:10-84 + SpaceHashClassImp.cpp:6-65 */
class ParticleManager {
   public:
    ParticleManager() : m_numParticles(0), m_hash(nullptr) { setSpatialHashGrid(64, 1.0f / 64.0f, zero3(), 0); }
    explicit ParticleManager(uint numParticles) : m_numParticles(numParticles), m_hash(nullptr) {
        setSpatialHashGrid(64, 1.0f / 64.0f, zero3(), numParticles);
    }
    ~ParticleManager() { endSpatialHash(); }
    int getNumParticles() const { return (int)m_numParticles; }
    uint3 getGridSize() { return make_uint3(m_g, m_g, m_g); }
    float3 getWorldOrigin() { return m_origin; }
    float3 getCellSize() { return make_float3(m_h, m_h, m_h); }
    uint getNumCells() { return m_g * m_g * m_g; }

    void initSpatialHash(uint numParticles, uint dimx, uint dimy, uint dimz) {
        (void)dimy;
        (void)dimz;  // cubic only, as SpaceHashClassImp.cpp:50
        endSpatialHash();
        m_numParticles = numParticles;
        setSpatialHashGrid(dimx, (float)(1.0 / (double)dimx), zero3(), numParticles);
        float o[3] = {0, 0, 0};
        pppm_hash_create(numParticles, dimx, m_h, o, &m_hash);
    }
    void endSpatialHash() {
        if (m_hash) pppm_hash_destroy(m_hash);
        m_hash = nullptr;
    }
    void setSpatialHashGrid(uint size0, float h, float3 bbmin, unsigned int numParticles) {
        m_g = size0;
        m_h = h;
        m_origin = bbmin;
        m_numParticles = numParticles;
    }
    void setHashParam() {
        float o[3] = {m_origin.x, m_origin.y, m_origin.z};
        if (m_hash) pppm_hash_set_grid(m_hash, m_g, m_h, o);
    }
    void doSpatialHash(float4* pos, uint numParticles) {
        if (!m_hash) {
            float o[3] = {m_origin.x, m_origin.y, m_origin.z};
            pppm_hash_create(numParticles, m_g, m_h, o, &m_hash);
        }
        setHashParam();
        m_numParticles = numParticles;
        pppm_hash_build(m_hash, (const float*)pos, numParticles, nullptr);
    }
    void reorderData(uint numParticles, void* data, void* dataReorder, unsigned int stride, uint precision) {
        pppm_hash_reorder(m_hash, numParticles, data, dataReorder, stride, (int32_t)precision, nullptr);
    }
    void deorderData(uint numParticles, void* data, void* dataReorder, unsigned int stride, uint precision) {
        pppm_hash_deorder(m_hash, numParticles, data, dataReorder, stride, (int32_t)precision, nullptr);
    }
    uint* getHashTable() { return table(0); }
    uint* getStartTable() { return table(2); }
    uint* getEndTable() { return table(3); }
    uint* getIndexTable() { return table(1); }

   private:
    static float3 zero3() { return make_float3(0.f, 0.f, 0.f); }
    uint* table(int which) {
        const uint32_t* t[4] = {nullptr, nullptr, nullptr, nullptr};
        if (!m_hash) return nullptr;
        pppm_hash_tables(m_hash, &t[0], &t[1], &t[2], &t[3]);
        return (uint*)t[which];
    }
    uint m_numParticles;
    uint m_g = 64;
    float m_h = 1.0f / 64.0f;
    float3 m_origin;
    pppm_hash* m_hash;
};

using namespace gf;

enum g_Components { COMPONENT_X, COMPONENT_Y, COMPONENT_Z, NUM_COMPONENTS };

/* BiotSavartSolver.h:16-89: PPPM velocity of vortex particles at evaluation
 * points.  The reference's settings: domain = 3x the vortices' bounding box
 * (setDomain, BiotSavartSolverImp.cpp:378-409), core a = 0.08, one FMG pass
 * of V(4,4) SOR 1.43f; FP64 strengths and velocities. */
class BiotSavartSolver {
   public:
    BiotSavartSolver() {
        for (int i = 0; i < NUM_COMPONENTS; i++) {
            m_particle_vort[i] = new GpuArrayd;
            m_particle_vort_Reorder[i] = new GpuArrayd;
            m_particle_U[i] = new GpuArrayd;
        }
        m_evalPos = new GpuArrayf4;
        m_p_vortPos = new GpuArrayf4;
        m_p_vortPos_Reorder = new GpuArrayf4;
    }
    ~BiotSavartSolver() {
        shutdown();
        for (int i = 0; i < NUM_COMPONENTS; i++) {
            delete m_particle_vort[i];
            delete m_particle_vort_Reorder[i];
            delete m_particle_U[i];
        }
        delete m_evalPos;
        delete m_p_vortPos;
        delete m_p_vortPos_Reorder;
    }

    bool initializeSolver(uint gdx, uint gdy, uint gdz, bool isVIC, int K, uint M, uint N) {
        if (gdx != gdy || gdx != gdz) return false;  // the hash is cubic (SpaceHashClass.h:46-54)
        m_isVIC = isVIC;
        m_K = K;
        if (m_ctx && gdx == m_grid && M == m_M && N == m_N) return true;
        shutdown();
        m_grid = gdx;
        m_M = M;
        m_N = N;
        pppm_params p;
        pppm_params_reference_preset(&p, gdx, isVIC ? -1 : K);
        if (pppm_create(&p, &m_ctx) != PPPM_OK) return false;
        for (int i = 0; i < NUM_COMPONENTS; i++) {
            m_particle_vort[i]->alloc(N);
            m_particle_vort_Reorder[i]->alloc(N);
            m_particle_U[i]->alloc(M);
        }
        m_evalPos->alloc(M);
        m_p_vortPos->alloc(N);
        m_p_vortPos_Reorder->alloc(N);
        return true;
    }
    bool setVortParameter(uint n, GpuArrayf4* pos, GpuArrayd* omega[]) {
        if (!m_ctx || n > m_p_vortPos->getSize()) return false;
        m_N = n;
        pppm_memcpy_dtod(m_p_vortPos->getDevicePtr(), pos->getDevicePtr(), sizeof(float4) * n);
        for (int i = 0; i < NUM_COMPONENTS; i++)
            pppm_memcpy_dtod(m_particle_vort[i]->getDevicePtr(), omega[i]->getDevicePtr(), sizeof(double) * n);
        return true;
    }
    bool setEvalParameter(uint m, GpuArrayf4* pos) {
        if (!m_ctx) return false;
        if (m > m_evalPos->getSize()) {
            m_evalPos->alloc(m);
            for (int i = 0; i < NUM_COMPONENTS; i++) m_particle_U[i]->alloc(m);
        }
        m_M = m;
        return pppm_memcpy_dtod(m_evalPos->getDevicePtr(), pos->getDevicePtr(), sizeof(float4) * m) == PPPM_OK;
    }
    /* BiotSavartSolverImp.cpp:755-816 */
    void computeFarFieldBuffer() {
        if (!sortVort()) return;
        m_have_far = pppm_compute_far_field(m_ctx) == PPPM_OK;
    }
    /* BiotSavartSolverImp.cpp:410-516; results in getU(c), caller order */
    bool evaluateVelocity(GpuArrayf4* another_end, uint is_segment) {
        if (!m_ctx || !m_have_far) return false;
        const float* other = (is_segment == 1 && another_end) ? (const float*)another_end->getDevicePtr() : nullptr;
        pppm_status s = pppm_evaluate(m_ctx, m_M, (const float*)m_evalPos->getDevicePtr(), other,
                                      m_particle_U[0]->getDevicePtr(), m_particle_U[1]->getDevicePtr(),
                                      m_particle_U[2]->getDevicePtr(), 1, PPPM_F64);
        return s == PPPM_OK && pppm_synchronize() == PPPM_OK;
    }
    GpuArrayd* getU(int i) { return m_particle_U[i]; }
    GpuArrayd* getSortVort(int i) { return m_particle_vort_Reorder[i]; }
    GpuArrayd* getVort(int i) { return m_particle_vort[i]; }
    GpuArrayf4* getSortPos() { return m_p_vortPos_Reorder; }
    uint* getVortStartTable() { return table(2); }
    uint* getVortEndTable() { return table(3); }
    /* BiotSavartSolverImp.cpp:733-753: domain from the vortices, hash, and the
     * sorted positions / strengths */
    bool sortVort() {
        if (!m_ctx || m_N == 0) return false;
        m_p_vortPos->copy(GpuArrayf4::DEVICE_TO_HOST, 0, m_N);
        double o[3], h;
        if (pppm_reference_domain((const float*)m_p_vortPos->getHostPtr(), m_N, m_grid, o, &h) != PPPM_OK) return false;
        if (pppm_set_domain(m_ctx, o, h) != PPPM_OK) return false;
        m_origin = make_double4(o[0], o[1], o[2], 0.0);
        m_cell_h = (double)(float)h;
        if (pppm_set_vortices(m_ctx, m_N, (const float*)m_p_vortPos->getDevicePtr(), m_particle_vort[0]->getDevicePtr(),
                              m_particle_vort[1]->getDevicePtr(), m_particle_vort[2]->getDevicePtr(), 1,
                              PPPM_F64) != PPPM_OK)
            return false;
        const uint32_t* idx = nullptr;
        void* st = nullptr;
        pppm_get_stream(m_ctx, &st);
        pppm_get_sorted_vortices(m_ctx, nullptr, &idx, nullptr, nullptr);
        pppm_gather(idx, m_p_vortPos->getDevicePtr(), m_p_vortPos_Reorder->getDevicePtr(), 4, PPPM_F32, m_N, st);
        for (int c = 0; c < NUM_COMPONENTS; c++)
            pppm_gather(idx, m_particle_vort[c]->getDevicePtr(), m_particle_vort_Reorder[c]->getDevicePtr(), 1,
                        PPPM_F64, m_N, st);
        return pppm_synchronize() == PPPM_OK;
    }
    /* BiotSavartSolverImp.cpp:818-827: sorted strengths back to caller order */
    void unsortVort() {
        const uint32_t* idx = nullptr;
        void* st = nullptr;
        if (!m_ctx || pppm_get_sorted_vortices(m_ctx, nullptr, &idx, nullptr, nullptr) != PPPM_OK) return;
        pppm_get_stream(m_ctx, &st);
        for (int c = 0; c < NUM_COMPONENTS; c++)
            pppm_scatter(idx, m_particle_vort_Reorder[c]->getDevicePtr(), m_particle_vort[c]->getDevicePtr(), 1,
                         PPPM_F64, m_N, st);
        pppm_synchronize();
    }
    double getCellh() { return m_cell_h; }
    double4 getOrigin() { return m_origin; }
    bool shutdown() {
        if (m_ctx) pppm_destroy(m_ctx);
        m_ctx = nullptr;
        m_have_far = false;
        return true;
    }
    pppm_ctx* context() { return m_ctx; }

   private:
    uint* table(int which) {
        const uint32_t *st = nullptr, *en = nullptr;
        if (!m_ctx || pppm_get_sorted_vortices(m_ctx, nullptr, nullptr, &st, &en) != PPPM_OK) return nullptr;
        return (uint*)(which == 2 ? st : en);
    }
    pppm_ctx* m_ctx = nullptr;
    bool m_isVIC = false, m_have_far = false;
    int m_K = 1;
    uint m_grid = 0, m_M = 0, m_N = 0;
    double m_cell_h = 0.0;
    double4 m_origin = make_double4(0, 0, 0, 0);
    GpuArrayf4* m_evalPos;
    GpuArrayf4* m_p_vortPos;
    GpuArrayf4* m_p_vortPos_Reorder;
    GpuArrayd* m_particle_vort[NUM_COMPONENTS];
    GpuArrayd* m_particle_vort_Reorder[NUM_COMPONENTS];
    GpuArrayd* m_particle_U[NUM_COMPONENTS];
};

#endif
