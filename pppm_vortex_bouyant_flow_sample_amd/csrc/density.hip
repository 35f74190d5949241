// Baroclinic vorticity source of the buoyant flow (VortexSimulationSystemImp.cpp:
// 369-465 addBaroclinicForce; kernels GPU_kernel.cu:3655-3898 AssignDensity,
// GradDensity, ParticleGetBaroclinic; the domain rule 320-367
// setDensityDomain).  The smoke particles are hashed on their own density
// grid (1.2 x their extent, cells of about 0.2, at most 200 per axis); the
// grid density is a normalised (1 - d^2)^4 average of the particles' values;
// its periodic central differences are interpolated at the vortices with hat
// weights, and w += dt_beta * 0.1 * (grad rho x g).
//
// The reference allocates and frees every buffer on each call; a density
// object here keeps them and grows them.  FP contraction is off in this file:
// every product and sum rounds as the reference writes it.
#include <algorithm>
#include <cmath>

#include "pppm_internal.hpp"

#pragma clang fp contract(off)

#define FAIL(code, msg)                                  \
    do {                                                 \
        pppm_set_last_error((msg), __FILE__, __LINE__);  \
        return (code);                                   \
    } while (0)
#define TRY(expr)                       \
    do {                                \
        pppm_status _s = (expr);        \
        if (_s != PPPM_OK) return _s;   \
    } while (0)

namespace pppm {
namespace {

// (1 - r)^4 on r = d^2 < 1, d in cells (GPU_kernel.cu poly_shape2 applied to
// the squared distance by interpolate_weights2)
__device__ inline double density_weight(double gx, double gy, double gz, double px, double py, double pz) {
    const double dx = gx - px, dy = gy - py, dz = gz - pz;
    const double r = fabs(dx * dx + dy * dy + dz * dz);
    if (!(r < 1)) return 0.0;
    const double r1 = 1 - r, r2 = r1 * r1;
    return r2 * r2;
}

// hat weight per axis, product over the axes (GPU_kernel.cu poly_shape,
// interpolate_weights)
__device__ inline double hat(double r) {
    const double a = fabs(r);
    return a < 1 ? 1 - a : 0.0;
}

// assign_density_kernel (GPU_kernel.cu:3655-3725), one thread per grid cell.
// The reference visits the hash cells within +-2; a particle two cells away
// is at least 1.5 cells from this cell's centre and weighs exactly zero, and
// an empty hash cell other than the cell's own adds a zero weight, so the
// +-1 cells in the same order give the same sums.  An empty hash cell takes
// the reference's `end < start` branch (its end table is never cleared):
// the cell's own adds weight 1 with density 0.
__global__ void __launch_bounds__(256) k_assign_density(const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ end,
                                                        const float4* __restrict__ pos,
                                                        const float* __restrict__ rho, float* __restrict__ grid,
                                                        uint32_t dim, double invh, uint32_t n, double ox, double oy,
                                                        double oz) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t G = dim * dim * dim;
    if (c >= G) return;
    const int d = (int)dim;
    const int i = (int)(c % dim), j = (int)((c / dim) % dim), k = (int)(c / (dim * dim));
    const double gx = (double)i + 0.5, gy = (double)j + 0.5, gz = (double)k + 0.5;
    double sum = 0, tw = 0;
    for (int kk = k - 1; kk <= k + 1; kk++)
        for (int jj = j - 1; jj <= j + 1; jj++)
            for (int ii = i - 1; ii <= i + 1; ii++) {
                if (kk < 0 || kk >= d || jj < 0 || jj >= d || ii < 0 || ii >= d) continue;
                const uint32_t q = ((uint32_t)kk * dim + (uint32_t)jj) * dim + (uint32_t)ii;
                const uint32_t s = start[q];
                if (s == 0xFFFFFFFFu) {
                    if (q == c) tw += 1.0;
                    continue;
                }
                const uint32_t e = end[q];
                for (uint32_t p = s; p < e; p++) {
                    if (p >= n) continue;
                    const float4 a = pos[p];
                    const float px = (float)((double)a.x - ox), py = (float)((double)a.y - oy),
                                pz = (float)((double)a.z - oz);
                    const double w = density_weight(gx, gy, gz, px * invh, py * invh, pz * invh);
                    sum += w * (double)rho[p];
                    tw += w;
                }
            }
    if (tw < 1e-12) tw = 1e+24;
    grid[c] = (float)(sum / tw);
}

// The same grid value with one workgroup per grid cell, for grids whose
// cells hold many smoke particles (the reference's rule gives 7^3 cells for a
// unit-sized cloud: a thread per cell would walk a seventh of the smoke
// alone).  1024 threads stride over the sorted smoke of each of the 27
// cells, and a fixed-order tree in LDS adds their partial sums, so the
// result is deterministic and within one float rounding of the sequential
// sum.
constexpr int kDensityBlock = 1024;

__global__ void __launch_bounds__(kDensityBlock) k_assign_density_block(
    const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, const float4* __restrict__ pos,
    const float* __restrict__ rho, float* __restrict__ grid, uint32_t dim, double invh, uint32_t n, double ox,
    double oy, double oz) {
    __shared__ double red[2][kDensityBlock];
    const uint32_t c = blockIdx.x;
    const int d = (int)dim;
    const int i = (int)(c % dim), j = (int)((c / dim) % dim), k = (int)(c / (dim * dim));
    const double gx = (double)i + 0.5, gy = (double)j + 0.5, gz = (double)k + 0.5;
    double sum = 0, tw = 0;
    bool own_empty = false;
    for (int kk = k - 1; kk <= k + 1; kk++)
        for (int jj = j - 1; jj <= j + 1; jj++) {
            if (kk < 0 || kk >= d || jj < 0 || jj >= d) continue;
            const uint32_t row = ((uint32_t)kk * dim + (uint32_t)jj) * dim;
            for (int ii = i - 1; ii <= i + 1; ii++) {
                if (ii < 0 || ii >= d) continue;
                const uint32_t q = row + (uint32_t)ii;
                const uint32_t s0 = start[q];
                if (s0 == 0xFFFFFFFFu) {
                    if (q == c) own_empty = true;
                    continue;
                }
                const uint32_t e0 = min(end[q], n);
                for (uint32_t p = s0 + threadIdx.x; p < e0; p += kDensityBlock) {
                    const float4 a = pos[p];
                    const float px = (float)((double)a.x - ox), py = (float)((double)a.y - oy),
                                pz = (float)((double)a.z - oz);
                    const double w = density_weight(gx, gy, gz, px * invh, py * invh, pz * invh);
                    sum += w * (double)rho[p];
                    tw += w;
                }
            }
        }
    red[0][threadIdx.x] = sum;
    red[1][threadIdx.x] = tw;
    __syncthreads();
    for (int s = kDensityBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double t = red[1][0] + (own_empty ? 1.0 : 0.0);
        if (t < 1e-12) t = 1e+24;
        grid[c] = (float)(red[0][0] / t);
    }
}

// grad_density_kernel (GPU_kernel.cu:3752-3791): periodic central
// differences, (r - l) * 0.5 * invh per axis, stored as float.
__global__ void __launch_bounds__(256) k_grad_density(const float* __restrict__ phi, float* __restrict__ grad,
                                                      uint32_t dim, double invh) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t G = dim * dim * dim;
    if (c >= G) return;
    const uint32_t i = c % dim, j = (c / dim) % dim, k = c / (dim * dim);
    const uint32_t im = (i + dim - 1) % dim, ip = (i + 1) % dim, jm = (j + dim - 1) % dim, jp = (j + 1) % dim;
    const uint32_t km = (k + dim - 1) % dim, kp = (k + 1) % dim;
    auto at = [&](uint32_t a, uint32_t b, uint32_t e) { return (double)phi[((size_t)e * dim + b) * dim + a]; };
    grad[3 * (size_t)c] = (float)((at(ip, j, k) - at(im, j, k)) * 0.5 * invh);
    grad[3 * (size_t)c + 1] = (float)((at(i, jp, k) - at(i, jm, k)) * 0.5 * invh);
    grad[3 * (size_t)c + 2] = (float)((at(i, j, kp) - at(i, j, km)) * 0.5 * invh);
}

// baroclinic_kernel (GPU_kernel.cu:3810-3880), one thread per vortex: the
// hat-weighted gradient over the +-1 grid cells (the +-2 the reference visits
// add exact zeros), then w += dt_beta * 0.1 * (sum x g).
__global__ void __launch_bounds__(256) k_baroclinic(const float4* __restrict__ vpos, double* __restrict__ wx,
                                                    double* __restrict__ wy, double* __restrict__ wz,
                                                    const float* __restrict__ grad, double dt_beta, uint32_t n,
                                                    uint32_t dim, double inv_h, double ox, double oy, double oz,
                                                    double g0, double g1, double g2) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float4 a = vpos[t];
    const float px = (float)((double)a.x - ox), py = (float)((double)a.y - oy), pz = (float)((double)a.z - oz);
    const double qx = px * inv_h, qy = py * inv_h, qz = pz * inv_h;
    const int d = (int)dim;
    auto cell = [&](double q) { return (int)fmin(fmax(floor(q), -2.0), (double)d + 1.0); };
    const int i = cell(qx), j = cell(qy), k = cell(qz);
    double sx = 0, sy = 0, sz = 0;
    for (int kk = k - 1; kk <= k + 1; kk++)
        for (int jj = j - 1; jj <= j + 1; jj++)
            for (int ii = i - 1; ii <= i + 1; ii++) {
                if (kk < 0 || kk >= d || jj < 0 || jj >= d || ii < 0 || ii >= d) continue;
                const size_t q = ((size_t)kk * dim + (size_t)jj) * dim + (size_t)ii;
                const double w = hat((double)ii + 0.5 - qx) * hat((double)jj + 0.5 - qy) * hat((double)kk + 0.5 - qz);
                sx += (double)grad[3 * q] * w;
                sy += (double)grad[3 * q + 1] * w;
                sz += (double)grad[3 * q + 2] * w;
            }
    // cross_uxv(sum, g) (GPU_kernel.cu:213-216)
    const double bx = sy * g2 - sz * g1, by = sz * g0 - sx * g2, bz = sx * g1 - sy * g0;
    wx[t] += dt_beta * 0.1 * bx;
    wy[t] += dt_beta * 0.1 * by;
    wz[t] += dt_beta * 0.1 * bz;
}

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + 255) / 256); }

// Smoke particles per grid cell from which a workgroup takes each cell.
constexpr uint64_t kDensityBlockMin = 64;

template <typename T>
pppm_status grow(T** p, size_t* cap, size_t count) {
    if (count <= *cap && *p) return PPPM_OK;
    std::lock_guard<std::mutex> lk(runtime_mutex());
    if (*p) (void)hipFree((void*)*p);
    *p = nullptr;
    *cap = 0;
    PPPM_HIP_TRY(hipMalloc((void**)p, (count ? count : 1) * sizeof(T)));
    *cap = count;
    return PPPM_OK;
}

template <typename T>
void release(T** p) {
    if (*p) (void)hipFree((void*)*p);
    *p = nullptr;
}

}  // namespace
}  // namespace pppm

using namespace pppm;

struct pppm_density {
    // smoke particles: keys and values of the sort, its scratch, sorted copies
    uint32_t *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
    uint8_t* sort_scr = nullptr;
    float4* spos = nullptr;
    float* srho = nullptr;
    size_t capN = 0, capK0 = 0, capV0 = 0, capK1 = 0, capV1 = 0, capScr = 0, capP = 0, capR = 0;
    // grid: start/end tables, density, gradient (x, y, z per cell)
    uint32_t *start = nullptr, *end = nullptr;
    float *rho = nullptr, *grad = nullptr;
    size_t capS = 0, capE = 0, capG = 0, capD = 0;
    float* bbox = nullptr;
    size_t capB = 0;
    // the last call's grid
    uint32_t dim = 0;
    double origin[3] = {0, 0, 0};
    double h = 0;

    ~pppm_density() {
        std::lock_guard<std::mutex> lk(runtime_mutex());
        release(&k0), release(&v0), release(&k1), release(&v1), release(&sort_scr), release(&spos),
            release(&srho), release(&start), release(&end), release(&rho), release(&grad), release(&bbox);
    }
};

extern "C" pppm_status pppm_density_create(pppm_density** out) {
    if (!out) FAIL(PPPM_ERR_INVALID, "null out");
    *out = new pppm_density();
    return PPPM_OK;
}

extern "C" void pppm_density_destroy(pppm_density* d) { delete d; }

extern "C" pppm_status pppm_density_grid(const pppm_density* d, const float** rho, const float** grad3,
                                         uint32_t* dim, double origin[3], double* h) {
    if (!d) FAIL(PPPM_ERR_INVALID, "null density");
    if (d->dim == 0) FAIL(PPPM_ERR_STATE, "no density grid yet");
    if (rho) *rho = d->rho;
    if (grad3) *grad3 = d->grad;
    if (dim) *dim = d->dim;
    if (origin)
        for (int q = 0; q < 3; q++) origin[q] = d->origin[q];
    if (h) *h = d->h;
    return PPPM_OK;
}

extern "C" pppm_status pppm_add_baroclinic(pppm_density* d, const float* heat_pos4, const float* heat_rho,
                                           uint32_t nheat, const float* cent4, double* wx, double* wy, double* wz,
                                           uint32_t n, double dt_beta, const double gvec[3], double cell_target,
                                           uint32_t max_dim, void* hip_stream) {
    if (!d || !gvec) FAIL(PPPM_ERR_INVALID, "null argument");
    if (nheat == 0) return PPPM_OK;  // no smoke: no source (the reference needs smoke to run this stage)
    if (!heat_pos4 || !heat_rho) FAIL(PPPM_ERR_INVALID, "null smoke array");
    if (n > 0 && (!cent4 || !wx || !wy || !wz)) FAIL(PPPM_ERR_INVALID, "null vortex array");
    if (!(cell_target > 0) || max_dim == 0 || max_dim > 1024) FAIL(PPPM_ERR_INVALID, "cell_target / max_dim");
    hipStream_t st = (hipStream_t)hip_stream;
    // setDensityDomain (VortexSimulationSystemImp.cpp:320-367) over the
    // smoke's bounding box, then the grid rule of addBaroclinicForce
    TRY(grow(&d->bbox, &d->capB, bbox_scratch_floats()));
    TRY(launch_bbox((const float4*)heat_pos4, nheat, d->bbox, st));
    float b[6];
    PPPM_HIP_TRY(hipMemcpyAsync(b, d->bbox + bbox_scratch_floats() - 6, sizeof b, hipMemcpyDeviceToHost, st));
    PPPM_HIP_TRY(hipStreamSynchronize(st));
    const double mnx = b[0], mny = b[1], mnz = b[2], mxx = b[3], mxy = b[4], mxz = b[5];
    const double L = 1.2 * std::max(std::max(mxx - mnx, mxy - mny), mxz - mnz);
    if (!(L > 0)) FAIL(PPPM_ERR_INVALID, "degenerate smoke extent");
    d->origin[0] = (mnx + 0.5 * (mxx - mnx)) - 0.5 * L;
    d->origin[1] = (mny + 0.5 * (mxy - mny)) - 0.5 * L;
    d->origin[2] = (mnz + 0.5 * (mxz - mnz)) - 0.5 * L;
    // floor(L / cell) + 1, capped at max_dim before the conversion
    const uint32_t dim = (uint32_t)std::min(std::floor(L / cell_target) + 1.0, (double)max_dim);
    const double h = L / (double)dim;
    d->dim = dim;
    d->h = h;
    const size_t G = (size_t)dim * dim * dim;
    // hash the smoke on the density grid (the reference's keys and tables)
    TRY(grow(&d->k0, &d->capK0, nheat));
    TRY(grow(&d->v0, &d->capV0, nheat));
    TRY(grow(&d->k1, &d->capK1, nheat));
    TRY(grow(&d->v1, &d->capV1, nheat));
    TRY(grow(&d->sort_scr, &d->capScr, sort_scratch_bytes(nheat)));
    TRY(grow(&d->spos, &d->capP, nheat));
    TRY(grow(&d->srho, &d->capR, nheat));
    TRY(grow(&d->start, &d->capS, G));
    TRY(grow(&d->end, &d->capE, G));
    TRY(grow(&d->rho, &d->capG, G));
    TRY(grow(&d->grad, &d->capD, 3 * G));
    HashGeom hg;
    hg.g = dim;
    hg.hf = (float)h;
    for (int q = 0; q < 3; q++) hg.of[q] = (float)d->origin[q];
    TRY(launch_calc_keys((const float4*)heat_pos4, nheat, hg, d->k0, nullptr, st));
    uint32_t *ks = nullptr, *vs = nullptr;
    TRY(radix_sort_pairs(d->k0, d->v0, d->k1, d->v1, nheat, key_bits_for(G), true, d->sort_scr, &ks, &vs, st));
    TRY(launch_cell_tables(ks, nheat, (uint32_t)G, d->start, d->end, st));
    TRY(launch_gather(vs, heat_pos4, d->spos, 4, PPPM_F32, nheat, st));
    TRY(launch_gather(vs, heat_rho, d->srho, 1, PPPM_F32, nheat, st));
    // AssignDensity, GradDensity, ParticleGetBaroclinic
    if ((uint64_t)nheat >= kDensityBlockMin * (uint64_t)G)
        hipLaunchKernelGGL(k_assign_density_block, dim3((uint32_t)G), dim3(kDensityBlock), 0, st, d->start, d->end,
                           d->spos, d->srho, d->rho, dim, 1.0 / h, nheat, d->origin[0], d->origin[1],
                           d->origin[2]);
    else
        hipLaunchKernelGGL(k_assign_density, dim3(blocks_for(G)), dim3(256), 0, st, d->start, d->end, d->spos,
                           d->srho, d->rho, dim, 1.0 / h, nheat, d->origin[0], d->origin[1], d->origin[2]);
    PPPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_grad_density, dim3(blocks_for(G)), dim3(256), 0, st, d->rho, d->grad, dim, 1.0 / h);
    PPPM_LAUNCH_CHECK();
    if (n > 0) {
        hipLaunchKernelGGL(k_baroclinic, dim3(blocks_for(n)), dim3(256), 0, st, (const float4*)cent4, wx, wy, wz,
                           d->grad, dt_beta, n, dim, 1.0 / (double)hg.hf, d->origin[0], d->origin[1],
                           d->origin[2], gvec[0], gvec[1], gvec[2]);
        PPPM_LAUNCH_CHECK();
    }
    return PPPM_OK;
}

// ---- the reference's free functions (ParticleMesh.h), null stream --------
extern "C" void AssignDensity(uint32_t* start, uint32_t* end, float4* pos, float* rho, float* vort_size,
                              double cell_h, float* gridval, uint3 griddim, uint3 hashdim, uint32_t num_particle,
                              double4 origin) {
    (void)vort_size;
    (void)hashdim;
    const size_t G = (size_t)griddim.x * griddim.x * griddim.x;
    if (G == 0) return;
    hipLaunchKernelGGL(k_assign_density, dim3(blocks_for(G)), dim3(256), 0, 0, start, end, pos, rho, gridval,
                       griddim.x, 1.0 / cell_h, num_particle, origin.x, origin.y, origin.z);
}

extern "C" void GradDensity(float* phi, float3* dphi, double cell_h, uint32_t dimx, uint32_t dimy, uint32_t dimz) {
    (void)dimy;
    (void)dimz;
    const size_t G = (size_t)dimx * dimx * dimx;
    if (G == 0) return;
    hipLaunchKernelGGL(k_grad_density, dim3(blocks_for(G)), dim3(256), 0, 0, phi, (float*)dphi, dimx, 1.0 / cell_h);
}

extern "C" void ParticleGetBaroclinic(float4* vortex_pos, double* vort_x, double* vort_y, double* vort_z,
                                      float3* grad_rho, double dt_beta, uint32_t num_vort, uint32_t dimx,
                                      uint32_t dimy, uint32_t dimz, double inv_h, double4 origin) {
    (void)dimy;
    (void)dimz;
    if (num_vort == 0) return;
    hipLaunchKernelGGL(k_baroclinic, dim3(blocks_for(num_vort)), dim3(256), 0, 0, vortex_pos, vort_x, vort_y, vort_z,
                       (const float*)grad_rho, dt_beta, num_vort, dimx, inv_h, origin.x, origin.y, origin.z, 0.0,
                       10.0, 0.0);
}
