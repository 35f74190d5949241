// M2P interpolation of the far-field buffer fused with the near-field
// Biot-Savart correction and the unsort to the caller's order
// (GPU_kernel.cu:2886-2964 BiotSavartInterpolate_far_field_kernel2,
//  2262-2436 BiotSavartPPCorrScaleMN_kernel, BiotSavartSolverImp.cpp:452-484).
#include "pppm_internal.hpp"

namespace pppm {

struct NearConst {
    float rbox;        // largest float r with (double)r < (K + 1/2) h
    float inv_a;       // 1 / core radius
    float inv_a3;      // 1 / a^3
    float inv4pi;      // 1 / (4 pi) with the reference's 4 pi
};

// q(rho) / (4 pi r^3) for the expcubic core, q = 1 - exp(-(r/a)^3)
// (GPU_kernel.cu:2389-2402), small-x series to avoid cancellation.
__device__ inline float coeff_expcubic(float r2, const NearConst& nc) {
    float rinv = __frsqrt_rn(r2);
    float r = r2 * rinv;
    float x = r2 * r * nc.inv_a3;
    float q;
    if (x < 0.1f) {
        q = x * (1.0f + x * (-0.5f + x * (1.0f / 6.0f + x * (-1.0f / 24.0f + x * (1.0f / 120.0f)))));
    } else {
        q = 1.0f - __expf(-x);
    }
    return q * (rinv * rinv * rinv) * nc.inv4pi;
}

// Gaussian core: q = erf(rho/sqrt2) - sqrt(2/pi) rho exp(-rho^2/2)
__device__ inline float coeff_gaussian(float r2, const NearConst& nc) {
    float rinv = __frsqrt_rn(r2);
    float r = r2 * rinv;
    float rho = r * nc.inv_a;
    float q;
    if (rho < 0.5f) {
        float p2 = rho * rho;
        float s = 1.0f / 3.0f + p2 * (-1.0f / 10.0f + p2 * (1.0f / 56.0f + p2 * (-1.0f / 432.0f + p2 * (1.0f / 3840.0f))));
        q = 0.7978845608f * rho * p2 * s;
    } else {
        q = erff(rho * 0.7071067812f) - 0.7978845608f * rho * __expf(-0.5f * rho * rho);
    }
    return q * (rinv * rinv * rinv) * nc.inv4pi;
}

// One lane per sorted evaluation point.  The vortices of a row of hash cells
// (fixed y, z; x from i-K-1 to i+K+1 around the point's cell i) are one
// contiguous sorted segment.  Lanes whose points share a cell run identical
// loops and read the same vortices (broadcast loads); a tighter per-point
// range (the point's box, 2K+2 cells per axis) measured slower because it
// breaks that lockstep.
template <int KERNEL, bool SEGMENT, typename TO>
__global__ void __launch_bounds__(256) k_m2p_near(const float4* __restrict__ epos, const float4* __restrict__ eoth,
                                                  const uint32_t* __restrict__ eidx, uint32_t m,
                                                  const uint32_t* __restrict__ off, const float4* __restrict__ spos,
                                                  const float4* __restrict__ somg, const float* __restrict__ far,
                                                  uint32_t g, int K, double invh, NearConst nc, TO* __restrict__ ux,
                                                  TO* __restrict__ uy, TO* __restrict__ uz, uint32_t u_stride) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int gi = (int)g;
    const size_t G = (size_t)g * g * g;
    const float4 a = epos[t];
    const int i = cell_of(a.x, invh), j = cell_of(a.y, invh), k = cell_of(a.z, invh);

    // ---- M2P: hat weights about cell centres, in-grid cells only
    float sx = 0.f, sy = 0.f, sz = 0.f;
    if (i >= 0 && i < gi && j >= 0 && j < gi && k >= 0 && k < gi) {
        double xh = (double)a.x * invh - 0.5, yh = (double)a.y * invh - 0.5, zh = (double)a.z * invh - 0.5;
        int i0 = (int)floor(xh), j0 = (int)floor(yh), k0 = (int)floor(zh);
        float fx = (float)(xh - i0), fy = (float)(yh - j0), fz = (float)(zh - k0);
        float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
#pragma unroll
        for (int c2 = 0; c2 < 2; c2++) {
            int kk = k0 + c2;
            if (kk < 0 || kk >= gi) continue;
#pragma unroll
            for (int b2 = 0; b2 < 2; b2++) {
                int jj = j0 + b2;
                if (jj < 0 || jj >= gi) continue;
#pragma unroll
                for (int a2 = 0; a2 < 2; a2++) {
                    int ii = i0 + a2;
                    if (ii < 0 || ii >= gi) continue;
                    size_t q = ((size_t)kk * g + jj) * g + ii;
                    float w = wx[a2] * wy[b2] * wz[c2];
                    sx += w * far[q];
                    sy += w * far[G + q];
                    sz += w * far[2 * G + q];
                }
            }
        }
    }

    // ---- near field over the hash rows of the point's box
    float vx = 0.f, vy = 0.f, vz = 0.f;
    if (K >= 0) {
        float4 bpos = make_float4(0.f, 0.f, 0.f, 0.f);
        if (SEGMENT) bpos = eoth[t];
        const int R = K + 1;
        const int xlo = max(i - R, 0), xhi = min(i + R, gi - 1);
        const int ylo = max(j - R, 0), yhi = min(j + R, gi - 1);
        const int zlo = max(k - R, 0), zhi = min(k + R, gi - 1);
        const uint32_t nsrc = off[G];
        if (xlo <= xhi)
            for (int kk = zlo; kk <= zhi; kk++) {
                for (int jj = ylo; jj <= yhi; jj++) {
                    const uint32_t rowc = ((uint32_t)kk * g + (uint32_t)jj) * g;
                    const uint32_t p0 = off[rowc + xlo], p1 = min(off[rowc + xhi + 1], nsrc);
                    for (uint32_t p = p0; p < p1; p++) {
                        const float4 s = spos[p];
                        const float dx = a.x - s.x, dy = a.y - s.y, dz = a.z - s.z;
                        if (!(fabsf(dx) <= nc.rbox && fabsf(dy) <= nc.rbox && fabsf(dz) <= nc.rbox)) continue;
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < 1e-24f) continue;
                        if (SEGMENT) {
                            const float ex = bpos.x - s.x, ey = bpos.y - s.y, ez = bpos.z - s.z;
                            if (ex * ex + ey * ey + ez * ez < 1e-24f) continue;
                        }
                        const float cf =
                            KERNEL == PPPM_KERNEL_GAUSSIAN ? coeff_gaussian(r2, nc) : coeff_expcubic(r2, nc);
                        const float4 o = somg[p];
                        vx += (o.y * dz - o.z * dy) * cf;
                        vy += (o.z * dx - o.x * dz) * cf;
                        vz += (o.x * dy - o.y * dx) * cf;
                    }
                }
            }
    }
    const uint32_t dt = eidx[t];
    if (dt >= m) return;
    const size_t dst = (size_t)dt * u_stride;
    ux[dst] = (TO)(sx + vx);
    uy[dst] = (TO)(sy + vy);
    uz[dst] = (TO)(sz + vz);
}

template <int KERNEL, bool SEGMENT>
static void launch_typed(const NearArgs& a, const NearConst& nc, hipStream_t st) {
    dim3 grid(div_up(a.m, 256)), blk(256);
    if (a.u_dtype == PPPM_F64)
        hipLaunchKernelGGL((k_m2p_near<KERNEL, SEGMENT, double>), grid, blk, 0, st, a.epos, a.eoth, a.eidx, a.m,
                           a.off, a.spos, a.somg, a.far, a.g, a.K, a.invh, nc, (double*)a.ux, (double*)a.uy,
                           (double*)a.uz, a.u_stride);
    else
        hipLaunchKernelGGL((k_m2p_near<KERNEL, SEGMENT, float>), grid, blk, 0, st, a.epos, a.eoth, a.eidx, a.m,
                           a.off, a.spos, a.somg, a.far, a.g, a.K, a.invh, nc, (float*)a.ux, (float*)a.uy,
                           (float*)a.uz, a.u_stride);
}

// largest float f with (double)f < x (x > 0; returns -1 otherwise)
static float float_below(double x) {
    if (!(x > 0)) return -1.0f;
    float f = (float)x;
    while ((double)f >= x) f = nextafterf(f, 0.0f);
    while ((double)nextafterf(f, 3.4e38f) < x) f = nextafterf(f, 3.4e38f);
    return f;
}

pppm_status launch_m2p_near(const NearArgs& a, hipStream_t st) {
    if (a.m == 0) return PPPM_OK;
    NearConst nc;
    const double h = 1.0 / a.invh;
    nc.rbox = float_below(((double)a.K + 0.5) / a.invh);
    nc.inv_a = (float)(1.0 / a.core_a);
    nc.inv_a3 = (float)(1.0 / (a.core_a * a.core_a * a.core_a));
    nc.inv4pi = (float)(1.0 / kRefFourPi);
    (void)h;
    const bool seg = a.eoth != nullptr;
    if (a.kernel == PPPM_KERNEL_GAUSSIAN) {
        if (seg) launch_typed<PPPM_KERNEL_GAUSSIAN, true>(a, nc, st);
        else launch_typed<PPPM_KERNEL_GAUSSIAN, false>(a, nc, st);
    } else {
        if (seg) launch_typed<PPPM_KERNEL_EXPCUBIC, true>(a, nc, st);
        else launch_typed<PPPM_KERNEL_EXPCUBIC, false>(a, nc, st);
    }
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

// Accepted near-field pairs (box test passed, r^2 >= 1e-24) of the evaluation
// points against the vortices, for FLOP accounting; one atomic per wave.
__global__ void __launch_bounds__(256) k_count_pairs(const float4* __restrict__ epos, uint32_t m,
                                                     const uint32_t* __restrict__ off, const float4* __restrict__ spos,
                                                     uint32_t g, int K, double invh, float rbox,
                                                     unsigned long long* __restrict__ total) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt = 0;
    if (t < m) {
        const int gi = (int)g;
        const size_t G = (size_t)g * g * g;
        const float4 a = epos[t];
        const int i = cell_of(a.x, invh), j = cell_of(a.y, invh), k = cell_of(a.z, invh);
        const int R = K + 1;
        const uint32_t nsrc = off[G];
        const int ilo = max(i - R, 0), ihi = min(i + R, gi - 1);
        if (K >= 0 && ilo <= ihi)
            for (int kk = max(k - R, 0); kk <= min(k + R, gi - 1); kk++)
                for (int jj = max(j - R, 0); jj <= min(j + R, gi - 1); jj++) {
                    const uint32_t rowc = ((uint32_t)kk * g + (uint32_t)jj) * g;
                    const uint32_t p0 = off[rowc + ilo], p1 = min(off[rowc + ihi + 1], nsrc);
                    for (uint32_t p = p0; p < p1; p++) {
                        const float4 s = spos[p];
                        const float dx = a.x - s.x, dy = a.y - s.y, dz = a.z - s.z;
                        if (fabsf(dx) <= rbox && fabsf(dy) <= rbox && fabsf(dz) <= rbox &&
                            dx * dx + dy * dy + dz * dz >= 1e-24f)
                            cnt++;
                    }
                }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(total, cnt);
}

pppm_status launch_count_pairs(const NearArgs& a, unsigned long long* d_total, hipStream_t st) {
    if (a.m == 0) return PPPM_OK;
    float rbox = float_below(((double)a.K + 0.5) / a.invh);
    hipLaunchKernelGGL(k_count_pairs, dim3(div_up(a.m, 256)), dim3(256), 0, st, a.epos, a.m, a.off, a.spos, a.g, a.K,
                       a.invh, rbox, d_total);
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

}  // namespace pppm
