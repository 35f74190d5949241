// Spatial hash for gfx950: cell keys, stable LSD radix sort of (key, index),
// cell start/end tables and gathers/scatters.  Restates the behaviour of
// particle_hash/particle_hash/particle_hash_kernel.cu:32-184 and
// SpaceHashFunctional.cu:53-145 (keys bit-exact; the reference's Thrust sort is
// replaced by a stable radix sort: order by (key, original index)).
#include "pppm_internal.hpp"

namespace pppm {

constexpr int kSortThreads = 256;  // 4 wave64s
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 keys per workgroup
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_calc_keys(const float4* __restrict__ pos, uint32_t n, uint32_t g, float hf,
                                                   float ox, float oy, float oz, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ idx, bool octant, bool clamp) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    keys[i] = clamp    ? eval_cell_key(p.x, p.y, p.z, g, hf, ox, oy, oz, octant)
              : octant ? hash_cell_octant_key(p.x, p.y, p.z, g, hf, ox, oy, oz)
                       : hash_cell_key(p.x, p.y, p.z, g, hf, ox, oy, oz);
    if (idx) idx[i] = i;
}

pppm_status launch_calc_keys(const float4* pos, uint32_t n, const HashGeom& hg, uint32_t* keys, uint32_t* idx,
                             hipStream_t st, bool octant, bool clamp) {
    if (n == 0) return PPPM_OK;
    if (octant && (uint64_t)hg.g * hg.g * hg.g * 8 > ((uint64_t)1 << 32)) return PPPM_ERR_INVALID;
    hipLaunchKernelGGL(k_calc_keys, dim3(div_up(n, 256)), dim3(256), 0, st, pos, n, hg.g, hg.hf, hg.of[0], hg.of[1],
                       hg.of[2], keys, idx, octant, clamp);
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

// ---------------------------------------------------------------------------
// Lanes of this wave whose digit equals mine (valid lanes only).
__device__ inline uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t bb = __ballot(valid && bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__device__ inline uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline uint32_t block_scan_256(uint32_t v, uint32_t* lw, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) lw[wave] = incl;
    __syncthreads();
    uint32_t wb = 0, tot = 0;
    for (int w = 0; w < 4; w++) {
        uint32_t s = lw[w];
        if (w < wave) wb += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wb + incl - v;
}

// Per-workgroup digit histogram, stored digit-major: hist[d * nblocks + b].
__global__ void __launch_bounds__(kSortThreads) k_radix_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                                int shift, uint32_t* __restrict__ hist,
                                                                uint32_t nblocks) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; j++) {
        uint32_t i = base + j * kSortThreads + threadIdx.x;
        bool valid = i < n;
        uint32_t d = valid ? (keys[i] >> shift) & (kRadix - 1) : 0u;
        uint64_t m = match_digit(d, valid);
        if (valid && (m & lanemask_lt()) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one 4096-key tile: wave w owns items [w*1024, w*1024+1024)
// of the tile, in 16 rounds of 64 consecutive items, so (wave, round, lane)
// order is item order.  Ranks come from ballots; the tile is staged in LDS in
// digit order and written out coalesced.
__global__ void __launch_bounds__(kSortThreads) k_radix_downsweep(const uint32_t* __restrict__ kin,
                                                                  const uint32_t* __restrict__ vin,
                                                                  uint32_t* __restrict__ kout,
                                                                  uint32_t* __restrict__ vout, uint32_t n, int shift,
                                                                  const uint32_t* __restrict__ hist_scan,
                                                                  uint32_t nblocks) {
    __shared__ uint32_t wcnt[4][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t sk[kSortTile];
    __shared__ uint32_t sv[kSortTile];
    __shared__ uint32_t lw[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int d = tid; d < 4 * kRadix; d += kSortThreads) (&wcnt[0][0])[d] = 0;
    __syncthreads();

    const uint32_t base = blockIdx.x * kSortTile;
    const uint32_t wbase = base + wave * (kSortTile / 4);
    uint32_t k[kSortItems], v[kSortItems], rank[kSortItems];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        uint32_t i = wbase + r * 64 + lane;
        bool valid = i < n;
        k[r] = valid ? kin[i] : 0u;
        v[r] = valid ? (vin ? vin[i] : i) : 0u;
        uint32_t d = (k[r] >> shift) & (kRadix - 1);
        uint64_t m = match_digit(d, valid);
        uint32_t below = (uint32_t)__popcll(m & lt);
        uint32_t before = valid ? wcnt[wave][d] : 0u;
        rank[r] = before + below;
        if (valid && below == 0) wcnt[wave][d] = before + (uint32_t)__popcll(m);
    }
    __syncthreads();
    {
        const int d = tid;  // kSortThreads == kRadix
        uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d], c3 = wcnt[3][d];
        uint32_t tot_all;
        uint32_t ex = block_scan_256(c0 + c1 + c2 + c3, lw, tot_all);
        wcnt[0][d] = 0;
        wcnt[1][d] = c0;
        wcnt[2][d] = c0 + c1;
        wcnt[3][d] = c0 + c1 + c2;
        dstart[d] = ex;
        gbase[d] = hist_scan[(size_t)d * nblocks + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        uint32_t i = wbase + r * 64 + lane;
        if (i < n) {
            uint32_t d = (k[r] >> shift) & (kRadix - 1);
            uint32_t lp = dstart[d] + wcnt[wave][d] + rank[r];
            if (lp < (uint32_t)kSortTile) {
                sk[lp] = k[r];
                sv[lp] = v[r];
            }
        }
    }
    __syncthreads();
    const uint32_t nvalid = (n - base) < (uint32_t)kSortTile ? (n - base) : (uint32_t)kSortTile;
    for (uint32_t j = tid; j < nvalid; j += kSortThreads) {
        uint32_t key = sk[j];
        uint32_t d = (key >> shift) & (kRadix - 1);
        uint32_t p = gbase[d] + (j - dstart[d]);
        if (p < n) {
            kout[p] = key;
            vout[p] = sv[j];
        }
    }
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sort_scratch_bytes(uint32_t n) {
    uint64_t nb = div_up64(n ? n : 1, kSortTile);
    uint64_t hw = nb * kRadix;
    return align256(hw * 4) * 2 + align256(scan_scratch_words(hw) * 4);
}

uint32_t key_bits_for(uint64_t ncells) {
    uint32_t b = 1;
    while (b < 32 && ((uint64_t)1 << b) < ncells) b++;
    return b;
}

pppm_status radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                             uint32_t key_bits, bool vals_identity, void* scratch, uint32_t** keys_out,
                             uint32_t** vals_out, hipStream_t st) {
    *keys_out = keys;
    *vals_out = vals;
    if (n == 0) return PPPM_OK;
    const uint32_t nb = div_up(n, kSortTile);
    const uint64_t hw = (uint64_t)nb * kRadix;
    if (hw > 0xFFFFFFFFull) return PPPM_ERR_INVALID;
    uint8_t* sp = (uint8_t*)scratch;
    uint32_t* hist = (uint32_t*)sp;
    uint32_t* hscan = (uint32_t*)(sp + align256(hw * 4));
    uint32_t* sscr = (uint32_t*)(sp + 2 * align256(hw * 4));
    int passes = (int)div_up(key_bits ? key_bits : 1, kRadixBits);
    uint32_t *ks = keys, *kd = keys_alt;
    uint32_t *vbuf = vals, *vd = vals_alt;
    const uint32_t* vread = vals_identity ? nullptr : vals;  // null: value = item index
    for (int p = 0; p < passes; p++) {
        int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_upsweep, dim3(nb), dim3(kSortThreads), 0, st, ks, n, shift, hist, nb);
        PPPM_LAUNCH_CHECK();
        pppm_status s = exclusive_scan_u32(hist, (uint32_t)hw, hscan, sscr, st);
        if (s != PPPM_OK) return s;
        hipLaunchKernelGGL(k_radix_downsweep, dim3(nb), dim3(kSortThreads), 0, st, ks, vread, kd, vd, n, shift,
                           hscan, nb);
        PPPM_LAUNCH_CHECK();
        uint32_t* t = ks;
        ks = kd;
        kd = t;
        t = vbuf;
        vbuf = vd;
        vd = t;
        vread = vbuf;
    }
    *keys_out = ks;
    *vals_out = vbuf;
    return PPPM_OK;
}

// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* __restrict__ a, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) a[i] = v;
}

// particle_hash_kernel.cu:74-126 (FindCellStartD), neighbour key read directly.
// cell = key >> shift
__global__ void __launch_bounds__(256) k_cell_bounds(const uint32_t* __restrict__ ks, uint32_t n, uint32_t ncells,
                                                     uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                     uint32_t shift) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = ks[i] >> shift;
    const uint32_t kp = i > 0 ? ks[i - 1] >> shift : 0u;
    if (i == 0 || k != kp) {
        if (k < ncells) start[k] = i;
        if (i > 0 && kp < ncells) end[kp] = i;
    }
    if (i == n - 1 && k < ncells) end[k] = n;
}

pppm_status launch_fill_u32(uint32_t* a, uint64_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return PPPM_OK;
    uint64_t blocks = div_up64(n, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, st, a, n, v);
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

pppm_status launch_cell_tables(const uint32_t* keys_sorted, uint32_t n, uint32_t ncells, uint32_t* start,
                               uint32_t* end, hipStream_t st, uint32_t shift) {
    pppm_status s = launch_fill_u32(start, ncells, kEmptyCell, st);
    if (s != PPPM_OK) return s;
    s = launch_fill_u32(end, ncells, kEmptyCell, st);
    if (s != PPPM_OK) return s;
    if (n == 0) return PPPM_OK;
    hipLaunchKernelGGL(k_cell_bounds, dim3(div_up(n, 256)), dim3(256), 0, st, keys_sorted, n, ncells, start, end,
                       shift);
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

// ---------------------------------------------------------------------------
// particle_hash_kernel.cu:128-184 (applyReorder / removeReorder)
template <typename T, bool kScatter>
__global__ void __launch_bounds__(256) k_permute(const uint32_t* __restrict__ idx, const T* __restrict__ in,
                                                 T* __restrict__ out, uint32_t stride, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t j = idx[i];
    if (j >= n) return;
    for (uint32_t c = 0; c < stride; c++) {
        if (kScatter)
            out[j * stride + c] = in[(size_t)i * stride + c];
        else
            out[(size_t)i * stride + c] = in[j * stride + c];
    }
}

template <bool kScatter>
static pppm_status permute(const uint32_t* idx, const void* in, void* out, uint32_t stride, int dtype, uint32_t n,
                           hipStream_t st) {
    if (n == 0) return PPPM_OK;
    if (dtype != PPPM_F32 && dtype != PPPM_F64) return PPPM_ERR_INVALID;
    if (dtype == PPPM_F32 && stride == 4) {
        hipLaunchKernelGGL((k_permute<float4, kScatter>), dim3(div_up(n, 256)), dim3(256), 0, st, idx,
                           (const float4*)in, (float4*)out, 1u, n);
    } else if (dtype == PPPM_F32) {
        hipLaunchKernelGGL((k_permute<float, kScatter>), dim3(div_up(n, 256)), dim3(256), 0, st, idx,
                           (const float*)in, (float*)out, stride, n);
    } else {
        hipLaunchKernelGGL((k_permute<double, kScatter>), dim3(div_up(n, 256)), dim3(256), 0, st, idx,
                           (const double*)in, (double*)out, stride, n);
    }
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

pppm_status launch_gather(const uint32_t* idx, const void* in, void* out, uint32_t stride, int dtype, uint32_t n,
                          hipStream_t st) {
    return permute<false>(idx, in, out, stride, dtype, n, st);
}
pppm_status launch_scatter(const uint32_t* idx, const void* in, void* out, uint32_t stride, int dtype, uint32_t n,
                           hipStream_t st) {
    return permute<true>(idx, in, out, stride, dtype, n, st);
}

// ---------------------------------------------------------------------------
// Deposit cell of a shifted coordinate (GPU_kernel.cu:36-40, 256-259): the
// cell i with (double)q*invh - (i + 0.5) in [-1/2, 1/2).  Equal to floor()
// except that values in (-2^-54, 0) round into cell 0.
__device__ inline int deposit_index(float q, double invh) {
    double x = (double)q * invh;
    int i = d2i_sat(floor(x));
    if (i == -1 && (x - 0.5) >= -0.5) i = 0;
    return i;
}

// Sorted vortex arrays (particle_hash_kernel.cu:128-184 applyReorder), by
// scatter: vortex j (caller order, read contiguously) goes to its sorted slot
// inv[j] (a gather from the caller's arrays reads whole 128-byte lines per
// 16- and 4-byte element).  spos = origin-shifted float coordinates with the
// deposit cell in .w; somg = strengths.  A stray is a vortex whose deposit
// cell is not its hash cell; the hash cell is recomputed from the position
// (the cell its sort key holds).
template <typename W>
__global__ void __launch_bounds__(256) k_reorder_vortices_scatter(
    const uint32_t* __restrict__ inv, const float4* __restrict__ pos, const W* __restrict__ wx,
    const W* __restrict__ wy, const W* __restrict__ wz, uint32_t w_stride, uint32_t n, uint32_t g, float hf,
    float hox, float hoy, float hoz, double ox, double oy, double oz, double invh, float4* __restrict__ spos,
    float4* __restrict__ somg, uint32_t* __restrict__ strays) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    bool stray = false;
    if (j < n) {
        const uint32_t i = inv[j];
        const float4 p = pos[j];
        const float qx = shift_coord(p.x, ox), qy = shift_coord(p.y, oy), qz = shift_coord(p.z, oz);
        const int ix = deposit_index(qx, invh), iy = deposit_index(qy, invh), iz = deposit_index(qz, invh);
        uint32_t dkey = kEmptyCell;
        if (ix >= 0 && ix < (int)g && iy >= 0 && iy < (int)g && iz >= 0 && iz < (int)g)
            dkey = ((uint32_t)iz * g + (uint32_t)iy) * g + (uint32_t)ix;
        stray = (dkey != kEmptyCell) && (dkey != hash_cell_key(p.x, p.y, p.z, g, hf, hox, hoy, hoz));
        const size_t js = (size_t)j * w_stride;
        const float4 w = make_float4((float)wx[js], (float)wy[js], (float)wz[js], 0.0f);
        if (i < n) {
            spos[i] = make_float4(qx, qy, qz, __uint_as_float(dkey));
            somg[i] = w;
        }
    }
    uint64_t m = __ballot(stray);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(strays, (uint32_t)__popcll(m));
}

// inv[idx[i]] = i
__global__ void __launch_bounds__(256) k_invert_perm(const uint32_t* __restrict__ idx, uint32_t n,
                                                     uint32_t* __restrict__ inv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    if (j < n) inv[j] = i;
}

pppm_status launch_reorder_vortices_scatter(const uint32_t* idx, uint32_t* inv, const float4* pos, const void* wx,
                                            const void* wy, const void* wz, uint32_t w_stride, int w_dtype,
                                            uint32_t n, const PathGeom& pg, float4* spos, float4* somg,
                                            uint32_t* strays, hipStream_t st) {
    if (n == 0) return PPPM_OK;
    dim3 grid(div_up(n, 256)), blk(256);
    hipLaunchKernelGGL(k_invert_perm, grid, blk, 0, st, idx, n, inv);
    PPPM_LAUNCH_CHECK();
    const HashGeom& hg = pg.hash;
    if (w_dtype == PPPM_F64)
        hipLaunchKernelGGL(k_reorder_vortices_scatter<double>, grid, blk, 0, st, inv, pos, (const double*)wx,
                           (const double*)wy, (const double*)wz, w_stride, n, hg.g, hg.hf, hg.of[0], hg.of[1],
                           hg.of[2], pg.origin[0], pg.origin[1], pg.origin[2], pg.invh, spos, somg, strays);
    else if (w_dtype == PPPM_F32)
        hipLaunchKernelGGL(k_reorder_vortices_scatter<float>, grid, blk, 0, st, inv, pos, (const float*)wx,
                           (const float*)wy, (const float*)wz, w_stride, n, hg.g, hg.hf, hg.of[0], hg.of[1],
                           hg.of[2], pg.origin[0], pg.origin[1], pg.origin[2], pg.invh, spos, somg, strays);
    else
        return PPPM_ERR_INVALID;
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}


__global__ void __launch_bounds__(256) k_reorder_eval(const uint32_t* __restrict__ idx, const float4* __restrict__ pos,
                                                      const float4* __restrict__ other, uint32_t m, double ox,
                                                      double oy, double oz, float4* __restrict__ epos,
                                                      float4* __restrict__ eoth) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    size_t j = idx[i];
    if (j >= m) return;
    float4 p = pos[j];
    epos[i] = make_float4(shift_coord(p.x, ox), shift_coord(p.y, oy), shift_coord(p.z, oz), 0.0f);
    if (other) {
        float4 b = other[j];
        eoth[i] = make_float4(shift_coord(b.x, ox), shift_coord(b.y, oy), shift_coord(b.z, oz), 0.0f);
    }
}

pppm_status launch_reorder_eval(const uint32_t* idx, const float4* pos, const float4* other, uint32_t m,
                                const PathGeom& pg, float4* epos, float4* eoth, hipStream_t st) {
    if (m == 0) return PPPM_OK;
    hipLaunchKernelGGL(k_reorder_eval, dim3(div_up(m, 256)), dim3(256), 0, st, idx, pos, other, m, pg.origin[0],
                       pg.origin[1], pg.origin[2], epos, eoth);
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

}  // namespace pppm
