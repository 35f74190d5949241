// Device-wide exclusive prefix sum over uint32 (reduce -> top-level scan ->
// apply), used by the radix sort's digit histograms and by the cell-offset
// table.  Three launches, no inter-workgroup waiting.
#include "pppm_internal.hpp"

namespace pppm {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_wave /*[4]*/, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) lds_wave[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (int w = 0; w < kScanThreads / 64; w++) {
        uint32_t s = lds_wave[w];
        if (w < wave) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + incl - v;
}

template <typename In>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(In in, uint32_t n, uint32_t* block_sums) {
    __shared__ uint32_t lw[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; j++) {
        uint64_t i = base + (uint64_t)j * kScanThreads + threadIdx.x;
        if (i < n) s += in(i);
    }
    uint32_t tot;
    block_exclusive_scan(s, lw, tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of up to kScanTile block sums, in place.
__global__ void __launch_bounds__(kScanThreads) k_scan_top(uint32_t* sums, uint32_t nb) {
    __shared__ uint32_t lw[4];
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; j++) {
        uint32_t i = threadIdx.x * kScanItems + j;
        v[j] = i < nb ? sums[i] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, lw, tot);
    for (int j = 0; j < kScanItems; j++) {
        uint32_t i = threadIdx.x * kScanItems + j;
        if (i < nb) sums[i] = run;
        run += v[j];
    }
}

template <typename In>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(In in, uint32_t n, const uint32_t* block_sums,
                                                             uint32_t* out, int write_total) {
    __shared__ uint32_t tile[kScanTile];
    __shared__ uint32_t lw[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    for (int j = 0; j < kScanItems; j++) {
        uint64_t i = base + (uint64_t)j * kScanThreads + threadIdx.x;
        tile[j * kScanThreads + threadIdx.x] = i < n ? in(i) : 0u;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; j++) {
        v[j] = tile[threadIdx.x * kScanItems + j];
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, lw, tot) + block_sums[blockIdx.x];
    for (int j = 0; j < kScanItems; j++) {
        tile[threadIdx.x * kScanItems + j] = run;
        run += v[j];
    }
    __syncthreads();
    for (int j = 0; j < kScanItems; j++) {
        uint64_t i = base + (uint64_t)j * kScanThreads + threadIdx.x;
        if (i < n) out[i] = tile[j * kScanThreads + threadIdx.x];
    }
    // out[n] = grand total, written by the workgroup owning element n-1
    if (write_total && threadIdx.x == kScanThreads - 1 && base < n && base + kScanTile >= n) {
        uint32_t last = (uint32_t)(n - 1 - base);
        out[n] = tile[last] + in(n - 1);
    }
}

struct InArray {
    const uint32_t* p;
    __device__ uint32_t operator()(uint64_t i) const { return p[i]; }
};

// Cell population from the reference start/end tables (empty cells: 0).
struct InCellCount {
    const uint32_t* start;
    const uint32_t* end;
    __device__ uint32_t operator()(uint64_t i) const {
        uint32_t s = start[i];
        return s == kEmptyCell ? 0u : end[i] - s;
    }
};

size_t scan_scratch_words(uint64_t n) {
    uint64_t nb = div_up64(n, kScanTile);
    return nb <= (uint64_t)kScanTile ? nb + 1 : nb + scan_scratch_words(nb);
}

template <typename In>
static pppm_status scan_impl(In in, uint32_t n, uint32_t* out, int write_total, uint32_t* scratch,
                             hipStream_t st);

template <typename In>
static pppm_status scan_impl(In in, uint32_t n, uint32_t* out, int write_total, uint32_t* scratch,
                             hipStream_t st) {
    if (n == 0) {
        if (write_total) PPPM_HIP_TRY(hipMemsetAsync(out, 0, sizeof(uint32_t), st));
        return PPPM_OK;
    }
    uint32_t nb = div_up(n, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce<In>, dim3(nb), dim3(kScanThreads), 0, st, in, n, scratch);
    PPPM_LAUNCH_CHECK();
    if (nb <= (uint32_t)kScanTile) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, st, scratch, nb);
        PPPM_LAUNCH_CHECK();
    } else {
        // block sums scanned in place by the same three-phase scheme
        pppm_status s2 = scan_impl(InArray{scratch}, nb, scratch, 0, scratch + nb, st);
        if (s2 != PPPM_OK) return s2;
    }
    hipLaunchKernelGGL(k_scan_apply<In>, dim3(nb), dim3(kScanThreads), 0, st, in, n, scratch, out, write_total);
    PPPM_LAUNCH_CHECK();
    return PPPM_OK;
}

pppm_status exclusive_scan_u32(const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* scratch,
                               hipStream_t st) {
    return scan_impl(InArray{in}, n, out, 0, scratch, st);
}

pppm_status cell_offsets_from_tables(const uint32_t* start, const uint32_t* end, uint32_t ncells,
                                     uint32_t* off, uint32_t* scratch, hipStream_t st) {
    return scan_impl(InCellCount{start, end}, ncells, off, 1, scratch, st);
}

}  // namespace pppm
