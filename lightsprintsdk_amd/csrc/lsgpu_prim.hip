// lsgpu_prim.hip -- device-wide primitives the build steps share (gfx950): an exclusive
// scan of 32-bit counts and a stable LSD radix sort of (32-bit key, 32-bit value) pairs.
//
// Scan: blocks of 2048 values (256 lanes x 8) scan in LDS with wave shuffles; the block
// totals are scanned the same way (recursively) and added back.  Sort: 8-bit digits per
// pass; each pass histograms the digits per 2048-key tile, scans the (digit, tile) counts
// and scatters every tile's keys in order, so equal keys keep their input order.
#include "lsgpu_internal.h"
#include <utility>

namespace lsg {

namespace {

constexpr uint32_t kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, o);
        if (lane >= o) v += u;
    }
    return v;
}

// exclusive scan of one tile per block; block totals into sums (when given)
__global__ __launch_bounds__(256) void k_scan_tile(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                                                   uint32_t* __restrict__ sums)
{
    __shared__ uint32_t wsum[4];
    const uint32_t tid = threadIdx.x, base = blockIdx.x * kScanTile + tid * kScanItems;
    uint32_t v[kScanItems], s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    const uint32_t incl = wave_incl_scan(s);
    if ((tid & 63u) == 63u) wsum[tid >> 6] = incl;
    __syncthreads();
    uint32_t before = incl - s;
    for (uint32_t w = 0; w < (tid >> 6); w++) before += wsum[w];
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = before;
        before += v[k];
    }
    if (sums && tid == 255) sums[blockIdx.x] = before;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ offs)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] += offs[i / kScanTile];
}

__global__ void k_scan_total(const uint32_t* __restrict__ in, const uint32_t* __restrict__ out, uint32_t n, uint32_t* __restrict__ total)
{
    if (threadIdx.x == 0) *total = n ? out[n - 1] + in[n - 1] : 0u;
}

inline uint32_t tiles(uint32_t n) { return (n + kScanTile - 1) / kScanTile; }

// ---- radix sort ------------------------------------------------------------------------
constexpr uint32_t kSortThreads = 256, kSortItems = 8, kSortTile = kSortThreads * kSortItems;

__global__ __launch_bounds__(256) void k_sort_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (uint32_t k = 0; k < kSortItems; k++) {
        const uint32_t i = base + k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];   // digit-major: tiles in order per digit
}

__global__ __launch_bounds__(256) void k_sort_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, int shift,
                                                      const uint32_t* __restrict__ offs)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    run[tid] = offs[tid * gridDim.x + blockIdx.x];
    for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (uint32_t k = 0; k < kSortItems; k++) {
        const uint32_t i = base + k * kSortThreads + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t dig = (key >> shift) & 255u;
        // lanes of this wave with the same digit, and this lane's rank among them
        uint64_t same = __ballot(valid);
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t b = __ballot((dig >> bit) & 1u);
            same &= ((dig >> bit) & 1u) ? b : ~b;
        }
        const uint64_t lower = same & ((1ull << lane) - 1ull);
        if (valid && lower == 0) wcnt[wave][dig] = (uint32_t)__popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[dig] + (uint32_t)__popcll(lower);
            for (uint32_t w = 0; w < wave; w++) pos += wcnt[w][dig];
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
        run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
        __syncthreads();
    }
}

} // namespace

size_t scan_tmp_words(uint32_t n)
{
    size_t w = 1;                       // the total
    for (uint32_t m = tiles(n); m > 1; m = tiles(m)) w += 2 * (size_t)m;
    return w + 2;
}

int scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, uint32_t* tmp, hipStream_t st)
{
    if (!n) {
        if (total) LSG_HIP(hipMemsetAsync(total, 0, 4, st));
        return LSG_OK;
    }
    const uint32_t nb = tiles(n);
    uint32_t* sums = nb > 1 ? tmp : nullptr;
    k_scan_tile<<<nb, 256, 0, st>>>(in, out, n, sums);
    if (nb > 1) {
        uint32_t* scanned = tmp + nb;
        int e = scan_exclusive_u32(sums, scanned, nb, nullptr, tmp + 2 * (size_t)nb, st);
        if (e) return e;
        k_scan_add<<<(n + 255) / 256, 256, 0, st>>>(out, n, scanned);
    }
    if (total) k_scan_total<<<1, 64, 0, st>>>(in, out, n, total);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

size_t sort_tmp_words(uint32_t n)
{
    const uint32_t nb = (n + kSortTile - 1) / kSortTile;
    const uint32_t h = 256 * (nb ? nb : 1);
    return 2 * (size_t)h + scan_tmp_words(h);
}

int sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int bits,
                   uint32_t* tmp, hipStream_t st)
{
    if (n < 2 || bits <= 0) return LSG_OK;
    const uint32_t nb = (n + kSortTile - 1) / kSortTile;
    const uint32_t h = 256 * nb;
    uint32_t* hist = tmp;
    uint32_t* offs = tmp + h;
    uint32_t* stmp = tmp + 2 * (size_t)h;
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    for (int shift = 0; shift < bits; shift += 8) {
        k_sort_hist<<<nb, 256, 0, st>>>(ki, n, shift, hist);
        int e = scan_exclusive_u32(hist, offs, h, nullptr, stmp, st);
        if (e) return e;
        k_sort_scatter<<<nb, 256, 0, st>>>(ki, vi, ko, vo, n, shift, offs);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (ki != keys) {
        LSG_HIP(hipMemcpyAsync(keys, ki, 4 * (size_t)n, hipMemcpyDeviceToDevice, st));
        LSG_HIP(hipMemcpyAsync(vals, vi, 4 * (size_t)n, hipMemcpyDeviceToDevice, st));
    }
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

} // namespace lsg
