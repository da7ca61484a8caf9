// sort.hip - stable LSD radix sort_by_key for uint32 keys/values (gfx950).
//
// Replaces thrust::sort_by_key of sortParticles (DamBreak.Core/fluidSystem.cu:94-99).
// Stability is what makes the reference's permutation reproducible: index
// starts as the identity every step, so the expected result is exactly the
// stable argsort of the hash.
//
// One pass per 8-bit digit, three launches per pass (reduce-then-scan):
//   rs_hist      per-tile digit histogram, written digit-major so that
//   scan         one exclusive scan gives every (digit, tile) its global base
//   rs_scatter   stable in-tile ranking with wave64 ballots (8 ballots give
//                the mask of lanes holding the same digit), a local sort of
//                the tile in LDS, then coalesced runs written per digit.
// No inter-workgroup waiting anywhere, so no launch order assumptions.
#include "common.hpp"
#include "kernels.hpp"

#include <mutex>
#include <string>

namespace cmag {

constexpr int RS_NT = 256;  // one thread per digit in the per-tile tables
static_assert(RS_NT == 256, "the histogram and scatter tables map thread t to digit t");
// keys per thread: 16 (4096-key tiles) for large sorts; 4 (1024-key tiles) up
// to RS_SMALL_N keys, where a pass is a handful of tiles and the in-tile
// ranking rounds, not bandwidth, set its duration
constexpr int RS_ITEMS = 16;
constexpr int RS_ITEMS_SMALL = 4;
constexpr uint32_t RS_SMALL_N = 1u << 18;
constexpr int SCAN_NT = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_NT * SCAN_ITEMS;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// exclusive prefix of v over the workgroup; lds needs NT/64 words
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    constexpr int NW = NT / 64;
    const int lane = __lane_id(), w = threadIdx.x / 64;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        uint32_t x = lds[k];
        pre += (k < w) ? x : 0u;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// lanes of this wave whose 8-bit digit equals mine (among valid lanes)
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint32_t bit = (d >> b) & 1u;
        uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// ---- exclusive scan of a uint32 array -------------------------------------
// The input is a plain array (ScanArray) or the cell counts end - start (0 for
// an empty cell, whose end is not read for its value) computed from the
// reorder's cell ranges as they are read (ScanCellCounts).
struct ScanArray {
    const uint32_t* in;
    __device__ __forceinline__ uint32_t at(size_t i) const { return in[i]; }
    __device__ __forceinline__ uint4 at4(size_t i) const { return *reinterpret_cast<const uint4*>(in + i); }
    bool aligned16() const { return (reinterpret_cast<uintptr_t>(in) & 15u) == 0; }
};

struct ScanCellCounts {
    const uint32_t* start;
    const uint32_t* end;
    __device__ static __forceinline__ uint32_t count(uint32_t st, uint32_t en) {
        return st == CMAG_CELL_EMPTY ? 0u : en - st;
    }
    __device__ __forceinline__ uint32_t at(size_t i) const { return count(start[i], end[i]); }
    __device__ __forceinline__ uint4 at4(size_t i) const {
        const uint4 a = *reinterpret_cast<const uint4*>(start + i), b = *reinterpret_cast<const uint4*>(end + i);
        return make_uint4(count(a.x, b.x), count(a.y, b.y), count(a.z, b.z), count(a.w, b.w));
    }
    bool aligned16() const {
        return ((reinterpret_cast<uintptr_t>(start) | reinterpret_cast<uintptr_t>(end)) & 15u) == 0;
    }
};

// the block total of one tile; each thread sums SCAN_ITEMS consecutive
// elements, as 16-byte words when vec and they lie inside len
template <typename Src>
__global__ __launch_bounds__(SCAN_NT) void scan_reduce_kernel(const Src src, size_t len,
                                                               uint32_t* __restrict__ partial, bool vec) {
    __shared__ uint32_t lds[SCAN_NT / 64];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t sum = 0;
    if (vec && base + SCAN_ITEMS <= len) {
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS / 4; ++q) {
            const uint4 x = src.at4(base + 4 * q);
            sum += (x.x + x.y) + (x.z + x.w);
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) sum += (base + k < len) ? src.at(base + k) : 0u;
    }
    uint32_t tot;
    block_exclusive_scan<SCAN_NT>(sum, lds, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// each thread owns SCAN_ITEMS consecutive elements, read and written as
// 16-byte words when vec (all pointers 16-byte aligned) and the thread's
// elements lie inside len; `carry` (may be null) holds the exclusive base of
// every block.  An in-place scan of a plain array is allowed: every thread
// reads its own elements before the block barrier and writes them after it.
template <typename Src>
__global__ __launch_bounds__(SCAN_NT) void scan_down_kernel(const Src src, uint32_t* out, size_t len,
                                                             const uint32_t* __restrict__ carry, bool vec) {
    static_assert(SCAN_ITEMS % 4 == 0, "16-byte words per thread");
    __shared__ uint32_t lds[SCAN_NT / 64];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    const bool whole = vec && base + SCAN_ITEMS <= len;
    uint32_t v[SCAN_ITEMS];
    if (whole) {
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS / 4; ++q) {
            const uint4 x = src.at4(base + 4 * q);
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) v[k] = (base + k < len) ? src.at(base + k) : 0u;
    }
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) sum += v[k];
    uint32_t tot;
    uint32_t run = block_exclusive_scan<SCAN_NT>(sum, lds, &tot);
    if (carry) run += carry[blockIdx.x];
    uint32_t o[SCAN_ITEMS];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        o[k] = run;
        run += v[k];
    }
    if (whole) {
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS / 4; ++q)
            *reinterpret_cast<uint4*>(out + base + 4 * q) = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
            if (base + k < len) out[base + k] = o[k];
    }
}

size_t scan_levels_words(size_t len) {
    size_t words = 0;
    while (len > (size_t)SCAN_TILE) {
        len = (len + SCAN_TILE - 1) / SCAN_TILE;
        words += len;
    }
    return words;
}

template <typename Src>
static int exclusive_scan(const Src& src, uint32_t* out, size_t len, uint32_t* ws, hipStream_t s) {
    if (len == 0) return CMAG_OK;
    const bool vec = src.aligned16() && (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
    if (len <= (size_t)SCAN_TILE) {
        hipLaunchKernelGGL(scan_down_kernel<Src>, dim3(1), dim3(SCAN_NT), 0, s, src, out, len,
                           (const uint32_t*)nullptr, vec);
        return check_launch("scan_down_kernel");
    }
    size_t blocks = (len + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t* partial = ws;
    hipLaunchKernelGGL(scan_reduce_kernel<Src>, dim3((uint32_t)blocks), dim3(SCAN_NT), 0, s, src, len, partial,
                       vec);
    CMAG_TRY(check_launch("scan_reduce_kernel"));
    // scan the partials in place (recursively for very long inputs)
    CMAG_TRY(exclusive_scan(ScanArray{partial}, partial, blocks, ws + blocks, s));
    hipLaunchKernelGGL(scan_down_kernel<Src>, dim3((uint32_t)blocks), dim3(SCAN_NT), 0, s, src, out, len,
                       (const uint32_t*)partial, vec);
    return check_launch("scan_down_kernel");
}

// out = exclusive_scan(in); workspace holds scan_levels_words(len) words
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t len, uint32_t* ws, hipStream_t s) {
    return exclusive_scan(ScanArray{in}, out, len, ws, s);
}

// first[c] = exclusive scan of the cell counts; workspace as above
int exclusive_scan_cell_counts(const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t* first,
                               size_t numCells, uint32_t* ws, hipStream_t s) {
    return exclusive_scan(ScanCellCounts{cellStart, cellEnd}, first, numCells, ws, s);
}

// ---- radix passes ----------------------------------------------------------
// per-tile digit counts: each wave counts into its own LDS histogram, and the
// block adds the wave histograms.  Keys in particle order are spatially
// coherent, so neighbouring lanes often share a digit: only the first lane of
// each run of equal digits adds, the run's length (lanes of one bin in
// different runs are serialised by the LDS atomics)
template <int ITEMS>
__global__ __launch_bounds__(RS_NT) void rs_hist_kernel(const uint32_t* __restrict__ keys, uint32_t n,
                                                         int shift, uint32_t* __restrict__ counts,
                                                         uint32_t numTiles) {
    constexpr int NW = RS_NT / 64;
    __shared__ uint32_t hist[NW][256];
    const int tid = threadIdx.x, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < NW; ++k) hist[k][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t)(RS_NT * ITEMS);
    uint32_t d[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t idx = base + k * RS_NT + tid;
        d[k] = idx < n ? (keys[idx] >> shift) & 255u : 256u;
    }
    const int lane = __lane_id();
    const uint64_t above = lane == 63 ? 0ull : (~0ull << (lane + 1));
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t prev = __shfl_up(d[k], 1);
        const bool head = lane == 0 || prev != d[k];
        const uint64_t heads = __ballot(head);
        const uint64_t later = heads & above;
        const uint32_t next = later ? (uint32_t)(__ffsll((unsigned long long)later) - 1) : 64u;
        if (head && d[k] < 256u) atomicAdd(&hist[w][d[k]], next - (uint32_t)lane);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) c += hist[k][tid];
    counts[(size_t)tid * numTiles + blockIdx.x] = c;
}

// Stable scatter of one tile: wave w ranks its own contiguous chunk of 64 *
// ITEMS keys in rounds of 64 (a key's rank is the number of keys of its digit
// before it in the chunk: earlier rounds, kept in the wave's LDS digit counts,
// plus lower lanes of its round), the waves' counts are turned into per-wave
// digit offsets, and the tile is reordered by digit in LDS and written out
// run by run.  Wave chunks, rounds and lanes follow the input order, so equal
// digits keep their order.
template <int ITEMS>
__global__ __launch_bounds__(RS_NT) void rs_scatter_kernel(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, int shift,
    const uint32_t* __restrict__ counts, const uint32_t* __restrict__ offsets, uint32_t numTiles) {
    constexpr int NW = RS_NT / 64;
    constexpr uint32_t TILE = RS_NT * ITEMS;
    constexpr uint32_t CHUNK = 64 * ITEMS;
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_wcnt[NW][256];
    __shared__ uint32_t s_scan[NW];

    const int tid = threadIdx.x, lane = __lane_id(), w = tid / 64;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * TILE;
    const uint32_t cnt = min(TILE, n - base);
    const uint32_t wbase = base + (uint32_t)w * CHUNK;

    uint32_t rk[ITEMS], rv[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t idx = wbase + k * 64 + lane;
        rk[k] = idx < n ? keys_in[idx] : 0u;
        rv[k] = idx < n ? vals_in[idx] : 0u;
    }
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) s_wcnt[ww][tid] = 0;
    const uint32_t c = counts[(size_t)tid * numTiles + tile];
    s_base[tid] = offsets[(size_t)tid * numTiles + tile];
    uint32_t tot;
    s_start[tid] = block_exclusive_scan<RS_NT>(c, s_scan, &tot);
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint32_t rank[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = wbase + k * 64 + lane < n;
        const uint32_t dg = (rk[k] >> shift) & 255u;
        const uint64_t m = match_digit(dg, valid);
        // LDS operations of one wave take effect in program order: every lane
        // reads its digit's count before the round's leaders advance it
        const uint32_t before = s_wcnt[w][dg];
        rank[k] = before + (uint32_t)__popcll(m & lt_mask);
        if (valid && (__ffsll((unsigned long long)m) - 1) == lane) s_wcnt[w][dg] = before + (uint32_t)__popcll(m);
    }
    __syncthreads();
    {
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) {
            const uint32_t x = s_wcnt[ww][tid];
            s_wcnt[ww][tid] = run;
            run += x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (wbase + k * 64 + lane < n) {
            const uint32_t dg = (rk[k] >> shift) & 255u;
            const uint32_t p = s_start[dg] + s_wcnt[w][dg] + rank[k];
            s_keys[p] = rk[k];
            s_vals[p] = rv[k];
        }
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t p = k * RS_NT + tid;
        if (p < cnt) {
            const uint32_t key = s_keys[p];
            const uint32_t dg = (key >> shift) & 255u;
            const uint32_t dst = s_base[dg] + (p - s_start[dg]);
            keys_out[dst] = key;
            vals_out[dst] = s_vals[p];
        }
    }
}

__global__ __launch_bounds__(256) void copy_u32_kernel(uint32_t* __restrict__ dst,
                                                        const uint32_t* __restrict__ src, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(copy_u32_kernel, dim3(blocks), dim3(256), 0, s, dst, src, n);
    return check_launch("copy_u32_kernel");
}

static uint32_t tile_keys(uint32_t n) { return RS_NT * (n <= RS_SMALL_N ? RS_ITEMS_SMALL : RS_ITEMS); }
static uint32_t num_tiles(uint32_t n) { return (uint32_t)(((uint64_t)n + tile_keys(n) - 1) / tile_keys(n)); }

size_t radix_sort_workspace_bytes(uint32_t n) {
    size_t len = (size_t)256 * num_tiles(n);
    size_t words = 2 * len + scan_levels_words(len);
    return (words * sizeof(uint32_t) + 255) & ~(size_t)255;
}

int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals,
                     uint32_t n, int key_bits, void* workspace, size_t workspace_bytes,
                     bool* result_in_alt, hipStream_t s) {
    *result_in_alt = false;
    if (key_bits < 0 || key_bits > 32) return fail(CMAG_E_INVALID, "key_bits outside [0,32]");
    if (n < 2 || key_bits == 0) return CMAG_OK;
    if (workspace_bytes < radix_sort_workspace_bytes(n))
        return fail(CMAG_E_SIZE, "radix sort workspace too small");
    const uint32_t tiles = num_tiles(n);
    const size_t len = (size_t)256 * tiles;
    uint32_t* counts = static_cast<uint32_t*>(workspace);
    uint32_t* offsets = counts + len;
    uint32_t* scan_ws = offsets + len;
    uint32_t *kin = keys, *vin = vals, *kout = alt_keys, *vout = alt_vals;
    const bool small = n <= RS_SMALL_N;
    for (int shift = 0; shift < key_bits; shift += 8) {
        if (small)
            hipLaunchKernelGGL(rs_hist_kernel<RS_ITEMS_SMALL>, dim3(tiles), dim3(RS_NT), 0, s, kin, n, shift,
                               counts, tiles);
        else
            hipLaunchKernelGGL(rs_hist_kernel<RS_ITEMS>, dim3(tiles), dim3(RS_NT), 0, s, kin, n, shift, counts,
                               tiles);
        CMAG_TRY(check_launch("rs_hist_kernel"));
        CMAG_TRY(exclusive_scan_u32(counts, offsets, len, scan_ws, s));
        if (small)
            hipLaunchKernelGGL(rs_scatter_kernel<RS_ITEMS_SMALL>, dim3(tiles), dim3(RS_NT), 0, s, kin, vin, kout,
                               vout, n, shift, counts, offsets, tiles);
        else
            hipLaunchKernelGGL(rs_scatter_kernel<RS_ITEMS>, dim3(tiles), dim3(RS_NT), 0, s, kin, vin, kout, vout,
                               n, shift, counts, offsets, tiles);
        CMAG_TRY(check_launch("rs_scatter_kernel"));
        std::swap(kin, kout);
        std::swap(vin, vout);
        *result_in_alt = !*result_in_alt;
    }
    return CMAG_OK;
}

// ---- scratch for callers that cannot allocate ------------------------------
// cmag_sort_by_key_noalloc keeps cmag_sort_by_key's scratch (alternate keys and
// values, then the per-tile digit tables) in this file-scope array, sized for
// CMAG_SORT_NOALLOC_MAX_KEYS keys.  Every GPU the library runs on holds its own
// copy until the process ends.  On the host the array's name is a stand-in, not
// a device address: host code only ever uses the address hipGetSymbolAddress
// returns for the current device, fetched once per device.
constexpr uint32_t NOALLOC_MAX_KEYS = CMAG_SORT_NOALLOC_MAX_KEYS;
static_assert(NOALLOC_MAX_KEYS > RS_SMALL_N && NOALLOC_MAX_KEYS % (RS_NT * RS_ITEMS) == 0,
              "the no-allocation limit sits in the large-tile regime, on a whole tile");

constexpr size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr size_t scan_levels_words_c(size_t len) {
    size_t words = 0;
    while (len > (size_t)SCAN_TILE) {
        len = (len + SCAN_TILE - 1) / SCAN_TILE;
        words += len;
    }
    return words;
}

// cmag_sort_scratch_bytes at a key count in the large-tile regime; the
// small-tile regime (n <= RS_SMALL_N) needs less than it does at the limit
constexpr size_t sort_scratch_bytes_large(size_t n) {
    return round256(n * 2 * sizeof(uint32_t)) +
           round256((2 * 256 * (n / (RS_NT * RS_ITEMS)) + scan_levels_words_c(256 * (n / (RS_NT * RS_ITEMS)))) *
                    sizeof(uint32_t));
}
constexpr size_t NOALLOC_BYTES = sort_scratch_bytes_large(NOALLOC_MAX_KEYS);
static_assert(sort_scratch_bytes_large(NOALLOC_MAX_KEYS) >=
                  round256((size_t)RS_SMALL_N * 8) +
                      round256((2 * 256 * (RS_SMALL_N / (RS_NT * RS_ITEMS_SMALL)) +
                                scan_levels_words_c(256 * (RS_SMALL_N / (RS_NT * RS_ITEMS_SMALL)))) *
                               sizeof(uint32_t)),
              "the array covers the small-tile regime too");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static void* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on the current device
static int noalloc_scratch(void** out) {
    int dev = -1;
    CMAG_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(CMAG_E_INVALID, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        CMAG_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CMAG_E_HIP, "hipGetSymbolAddress returned a null address for the sort scratch");
        noalloc_ptr[dev] = p;
    }
    *out = noalloc_ptr[dev];
    return CMAG_OK;
}

}  // namespace cmag

using namespace cmag;

extern "C" {

size_t cmag_sort_scratch_bytes(uint32_t n, int key_bits) {
    (void)key_bits;
    size_t pairs = ((size_t)n * 2 * sizeof(uint32_t) + 255) & ~(size_t)255;
    return pairs + radix_sort_workspace_bytes(n);
}

int cmag_sort_by_key(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits, void* scratch,
                     size_t scratch_bytes, cmag_stream_t stream) {
    if (n && (!keys || !vals)) return fail(CMAG_E_INVALID, "cmag_sort_by_key: null pointer");
    if (n < 2 || key_bits == 0) return CMAG_OK;
    if (!scratch || scratch_bytes < cmag_sort_scratch_bytes(n, key_bits))
        return fail(CMAG_E_SIZE, "cmag_sort_by_key: scratch smaller than cmag_sort_scratch_bytes");
    hipStream_t s = as_stream(stream);
    uint32_t* alt_keys = static_cast<uint32_t*>(scratch);
    uint32_t* alt_vals = alt_keys + n;
    size_t pairs = ((size_t)n * 2 * sizeof(uint32_t) + 255) & ~(size_t)255;
    void* ws = static_cast<char*>(scratch) + pairs;
    bool in_alt = false;
    CMAG_TRY(radix_sort_pairs(keys, vals, alt_keys, alt_vals, n, key_bits, ws,
                              scratch_bytes - pairs, &in_alt, s));
    if (in_alt) {
        CMAG_TRY(launch_copy_u32(keys, alt_keys, n, s));
        CMAG_TRY(launch_copy_u32(vals, alt_vals, n, s));
    }
    return CMAG_OK;
}

size_t cmag_sort_noalloc_scratch_bytes(void) { return NOALLOC_BYTES; }

int cmag_sort_by_key_noalloc(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits, cmag_stream_t stream) {
    if (n && (!keys || !vals)) return fail(CMAG_E_INVALID, "cmag_sort_by_key_noalloc: null pointer");
    if (n > NOALLOC_MAX_KEYS)
        return fail(CMAG_E_SIZE, "cmag_sort_by_key_noalloc: n = " + std::to_string(n) +
                                     " keys is above the limit of " + std::to_string(NOALLOC_MAX_KEYS) +
                                     " (CMAG_SORT_NOALLOC_MAX_KEYS); use cmag_sort_by_key with a scratch buffer");
    if (n < 2 || key_bits == 0) return CMAG_OK;
    if (cmag_sort_scratch_bytes(n, key_bits) > NOALLOC_BYTES)
        return fail(CMAG_E_SIZE, "cmag_sort_by_key_noalloc: scratch array smaller than the sort needs");
    void* ws = nullptr;
    CMAG_TRY(noalloc_scratch(&ws));
    return cmag_sort_by_key(keys, vals, n, key_bits, ws, NOALLOC_BYTES, stream);
}

}  // extern "C"
