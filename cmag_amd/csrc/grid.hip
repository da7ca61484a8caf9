// grid.hip - cell hash and reorder/cellStart kernels (gfx950, wave64).
//
// calcHash            DamBreak.Core/fluid_kernel.cu:34-48
// reorder+cellStart   DamBreak.Core/fluid_kernel.cu:50-97 (+ the memset of
//                     its launcher, fluidSystem.cu:164)
// The reference staged the previous hash through shared memory; on CDNA4 the
// previous key is a plain coalesced load that hits the same cache line, and
// the gather of pos/velLeapFrog is a float4 (16 B) load per particle.
#include "common.hpp"
#include "kernels.hpp"

namespace cmag {

__global__ __launch_bounds__(256) void fill_u32x4_kernel(uint4* __restrict__ p, uint32_t v, size_t n4) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    uint4 v4 = make_uint4(v, v, v, v);
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n4; k += stride) p[k] = v4;
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) p[k] = v;
}

// a fill kernel instead of hipMemsetAsync, so the step stays graph-capturable
int launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    size_t head = 0;
    if (reinterpret_cast<uintptr_t>(p) % 16 == 0 && n >= 4) {
        size_t n4 = n / 4;
        uint32_t blocks = (uint32_t)std::min<size_t>((n4 + 255) / 256, 16384);
        hipLaunchKernelGGL(fill_u32x4_kernel, dim3(blocks), dim3(256), 0, s,
                           reinterpret_cast<uint4*>(p), v, n4);
        head = n4 * 4;
    }
    if (head < n) {
        size_t rest = n - head;
        uint32_t blocks = (uint32_t)std::min<size_t>((rest + 255) / 256, 16384);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p + head, v, rest);
    }
    return check_launch("fill_u32_kernel");
}

__global__ __launch_bounds__(256) void calc_hash_kernel(cmag_params P, const float4* __restrict__ pos,
                                                         uint32_t* __restrict__ hash,
                                                         uint32_t* __restrict__ index, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    Cell c = grid_pos(P, p.x, p.y, p.z);
    hash[i] = grid_hash(P, c.x, c.y, c.z);
    index[i] = i;
}

int launch_calc_hash(const cmag_params& P, const float* pos4, uint32_t* hash, uint32_t* index,
                     uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    hipLaunchKernelGGL(calc_hash_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<const float4*>(pos4), hash, index, n);
    return check_launch("calc_hash_kernel");
}

// Run heads of the sorted keys mark cellStart; the element before a head ends
// the previous cell; the last element ends its cell.  Gathers pos and the
// leapfrog velocity into sorted order.
// BUCKET: also the sub-cell bucket of every sorted row (DamBreak: fluid
// eighth 0..7 of the cell along x, walls 8), one byte per row
template <bool BUCKET>
__global__ __launch_bounds__(256) void reorder_kernel(uint32_t* __restrict__ cellStart,
                                                       uint32_t* __restrict__ cellEnd,
                                                       float4* __restrict__ sortedPos,
                                                       float4* __restrict__ sortedVel,
                                                       const uint32_t* __restrict__ hash,
                                                       const uint32_t* __restrict__ index,
                                                       const float4* __restrict__ oldPos,
                                                       const float4* __restrict__ oldVel,
                                                       uint32_t n, uint32_t numCells, const cmag_params P,
                                                       uint8_t* __restrict__ bucket) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = hash[i];
    // keys outside [0, numCells) cannot come from a matching hash; they are
    // skipped rather than written out of bounds
    if (i == 0) {
        if (h < numCells) cellStart[h] = 0;
    } else {
        uint32_t prev = hash[i - 1];
        if (h != prev) {
            if (h < numCells) cellStart[h] = i;
            if (prev < numCells) cellEnd[prev] = i;
        }
    }
    if (i == n - 1 && h < numCells) cellEnd[h] = n;
    const uint32_t k = index[i];
    if (k < n) {  // a valid permutation always passes; anything else is skipped
        const float4 p = oldPos[k];
        sortedPos[i] = p;
        sortedVel[i] = oldVel[k];
        if (BUCKET) bucket[i] = (uint8_t)(p.w != CMAG_DB_FLUID ? X_SUBS : x_sub(P, p.x));
    }
}

int launch_reorder(uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos, float* sortedVel,
                   const uint32_t* hash, const uint32_t* index, const float* oldPos,
                   const float* oldVel, uint32_t n, uint32_t numCells, hipStream_t s, const cmag_params* P,
                   uint8_t* bucket) {
    CMAG_TRY(launch_fill_u32(cellStart, CMAG_CELL_EMPTY, numCells, s));
    if (n == 0) return CMAG_OK;
    auto f4 = [](float* q) { return reinterpret_cast<float4*>(q); };
    auto c4 = [](const float* q) { return reinterpret_cast<const float4*>(q); };
    if (bucket) {
        if (!P) return fail(CMAG_E_INVALID, "reorder: the buckets need the parameter block");
        hipLaunchKernelGGL(reorder_kernel<true>, dim3(div_up(n, 256)), dim3(256), 0, s, cellStart, cellEnd,
                           f4(sortedPos), f4(sortedVel), hash, index, c4(oldPos), c4(oldVel), n, numCells, *P,
                           bucket);
    } else {
        hipLaunchKernelGGL(reorder_kernel<false>, dim3(div_up(n, 256)), dim3(256), 0, s, cellStart, cellEnd,
                           f4(sortedPos), f4(sortedVel), hash, index, c4(oldPos), c4(oldVel), n, numCells,
                           cmag_params{}, nullptr);
    }
    return check_launch("reorder_kernel");
}

__global__ void set_word_kernel(uint32_t* p, uint32_t v) { *p = v; }

// compact cell offsets for the neighbour walkers: a run of cells with
// consecutive keys is then one range [first[k0], first[k1 + 1]) without
// any empty-cell test
int launch_cell_offsets(const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t* first,
                        uint32_t numCells, uint32_t n, uint32_t* ws, hipStream_t s) {
    CMAG_TRY(exclusive_scan_cell_counts(cellStart, cellEnd, first, numCells, ws, s));
    hipLaunchKernelGGL(set_word_kernel, dim3(1), dim3(1), 0, s, first + numCells, n);
    return check_launch("set_word_kernel");
}

}  // namespace cmag

using namespace cmag;

extern "C" {

int cmag_calc_hash(const cmag_params* p, const float* pos4, uint32_t* hash, uint32_t* index,
                   uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!pos4 || !hash || !index)) return fail(CMAG_E_INVALID, "cmag_calc_hash: null pointer");
    return launch_calc_hash(*p, pos4, hash, index, n, as_stream(stream));
}

int cmag_reorder_find_cell_start(uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos,
                                 float* sortedVel, const uint32_t* hash, const uint32_t* index,
                                 const float* oldPos, const float* oldVel, uint32_t n,
                                 uint32_t numCells, cmag_stream_t stream) {
    if (!cellStart || !cellEnd || (n && (!sortedPos || !sortedVel || !hash || !index || !oldPos || !oldVel)))
        return fail(CMAG_E_INVALID, "cmag_reorder_find_cell_start: null pointer");
    return launch_reorder(cellStart, cellEnd, sortedPos, sortedVel, hash, index, oldPos, oldVel, n,
                          numCells, as_stream(stream));
}

}  // extern "C"
