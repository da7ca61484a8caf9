// batch.hip - many independent systems stepped together (gfx950).
//
// The reference holds its parameters in a process-global __constant__ and
// reuses its launcher names across libraries, so one process runs one system
// (fluid_kernel.cu:17, fluidSystem.cu:17-19).  Here m instances of uneven size
// share one arena: instance k owns particles [part0, part1) and cell keys
// [cell0, cell0 + its cell count).  One step of all instances is
//
//   prologue   per-instance time state on the device (PeristalsisSystem.cpp:
//              330-343 for the moving walls), so the step needs no host input
//   hash       global key = cell0 + local cell hash; the boundary
//              configuration of Peristalsis walls is fused in front of it
//   sort       one stable radix sort of every instance's keys: instance k's
//              particles stay in [part0, part1), ordered as its own sort orders
//              them
//   reorder    one gather + cell ranges + compact offsets over all cells
//   density, force, integrate
//              persistent kernels over a table of (instance, first particle)
//              tiles; a workgroup runs one contiguous chunk of tiles and reads
//              its instance's parameter block from device memory
//
// Every per-particle body is the one the single-system kernels run
// (sph_device.hpp, integrate_device.hpp), with the instance's cells seen
// through cellFirst + cell0, so each instance's results equal a single-system
// run of the same state bit for bit.  A step is a fixed kernel sequence with
// no host decisions: the graph path captures several steps at once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "integrate_device.hpp"
#include "kernels.hpp"
#include "sph_device.hpp"
#include "step_state.hpp"

namespace cmag {

struct BatchInst {
    uint32_t part0, part1;  // particle range (unsorted and sorted order alike)
    uint32_t cell0, cells;  // first global cell key, number of cells
};

// the chunk [t0, t1) of tiles this workgroup runs: contiguous per workgroup,
// and contiguous per XCD through the XCD-aware renumbering
__device__ __forceinline__ void tile_chunk(uint32_t num_tiles, uint32_t& t0, uint32_t& t1) {
    const uint32_t per = (num_tiles + gridDim.x - 1u) / gridDim.x;
    t0 = min(num_tiles, xcd_block() * per);
    t1 = min(num_tiles, t0 + per);
}

// PeristalsisSystem::Update's state machine (PeristalsisSystem.cpp:330-343),
// one thread per instance; other systems only advance their clock
__global__ __launch_bounds__(64) void batch_prologue_kernel(const cmag_params* __restrict__ params,
                                                            StepState* __restrict__ state, uint32_t m) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    StepState s = state[k];
    advance_step_state(params[k], s);
    state[k] = s;
}

__global__ __launch_bounds__(256) void batch_hash_kernel(const cmag_params* __restrict__ params,
                                                         const BatchInst* __restrict__ inst,
                                                         const StepState* __restrict__ state,
                                                         const uint2* __restrict__ tiles, uint32_t num_tiles,
                                                         float4* __restrict__ pos, uint32_t* __restrict__ hash,
                                                         uint32_t* __restrict__ index) {
    uint32_t t0, t1;
    tile_chunk(num_tiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 tile = tiles[t];
        const BatchInst I = inst[tile.x];
        const uint32_t i = tile.y + threadIdx.x;
        if (i >= I.part1) continue;
        const cmag_params& P = params[tile.x];
        float4 p = pos[i];
        if (P.system == CMAG_SYS_PERISTALSIS && state[tile.x].configure_now && (p.w > 0.0f || p.w < 0.0f)) {
            // configureBoundaryD, PeristalsisKernel.cu:144-171
            p = make_float4(p.x, wall_y(P, p.x, p.w, state[tile.x].configure_height, 0.0f), p.z, p.w);
            pos[i] = p;
        }
        const Cell c = grid_pos(P, p.x, p.y, p.z);
        hash[i] = I.cell0 + grid_hash(P, c.x, c.y, c.z);
        index[i] = i;
    }
}

template <int DIM, int HASH>
__global__ __launch_bounds__(256) void batch_density_kernel(const cmag_params* __restrict__ params,
                                                            const BatchInst* __restrict__ inst,
                                                            const uint2* __restrict__ tiles, uint32_t num_tiles,
                                                            float4* __restrict__ measures,
                                                            const float4* __restrict__ sortedPos,
                                                            const uint32_t* __restrict__ cellFirst) {
    uint32_t t0, t1;
    tile_chunk(num_tiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 tile = tiles[t];
        const BatchInst I = inst[tile.x];
        const uint32_t i = tile.y + threadIdx.x;
        if (i >= I.part1) continue;
        const cmag_params& P = params[tile.x];
        const FirstCells cells{cellFirst + I.cell0};
        if (P.system == CMAG_SYS_DAMBREAK)
            density_body<CMAG_SYS_DAMBREAK, DIM, HASH>(P, i, measures, sortedPos, cells);
        else if (P.system == CMAG_SYS_POISEUILLE)
            density_body<CMAG_SYS_POISEUILLE, DIM, HASH>(P, i, measures, sortedPos, cells);
        else
            density_body<CMAG_SYS_PERISTALSIS, DIM, HASH>(P, i, measures, sortedPos, cells);
    }
}

template <int DIM, int HASH, int CAP>
__global__ __launch_bounds__(256) void batch_force_kernel(const cmag_params* __restrict__ params,
                                                          const BatchInst* __restrict__ inst,
                                                          const uint2* __restrict__ tiles, uint32_t num_tiles,
                                                          float4* __restrict__ acc, float4* __restrict__ presF,
                                                          const float4* __restrict__ measures,
                                                          const float4* __restrict__ sortedPos,
                                                          const float4* __restrict__ sortedVel,
                                                          const uint32_t* __restrict__ index,
                                                          const uint32_t* __restrict__ cellFirst, uint32_t n) {
    __shared__ uint32_t nbl[256 / 64][CAP][64];
    uint32_t(*mine)[64] = nbl[threadIdx.x >> 6];
    uint32_t t0, t1;
    tile_chunk(num_tiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 tile = tiles[t];
        const BatchInst I = inst[tile.x];
        const uint32_t i = tile.y + threadIdx.x;
        if (i >= I.part1) continue;
        const cmag_params& P = params[tile.x];
        const FirstCells cells{cellFirst + I.cell0};
        if (P.system == CMAG_SYS_DAMBREAK)
            force_body<CMAG_SYS_DAMBREAK, DIM, HASH, FirstCells, CAP>(P, i, acc, measures, sortedPos, sortedVel,
                                                                      index, cells, n, mine);
        else if (P.system == CMAG_SYS_POISEUILLE)
            force_body<CMAG_SYS_POISEUILLE, DIM, HASH, FirstCells, CAP>(P, i, acc, measures, sortedPos,
                                                                        sortedVel, index, cells, n, mine);
        else
            peristalsis_force_body<DIM, HASH, true, true, FirstCells, CAP>(P, i, acc, presF, measures, sortedPos,
                                                                           sortedVel, index, cells, n, mine);
    }
}

__global__ __launch_bounds__(256) void batch_integrate_kernel(const cmag_params* __restrict__ params,
                                                              const BatchInst* __restrict__ inst,
                                                              const StepState* __restrict__ state,
                                                              const uint2* __restrict__ tiles, uint32_t num_tiles,
                                                              float4* __restrict__ pos, float4* __restrict__ vel,
                                                              float4* __restrict__ velLF,
                                                              const float4* __restrict__ acc,
                                                              const float4* __restrict__ presF) {
    uint32_t t0, t1;
    tile_chunk(num_tiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 tile = tiles[t];
        const BatchInst I = inst[tile.x];
        const uint32_t i = tile.y + threadIdx.x;
        if (i >= I.part1) continue;
        const cmag_params& P = params[tile.x];
        const float t_w = state[tile.x].t_w;
        const int cfg = state[tile.x].is_config;
        if (P.system == CMAG_SYS_DAMBREAK)
            integrate_body<CMAG_SYS_DAMBREAK>(P, i, pos, vel, velLF, acc, presF, t_w, cfg);
        else if (P.system == CMAG_SYS_POISEUILLE)
            integrate_body<CMAG_SYS_POISEUILLE>(P, i, pos, vel, velLF, acc, presF, t_w, cfg);
        else
            integrate_body<CMAG_SYS_PERISTALSIS>(P, i, pos, vel, velLF, acc, presF, t_w, cfg);
    }
}

}  // namespace cmag

using namespace cmag;

struct cmag_batch {
    std::vector<cmag_params> P;
    std::vector<BatchInst> I;
    uint32_t m = 0, n = 0, totalCells = 0;
    int key_bits = 0, dim = 2, hash_order = CMAG_HASH_LINEAR;
    int num_cus = 256;
    cmag_batch_tuning tune{256u, (uint32_t)NB_CAP, 0u, 10u};
    hipStream_t stream = nullptr;

    cmag_params* dP = nullptr;
    BatchInst* dI = nullptr;
    StepState* dS = nullptr;
    uint2* dTiles = nullptr;
    uint32_t numTiles = 0, tilesFor = 0;  // tiles built for block size tilesFor

    float *pos = nullptr, *vel = nullptr, *velLF = nullptr, *acc = nullptr, *presF = nullptr;
    float *measures = nullptr, *sortedPos = nullptr, *sortedVel = nullptr;
    uint32_t *hash = nullptr, *index = nullptr, *hashAlt = nullptr, *indexAlt = nullptr;
    uint32_t *cellStart = nullptr, *cellEnd = nullptr, *cellFirst = nullptr, *scanWs = nullptr;
    void* sortWs = nullptr;
    size_t sortWsBytes = 0;

    hipGraphExec_t graph = nullptr;
    uint32_t graph_steps = 0;

    bool timing = false, pending = false;
    hipEvent_t ev[CMAG_NUM_STAGES + 1] = {};
    double stage_ms[CMAG_NUM_STAGES] = {};
    uint32_t stage_count[CMAG_NUM_STAGES] = {};

    ~cmag_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        void* bufs[] = {dP, dI, dS, dTiles, pos, vel, velLF, acc, presF, measures, sortedPos, sortedVel,
                        hash, index, hashAlt, indexAlt, cellStart, cellEnd, cellFirst, scanWs, sortWs};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int alloc() {
        int dev = 0;
        CMAG_HIP_TRY(hipGetDevice(&dev));
        CMAG_HIP_TRY(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
        CMAG_HIP_TRY(hipStreamCreate(&stream));
        const size_t f4 = sizeof(float) * 4 * (size_t)n;
        const size_t u = sizeof(uint32_t) * (size_t)n;
        float** fb[] = {&pos, &vel, &velLF, &acc, &presF, &measures, &sortedPos, &sortedVel};
        for (float** b : fb) {
            CMAG_HIP_TRY(hipMalloc((void**)b, f4));
            CMAG_HIP_TRY(hipMemsetAsync(*b, 0, f4, stream));
        }
        uint32_t** ub[] = {&hash, &index, &hashAlt, &indexAlt};
        for (uint32_t** b : ub) {
            CMAG_HIP_TRY(hipMalloc((void**)b, u));
            CMAG_HIP_TRY(hipMemsetAsync(*b, 0, u, stream));
        }
        CMAG_HIP_TRY(hipMalloc((void**)&cellStart, sizeof(uint32_t) * (size_t)totalCells));
        CMAG_HIP_TRY(hipMalloc((void**)&cellEnd, sizeof(uint32_t) * (size_t)totalCells));
        CMAG_HIP_TRY(hipMalloc((void**)&cellFirst, sizeof(uint32_t) * ((size_t)totalCells + 1)));
        CMAG_HIP_TRY(hipMalloc((void**)&scanWs, sizeof(uint32_t) * (scan_levels_words(totalCells) + 1)));
        sortWsBytes = radix_sort_workspace_bytes(n);
        CMAG_HIP_TRY(hipMalloc(&sortWs, sortWsBytes));
        CMAG_HIP_TRY(hipMalloc((void**)&dP, sizeof(cmag_params) * m));
        CMAG_HIP_TRY(hipMalloc((void**)&dI, sizeof(BatchInst) * m));
        CMAG_HIP_TRY(hipMalloc((void**)&dS, sizeof(StepState) * m));
        CMAG_HIP_TRY(hipMemcpyAsync(dP, P.data(), sizeof(cmag_params) * m, hipMemcpyHostToDevice, stream));
        CMAG_HIP_TRY(hipMemcpyAsync(dI, I.data(), sizeof(BatchInst) * m, hipMemcpyHostToDevice, stream));
        CMAG_TRY(reset_clocks());
        CMAG_TRY(build_tiles());
        CMAG_HIP_TRY(hipStreamSynchronize(stream));
        return CMAG_OK;
    }

    // elapsed time 0 and the Peristalsis configuration state of a fresh system
    // (PeristalsisSystem::Reset, PeristalsisSystem.cpp:213-237)
    int reset_clocks() {
        std::vector<StepState> S(m);
        for (uint32_t k = 0; k < m; ++k) S[k] = initial_step_state(P[k], 0.0f);
        CMAG_HIP_TRY(hipMemcpyAsync(dS, S.data(), sizeof(StepState) * m, hipMemcpyHostToDevice, stream));
        CMAG_HIP_TRY(hipStreamSynchronize(stream));
        return CMAG_OK;
    }

    int build_tiles() {
        const uint32_t B = tune.block;
        if (tilesFor == B && dTiles) return CMAG_OK;
        std::vector<uint2> T;
        for (uint32_t k = 0; k < m; ++k)
            for (uint32_t f = I[k].part0; f < I[k].part1; f += B) T.push_back(make_uint2(k, f));
        if (dTiles) CMAG_HIP_TRY(hipFree(dTiles));
        dTiles = nullptr;
        numTiles = (uint32_t)T.size();
        CMAG_HIP_TRY(hipMalloc((void**)&dTiles, sizeof(uint2) * (T.empty() ? 1 : T.size())));
        if (!T.empty())
            CMAG_HIP_TRY(hipMemcpy(dTiles, T.data(), sizeof(uint2) * T.size(), hipMemcpyHostToDevice));
        tilesFor = B;
        return CMAG_OK;
    }

    uint32_t grid_blocks() const {
        if (tune.blocks_per_cu == 0) return numTiles;
        return std::max(1u, std::min(numTiles, tune.blocks_per_cu * (uint32_t)num_cus));
    }

    int mark(int k) {
        if (timing) CMAG_HIP_TRY(hipEventRecord(ev[k], stream));
        return CMAG_OK;
    }

    int collect() {
        if (!timing || !pending) return CMAG_OK;
        CMAG_HIP_TRY(hipEventSynchronize(ev[CMAG_NUM_STAGES]));
        for (int k = 0; k < CMAG_NUM_STAGES; ++k) {
            float ms = 0.0f;
            CMAG_HIP_TRY(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            stage_ms[k] += ms;
            stage_count[k] += 1;
        }
        pending = false;
        return CMAG_OK;
    }

    template <int DIM, int HASH>
    int launch_force_cap(dim3 g, dim3 b) {
        auto f4 = [](float* q) { return reinterpret_cast<float4*>(q); };
        auto c4 = [](const float* q) { return reinterpret_cast<const float4*>(q); };
#define CMAG_BATCH_FORCE(CAP)                                                                              \
    hipLaunchKernelGGL((batch_force_kernel<DIM, HASH, CAP>), g, b, 0, stream, dP, dI, dTiles, numTiles,     \
                       f4(acc), f4(presF), c4(measures), c4(sortedPos), c4(sortedVel), index, cellFirst, n)
        if (tune.nb_cap == 8) CMAG_BATCH_FORCE(8);
        else if (tune.nb_cap == 32) CMAG_BATCH_FORCE(32);
        else CMAG_BATCH_FORCE(16);
#undef CMAG_BATCH_FORCE
        return check_launch("batch_force_kernel");
    }

    template <int DIM, int HASH>
    int launch_passes(dim3 g, dim3 b) {
        hipLaunchKernelGGL((batch_density_kernel<DIM, HASH>), g, b, 0, stream, dP, dI, dTiles, numTiles,
                           reinterpret_cast<float4*>(measures), reinterpret_cast<const float4*>(sortedPos),
                           cellFirst);
        CMAG_TRY(check_launch("batch_density_kernel"));
        CMAG_TRY(mark(CMAG_STAGE_FORCE));
        return launch_force_cap<DIM, HASH>(g, b);
    }

    // one step of every instance
    int enqueue_step() {
        if (n == 0) return CMAG_OK;
        const dim3 g(grid_blocks()), b(tune.block);
        const int passes = (key_bits + 7) / 8;
        const bool odd = passes % 2 == 1;
        uint32_t* hk = odd ? hashAlt : hash;
        uint32_t* hv = odd ? indexAlt : index;
        CMAG_TRY(mark(CMAG_STAGE_HASH));
        hipLaunchKernelGGL(batch_prologue_kernel, dim3(div_up(m, 64)), dim3(64), 0, stream, dP, dS, m);
        CMAG_TRY(check_launch("batch_prologue_kernel"));
        hipLaunchKernelGGL(batch_hash_kernel, g, b, 0, stream, dP, dI, dS, dTiles, numTiles,
                           reinterpret_cast<float4*>(pos), hk, hv);
        CMAG_TRY(check_launch("batch_hash_kernel"));
        CMAG_TRY(mark(CMAG_STAGE_SORT));
        bool in_alt = false;
        CMAG_TRY(radix_sort_pairs(hk, hv, odd ? hash : hashAlt, odd ? index : indexAlt, n, key_bits, sortWs,
                                  sortWsBytes, &in_alt, stream));
        const bool moved = n >= 2 && key_bits > 0;
        if (moved && in_alt != odd) return fail(CMAG_E_STATE, "radix sort ended in an unexpected buffer");
        if (!moved && odd) {
            CMAG_TRY(launch_copy_u32(hash, hk, n, stream));
            CMAG_TRY(launch_copy_u32(index, hv, n, stream));
        }
        CMAG_TRY(mark(CMAG_STAGE_REORDER));
        CMAG_TRY(launch_reorder(cellStart, cellEnd, sortedPos, sortedVel, hash, index, pos, velLF, n, totalCells,
                                stream));
        CMAG_TRY(launch_cell_offsets(cellStart, cellEnd, cellFirst, totalCells, n, scanWs, stream));
        CMAG_TRY(mark(CMAG_STAGE_DENSITY));
        if (dim == 3) {
            if (hash_order == CMAG_HASH_MORTON) CMAG_TRY((launch_passes<3, 1>(g, b)));
            else CMAG_TRY((launch_passes<3, 0>(g, b)));
        } else {
            if (hash_order == CMAG_HASH_MORTON) CMAG_TRY((launch_passes<2, 1>(g, b)));
            else CMAG_TRY((launch_passes<2, 0>(g, b)));
        }
        CMAG_TRY(mark(CMAG_STAGE_INTEGRATE));
        hipLaunchKernelGGL(batch_integrate_kernel, g, b, 0, stream, dP, dI, dS, dTiles, numTiles,
                           reinterpret_cast<float4*>(pos), reinterpret_cast<float4*>(vel),
                           reinterpret_cast<float4*>(velLF), reinterpret_cast<const float4*>(acc),
                           reinterpret_cast<const float4*>(presF));
        CMAG_TRY(check_launch("batch_integrate_kernel"));
        CMAG_TRY(mark(CMAG_NUM_STAGES));
        if (timing) pending = true;
        return CMAG_OK;
    }

    int update() {
        CMAG_TRY(collect());
        return enqueue_step();
    }

    void invalidate_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_steps = 0;
    }

    int build_graph(uint32_t steps) {
        invalidate_graph();
        hipGraph_t gr = nullptr;
        const bool was_timing = timing;
        timing = false;
        CMAG_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = CMAG_OK;
        for (uint32_t k = 0; k < steps && rc == CMAG_OK; ++k) rc = enqueue_step();
        timing = was_timing;
        hipError_t e = hipStreamEndCapture(stream, &gr);
        if (rc != CMAG_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        graph_steps = steps;
        return CMAG_OK;
    }

    int run(uint32_t steps, bool use_graph) {
        if (!use_graph) {
            for (uint32_t k = 0; k < steps; ++k) CMAG_TRY(update());
            return CMAG_OK;
        }
        CMAG_TRY(collect());
        const uint32_t S = std::max(1u, tune.steps_per_graph);
        if (steps >= S) {
            if (graph_steps != S) CMAG_TRY(build_graph(S));
            for (uint32_t k = 0; k + S <= steps; k += S) CMAG_HIP_TRY(hipGraphLaunch(graph, stream));
        }
        for (uint32_t k = 0; k < steps % S; ++k) CMAG_TRY(update());
        return CMAG_OK;
    }

    float* farray(int a) {
        switch (a) {
        case CMAG_ARRAY_POSITION: return pos;
        case CMAG_ARRAY_VELOCITY: return vel;
        case CMAG_ARRAY_MEASURES: return measures;
        case CMAG_ARRAY_ACCELERATION: return acc;
        case CMAG_ARRAY_VELOCITYLEAPFROG: return velLF;
        case CMAG_ARRAY_PRESSUREFORCE: return presF;
        case CMAG_ARRAY_SORTEDPOS: return sortedPos;
        case CMAG_ARRAY_SORTEDVEL: return sortedVel;
        default: return nullptr;
        }
    }

    int set_tuning(const cmag_batch_tuning& t) {
        if (t.block != 64 && t.block != 128 && t.block != 256)
            return fail(CMAG_E_INVALID, "batch tuning: block must be 64, 128 or 256");
        if (t.nb_cap != 8 && t.nb_cap != 16 && t.nb_cap != 32)
            return fail(CMAG_E_INVALID, "batch tuning: nb_cap must be 8, 16 or 32");
        if (t.blocks_per_cu > 64) return fail(CMAG_E_INVALID, "batch tuning: blocks_per_cu above 64");
        if (t.steps_per_graph == 0 || t.steps_per_graph > 1000)
            return fail(CMAG_E_INVALID, "batch tuning: steps_per_graph must be in 1..1000");
        CMAG_HIP_TRY(hipStreamSynchronize(stream));
        tune = t;
        invalidate_graph();
        return build_tiles();
    }

    // time `steps` eager steps on the GPU clock (events on the batch stream)
    int time_steps(uint32_t steps, float* ms) {
        hipEvent_t a = nullptr, b = nullptr;
        CMAG_HIP_TRY(hipEventCreate(&a));
        hipError_t e = hipEventCreate(&b);
        if (e != hipSuccess) {
            (void)hipEventDestroy(a);
            return hip_fail(e, "hipEventCreate");
        }
        int rc = CMAG_OK;
        if (hipEventRecord(a, stream) != hipSuccess) rc = fail(CMAG_E_HIP, "hipEventRecord");
        for (uint32_t k = 0; k < steps && rc == CMAG_OK; ++k) rc = enqueue_step();
        if (rc == CMAG_OK && hipEventRecord(b, stream) != hipSuccess) rc = fail(CMAG_E_HIP, "hipEventRecord");
        if (rc == CMAG_OK && hipEventSynchronize(b) != hipSuccess) rc = fail(CMAG_E_HIP, "hipEventSynchronize");
        if (rc == CMAG_OK && hipEventElapsedTime(ms, a, b) != hipSuccess) rc = fail(CMAG_E_HIP, "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return rc;
    }
};

namespace {

// the arrays a step reads before writing them, plus the clocks: what an
// autotuning trial must put back
struct Snapshot {
    std::vector<void*> bufs;
    ~Snapshot() {
        for (void* b : bufs)
            if (b) (void)hipFree(b);
    }
};

int snapshot(cmag_batch* b, Snapshot& s, bool restore) {
    float* arrays[] = {b->pos, b->vel, b->velLF};
    const size_t f4 = sizeof(float) * 4 * (size_t)b->n;
    if (!restore) {
        for (int k = 0; k < 4; ++k) {
            void* q = nullptr;
            CMAG_HIP_TRY(hipMalloc(&q, k < 3 ? f4 : sizeof(StepState) * b->m));
            s.bufs.push_back(q);
        }
    }
    for (int k = 0; k < 3; ++k) {
        void* dst = restore ? (void*)arrays[k] : s.bufs[k];
        const void* src = restore ? s.bufs[k] : (const void*)arrays[k];
        CMAG_HIP_TRY(hipMemcpyAsync(dst, src, f4, hipMemcpyDeviceToDevice, b->stream));
    }
    void* dst = restore ? (void*)b->dS : s.bufs[3];
    const void* src = restore ? s.bufs[3] : (const void*)b->dS;
    CMAG_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(StepState) * b->m, hipMemcpyDeviceToDevice, b->stream));
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    return CMAG_OK;
}

}  // namespace

extern "C" {

int cmag_batch_create(const cmag_params* params, const uint32_t* counts, uint32_t m, cmag_batch** out) {
    if (!out) return fail(CMAG_E_INVALID, "null output handle");
    *out = nullptr;
    if (m == 0 || !params || !counts) return fail(CMAG_E_INVALID, "a batch needs at least one instance");
    uint64_t n = 0, cells = 0;
    for (uint32_t k = 0; k < m; ++k) {
        CMAG_TRY(validate_params(&params[k]));
        if (params[k].dim != params[0].dim || params[k].hash_order != params[0].hash_order)
            return fail(CMAG_E_INVALID, "instances of a batch share the dimension and the hash order");
        n += counts[k];
        cells += cmag_params_num_cells(&params[k]);
    }
    if (n >= (1u << NB_SHIFT_BITS)) return fail(CMAG_E_INVALID, "a batch holds fewer than 2^27 particles");
    if (cells > 0xFFFFFFFEull) return fail(CMAG_E_INVALID, "a batch holds fewer than 2^32 - 1 cells");
    cmag_batch* b = new (std::nothrow) cmag_batch();
    if (!b) return fail(CMAG_E_INVALID, "out of host memory");
    b->m = m;
    b->n = (uint32_t)n;
    b->totalCells = (uint32_t)cells;
    b->dim = params[0].dim;
    b->hash_order = params[0].hash_order;
    b->P.assign(params, params + m);
    uint32_t p0 = 0, c0 = 0;
    for (uint32_t k = 0; k < m; ++k) {
        const uint32_t c = cmag_params_num_cells(&params[k]);
        b->I.push_back(BatchInst{p0, p0 + counts[k], c0, c});
        p0 += counts[k];
        c0 += c;
    }
    int bits = 0;
    while (bits < 32 && ((uint64_t)1 << bits) < cells) ++bits;
    b->key_bits = bits;
    int rc = b->alloc();
    if (rc != CMAG_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return CMAG_OK;
}

int cmag_batch_destroy(cmag_batch* b) {
    delete b;
    return CMAG_OK;
}

uint32_t cmag_batch_num_instances(const cmag_batch* b) { return b ? b->m : 0u; }
uint32_t cmag_batch_num_particles(const cmag_batch* b) { return b ? b->n : 0u; }
uint32_t cmag_batch_num_cells(const cmag_batch* b) { return b ? b->totalCells : 0u; }

int cmag_batch_instance_range(const cmag_batch* b, uint32_t inst, uint32_t* first, uint32_t* count) {
    if (!b || inst >= b->m) return fail(CMAG_E_INVALID, "bad batch or instance");
    if (first) *first = b->I[inst].part0;
    if (count) *count = b->I[inst].part1 - b->I[inst].part0;
    return CMAG_OK;
}

int cmag_batch_get_params(const cmag_batch* b, uint32_t inst, cmag_params* out) {
    if (!b || inst >= b->m || !out) return fail(CMAG_E_INVALID, "bad batch, instance or output");
    *out = b->P[inst];
    return CMAG_OK;
}

int cmag_batch_set_array(cmag_batch* b, uint32_t inst, int array, const float* data, uint32_t count) {
    if (!b || inst >= b->m) return fail(CMAG_E_INVALID, "bad batch or instance");
    float* d = b->farray(array);
    if (!d) return fail(CMAG_E_INVALID, "cmag_batch_set_array: unknown array");
    const BatchInst& I = b->I[inst];
    if (count != I.part1 - I.part0) return fail(CMAG_E_INVALID, "cmag_batch_set_array: count is not the instance size");
    if (count && !data) return fail(CMAG_E_INVALID, "cmag_batch_set_array: null data");
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    if (count)
        CMAG_HIP_TRY(hipMemcpy(d + 4 * (size_t)I.part0, data, sizeof(float) * 4 * (size_t)count,
                               hipMemcpyHostToDevice));
    return CMAG_OK;
}

int cmag_batch_get_array(cmag_batch* b, uint32_t inst, int array, float* data, uint32_t count) {
    if (!b || inst >= b->m) return fail(CMAG_E_INVALID, "bad batch or instance");
    float* d = b->farray(array);
    if (!d) return fail(CMAG_E_INVALID, "cmag_batch_get_array: unknown array");
    const BatchInst& I = b->I[inst];
    if (count != I.part1 - I.part0) return fail(CMAG_E_INVALID, "cmag_batch_get_array: count is not the instance size");
    if (count && !data) return fail(CMAG_E_INVALID, "cmag_batch_get_array: null data");
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    if (count)
        CMAG_HIP_TRY(hipMemcpy(data, d + 4 * (size_t)I.part0, sizeof(float) * 4 * (size_t)count,
                               hipMemcpyDeviceToHost));
    return CMAG_OK;
}

void* cmag_batch_device_ptr(cmag_batch* b, int array) {
    if (!b) return nullptr;
    switch (array) {
    case CMAG_ARRAY_HASH: return b->hash;
    case CMAG_ARRAY_INDEX: return b->index;
    case CMAG_ARRAY_CELLSTART: return b->cellStart;
    case CMAG_ARRAY_CELLEND: return b->cellEnd;
    default: return b->farray(array);
    }
}

int cmag_batch_elapsed_time(cmag_batch* b, uint32_t inst, float* out) {
    if (!b || inst >= b->m || !out) return fail(CMAG_E_INVALID, "bad batch, instance or output");
    StepState s;
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    CMAG_HIP_TRY(hipMemcpy(&s, b->dS + inst, sizeof(s), hipMemcpyDeviceToHost));
    *out = s.elapsed;
    return CMAG_OK;
}

int cmag_batch_reset_clocks(cmag_batch* b) {
    if (!b) return fail(CMAG_E_INVALID, "null batch");
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    return b->reset_clocks();
}

int cmag_batch_update(cmag_batch* b) {
    if (!b) return fail(CMAG_E_INVALID, "null batch");
    return b->update();
}

int cmag_batch_run(cmag_batch* b, uint32_t steps, int use_graph) {
    if (!b) return fail(CMAG_E_INVALID, "null batch");
    return b->run(steps, use_graph != 0);
}

int cmag_batch_synchronize(cmag_batch* b) {
    if (!b) return fail(CMAG_E_INVALID, "null batch");
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    return CMAG_OK;
}

int cmag_batch_get_tuning(const cmag_batch* b, cmag_batch_tuning* out) {
    if (!b || !out) return fail(CMAG_E_INVALID, "null batch or output");
    *out = b->tune;
    return CMAG_OK;
}

int cmag_batch_set_tuning(cmag_batch* b, const cmag_batch_tuning* t) {
    if (!b || !t) return fail(CMAG_E_INVALID, "null batch or tuning");
    return b->set_tuning(*t);
}

int cmag_batch_autotune(cmag_batch* b, uint32_t steps_per_trial, cmag_batch_tuning* best,
                        float* best_ms_per_step) {
    if (!b) return fail(CMAG_E_INVALID, "null batch");
    if (steps_per_trial == 0) return fail(CMAG_E_INVALID, "autotune needs at least one step per trial");
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    const bool was_timing = b->timing;
    b->timing = false;
    Snapshot snap;
    CMAG_TRY(snapshot(b, snap, false));
    const cmag_batch_tuning start = b->tune;
    cmag_batch_tuning win = start;
    float win_ms = INFINITY;
    const uint32_t blocks[] = {64u, 128u, 256u};
    const uint32_t caps[] = {8u, 16u, 32u};
    const uint32_t bpcs[] = {0u, 2u, 4u, 8u};
    int rc = CMAG_OK;
    for (uint32_t blk : blocks)
        for (uint32_t cap : caps)
            for (uint32_t bpc : bpcs) {
                if (rc != CMAG_OK) break;
                cmag_batch_tuning t{blk, cap, bpc, start.steps_per_graph};
                float ms = 0.0f;
                rc = b->set_tuning(t);
                if (rc == CMAG_OK) rc = b->time_steps(1, &ms);  // warm-up
                if (rc == CMAG_OK) rc = b->time_steps(steps_per_trial, &ms);
                if (rc == CMAG_OK) rc = snapshot(b, snap, true);
                if (rc == CMAG_OK && ms / steps_per_trial < win_ms) {
                    win_ms = ms / steps_per_trial;
                    win = t;
                }
            }
    b->timing = was_timing;
    if (rc != CMAG_OK) {
        (void)snapshot(b, snap, true);
        (void)b->set_tuning(start);
        return rc;
    }
    CMAG_TRY(b->set_tuning(win));
    if (best) *best = win;
    if (best_ms_per_step) *best_ms_per_step = win_ms;
    return CMAG_OK;
}

int cmag_batch_enable_timing(cmag_batch* b, int on) {
    if (!b) return fail(CMAG_E_INVALID, "null batch");
    CMAG_HIP_TRY(hipStreamSynchronize(b->stream));
    if (on && !b->ev[0])
        for (hipEvent_t& e : b->ev) CMAG_HIP_TRY(hipEventCreate(&e));
    b->timing = on != 0;
    b->pending = false;
    return CMAG_OK;
}

int cmag_batch_stage_times(cmag_batch* b, double* ms, uint32_t* counts, int max_stages) {
    if (!b) return fail(CMAG_E_INVALID, "null batch");
    CMAG_TRY(b->collect());
    for (int k = 0; k < CMAG_NUM_STAGES && k < max_stages; ++k) {
        if (ms) ms[k] = b->stage_ms[k];
        if (counts) counts[k] = b->stage_count[k];
        b->stage_ms[k] = 0.0;
        b->stage_count[k] = 0;
    }
    return CMAG_OK;
}

}  // extern "C"
