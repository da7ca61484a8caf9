// system.cpp - System objects behind the C ABI: device buffers, reset(),
// update() in the reference stage order, setArray/getArray, HIP-graph replay.
//
// DamBreakSystem       DamBreak.Core/fluidSystem.cpp
// PoiseuilleFlowSystem Poiseuille.Core/poiseuilleFlowSystem.cpp
// PeristalsisSystem    Peristalsis.Core/PeristalsisSystem.cpp
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "step_state.hpp"

using namespace cmag;

enum RefKind { REF_NONE = -1 };

struct cmag_system {
    cmag_params P{};
    int ref_kind = REF_NONE;  // system kind when built by a reference constructor
    uint32_t n = 0, numCells = 0;
    int key_bits = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;

    float *pos = nullptr, *vel = nullptr, *velLF = nullptr, *acc = nullptr, *presF = nullptr;
    float *measures = nullptr, *sortedPos = nullptr, *sortedVel = nullptr, *predicted = nullptr;
    uint32_t *hash = nullptr, *index = nullptr, *hashAlt = nullptr, *indexAlt = nullptr;
    uint32_t *cellStart = nullptr, *cellEnd = nullptr;
    uint32_t *cellFirst = nullptr, *scanWs = nullptr;  // compact cell offsets for the walkers
    void* sortWs = nullptr;
    size_t sortWsBytes = 0;
    // adjoint scratch (sorted order), allocated by the first vjp_step
    float *gaccU = nullptr, *gaccS = nullptr, *gposS = nullptr, *gvelS = nullptr, *grho = nullptr;
    // sub-cell walk order (DamBreak): the neighbour passes run over copies of
    // the sorted rows in bucket order inside each cell; the density pass writes
    // measures in the reference order, so every array a caller sees keeps its
    // meaning, and -rho / p into the copies' w for the force pass
    bool subwalk = false;
    float *posW = nullptr, *velW = nullptr;
    uint32_t *indexW = nullptr, *w2r = nullptr, *preW = nullptr;
    uint8_t* bucketR = nullptr;
    // in-range masks the density pass leaves for the force pass (staged blocks
    // of the sub-cell walk): 9 rows x 64 bits per particle
    static constexpr size_t MASK_WORDS = 18;
    bool use_masks = true;
    uint32_t* masks = nullptr;
    // the demos' colour buffer (float4 RGBA per particle), made on first use
    float* color = nullptr;

    float elapsed = 0.0f;
    // Peristalsis boundary-configuration state (PeristalsisSystem.cpp:330-343):
    // dState drives the kernels, hstate mirrors it on the host (same function,
    // same arithmetic) for queries; state_dirty sends hstate to the device in
    // front of the next step after the host changed it
    StepState hstate{};
    StepState* dState = nullptr;
    bool state_dirty = true;

    hipGraphExec_t graph = nullptr;
    bool graph_valid = false;

    // per-stage event timing: ev[k] recorded before stage k, ev[NUM] after the last
    bool timing = false;
    hipEvent_t ev[CMAG_NUM_STAGES + 1] = {};
    double stage_ms[CMAG_NUM_STAGES] = {};
    uint32_t stage_count[CMAG_NUM_STAGES] = {};
    bool pending = false;

    int mark(int k) {
        if (timing) CMAG_HIP_TRY(hipEventRecord(ev[k], stream));
        return CMAG_OK;
    }

    // fold the previous step's events into the totals (waits for that step)
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

    ~cmag_system() {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        void* bufs[] = {pos, vel, velLF, acc, presF, measures, sortedPos, sortedVel, predicted,
                        hash, index, hashAlt, indexAlt, cellStart, cellEnd, cellFirst, scanWs, sortWs,
                        gaccU, gaccS, gposS, gvelS, grho, dState, posW, velW, indexW, w2r, preW, bucketR,
                        masks, color};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }

    int alloc() {
        const size_t f4 = sizeof(float) * 4 * (size_t)(n ? n : 1);
        const size_t u = sizeof(uint32_t) * (size_t)(n ? n : 1);
        CMAG_HIP_TRY(hipStreamCreate(&own_stream));
        stream = own_stream;
        float** fb[] = {&pos, &vel, &velLF, &acc, &measures, &sortedPos, &sortedVel};
        for (float** b : fb) CMAG_HIP_TRY(hipMalloc((void**)b, f4));
        if (P.system == CMAG_SYS_PERISTALSIS) {
            CMAG_HIP_TRY(hipMalloc((void**)&presF, f4));
            CMAG_HIP_TRY(hipMalloc((void**)&predicted, f4));
            CMAG_HIP_TRY(hipMalloc((void**)&dState, sizeof(StepState)));
        }
        uint32_t** ub[] = {&hash, &index, &hashAlt, &indexAlt};
        for (uint32_t** b : ub) CMAG_HIP_TRY(hipMalloc((void**)b, u));
        CMAG_HIP_TRY(hipMalloc((void**)&cellStart, sizeof(uint32_t) * (size_t)numCells));
        CMAG_HIP_TRY(hipMalloc((void**)&cellEnd, sizeof(uint32_t) * (size_t)numCells));
        CMAG_HIP_TRY(hipMalloc((void**)&cellFirst, sizeof(uint32_t) * ((size_t)numCells + 1)));
        CMAG_HIP_TRY(hipMalloc((void**)&scanWs, sizeof(uint32_t) * (scan_levels_words(numCells) + 1)));
        sortWsBytes = radix_sort_workspace_bytes(n);
        CMAG_HIP_TRY(hipMalloc(&sortWs, sortWsBytes));
        float* zb[] = {pos, vel, velLF, acc, measures, sortedPos, sortedVel, presF, predicted};
        for (float* b : zb)
            if (b) CMAG_HIP_TRY(hipMemsetAsync(b, 0, f4, stream));
        for (uint32_t* b : {hash, index, hashAlt, indexAlt}) CMAG_HIP_TRY(hipMemsetAsync(b, 0, u, stream));
        CMAG_HIP_TRY(hipMemsetAsync(cellStart, 0xff, sizeof(uint32_t) * (size_t)numCells, stream));
        CMAG_HIP_TRY(hipMemsetAsync(cellEnd, 0, sizeof(uint32_t) * (size_t)numCells, stream));
        CMAG_HIP_TRY(hipStreamSynchronize(stream));
        return CMAG_OK;
    }

    int passes() const { return (key_bits + 7) / 8; }

    // the sub-cell walk needs DamBreak's DIV_CELL cells, no periodic images,
    // rows whose first and last cells are distinct, and a stencil that reaches
    // 2h along x (what it drops from the outer cells lies beyond cellcount cells)
    bool subwalk_eligible() const {
        return P.system == CMAG_SYS_DAMBREAK && P.gridpos_rule == CMAG_GRIDPOS_DIV_CELL && P.periodic_mask == 0 &&
               P.grid_size[0] >= 2u * (uint32_t)P.cellcount + 2u &&
               (double)P.cellcount * P.cell_size[0] >= 2.0 * P.smoothing_radius;
    }

    int set_subwalk(bool on) {
        if (on && !subwalk_eligible())
            return fail(CMAG_E_INVALID, "the sub-cell walk order needs a DamBreak system with DIV_CELL cells, no "
                                        "periodic images, more than 2*cellcount+1 cells along x and "
                                        "cellcount * cell_size.x >= 2h");
        if (on && !posW) {
            const size_t f4 = sizeof(float) * 4 * (size_t)(n ? n : 1);
            const size_t u = sizeof(uint32_t) * (size_t)(n ? n : 1);
            // padding rows: the walk reads up to two rows past a run
            for (float** b : {&posW, &velW}) CMAG_HIP_TRY(hipMalloc((void**)b, f4 + 64));
            for (uint32_t** b : {&indexW, &w2r}) CMAG_HIP_TRY(hipMalloc((void**)b, u));
            // padded to whole 16-byte words: the move kernel reads buckets 16 at a time
            CMAG_HIP_TRY(hipMalloc((void**)&bucketR, ((size_t)n + 31) & ~(size_t)15));
            CMAG_HIP_TRY(hipMalloc((void**)&preW, sizeof(uint32_t) * 4 * (size_t)numCells));
            CMAG_HIP_TRY(hipMalloc((void**)&masks, sizeof(uint32_t) * MASK_WORDS * (size_t)(n ? n : 1)));
        }
        if (on != subwalk) invalidate_graph();
        subwalk = on;
        return CMAG_OK;
    }

    // hash -> sort -> reorder/cellStart -> density: the neighbour structure
    // and densities of the current state
    int enqueue_front() {
        // the hash goes to whichever buffer makes the sort end in (hash, index)
        const bool odd = passes() % 2 == 1;
        uint32_t* hk = odd ? hashAlt : hash;
        uint32_t* hv = odd ? indexAlt : index;
        CMAG_TRY(mark(CMAG_STAGE_HASH));
        CMAG_TRY(launch_calc_hash(P, pos, hk, hv, n, stream));
        CMAG_TRY(mark(CMAG_STAGE_SORT));
        bool in_alt = false;
        CMAG_TRY(radix_sort_pairs(hk, hv, odd ? hash : hashAlt, odd ? index : indexAlt, n, key_bits,
                                  sortWs, sortWsBytes, &in_alt, stream));
        const bool sorted_moved = n >= 2 && key_bits > 0;
        if (sorted_moved && in_alt != odd)
            return fail(CMAG_E_STATE, "radix sort ended in an unexpected buffer");
        if (!sorted_moved && odd) {
            // no pass ran: move the keys from where the hash put them
            CMAG_TRY(launch_copy_u32(hash, hk, n, stream));
            CMAG_TRY(launch_copy_u32(index, hv, n, stream));
        }
        CMAG_TRY(mark(CMAG_STAGE_REORDER));
        CMAG_TRY(launch_reorder(cellStart, cellEnd, sortedPos, sortedVel, hash, index, pos, velLF, n,
                                numCells, stream, &P, subwalk ? bucketR : nullptr));
        CMAG_TRY(launch_cell_offsets(cellStart, cellEnd, cellFirst, numCells, n, scanWs, stream));
        if (subwalk)
            CMAG_TRY(launch_subcell_build(P, hash, index, sortedPos, sortedVel, cellFirst, n, numCells, preW, bucketR,
                                          posW, velW, indexW, w2r, stream));
        CMAG_TRY(mark(CMAG_STAGE_DENSITY));
        if (subwalk)
            CMAG_TRY(launch_density_sub(P, measures, w2r, posW, velW, cellFirst, preW, n, stream,
                                        use_masks ? masks : nullptr));
        else
            CMAG_TRY(launch_density(P, measures, sortedPos, cellStart, cellEnd, n, stream, cellFirst));
        return CMAG_OK;
    }

    // one update(): hash -> sort -> reorder/cellStart -> density -> forces -> integrate
    int enqueue_step() {
        CMAG_TRY(enqueue_front());
        CMAG_TRY(mark(CMAG_STAGE_FORCE));
        if (P.system == CMAG_SYS_PERISTALSIS) {
            CMAG_TRY(launch_peristalsis_forces(P, acc, presF, measures, sortedPos, sortedVel, index,
                                               cellStart, cellEnd, n, stream, cellFirst));
        } else if (subwalk) {
            CMAG_TRY(launch_force_sub(P, acc, posW, velW, indexW, cellFirst, preW, n, stream,
                                      use_masks ? masks : nullptr));
        } else {
            CMAG_TRY(launch_force(P, acc, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n,
                                  stream, cellFirst));
        }
        CMAG_TRY(mark(CMAG_STAGE_INTEGRATE));
        CMAG_TRY(launch_integrate(P, pos, vel, velLF, acc, presF, 0.0f, n, stream, dState));
        CMAG_TRY(mark(CMAG_NUM_STAGES));
        if (timing) pending = true;
        return CMAG_OK;
    }

    // one update() on the stream: for Peristalsis the state record advances on
    // the device (PeristalsisSystem::Update, PeristalsisSystem.cpp:330-343) and
    // the walls are configured when it says so, in front of the step
    int enqueue_update() {
        if (P.system == CMAG_SYS_PERISTALSIS) {
            CMAG_TRY(launch_step_prologue(P, dState, stream));
            CMAG_TRY(launch_configure_boundary_from_state(P, pos, dState, n, stream));
        }
        return enqueue_step();
    }

    // send the host's state record to the device if the host changed it
    int upload_state() {
        if (P.system != CMAG_SYS_PERISTALSIS || !state_dirty) return CMAG_OK;
        hstate.elapsed = elapsed;
        hstate.wave_height_inc = P.delta_time * powf(10.0f, -3.0f);
        hstate.is_config = P.is_boundary_configuration ? 1 : 0;
        CMAG_HIP_TRY(hipMemcpyAsync(dState, &hstate, sizeof(StepState), hipMemcpyHostToDevice, stream));
        CMAG_HIP_TRY(hipStreamSynchronize(stream));
        state_dirty = false;
        return CMAG_OK;
    }

    // the host side of one step: the mirror takes the step the device took
    void advance_clock() {
        if (P.system == CMAG_SYS_PERISTALSIS) {
            advance_step_state(P, hstate);
            P.is_boundary_configuration = hstate.is_config;
            elapsed = hstate.elapsed;
        } else {
            elapsed += P.delta_time;
        }
    }

    int update() {
        CMAG_TRY(collect());
        CMAG_TRY(upload_state());
        CMAG_TRY(enqueue_update());
        advance_clock();
        return CMAG_OK;
    }

    void invalidate_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_valid = false;
    }

    int build_graph() {
        invalidate_graph();
        if (stream == nullptr) return fail(CMAG_E_STATE, "graph capture needs a non-default stream");
        hipGraph_t g = nullptr;
        const bool was_timing = timing;
        timing = false;  // no event nodes inside the graph
        CMAG_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_update();
        timing = was_timing;
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc != CMAG_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        graph_valid = true;
        return CMAG_OK;
    }

    int run(uint32_t steps, bool use_graph) {
        if (!use_graph) {
            for (uint32_t k = 0; k < steps; ++k) CMAG_TRY(update());
            return CMAG_OK;
        }
        CMAG_TRY(collect());
        CMAG_TRY(upload_state());
        if (!graph_valid) CMAG_TRY(build_graph());
        for (uint32_t k = 0; k < steps; ++k) {
            CMAG_HIP_TRY(hipGraphLaunch(graph, stream));
            advance_clock();
        }
        return CMAG_OK;
    }

    // adjoint of one update() at the current state (pos, vel, velLF), which
    // is left unchanged: (gpos, gvel, gvelLF) hold the adjoints of the step's
    // outputs on entry and of its inputs on return; gmu_acc[n] and
    // ggrav_acc[4n] (FP64, optional) accumulate the parameter adjoints
    int vjp_step(float* gpos, float* gvel, float* gvelLF, double* gmu_acc, double* ggrav_acc) {
        if (P.system != CMAG_SYS_POISEUILLE)
            return fail(CMAG_E_INVALID, "vjp_step: the adjoint covers the Poiseuille system");
        if (n && (!gpos || !gvel || !gvelLF)) return fail(CMAG_E_INVALID, "vjp_step: null adjoint");
        if (!gaccU) {
            const size_t f4 = sizeof(float) * 4 * (size_t)(n ? n : 1);
            float** fb[] = {&gaccU, &gaccS, &gposS, &gvelS};
            for (float** b : fb) CMAG_HIP_TRY(hipMalloc((void**)b, f4));
            CMAG_HIP_TRY(hipMalloc((void**)&grho, sizeof(float) * (size_t)(n ? n : 1)));
        }
        invalidate_graph();
        const bool was_timing = timing;
        timing = false;
        int rc = enqueue_front();
        timing = was_timing;
        CMAG_TRY(rc);
        CMAG_TRY(launch_integrate_vjp(P, pos, gpos, gvel, gvelLF, gaccU, ggrav_acc, n, stream));
        CMAG_TRY(launch_gather_f4(gaccS, gaccU, index, n, stream));
        CMAG_TRY(launch_force_vjp(P, gposS, gvelS, grho, nullptr, gmu_acc, gaccS, measures, sortedPos, sortedVel,
                                  cellStart, cellEnd, cellFirst, n, stream));
        CMAG_TRY(launch_density_vjp(P, gposS, grho, sortedPos, cellStart, cellEnd, cellFirst, n, index, gvelS,
                                    gpos, gvelLF, stream));
        return CMAG_OK;
    }

    // device-to-device copies, ordered on the system's stream
    int copy_array_device(int a, float* ext, int start, int count, bool to_system) {
        float* d = farray(a);
        if (!d) return fail(CMAG_E_INVALID, "array copy: unknown array for this system");
        if (start < 0 || count < 0 || (uint64_t)start + (uint64_t)count > n)
            return fail(CMAG_E_INVALID, "array copy: range outside the particle arrays");
        if (count == 0) return CMAG_OK;
        if (!ext) return fail(CMAG_E_INVALID, "array copy: null device pointer");
        float* sys = d + 4 * (size_t)start;
        CMAG_HIP_TRY(hipMemcpyAsync(to_system ? sys : ext, to_system ? ext : sys,
                                    sizeof(float) * 4 * (size_t)count, hipMemcpyDeviceToDevice, stream));
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
        case CMAG_ARRAY_PREDICTEDPOSITION: return predicted;
        default: return nullptr;
        }
    }
    uint32_t* uarray(int a, uint32_t* len) {
        switch (a) {
        case CMAG_ARRAY_HASH: *len = n; return hash;
        case CMAG_ARRAY_INDEX: *len = n; return index;
        case CMAG_ARRAY_CELLSTART: *len = numCells; return cellStart;
        case CMAG_ARRAY_CELLEND: *len = numCells; return cellEnd;
        default: *len = 0; return nullptr;
        }
    }

    int set_array(int a, const float* data, int start, int count) {
        float* d = farray(a);
        if (!d) return fail(CMAG_E_INVALID, "setArray: unknown array for this system");
        if (start < 0 || count < 0 || (uint64_t)start + (uint64_t)count > n)
            return fail(CMAG_E_INVALID, "setArray: range outside the particle arrays");
        if (count == 0) return CMAG_OK;
        if (!data) return fail(CMAG_E_INVALID, "setArray: null data");
        CMAG_HIP_TRY(hipMemcpyAsync(d + 4 * (size_t)start, data, sizeof(float) * 4 * (size_t)count,
                                    hipMemcpyHostToDevice, stream));
        CMAG_HIP_TRY(hipStreamSynchronize(stream));
        return CMAG_OK;
    }

    int get_array(int a, float* data, int start, int count) {
        float* d = farray(a);
        if (!d) return fail(CMAG_E_INVALID, "getArray: unknown array for this system");
        if (start < 0 || count < 0 || (uint64_t)start + (uint64_t)count > n)
            return fail(CMAG_E_INVALID, "getArray: range outside the particle arrays");
        if (count == 0) return CMAG_OK;
        if (!data) return fail(CMAG_E_INVALID, "getArray: null data");
        CMAG_HIP_TRY(hipMemcpyAsync(data, d + 4 * (size_t)start, sizeof(float) * 4 * (size_t)count,
                                    hipMemcpyDeviceToHost, stream));
        CMAG_HIP_TRY(hipStreamSynchronize(stream));
        return CMAG_OK;
    }
};

namespace cmag {
// reference reset() initial states (shared with nothing else; the oracle has
// its own restatement for the tests)
void init_dambreak(const cmag_params& P, float* pos4, uint32_t n);
void init_poiseuille(const cmag_params& P, float* pos4, uint32_t n);
void init_peristalsis(const cmag_params& P, float* pos4, uint32_t n);
float peristalsis_lattice_sum(const cmag_params& P, const float* pos4, uint32_t n);
}  // namespace cmag

static int make_system(const cmag_params& P, uint32_t n, int ref_kind, cmag_system** out) {
    if (!out) return fail(CMAG_E_INVALID, "null output handle");
    *out = nullptr;
    CMAG_TRY(validate_params(&P));
    cmag_system* s = new (std::nothrow) cmag_system();
    if (!s) return fail(CMAG_E_INVALID, "out of host memory");
    s->P = P;
    s->ref_kind = ref_kind;
    s->n = n;
    s->numCells = cmag_params_num_cells(&P);
    s->key_bits = cmag_params_key_bits(&P);
    s->hstate = initial_step_state(P, 0.0f);
    int rc = s->alloc();
    if (rc == CMAG_OK && s->subwalk_eligible()) rc = s->set_subwalk(true);
    if (rc != CMAG_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return CMAG_OK;
}

extern "C" {

int cmag_dambreak_create(uint32_t fx, uint32_t fy, uint32_t fz, int32_t bo, uint32_t gx, uint32_t gy,
                         uint32_t gz, float r, cmag_system** out) {
    cmag_params P;
    CMAG_TRY(cmag_params_dambreak(&P, fx, fy, fz, bo, gx, gy, gz, r));
    if (bo < 1) return fail(CMAG_E_INVALID, "DamBreak needs boundaryOffset >= 1");
    // fluidSystem.cpp:34-37
    uint32_t n = fx * fy * fz + gx * (uint32_t)bo + 2u * (gy - (uint32_t)bo) * (uint32_t)bo;
    return make_system(P, n, CMAG_SYS_DAMBREAK, out);
}

int cmag_poiseuille_create(uint32_t fx, uint32_t fy, uint32_t fz, int32_t bo, uint32_t gx, uint32_t gy,
                           uint32_t gz, float r, cmag_system** out) {
    cmag_params P;
    CMAG_TRY(cmag_params_poiseuille(&P, fx, fy, fz, bo, gx, gy, gz, r));
    // poiseuilleFlowSystem.cpp:36-37
    uint32_t n = fx * fy * fz + 2u * gx * (uint32_t)bo;
    return make_system(P, n, CMAG_SYS_POISEUILLE, out);
}

int cmag_peristalsis_create(float dt, uint32_t fx, uint32_t fy, uint32_t fz, float amplitude,
                            float wave_speed, float soundspeed, float gravx, float gravy, float gravz,
                            int32_t bo, uint32_t gx, uint32_t gy, uint32_t gz, float radius,
                            cmag_system** out) {
    cmag_params P;
    CMAG_TRY(cmag_params_peristalsis(&P, dt, fx, fy, fz, amplitude, wave_speed, soundspeed, gravx, gravy,
                                     gravz, bo, gx, gy, gz, radius));
    // PeristalsisSystem.cpp:36-37
    uint32_t n = fx * fy * fz + 2u * gx * (uint32_t)bo;
    return make_system(P, n, CMAG_SYS_PERISTALSIS, out);
}

int cmag_system_create(const cmag_params* p, uint32_t n, cmag_system** out) {
    if (!p) return fail(CMAG_E_INVALID, "null parameter block");
    return make_system(*p, n, REF_NONE, out);
}

int cmag_reference_initial_state(const cmag_params* p, uint32_t n, float* pos4) {
    CMAG_TRY(validate_params(p));
    if (n && !pos4) return fail(CMAG_E_INVALID, "null output");
    if (p->system == CMAG_SYS_DAMBREAK) init_dambreak(*p, pos4, n);
    else if (p->system == CMAG_SYS_POISEUILLE) init_poiseuille(*p, pos4, n);
    else init_peristalsis(*p, pos4, n);
    return CMAG_OK;
}

float cmag_peristalsis_lattice_sum(const cmag_params* p, const float* pos4, uint32_t n) {
    if (!p || !pos4 || n == 0) return 0.0f;
    return peristalsis_lattice_sum(*p, pos4, n);
}

int cmag_system_destroy(cmag_system* s) {
    if (s) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        delete s;
    }
    return CMAG_OK;
}

int cmag_system_reset(cmag_system* s) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    if (s->ref_kind == REF_NONE)
        return fail(CMAG_E_STATE, "reset() exists only for systems built by a reference constructor");
    s->invalidate_graph();
    const uint32_t n = s->n;
    std::vector<float> hpos(4 * (size_t)n, 0.0f), zero(4 * (size_t)n, 0.0f), meas(4 * (size_t)n, 0.0f);
    s->elapsed = 0.0f;
    if (s->ref_kind == CMAG_SYS_PERISTALSIS) {
        // PeristalsisSystem::Reset, PeristalsisSystem.cpp:213-237
        s->P.is_boundary_configuration = 1;
        s->hstate = initial_step_state(s->P, 0.0f);
        s->state_dirty = true;
        init_peristalsis(s->P, hpos.data(), n);
        CMAG_TRY(s->set_array(CMAG_ARRAY_POSITION, hpos.data(), 0, (int)n));
        CMAG_TRY(s->set_array(CMAG_ARRAY_VELOCITY, zero.data(), 0, (int)n));
        CMAG_TRY(s->set_array(CMAG_ARRAY_MEASURES, zero.data(), 0, (int)n));
        CMAG_TRY(s->set_array(CMAG_ARRAY_ACCELERATION, zero.data(), 0, (int)n));
        CMAG_TRY(s->set_array(CMAG_ARRAY_PRESSUREFORCE, zero.data(), 0, (int)n));
        CMAG_TRY(s->set_array(CMAG_ARRAY_VELOCITYLEAPFROG, zero.data(), 0, (int)n));
        CMAG_TRY(s->set_array(CMAG_ARRAY_PREDICTEDPOSITION, zero.data(), 0, (int)n));
        s->P.particle_mass = s->P.rest_density / peristalsis_lattice_sum(s->P, hpos.data(), n);
        if (s->color) CMAG_TRY(launch_fill_colors(s->P, s->pos, s->color, n, s->stream));
        return CMAG_OK;
    }
    // DamBreakSystem::reset fluidSystem.cpp:361-380; PoiseuilleFlowSystem::reset
    // poiseuilleFlowSystem.cpp:311-323.  hMeasures starts as (restDensity, 0, 0, 0).
    if (s->ref_kind == CMAG_SYS_DAMBREAK) init_dambreak(s->P, hpos.data(), n);
    else init_poiseuille(s->P, hpos.data(), n);
    for (uint32_t i = 0; i < n; ++i) meas[4 * (size_t)i] = s->P.rest_density;
    CMAG_TRY(s->set_array(CMAG_ARRAY_POSITION, hpos.data(), 0, (int)n));
    CMAG_TRY(s->set_array(CMAG_ARRAY_VELOCITY, zero.data(), 0, (int)n));
    CMAG_TRY(s->set_array(CMAG_ARRAY_MEASURES, meas.data(), 0, (int)n));
    CMAG_TRY(s->set_array(CMAG_ARRAY_ACCELERATION, zero.data(), 0, (int)n));
    CMAG_TRY(s->set_array(CMAG_ARRAY_VELOCITYLEAPFROG, zero.data(), 0, (int)n));
    if (s->ref_kind == CMAG_SYS_DAMBREAK) {
        const cmag_params& P = s->P;
        s->P.right_boundary = P.world_origin[0] + (P.boundary_offset + P.fluid_size[0]) * 2 * P.particle_radius;
    }
    if (s->color) CMAG_TRY(launch_fill_colors(s->P, s->pos, s->color, n, s->stream));
    return CMAG_OK;
}

int cmag_system_update(cmag_system* s) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    return s->update();
}

int cmag_system_update_dt(cmag_system* s, float dt) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    if (!(dt > 0.0f)) return fail(CMAG_E_INVALID, "update(dt): dt must be positive");
    if (dt != s->P.delta_time) {
        s->P.delta_time = dt;
        s->state_dirty = true;
        s->invalidate_graph();
    }
    return s->update();
}

int cmag_system_run(cmag_system* s, uint32_t steps, int use_graph) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    return s->run(steps, use_graph != 0);
}

int cmag_system_set_array(cmag_system* s, int array, const float* data, int start, int count) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    return s->set_array(array, data, start, count);
}

int cmag_system_get_array(cmag_system* s, int array, float* data, int start, int count) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    return s->get_array(array, data, start, count);
}

int cmag_system_get_uint_array(cmag_system* s, int array, uint32_t* data, uint32_t start, uint32_t count) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    uint32_t len = 0;
    uint32_t* d = s->uarray(array, &len);
    if (!d) return fail(CMAG_E_INVALID, "unknown uint array");
    if ((uint64_t)start + count > len) return fail(CMAG_E_INVALID, "range outside the array");
    if (count == 0) return CMAG_OK;
    if (!data) return fail(CMAG_E_INVALID, "null data");
    CMAG_HIP_TRY(hipMemcpyAsync(data, d + start, sizeof(uint32_t) * (size_t)count, hipMemcpyDeviceToHost,
                                s->stream));
    CMAG_HIP_TRY(hipStreamSynchronize(s->stream));
    return CMAG_OK;
}

void* cmag_system_device_ptr(cmag_system* s, int array) {
    if (!s) return nullptr;
    float* f = s->farray(array);
    if (f) return f;
    uint32_t len;
    return s->uarray(array, &len);
}

uint32_t cmag_system_num_particles(const cmag_system* s) { return s ? s->n : 0u; }
uint32_t cmag_system_num_cells(const cmag_system* s) { return s ? s->numCells : 0u; }
float cmag_system_elapsed_time(const cmag_system* s) { return s ? s->elapsed : 0.0f; }

int cmag_system_get_params(const cmag_system* s, cmag_params* out) {
    if (!s || !out) return fail(CMAG_E_INVALID, "null argument");
    *out = s->P;
    return CMAG_OK;
}

int cmag_system_set_params(cmag_system* s, const cmag_params* p) {
    if (!s || !p) return fail(CMAG_E_INVALID, "null argument");
    CMAG_TRY(validate_params(p));
    for (int a = 0; a < 3; ++a)
        if (p->grid_size[a] != s->P.grid_size[a]) return fail(CMAG_E_INVALID, "set_params may not change the grid");
    if (p->system != s->P.system) return fail(CMAG_E_INVALID, "set_params may not change the system kind");
    s->P = *p;
    s->key_bits = cmag_params_key_bits(p);
    s->state_dirty = true;
    s->invalidate_graph();
    if (s->subwalk && !s->subwalk_eligible()) s->subwalk = false;
    return CMAG_OK;
}

int cmag_system_set_subcell_walk(cmag_system* s, int on) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    CMAG_TRY(s->collect());
    return s->set_subwalk(on != 0);
}

int cmag_system_subcell_walk(const cmag_system* s) { return s && s->subwalk ? 1 : 0; }

int cmag_system_set_inrange_masks(cmag_system* s, int on) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    CMAG_TRY(s->collect());
    if ((on != 0) != s->use_masks) s->invalidate_graph();
    s->use_masks = on != 0;
    return CMAG_OK;
}

int cmag_system_inrange_masks(const cmag_system* s) { return s && s->use_masks ? 1 : 0; }

int cmag_system_set_stream(cmag_system* s, cmag_stream_t stream) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    CMAG_HIP_TRY(hipStreamSynchronize(s->stream));
    s->stream = stream ? as_stream(stream) : s->own_stream;
    s->invalidate_graph();
    return CMAG_OK;
}

int cmag_system_vjp_step(cmag_system* s, float* gpos, float* gvel, float* gvelLF, double* gmu_acc,
                          double* ggrav_acc) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    CMAG_TRY(s->collect());
    return s->vjp_step(gpos, gvel, gvelLF, gmu_acc, ggrav_acc);
}

int cmag_system_copy_array_device(cmag_system* s, int array, float* device_data, int start, int count,
                                  int to_system) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    return s->copy_array_device(array, device_data, start, count, to_system != 0);
}

int cmag_system_synchronize(cmag_system* s) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    CMAG_HIP_TRY(hipStreamSynchronize(s->stream));
    return CMAG_OK;
}

int cmag_system_enable_timing(cmag_system* s, int on) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    if (on && !s->ev[0])
        for (hipEvent_t& e : s->ev) CMAG_HIP_TRY(hipEventCreate(&e));
    CMAG_TRY(s->collect());
    s->timing = on != 0;
    return CMAG_OK;
}

int cmag_system_stage_times(cmag_system* s, double* ms, uint32_t* counts, int max_stages) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    CMAG_TRY(s->collect());
    for (int k = 0; k < CMAG_NUM_STAGES && k < max_stages; ++k) {
        if (ms) ms[k] = s->stage_ms[k];
        if (counts) counts[k] = s->stage_count[k];
        s->stage_ms[k] = 0.0;
        s->stage_count[k] = 0;
    }
    return CMAG_OK;
}

// DamBreakSystem::changeRightBoundary, fluidSystem.cpp:239-255
int cmag_system_change_right_boundary(cmag_system* s) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    if (s->P.system != CMAG_SYS_DAMBREAK) return fail(CMAG_E_STATE, "DamBreak only");
    s->P.right_boundary += s->P.fluid_size[0] * 2 * s->P.particle_radius;
    s->invalidate_graph();
    return launch_shift_right_boundary(s->P, s->pos, s->n, s->stream);
}

// DamBreakSystem::removeRightBoundary, fluidSystem.cpp:222-237
int cmag_system_remove_right_boundary(cmag_system* s) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    if (s->P.system != CMAG_SYS_DAMBREAK) return fail(CMAG_E_STATE, "DamBreak only");
    s->P.right_boundary = (float)0xffffffffu;
    s->invalidate_graph();
    CMAG_TRY(launch_remove_right_boundary(s->P, s->pos, s->n, s->stream));
    s->elapsed = 0.0f;
    return CMAG_OK;
}

// PoiseuilleFlowSystem::changeGravity, poiseuilleFlowSystem.cpp:207-210
int cmag_system_color_buffer(cmag_system* s, void** dev_color4) {
    if (!s || !dev_color4) return fail(CMAG_E_INVALID, "cmag_system_color_buffer: null pointer");
    if (!s->color) {
        CMAG_HIP_TRY(hipMalloc((void**)&s->color, sizeof(float) * 4 * (size_t)(s->n ? s->n : 1)));
        CMAG_TRY(launch_fill_colors(s->P, s->pos, s->color, s->n, s->stream));
    }
    *dev_color4 = s->color;
    return CMAG_OK;
}

int cmag_system_coloring(cmag_system* s) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    if (s->P.system != CMAG_SYS_PERISTALSIS) return fail(CMAG_E_INVALID, "Coloring() is PeristalsisSystem's");
    void* c = nullptr;
    if (s->color) return launch_fill_colors(s->P, s->pos, s->color, s->n, s->stream);
    return cmag_system_color_buffer(s, &c);
}

int cmag_system_change_gravity(cmag_system* s) {
    if (!s) return fail(CMAG_E_INVALID, "null system");
    s->P.gravity[1] *= -1.0f;
    s->invalidate_graph();
    return CMAG_OK;
}

}  // extern "C"
