// sph_device.hpp - per-particle bodies of the neighbour passes (gfx950, wave64,
// FP32), shared by the single-system kernels (sph.hip) and the batched
// persistent kernels (batch.hip).
//
// The stencil of (2*cellcount+1)^dim cells is walked in the reference's z, y, x
// order and each cell's particles in sorted order, so sums accumulate in the
// reference's order.  Candidates are culled on r^2 < 4h^2 before any square
// root; inner-loop reciprocals use the hardware v_rcp/v_sqrt (1 ulp), the
// per-particle divisions stay IEEE.
//
//   density       fluid_kernel.cu:99-169, poiseuilleFlowKernel.cu:99-182,
//                 PeristalsisDensity.cu:3-79
//   force         fluid_kernel.cu:172-292 (DamBreak: pressure + Monaghan AV,
//                 Lennard-Jones walls), poiseuilleFlowKernel.cu:184-309
//                 (pressure + Morris viscosity with wall extrapolation)
//   pressure /    PeristalsisPressureForce.cu:3-90,
//   viscous       PeristalsisViscousForce.cu:143-241
#pragma once

#include "common.hpp"

// Every fused multiply-add below is written out (fmaf) and no other is formed:
// the same sum then rounds the same way in every kernel that inlines it
// (walker or neighbour list, single system or batch), which keeps those
// paths bit-identical.
#pragma clang fp contract(off)

namespace cmag {

// |r|^2 as (x*x) then two fused steps
__device__ __forceinline__ float norm2(float rx, float ry, float rz) {
    return fmaf(rz, rz, fmaf(ry, ry, rx * rx));
}


__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Renumbering them so XCD k runs one contiguous block range keeps the
// neighbour particles a workgroup reads in the L2 of the XCD that reads them.
__device__ __forceinline__ uint32_t xcd_block() {
    constexpr uint32_t XCDS = 8;
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t k = b % XCDS, slot = b / XCDS;
    const uint32_t base = nb / XCDS, extra = nb % XCDS;
    return k * base + min(k, extra) + slot;
}

// integer power by squaring (gamma is an integer in [1,15])
__device__ __forceinline__ float pow_int(float x, int e) {
    float r = 1.0f, b = x;
    while (e) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return r;
}

__device__ __forceinline__ float eos(const cmag_params& P, float dens) {
    if (P.eos == CMAG_EOS_TAIT) return P.B * (pow_int(dens / P.rest_density, (int)P.gamma) - 1.0f);
    if (P.eos == CMAG_EOS_LINEAR_C2) return P.soundspeed * P.soundspeed * dens;
    return 50.0f * P.soundspeed * dens;
}

// Range policies of the stencil walker.  A run holds the cells with keys
// k0..k1 (consecutive keys are contiguous in the sorted arrays).
//   RefCells   the reference arrays: first non-empty cellStart and last
//              cellEnd of the run (empty cells hold 0xffffffff)
//   FirstCells compact offsets first[c] (cells before c), one pair of loads
struct RefCells {
    static constexpr bool SUB = false;
    const uint32_t* __restrict__ cellStart;
    const uint32_t* __restrict__ cellEnd;
    uint32_t n;
    __device__ __forceinline__ bool range(uint32_t k0, uint32_t k1, uint32_t& s, uint32_t& e) const {
        s = CMAG_CELL_EMPTY;
        e = 0;
        for (uint32_t k = k0; k <= k1; ++k) {
            const uint32_t st = cellStart[k];
            if (st == CMAG_CELL_EMPTY) continue;
            if (s == CMAG_CELL_EMPTY) s = st;
            // clamped to the particle count: consistent cell arrays never
            // need it, inconsistent ones cannot read past n
            e = min(cellEnd[k], n);
        }
        return s != CMAG_CELL_EMPTY && s < e;
    }
};

struct FirstCells {
    static constexpr bool SUB = false;
    const uint32_t* __restrict__ first;
    __device__ __forceinline__ bool range(uint32_t k0, uint32_t k1, uint32_t& s, uint32_t& e) const {
        s = first[k0];
        e = first[k1 + 1];
        return s < e;
    }
};

// ---- sub-cell walk order (DamBreak) ---------------------------------------------
// Inside each cell the particles can be kept in a second, internal order: by
// bucket, then in reference order, with buckets 0..7 the fluid particles of
// each eighth of the cell along x and bucket 8 the wall particles.  A target
// in eighth q of its cell needs, of the stencil's first cell along x, only the
// fluid of eighths >= q (the rest lies more than cellcount cells away, beyond
// 2h), and of the last cell only the eighths <= q.  Wall particles sit in
// their own bucket and are always walked: DamBreak's Lennard-Jones term acts
// on every wall particle of the stencil.  x_sub (common.hpp) is the target's /
// particle's eighth.

// SubCells: cell offsets first[] as FirstCells plus, per cell, the running
// counts pre[c]: eight 16-bit counts, #buckets <= t for t = 0..7 (t = 7: the
// cell's fluid), two to a word, low half first; cells without particles hold
// zeros
struct SubCells {
    static constexpr bool SUB = true;
    const uint32_t* __restrict__ first;
    const uint4* __restrict__ pre;
    // count t (0..7) of a cell's entry, read as one 16-bit value (the eighth
    // differs between the lanes of a wave)
    __device__ __forceinline__ static uint32_t count(const uint4* c, int t) {
        return reinterpret_cast<const uint16_t*>(c)[t];
    }
    // the fluid of eighths <= q, of eighths < q, and all of it
    __device__ __forceinline__ static uint32_t through(const uint4* c, int q) { return count(c, q); }
    __device__ __forceinline__ static uint32_t below(const uint4* c, int q) {
        const uint32_t v = count(c, q > 0 ? q - 1 : 0);
        return q > 0 ? v : 0u;
    }
    __device__ __forceinline__ static uint32_t fluid(const uint4* c) { return count(c, X_SUBS - 1); }
    // the run k0..k1 trimmed for a target in eighth q: lo (k0 is the
    // stencil's first cell along x) drops k0's fluid eighths < q, hi (k1 is
    // the last) ends at k1's eighth q and returns k1's walls as [ws, we)
    __device__ __forceinline__ void range_sub(uint32_t k0, uint32_t k1, bool lo, bool hi, int q, uint32_t& s,
                                              uint32_t& e, uint32_t& ws, uint32_t& we) const {
        s = first[k0];
        if (lo) s += below(pre + k0, q);
        ws = we = 0;
        if (hi) {
            const uint32_t f1 = first[k1];
            e = f1 + through(pre + k1, q);
            ws = f1 + fluid(pre + k1);
            we = first[k1 + 1];
        } else {
            e = first[k1 + 1];
        }
    }
};

// A run of key-consecutive cells: a contiguous range of the sorted arrays,
// with the target position minus the run's periodic image offset and the
// image-shift code (sx+1) + 3(sy+1) + 9(sz+1) (13: no shift).
struct Run {
    uint32_t s, len;
    float px;
    int code;
};

// The order in which the force pass takes the stencil rows (row r = (dz + 1)
// * 3 + (dy + 1)): rows of opposite offsets in pairs -- the diagonal corners
// (y - 1, z - 1) with (y + 1, z + 1) and (y + 1, z - 1) with (y - 1, z + 1),
// then y - 1 with y + 1 and z - 1 with z + 1 -- and the centre row last.  A
// lane's in-range neighbours are spread unevenly between opposite rows (a
// target near the top of its cell has most of them above), so taking a pair's
// mask bits as one sequence per lane evens the lanes' work out; the walks
// without masks take the rows in the same order, so the sums are the same.
__device__ __forceinline__ int force_row(int dim, int k) {
    // 3-D: 0 8 2 6 3 5 1 7 4; 2-D: 0 2 1
    return dim == 3 ? (int)((0x471536280ull >> (4 * k)) & 15u) : (k == 0 ? 0 : k == 1 ? 2 : 1);
}

// Walk the stencil of gp in z, y, x order and call cand(j, qx, qy, qz, code)
// for every candidate j in the reference's order.  Along x, cells with
// consecutive keys form one run (linear order: the whole row; brick Z order:
// the part inside one brick; plain Z order: an even/odd pair); a row has at
// most two runs except at periodic seams.  The runs of a row are gathered
// first and then walked as one loop, so every lane of a wave spends the row in
// the same loop whatever its x position -- lanes whose row splits at a
// different x would otherwise serialise the wave.
// PERIODIC = false compiles the image shifts out (DamBreak has none).
//
// walk_rows does the walk and hands each row's runs to row(a, b, qy, qz);
// walk_stencil and walk_stencil_pf turn them into candidate calls.  With
// FORCE_ORDER and one cell per stencil side, the rows come in force_row order.
template <int DIM, int HASH, bool PERIODIC = true, bool FORCE_ORDER = false, typename Cells, typename R>
__device__ __forceinline__ void walk_rows(const cmag_params& P, const float4& pi, const Cell& gp,
                                          const Cells& cells, R&& run_pair) {
    const int cc = P.cellcount;
    const int zlo = DIM == 2 ? 0 : -cc, zhi = DIM == 2 ? 0 : cc;
    const int sub = Cells::SUB ? x_sub(P, pi.x) : 0;
    const uint32_t gxm = P.grid_size[0] - 1u, gym = P.grid_size[1] - 1u, gzm = P.grid_size[2] - 1u;
    const uint32_t Mx = P.morton_mask[0], My = P.morton_mask[1], Mz = P.morton_mask[2];
    const bool perx = PERIODIC && (P.periodic_mask & 1), pery = PERIODIC && (P.periodic_mask & 2),
               perz = PERIODIC && (P.periodic_mask & 4);
    uint32_t kx0 = 0, ky0 = 0, kz0 = 0;
    if (HASH == CMAG_HASH_MORTON) {
        const uint32_t own = grid_hash(P, gp.x, gp.y, gp.z);
        kx0 = own & Mx;
        ky0 = own & My;
        kz0 = own & Mz;
        for (int k = 0; k < cc; ++k) {
            kx0 = dilated_dec(kx0, Mx);
            ky0 = dilated_dec(ky0, My);
            if (DIM == 3) kz0 = dilated_dec(kz0, Mz);
        }
    }
    // one stencil row (dy, dz); ky, kz: its Morton key parts
    auto row_body = [&](int dy, int dz, uint32_t ky, uint32_t kz) {
        const int nz = gp.z + dz;
        const int sz = perz ? image_shift(nz, P.grid_size[2]) : 0;
        const float pz = perz ? pi.z - (float)sz * P.world_size[2] : pi.z;
        const int ny = gp.y + dy;
        const int sy = pery ? image_shift(ny, P.grid_size[1]) : 0;
        const float py = pery ? pi.y - (float)sy * P.world_size[1] : pi.y;
        const int code_yz = 3 * (sy + 1) + 9 * (sz + 1);
        const uint32_t row = (HASH == CMAG_HASH_LINEAR)
                                 ? (((uint32_t)nz & gzm) * P.grid_size[1] + ((uint32_t)ny & gym)) * P.grid_size[0]
                                 : (kz | ky);
        Run a{0u, 0u, pi.x, 13}, b{0u, 0u, pi.x, 13};
        bool have_a = false;
        auto push = [&](uint32_t s, uint32_t e, float px, int code) {
            const Run r{s, e - s, px, code};
            if (!have_a) {
                a = r;
                have_a = true;
            } else if (b.len == 0) {
                b = r;
            } else {  // a third run in one row (periodic seam with plain Z order, or walls)
                run_pair(a, b, py, pz);
                a = r;
                b.len = 0;
            }
        };
        // lo / hi: the run starts at the stencil's first cell / ends at its last
        auto emit = [&](uint32_t k0, uint32_t k1, float px, int code, bool lo, bool hi) {
            uint32_t s, e;
            if constexpr (Cells::SUB) {
                uint32_t ws, we;
                cells.range_sub(k0, k1, lo, hi, sub, s, e, ws, we);
                if (s < e) push(s, e, px, code);
                if (ws < we) push(ws, we, px, code);
            } else {
                if (cells.range(k0, k1, s, e)) push(s, e, px, code);
            }
        };
        uint32_t k0 = 0, k1 = 0;
        float run_px = pi.x;
        int run_sx = 0;
        bool run_lo = true;
        uint32_t kx = kx0;
        for (int dx = -cc; dx <= cc; ++dx) {
            const int nx = gp.x + dx;
            const int sx = perx ? image_shift(nx, P.grid_size[0]) : 0;
            const float px = perx ? pi.x - (float)sx * P.world_size[0] : pi.x;
            const uint32_t key = (HASH == CMAG_HASH_LINEAR) ? row + ((uint32_t)nx & gxm) : (row | kx);
            if (dx == -cc) {
                k0 = k1 = key;
                run_px = px;
                run_sx = sx;
            } else if (key == k1 + 1u && sx == run_sx) {
                k1 = key;
            } else {
                emit(k0, k1, run_px, (run_sx + 1) + code_yz, run_lo, false);
                k0 = k1 = key;
                run_px = px;
                run_sx = sx;
                run_lo = false;
            }
            if (HASH == CMAG_HASH_MORTON) kx = dilated_inc(kx, Mx);
        }
        emit(k0, k1, run_px, (run_sx + 1) + code_yz, run_lo, true);
        run_pair(a, b, py, pz);
    };
    if (FORCE_ORDER && cc == 1) {
        // the force pass's row order (force_row), as its staged blocks take it;
        // the Morton key parts are stepped from the stencil's first row each time
        constexpr int NR = DIM == 3 ? 9 : 3;
        for (int k = 0; k < NR; ++k) {
            const int r = force_row(DIM, k);
            const int iy = r % 3, iz = DIM == 3 ? r / 3 : 0;
            uint32_t ky = ky0, kz = kz0;
            if (HASH == CMAG_HASH_MORTON) {
                for (int t = 0; t < iy; ++t) ky = dilated_inc(ky, My);
                for (int t = 0; t < iz; ++t) kz = dilated_inc(kz, Mz);
            }
            row_body(iy - 1, DIM == 3 ? iz - 1 : 0, ky, kz);
        }
        return;
    }
    uint32_t kz = kz0;
    for (int dz = zlo; dz <= zhi; ++dz) {
        uint32_t ky = ky0;
        for (int dy = -cc; dy <= cc; ++dy) {
            row_body(dy, dz, ky, kz);
            if (HASH == CMAG_HASH_MORTON) ky = dilated_inc(ky, My);
        }
        if (HASH == CMAG_HASH_MORTON) kz = dilated_inc(kz, Mz);
    }
}

// row j of a float4 array through a 32-bit byte offset (j < 2^28), which
// lets the load take the array base from scalar registers
__device__ __forceinline__ float4 row4(const float4* __restrict__ base, uint32_t j) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + (j << 4));
}

template <int DIM, int HASH, bool PERIODIC = true, typename Cells, typename C>
__device__ __forceinline__ void walk_stencil(const cmag_params& P, const float4& pi, const Cell& gp,
                                             const Cells& cells, C&& cand) {
    walk_rows<DIM, HASH, PERIODIC>(P, pi, gp, cells, [&](const Run& a, const Run& b, float py, float pz) {
        const uint32_t total = a.len + b.len;
        for (uint32_t t = 0; t < total; ++t) {
            const bool second = t >= a.len;
            const uint32_t j = second ? b.s + (t - a.len) : a.s + t;
            cand(j, second ? b.px : a.px, py, pz, second ? b.code : a.code);
        }
    });
}

// walk_stencil with the candidates' sorted positions loaded ahead:
// cand(j, pos[j], valid, qx, qy, qz, code).  The loop is unrolled by two; the
// load of candidate t + 1 is issued before candidate t is evaluated and the
// load for the next pass before candidate t + 1, so each wait covers loads
// that were issued a candidate earlier.  Loads and calls are unconditional
// (indices clamped to the run's last row, and the second call flagged
// valid = false when it lies past the run: the candidate must then contribute
// nothing), which keeps the loads out of divergent branches.
// the rows [s, e) of a sub-cell copy as one loop over a byte offset, two rows
// per iteration, each load issued one candidate ahead of its use; it reads up
// to two rows past the run (the copies carry padding rows)
template <typename C>
__device__ __forceinline__ void walk_run_sub(const float4* __restrict__ pos, uint32_t s, uint32_t e, float px,
                                             float py, float pz, int code, C&& cand) {
    const char* base = reinterpret_cast<const char*>(pos);
    auto ld = [&](uint32_t off) { return *reinterpret_cast<const float4*>(base + off); };
    const uint32_t end = e << 4, end1 = end - 16u;
    uint32_t off = s << 4;
    if (off >= end) return;
    float4 p0 = ld(off);
    do {
        const float4 p1 = ld(off + 16u);
        cand(off >> 4, p0, true, px, py, pz, code);
        p0 = ld(off + 32u);
        cand((off >> 4) + 1u, p1, off < end1, px, py, pz, code);
        off += 32u;
    } while (off < end);
}

template <int DIM, int HASH, bool PERIODIC = true, bool FORCE_ORDER = false, typename Cells, typename C>
__device__ __forceinline__ void walk_stencil_pf(const cmag_params& P, const float4& pi, const Cell& gp,
                                                const Cells& cells, const float4* __restrict__ pos, C&& cand) {
    if constexpr (Cells::SUB) {
        walk_rows<DIM, HASH, PERIODIC, FORCE_ORDER>(P, pi, gp, cells,
                                                    [&](const Run& a, const Run& b, float py, float pz) {
            walk_run_sub(pos, a.s, a.s + a.len, a.px, py, pz, a.code, cand);
            walk_run_sub(pos, b.s, b.s + b.len, b.px, py, pz, b.code, cand);
        });
        return;
    }
    walk_rows<DIM, HASH, PERIODIC>(P, pi, gp, cells, [&](const Run& a, const Run& b, float py, float pz) {
        const uint32_t total = a.len + b.len;
        if (total == 0) return;
        const uint32_t last = total - 1u;
        const uint32_t off_b = b.s - a.len;
        auto row = [&](uint32_t t) { return (t >= a.len ? off_b : a.s) + t; };
        auto visit = [&](uint32_t t, uint32_t j, const float4& pj, bool valid) {
            const bool second = t >= a.len;
            cand(j, pj, valid, second ? b.px : a.px, py, pz, second ? b.code : a.code);
        };
        uint32_t j0 = row(0);
        float4 p0 = row4(pos, j0);
        for (uint32_t t = 0; t < total; t += 2) {
            const uint32_t t1 = min(t + 1u, last);
            const uint32_t j1 = row(t1);
            const float4 p1 = row4(pos, j1);
            visit(t, j0, p0, true);
            j0 = row(min(t + 2u, last));
            p0 = row4(pos, j0);
            visit(t1, j1, p1, t + 1u < total);
        }
    });
}

// ---- cell lookups staged in LDS (sub-cell walk) -----------------------------
// A block whose targets lie in one row of cells along x, or in two rows (the
// end of one and the start of the next in key order), with each row's stencil
// x-range [cx0 - 1, cx1 + 1] inside the grid, on consecutive keys and at most
// LK_W cells wide (both rows together when there are two), copies the compact offsets
// and eighth counts of those cells in the 9 (2-D: 3) stencil rows of each of
// its rows into LDS once; its targets then find their runs there instead of
// with five loads per stencil row.  The runs and their order are the global
// walk's, so the sums are the same bit for bit.
constexpr int LK_W = 32;

struct LookupPlan {
    uint32_t key[2][9];  // key of cell cx0 - 1 in each stencil row of each part
    uint32_t width[2];   // cells per stencil row segment
    int cx0[2];
    int ry[2], rz[2];    // each part's row of cells (wrapped)
    int parts;           // 0: not staged, 1 or 2 rows
    int split;           // first target (block thread) of the second row
    int bad;             // a target outside the two rows
};

// part p (width w_p) keeps its stencil rows r at first[f_p + r * (w_p + 1) + c]
// and pre[q_p + r * w_p + c]: f_0 = q_0 = 0, part 1 after part 0 (f_1 =
// R (w_0 + 1), q_1 = R w_0); w_0 + w_1 <= LK_W
struct LookupTables {
    uint32_t first[18 * (LK_W / 2 + 1)];
    uint4 pre[9 * LK_W];
};

// a target's view of the staged tables: entries from its cell x - 1 on, per
// stencil row
struct LookupView {
    const uint32_t* first;
    const uint4* pre;
    uint32_t fs, ps;
    __device__ __forceinline__ uint32_t F(int r, uint32_t c) const { return first[(uint32_t)r * fs + c]; }
    __device__ __forceinline__ const uint4* Pr(int r, uint32_t c) const { return pre + (uint32_t)r * ps + c; }
};

template <int DIM>
__device__ __forceinline__ LookupView lookup_view(const cmag_params& P, const LookupPlan& lp, const LookupTables& lt,
                                                  const Cell& gp) {
    constexpr int R = DIM == 3 ? 9 : 3;
    const int wx = gp.x & ((int)P.grid_size[0] - 1), wy = gp.y & ((int)P.grid_size[1] - 1),
              wz = gp.z & ((int)P.grid_size[2] - 1);
    const int part = (lp.parts == 2 && wy == lp.ry[1] && wz == lp.rz[1]) ? 1 : 0;
    const uint32_t w0 = lp.width[0], w = part ? lp.width[1] : w0;
    const uint32_t dx = (uint32_t)(wx - lp.cx0[part]);
    LookupView v;
    v.fs = w + 1u;
    v.ps = w;
    v.first = lt.first + (part ? (uint32_t)R * (w0 + 1u) : 0u) + dx;
    v.pre = lt.pre + (part ? (uint32_t)R * w0 : 0u) + dx;
    return v;
}

// all threads of the block call this; returns whether the block's lookups are staged
template <int DIM>
__device__ __forceinline__ bool stage_lookups(const cmag_params& P, const float4* __restrict__ pos,
                                              const SubCells& cells, uint32_t n, uint32_t b0, LookupPlan& lp,
                                              LookupTables& lt) {
    constexpr int R = DIM == 3 ? 9 : 3;
    const int gx = (int)P.grid_size[0], gy = (int)P.grid_size[1], gz = (int)P.grid_size[2];
    const bool eligible = b0 < n && P.cellcount == 1;
    const uint32_t b1 = eligible ? min(b0 + blockDim.x, n) - 1u : 0u;
    auto wrapped = [&](uint32_t k) {
        const float4 p = pos[k];
        Cell c = grid_pos(P, p.x, p.y, p.z);
        c.x &= gx - 1;
        c.y &= gy - 1;
        c.z &= gz - 1;
        return c;
    };
    if (threadIdx.x == 0) {
        lp.split = 0x7fffffff;
        lp.bad = 0;
        if (eligible) {
            const Cell a = wrapped(b0);
            lp.ry[0] = a.y;
            lp.rz[0] = a.z;
        }
    }
    __syncthreads();
    const uint32_t i = b0 + threadIdx.x;
    Cell mine{0, 0, 0};
    if (eligible && i <= b1) {
        mine = wrapped(i);
        if (mine.y != lp.ry[0] || mine.z != lp.rz[0]) atomicMin(&lp.split, (int)threadIdx.x);
    }
    __syncthreads();
    if (threadIdx.x == 0 && eligible && lp.split != 0x7fffffff) {
        const Cell c = wrapped(b0 + (uint32_t)lp.split);
        lp.ry[1] = c.y;
        lp.rz[1] = c.z;
    }
    __syncthreads();
    if (eligible && i <= b1 && (int)threadIdx.x > lp.split && (mine.y != lp.ry[1] || mine.z != lp.rz[1]))
        lp.bad = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        int parts = 0;
        if (eligible && !lp.bad) {
            const int np = lp.split == 0x7fffffff ? 1 : 2;
            uint32_t wsum = 0;
            // each part's targets run from cell (ax) to cell (bx) of its row
            const uint32_t first_t[2] = {b0, np == 2 ? b0 + (uint32_t)lp.split : 0u};
            const uint32_t last_t[2] = {np == 2 ? b0 + (uint32_t)lp.split - 1u : b1, b1};
            bool ok = true;
            for (int pt = 0; pt < np && ok; ++pt) {
                const Cell a = wrapped(first_t[pt]), b = wrapped(last_t[pt]);
                const int ax = a.x, bx = b.x, ay = a.y, az = a.z;
                ok = ax >= 1 && bx >= ax && bx + 1 < gx && bx - ax + 3 <= LK_W;
                if (!ok) break;
                const uint32_t width = (uint32_t)(bx - ax) + 3u;
                wsum += width;
                ok = wsum <= (uint32_t)LK_W;
                if (!ok) break;
                ok = grid_hash(P, bx + 1, ay, az) - grid_hash(P, ax - 1, ay, az) == width - 1u;
                if (!ok) break;
                int r = 0;
                for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
                    for (int dy = -1; dy <= 1; ++dy, ++r) lp.key[pt][r] = grid_hash(P, ax - 1, ay + dy, az + dz);
                lp.width[pt] = width;
                lp.cx0[pt] = ax;
            }
            if (ok) parts = np;
        }
        lp.parts = parts;
    }
    __syncthreads();
    if (!lp.parts) return false;
    // part 0's entries first, then part 1's; entry t of first[] is t
    const uint32_t w0 = lp.width[0], w1 = lp.parts == 2 ? lp.width[1] : 0u;
    const uint32_t n0 = (uint32_t)R * (w0 + 1u), total = n0 + (lp.parts == 2 ? (uint32_t)R * (w1 + 1u) : 0u);
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
        const uint32_t pt = t >= n0 ? 1u : 0u, loc = pt ? t - n0 : t, w = pt ? w1 : w0;
        const uint32_t r = loc / (w + 1u), c = loc - r * (w + 1u);
        const uint32_t key = lp.key[pt][r];
        lt.first[t] = cells.first[key + c];
        if (c < w) lt.pre[(pt ? (uint32_t)R * w0 : 0u) + r * w + c] = cells.pre[key + c];
    }
    __syncthreads();
    return true;
}

// The density pass also stages candidate rows: after the lookups, the
// positions of the rows of four stencil rows of cells (r = 3..6: the three of
// dz = 0 and the first of dz = +1 in 3-D), up to ROW_STAGE_CAP rows, are copied
// into LDS as (x, y) pairs and z (12 B a row, 16 KB) and read from there.
// 16 KB keeps seven blocks per CU; more rows cost more in occupancy than they
// save in loads, and the force pass has no LDS to spare (its queues take 16 KB).
constexpr int ROW_STAGE_FIRST = 3, ROW_STAGE_ROWS = 4, ROW_STAGE_CAP = 1365;

struct RowStage {
    int32_t off[ROW_STAGE_ROWS];  // LDS row of sorted row j of staged stencil row r: j + off
    int ok;
};

// ROW_STAGE_CAP + 1 rows: the walk reads one row past a run
struct RowLds {
    float2 xy[ROW_STAGE_CAP + 1];
    float z[ROW_STAGE_CAP + 1];
};

// all threads of a staged block call this after stage_lookups
template <int DIM>
__device__ __forceinline__ void stage_rows(const LookupPlan& lp, const LookupTables& lt,
                                           const float4* __restrict__ pos, RowLds& lds, RowStage& rs) {
    // blocks staged over one row of cells only (its rows are at stride width + 1)
    const uint32_t FS = lp.width[0] + 1u;
    if (threadIdx.x == 0) {
        int32_t tot = 0;
        for (int k = 0; k < ROW_STAGE_ROWS; ++k) {
            const uint32_t r = (uint32_t)(ROW_STAGE_FIRST + k);
            const uint32_t s = lt.first[r * FS], e = lt.first[r * FS + lp.width[0]];
            rs.off[k] = tot - (int32_t)s;
            tot += (int32_t)(e - s);
        }
        rs.ok = DIM == 3 && lp.parts == 1 && tot <= ROW_STAGE_CAP;
    }
    __syncthreads();
    if (!rs.ok) return;
    for (int k = 0; k < ROW_STAGE_ROWS; ++k) {
        const uint32_t r = (uint32_t)(ROW_STAGE_FIRST + k);
        const uint32_t s = lt.first[r * FS], e = lt.first[r * FS + lp.width[0]];
        const int32_t d0 = (int32_t)s + rs.off[k];
        for (uint32_t t = threadIdx.x; t < e - s; t += blockDim.x) {
            const float4 v = pos[s + t];
            lds.xy[d0 + t] = make_float2(v.x, v.y);
            lds.z[d0 + t] = v.z;
        }
    }
    __syncthreads();
}

// the candidates of target pi (cell x wx, wrapped) through the staged
// lookups, as walk_stencil_pf hands them; with ROWS, the stencil rows that
// stage_rows copied are read from LDS; with FORCE_ORDER, rows in force_row order
template <int DIM, bool ROWS = false, bool FORCE_ORDER = false, typename C>
__device__ __forceinline__ void walk_lookup(const cmag_params& P, const float4& pi, const Cell& gp,
                                            const LookupPlan& lp, const LookupTables& lt,
                                            const float4* __restrict__ pos, C&& cand, const RowStage* rs = nullptr,
                                            const RowLds* lds = nullptr) {
    constexpr int R = DIM == 3 ? 9 : 3;
    const int q = x_sub(P, pi.x);
    const LookupView v = lookup_view<DIM>(P, lp, lt, gp);
    auto run_lds = [&](uint32_t s, uint32_t e, int32_t off) {
        for (uint32_t j = s; j < e; j += 2u) {
            const int32_t l = (int32_t)j + off;  // in [0, staged rows]
            const float2 a0 = lds->xy[l], a1 = lds->xy[l + 1];
            const float4 p0 = make_float4(a0.x, a0.y, lds->z[l], 0.0f);
            const float4 p1 = make_float4(a1.x, a1.y, lds->z[l + 1], 0.0f);
            cand(j, p0, true, pi.x, pi.y, pi.z, 13);
            cand(j + 1u, p1, j + 1u < e, pi.x, pi.y, pi.z, 13);
        }
    };
    for (int k = 0; k < R; ++k) {
        const int r = FORCE_ORDER ? force_row(DIM, k) : k;
        const uint32_t s = v.F(r, 0) + SubCells::below(v.Pr(r, 0), q);
        const uint32_t f1 = v.F(r, 2);
        const uint4* c1 = v.Pr(r, 2);
        const uint32_t e = f1 + SubCells::through(c1, q), ws = f1 + SubCells::fluid(c1), we = v.F(r, 3);
        if (ROWS && rs->ok && r >= ROW_STAGE_FIRST && r < ROW_STAGE_FIRST + ROW_STAGE_ROWS) {
            run_lds(s, e, rs->off[r - ROW_STAGE_FIRST]);
            run_lds(ws, we, rs->off[r - ROW_STAGE_FIRST]);
        } else {
            walk_run_sub(pos, s, e, pi.x, pi.y, pi.z, 13, cand);
            walk_run_sub(pos, ws, we, pi.x, pi.y, pi.z, 13, cand);
        }
    }
}

// ---- in-range masks (staged blocks of the sub-cell walk) ---------------------
// The density pass tests every candidate of a staged block anyway; it leaves,
// per target and stencil row r, a 64-bit mask of which fluid candidates of the
// row's main run [s, e) lie within 2h (r^2 < 4h^2, the force pass's own test
// on the same float operations; the walls in the run, which are those of its
// first two cells, are cleared).  The force pass then pushes the row's in-range
// fluid neighbours straight from the mask, in the walk's order, without loading
// or testing those candidates, and walks the walls apart.  Masks are stored
// row-major, [r][target] (uint2: low word first); a row whose run is longer
// than MASK_BITS is walked as before.
// Layout: the run is scanned in pairs (t, t + 1) and each pair shifts two bits
// in from the right, t's bit above t + 1's, so after the K = ceil(L / 2) pairs
// of a run of L candidates, candidate t sits at bit 2K - 1 - t (a pair past the
// run's end adds a zero bit).  Bit-reversed, t sits at 64 - 2K + t.
constexpr int MASK_BITS = 64;
constexpr int MASK_ROWS = 9;

// the mask bits of run candidates t in [lo, hi) for a run of K2 = 2K <= 64
// shifted-in candidates (t at bit K2 - 1 - t)
__device__ __forceinline__ uint64_t mask_bits(uint32_t K2, uint32_t lo, uint32_t hi) {
    if (hi <= lo) return 0ull;
    const uint32_t len = hi - lo;
    const uint64_t ones = len >= 64u ? ~0ull : ((1ull << len) - 1ull);
    return ones << (K2 - hi);
}

__device__ __forceinline__ uint64_t mask_shift_in(uint64_t m, float r2a, float r2b, float r2max) {
    const uint32_t b = (r2a < r2max ? 2u : 0u) | (r2b < r2max ? 1u : 0u);
    return (m << 2) + (uint64_t)b;
}

// density of the target pi over a staged block's lookups, as
// density_sum_lookup, writing the in-range masks of its rows' main runs to
// masks[r * stride + i] (fluid targets only)
template <int DIM>
__device__ __forceinline__ float density_sum_masks(const cmag_params& P, const float4& pi, const LookupPlan& lp,
                                                   const LookupTables& lt, const RowStage& rs, const RowLds* lds,
                                                   const float4* __restrict__ pos, uint2* __restrict__ masks,
                                                   uint32_t i, uint32_t stride) {
    constexpr int R = DIM == 3 ? 9 : 3;
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const int q = x_sub(P, pi.x);
    const LookupView v = lookup_view<DIM>(P, lp, lt, gp);
    const float half_inv_h = -0.5f / P.smoothing_radius;
    const float r2max = 4.0f * P.smoothing_radius * P.smoothing_radius;
    // the force pass reads the masks of fluid targets only
    const bool fluid = pi.w == CMAG_DB_FLUID;
    float sum = 0.0f;
    auto term = [&](float r2) {
        const float t = fminf(fmaxf(fmaf(half_inv_h, fsqrt(r2), 1.0f), 0.0f), 1.0f);
        const float t2 = t * t;
        sum = fmaf(t2 * t2, fmaf(-4.0f, t, 5.0f), sum);
    };
    auto d2 = [&](const float4& pj) { return norm2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z); };
    auto cand = [&](uint32_t, const float4& pj, bool valid, float, float, float, int) {
        term(valid ? d2(pj) : 1e30f);
    };
    for (int r = 0; r < R; ++r) {
        const uint4* c0 = v.Pr(r, 0);
        const uint32_t f0 = v.F(r, 0);
        const uint32_t s = f0 + SubCells::below(c0, q);
        const uint32_t f1 = v.F(r, 2);
        const uint4* c1 = v.Pr(r, 2);
        const uint32_t e = f1 + SubCells::through(c1, q), ws = f1 + SubCells::fluid(c1), we = v.F(r, 3);
        uint64_t m = 0;
        const bool in_lds = DIM == 3 && rs.ok && r >= ROW_STAGE_FIRST && r < ROW_STAGE_FIRST + ROW_STAGE_ROWS;
        if (in_lds) {
            const int32_t off = rs.off[r - ROW_STAGE_FIRST];
            for (uint32_t j = s; j < e; j += 2u) {
                const int32_t l = (int32_t)j + off;  // in [0, staged rows]
                const float2 a0 = lds->xy[l], a1 = lds->xy[l + 1];
                const float ra = d2(make_float4(a0.x, a0.y, lds->z[l], 0.0f));
                const float rb = j + 1u < e ? d2(make_float4(a1.x, a1.y, lds->z[l + 1], 0.0f)) : 1e30f;
                term(ra);
                term(rb);
                m = mask_shift_in(m, ra, rb, r2max);
            }
        } else {
            // two rows per iteration over a byte offset
            const char* base = reinterpret_cast<const char*>(pos);
            auto ld = [&](uint32_t off) { return *reinterpret_cast<const float4*>(base + off); };
            const uint32_t end = e << 4, end1 = end - 16u;
            for (uint32_t off = s << 4; off < end; off += 32u) {
                const float4 p0 = ld(off), p1 = ld(off + 16u);
                const float ra = d2(p0);
                const float rb = off < end1 ? d2(p1) : 1e30f;
                term(ra);
                term(rb);
                m = mask_shift_in(m, ra, rb, r2max);
            }
        }
        // the walls of the run's first two cells are walked by the force pass
        // (every wall of the stencil acts): their bits are cleared
        const uint32_t fx = v.F(r, 1), cxw = SubCells::fluid(v.Pr(r, 1));
        const uint32_t K2 = (e - s + 1u) & ~1u;
        if ((f1 - f0) - SubCells::fluid(c0) - cxw != 0u && K2 <= (uint32_t)MASK_BITS)
            m &= ~(mask_bits(K2, f0 + SubCells::fluid(c0) - s, fx - s) | mask_bits(K2, fx + cxw - s, f1 - s));
        if (fluid) masks[(size_t)r * stride + i] = make_uint2((uint32_t)m, (uint32_t)(m >> 32));
        if (in_lds) {
            const int32_t off = rs.off[r - ROW_STAGE_FIRST];
            for (uint32_t j = ws; j < we; j += 2u) {
                const int32_t l = (int32_t)j + off;
                const float2 a0 = lds->xy[l], a1 = lds->xy[l + 1];
                cand(j, make_float4(a0.x, a0.y, lds->z[l], 0.0f), true, 0.f, 0.f, 0.f, 13);
                cand(j + 1u, make_float4(a1.x, a1.y, lds->z[l + 1], 0.0f), j + 1u < we, 0.f, 0.f, 0.f, 13);
            }
        } else {
            walk_run_sub(pos, ws, we, pi.x, pi.y, pi.z, 13, cand);
        }
    }
    return sum * P.kernel_coeff * P.particle_mass;
}

// queue depth and drain size (compile-time; -D overrides for tuning builds)
#ifndef CMAG_NB_CAP
#define CMAG_NB_CAP 16
#endif
#ifndef CMAG_NB_DRAIN
#define CMAG_NB_DRAIN 2
#endif
constexpr int NB_CAP = CMAG_NB_CAP;
constexpr int NB_SHIFT_BITS = 27;  // entry = j | (image-shift code << 27)

struct NbEntry {
    __device__ static __forceinline__ uint32_t pack(uint32_t j, int code) {
        return j | ((uint32_t)code << NB_SHIFT_BITS);
    }
    __device__ static __forceinline__ uint32_t index(uint32_t e) { return e & ((1u << NB_SHIFT_BITS) - 1u); }
    __device__ static __forceinline__ int code(uint32_t e) { return (int)(e >> NB_SHIFT_BITS); }
};

// A per-lane queue of CAP (a power of two) neighbour entries in LDS, one
// [CAP][64] block per wave.  Entries leave in the order they came, so pair
// terms are summed in walk order whatever the queue depth and drain sizes.
// The walk pushes; when any lane's queue is full every active lane drains its
// NB_DRAIN oldest entries (fewer if it has fewer), which keeps all lanes busy
// in the pair loop instead of waiting for the fullest one; a final drain
// empties the queues.  push() has no room check: every site that pushes
// leaves each queue below CAP before the next site may push.
constexpr int NB_DRAIN = CMAG_NB_DRAIN;

template <int CAP>
struct NbQueue {
    static_assert((CAP & (CAP - 1)) == 0 && CAP >= 2, "queue depth: a power of two");
    uint32_t (*q)[64];
    int lane;
    int head = 0, count = 0;
    __device__ __forceinline__ NbQueue(uint32_t (*q_)[64], int lane_) : q(q_), lane(lane_) {}
    __device__ __forceinline__ void push(uint32_t e) {
        q[(head + count) & (CAP - 1)][lane] = e;
        ++count;
    }
    __device__ __forceinline__ bool full() const { return count == CAP; }
    // f(e0, e1, second): the oldest entries two at a time; a lane with an odd
    // number left repeats its last entry with second = false (contributing
    // nothing), so both entries' rows can be loaded before either is used
    template <typename F>
    __device__ __forceinline__ void drain(int upto, F&& f) {
        for (int k = 0; k < upto; k += 2) {
            if (k >= count) break;
            const uint32_t e0 = q[(head + k) & (CAP - 1)][lane];
            const uint32_t e1 = q[(head + min(k + 1, count - 1)) & (CAP - 1)][lane];
            f(e0, e1, k + 1 < count);
        }
        const int d = min(upto, count);
        head = (head + d) & (CAP - 1);
        count -= d;
    }
};

// ---------------------------------------------------------------------------
// the density of the particle at pi: the mass-weighted kernel sum over the
// stencil (the particle itself included)
template <int SYS, int DIM, int HASH, typename Cells>
__device__ __forceinline__ float density_sum(const cmag_params& P, const float4& pi,
                                             const float4* __restrict__ sortedPos, const Cells& cells) {
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float half_inv_h = -0.5f / P.smoothing_radius;
    float sum = 0.0f;
    walk_stencil_pf<DIM, HASH, SYS != CMAG_SYS_DAMBREAK>(P, pi, gp, cells, sortedPos,
                                                         [&](uint32_t, const float4& pj, bool valid, float px,
                                                             float py, float pz, int) {
        // a candidate past the run sits at r^2 = 1e30: t = 0 and it adds +0
        const float r2 = valid ? norm2(px - pj.x, py - pj.y, pz - pj.z) : 1e30f;
        // t = 1 - q/2 with q = r/h, clamped at 0 (nothing beyond 2h, no
        // branch); 1 - q/2 <= 1, so the clamp to [0, 1] is max(0, .) and folds
        // into the fma.  (1 - q/2)^4 (2q + 1) = t^4 (5 - 4t)
        const float t = fminf(fmaxf(fmaf(half_inv_h, fsqrt(r2), 1.0f), 0.0f), 1.0f);
        const float t2 = t * t;
        sum = fmaf(t2 * t2, fmaf(-4.0f, t, 5.0f), sum);
    });
    return sum * P.kernel_coeff * P.particle_mass;
}

template <int DIM>
__device__ __forceinline__ float density_sum_lookup(const cmag_params& P, const float4& pi, const LookupPlan& lp,
                                                    const LookupTables& lt, const RowStage& rs, const RowLds* lds,
                                                    const float4* __restrict__ pos) {
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float half_inv_h = -0.5f / P.smoothing_radius;
    float sum = 0.0f;
    walk_lookup<DIM, true>(P, pi, gp, lp, lt, pos,
                           [&](uint32_t, const float4& pj, bool valid, float px, float py, float pz, int) {
        const float r2 = valid ? norm2(px - pj.x, py - pj.y, pz - pj.z) : 1e30f;
        const float t = fminf(fmaxf(fmaf(half_inv_h, fsqrt(r2), 1.0f), 0.0f), 1.0f);
        const float t2 = t * t;
        sum = fmaf(t2 * t2, fmaf(-4.0f, t, 5.0f), sum);
    }, &rs, lds);
    return sum * P.kernel_coeff * P.particle_mass;
}

template <int SYS, int DIM, int HASH, typename Cells>
__device__ __forceinline__ void density_body(const cmag_params& P, uint32_t i, float4* __restrict__ measures,
                                             const float4* __restrict__ sortedPos, const Cells& cells) {
    const float4 pi = sortedPos[i];
    float2* m2 = reinterpret_cast<float2*>(measures);
    if (SYS == CMAG_SYS_POISEUILLE && pi.w == CMAG_PO_WALL) {
        m2[2 * (size_t)i] = make_float2(P.rest_density, eos(P, P.rest_density));
        return;
    }
    const float dens = density_sum<SYS, DIM, HASH>(P, pi, sortedPos, cells);
    if (SYS == CMAG_SYS_PERISTALSIS) {
        measures[i] = make_float4(dens, eos(P, dens), 0.0f, pi.w);
    } else {
        m2[2 * (size_t)i] = make_float2(dens, eos(P, dens));
    }
}

// getVelocityDiff, poiseuilleFlowKernel.cu:184-208
__device__ __forceinline__ float3 vel_diff_poiseuille(const cmag_params& P, const float4& vi,
                                                      const float4& pi, const float4& vj,
                                                      const float4& pj) {
    if (pj.w == CMAG_PO_WALL && pj.y > P.wall_top) {
        float beta = fminf(P.beta_max, 1.0f + (pj.y - P.wall_top) / (P.wall_top - pi.y));
        return make_float3(beta * vi.x, beta * vi.y, beta * vi.z);
    }
    if (pj.w == CMAG_PO_WALL && pj.y < P.wall_bottom) {
        float beta = fminf(P.beta_max, 1.0f + (P.wall_bottom - pj.y) / (pi.y - P.wall_bottom));
        return make_float3(beta * vi.x, beta * vi.y, beta * vi.z);
    }
    return make_float3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z);
}

// target position minus the periodic image offset named by code, computed with
// the same float operations as walk_stencil
template <bool PERIODIC = true>
__device__ __forceinline__ float3 image_target(const cmag_params& P, const float4& pi, int code) {
    if (!PERIODIC) return make_float3(pi.x, pi.y, pi.z);
    const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
    float3 q;
    q.x = (P.periodic_mask & 1) ? pi.x - (float)sx * P.world_size[0] : pi.x;
    q.y = (P.periodic_mask & 2) ? pi.y - (float)sy * P.world_size[1] : pi.y;
    q.z = (P.periodic_mask & 4) ? pi.z - (float)sz * P.world_size[2] : pi.z;
    return q;
}

// ---- fused force (DamBreak, Poiseuille) -------------------------------------------
// Per-target state of the force sum.  Pair terms accumulate in f in walk
// order; DamBreak's wall terms go to w as they are met, and f + w is formed
// once at the end, so the result does not depend on how the neighbours were
// staged (list depth, which particles share the wave, lists from the density
// pass).
template <int SYS>
struct ForceSum {
    const cmag_params& P;
    float4 pi, vi;
    float dens, inv_dens, pterm_i, inv_h, m, gc, av_k, av_eps, visc_k, t_k, w_k;
    float fx = 0.f, fy = 0.f, fz = 0.f, wx = 0.f, wy = 0.f, wz = 0.f;

    // rho, pterm: the target's density and p / rho^2
    __device__ __forceinline__ ForceSum(const cmag_params& P_, const float4& pi_, const float4& vi_, float rho,
                                        float pterm)
        : P(P_), pi(pi_), vi(vi_) {
        dens = rho;
        inv_dens = 1.0f / dens;
        pterm_i = pterm;
        const float h = P.smoothing_radius;
        inv_h = 1.0f / h;
        m = P.particle_mass;
        gc = P.grad_coeff;
        av_k = 2.0f * P.alpha * h * P.soundspeed;  // nu * (rho_i + rho_j)
        av_eps = 0.001f * h * h;
        visc_k = m * (P.mu + P.mu);
        t_k = -0.5f * inv_h;
        w_k = -5.0f * m * gc;
    }

    // Lennard-Jones wall term (DamBreak), fluid_kernel.cu:199-208
    __device__ __forceinline__ void wall(float rx, float ry, float rz, float r2) {
        const float inv_d = frsq(r2);
        const float ad = P.lj_a * inv_d;
        const float ad2 = ad * ad;
        const float ad6 = ad2 * ad2 * ad2;
        const float sc = P.lj_D * (ad6 * ad6 - ad6) * (inv_d * inv_d);
        wx = fmaf(sc, rx, wx); wy = fmaf(sc, ry, wy); wz = fmaf(sc, rz, wz);
    }

    // DamBreak pressure + artificial viscosity of an in-range fluid neighbour
    // (fluid_kernel.cu:210-262) with density rho_j and pterm_j = p_j / rho_j^2.
    // With t = 1 - q/2 the kernel gradient (1-q/2)^4 - (1-q/2)^3 (2q+1) is
    // 5 t^3 (t - 1); the -m gc factors fold into one constant.  The viscosity
    // term is branch-free (min(v.r, 0) makes it +0 for separating pairs), which
    // also keeps rho_j's load with the rest of its row.  One reciprocal square
    // root and one reciprocal per pair.  valid = false adds +0: the term is
    // scaled by 0 rather than skipped (a skipped term would move its row loads
    // into a branch); an invalid entry repeats a valid one, so its values are
    // finite.
    __device__ __forceinline__ void pair_db(const float3& q, const float4& pj, const float4& vj, float rho_j,
                                            float pterm_j, bool valid = true) {
        const float rx = q.x - pj.x, ry = q.y - pj.y, rz = q.z - pj.z;
        const float r2 = norm2(rx, ry, rz);
        const float inv_d = frsq(r2);
        const float t = fmaf(t_k, r2 * inv_d, 1.0f);
        const float t3 = t * t * t;
        const float vr = fmaf(vi.z - vj.z, rz, fmaf(vi.y - vj.y, ry, (vi.x - vj.x) * rx));
        const float av = -av_k * fminf(vr, 0.0f) * frcp((dens + rho_j) * (r2 + av_eps));
        const float s = (pterm_i + pterm_j) + av;
        const float kk = s * inv_d * t3 * fmaf(w_k, t, -w_k) * (valid ? 1.0f : 0.0f);
        fx = fmaf(kk, rx, fx); fy = fmaf(kk, ry, fy); fz = fmaf(kk, rz, fz);
    }

    // Poiseuille pressure + Morris viscosity of an in-range neighbour
    // (poiseuilleFlowKernel.cu:184-270); one reciprocal square root and one
    // reciprocal per pair; valid = false adds nothing
    __device__ __forceinline__ void pair(const float3& q, const float4& pj, const float4& mj, const float4& vj,
                                         bool valid = true) {
        const float rx = q.x - pj.x, ry = q.y - pj.y, rz = q.z - pj.z;
        const float r2 = norm2(rx, ry, rz);
        const float inv_d = frsq(r2);
        const float dist = r2 * inv_d;
        const float qq = dist * inv_h;
        const float t = fmaf(-0.5f, qq, 1.0f);
        const float t2 = t * t;
        const float temp = gc * fmaf(t2, t2, -(t2 * t) * fmaf(2.0f, qq, 1.0f));
        const float inv_dj = frcp(mj.x);
        const float s = fmaf(mj.y * inv_dj, inv_dj, pterm_i);
        const float3 V = vel_diff_poiseuille(P, vi, pi, vj, pj);
        const float kk = valid ? -m * s * inv_d * temp : 0.0f;
        const float kv = valid ? visc_k * inv_dens * inv_dj * inv_d * temp : 0.0f;
        fx += fmaf(kk, rx, kv * V.x);
        fy += fmaf(kk, ry, kv * V.y);
        fz += fmaf(kk, rz, kv * V.z);
    }
};

// ---- force passes: two phases per target ------------------------------------
// Phase 1 walks the stencil and only tests candidates (r^2 < 4h^2), appending
// the in-range ones to a per-lane list in LDS; phase 2 evaluates the pair
// terms of the listed neighbours.  Without the split, the ~50-instruction pair
// term would run for every candidate (some lane of the wave is almost always
// in range), i.e. ~475 times per particle instead of ~74.  The list is an
// NbQueue of CAP entries per lane (NB_CAP unless a caller picks another
// depth); pair terms take entries two at a time with all their row loads
// (four in the sub-cell walk, six otherwise) issued before either term.
// nbl: this wave's queues in LDS, [CAP][64]
// In the sub-cell walk (Cells::SUB, DamBreak) the row copies carry the
// density pass's results: a fluid row holds -rho in pos.w, sign bit set (a wall
// row keeps its type flag, 0..3, sign bit clear) and p / rho^2 in vel.w, so a
// pair term gathers two rows instead of three and measures is not read.
// With masks (staged blocks of the sub-cell walk), the in-range fluid
// neighbours of a row whose main run holds at most MASK_BITS candidates come
// from the density pass's mask instead of a walk, and their pair terms are
// evaluated at once, two at a time, after whatever earlier rows left in the
// queue; the walls of the row's first two cells are walked on their own.  Pair
// terms keep the walk's order and wall terms accumulate apart, so the sums are
// the same bit for bit.
template <int SYS, int DIM, int HASH, typename Cells, int CAP = NB_CAP>
__device__ __forceinline__ void force_body(const cmag_params& P, uint32_t i, float4* __restrict__ acc,
                                           const float4* __restrict__ measures,
                                           const float4* __restrict__ sortedPos,
                                           const float4* __restrict__ sortedVel,
                                           const uint32_t* __restrict__ index, const Cells& cells, uint32_t n,
                                           uint32_t (*nbl)[64], const LookupPlan* lp = nullptr,
                                           const LookupTables* lt = nullptr,
                                           const uint2* __restrict__ masks = nullptr, uint32_t stride = 0) {
    const int lane = threadIdx.x & 63;
    constexpr bool PACKED = SYS == CMAG_SYS_DAMBREAK && Cells::SUB;
    auto is_wall = [](float w) { return PACKED ? __float_as_int(w) >= 0 : w != CMAG_DB_FLUID; };
    const float4 pi = sortedPos[i];
    if (SYS == CMAG_SYS_DAMBREAK && is_wall(pi.w)) return;
    const float4 vi = sortedVel[i];
    float rho_i, pterm_i;
    if (PACKED) {
        rho_i = -pi.w;
        pterm_i = vi.w;
    } else {
        const float4 mi = measures[i];
        rho_i = mi.x;
        pterm_i = mi.y / (mi.x * mi.x);
    }
    ForceSum<SYS> F(P, pi, vi, rho_i, pterm_i);
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float r2max = 4.0f * P.smoothing_radius * P.smoothing_radius;
    NbQueue<CAP> nq(nbl, lane);
    auto pairs = [&](uint32_t e0, uint32_t e1, bool second) {
        const uint32_t j0 = NbEntry::index(e0), j1 = NbEntry::index(e1);
        const float4 p0 = row4(sortedPos, j0), v0 = row4(sortedVel, j0);
        const float4 p1 = row4(sortedPos, j1), v1 = row4(sortedVel, j1);
        if (SYS == CMAG_SYS_DAMBREAK) {
            const float3 q = make_float3(pi.x, pi.y, pi.z);
            if (PACKED) {
                F.pair_db(q, p0, v0, -p0.w, v0.w);
                F.pair_db(q, p1, v1, -p1.w, v1.w, second);
            } else {
                const float4 m0 = row4(measures, j0), m1 = row4(measures, j1);
                const float i0 = frcp(m0.x), i1 = frcp(m1.x);
                F.pair_db(q, p0, v0, m0.x, m0.y * i0 * i0);
                F.pair_db(q, p1, v1, m1.x, m1.y * i1 * i1, second);
            }
        } else {
            const float4 m0 = row4(measures, j0), m1 = row4(measures, j1);
            F.pair(image_target(P, pi, NbEntry::code(e0)), p0, m0, v0);
            F.pair(image_target(P, pi, NbEntry::code(e1)), p1, m1, v1, second);
        }
    };

    auto cand = [&](uint32_t j, const float4& pj, bool valid, float px, float py, float pz, int code) {
        const float rx = px - pj.x, ry = py - pj.y, rz = pz - pj.z;
        const float r2 = norm2(rx, ry, rz);
        if (valid && j != i) {
            if (SYS == CMAG_SYS_DAMBREAK && is_wall(pj.w)) {
                // Lennard-Jones wall term: every wall particle of the stencil acts
                F.wall(rx, ry, rz, r2);
            } else if (r2 < r2max) {
                nq.push(NbEntry::pack(j, code));
            }
        }
        if (__ballot(nq.full())) nq.drain(NB_DRAIN, pairs);
    };
    // a run of wall rows only: every one acts (Lennard-Jones), none is queued
    auto wall_only = [&](uint32_t, const float4& pj, bool valid, float px, float py, float pz, int) {
        if (valid) {
            const float rx = px - pj.x, ry = py - pj.y, rz = pz - pj.z;
            F.wall(rx, ry, rz, norm2(rx, ry, rz));
        }
    };
    if constexpr (Cells::SUB) {
        if (lp && masks) {
            constexpr int R = DIM == 3 ? 9 : 3;
            const int q = x_sub(P, pi.x);
            const LookupView v = lookup_view<DIM>(P, *lp, *lt, gp);
            // the main run [s, e) of row r: cells x - 1 (from eighth q) to
            // x + 1 (through eighth q)
            auto main_run = [&](int r, uint32_t& s, uint32_t& e) {
                s = v.F(r, 0) + SubCells::below(v.Pr(r, 0), q);
                e = v.F(r, 2) + SubCells::through(v.Pr(r, 2), q);
            };
            // the walls of a main run (those of its first two cells, in run
            // order) when the run itself is not walked: the mask holds fluid only
            auto run_walls = [&](int r) {
                const uint4* c0 = v.Pr(r, 0);
                const uint32_t f0 = v.F(r, 0), fx = v.F(r, 1), f1 = v.F(r, 2), cxw = SubCells::fluid(v.Pr(r, 1));
                if ((f1 - f0) - SubCells::fluid(c0) - cxw != 0u) {
                    walk_run_sub(sortedPos, f0 + SubCells::fluid(c0), fx, pi.x, pi.y, pi.z, 13, wall_only);
                    walk_run_sub(sortedPos, fx + cxw, f1, pi.x, pi.y, pi.z, 13, wall_only);
                }
            };
            // the last cell's walls [ws, we)
            auto last_walls = [&](int r) {
                walk_run_sub(sortedPos, v.F(r, 2) + SubCells::fluid(v.Pr(r, 2)), v.F(r, 3), pi.x, pi.y, pi.z, 13, cand);
            };
            // row r's mask, bit-reversed: candidate t of the run sits at bit
            // 64 - 2K + t, i.e. candidate j at bit j - s0; the target's own bit
            // is cleared
            auto row_mask = [&](const uint2 w, uint32_t s, uint32_t e, uint32_t& s0) {
                uint64_t m = __builtin_bitreverse64(((uint64_t)w.y << 32) | w.x);
                s0 = s - (64u - ((e - s + 1u) & ~1u));
                const uint32_t d = i - s0;
                return m & ~(d < 64u ? (1ull << d) : 0ull);
            };
            // pair terms straight from masks, two neighbours per iteration in
            // walk order (a missing second repeats the first and adds nothing),
            // without the queue: first mA's bits, then mB's
            auto mask_pairs = [&](uint64_t mA, uint32_t sA0, uint64_t mB, uint32_t sB0) {
                // row b's bits follow once row a's are used up
                auto next = [&]() {
                    if (mA == 0u) {
                        mA = mB;
                        sA0 = sB0;
                        mB = 0u;
                    }
                };
                next();
                while (mA != 0u) {
                    const uint32_t j0 = sA0 + (uint32_t)__builtin_ctzll(mA);
                    mA &= mA - 1u;
                    const bool two = mA != 0u;
                    const uint32_t j1 = two ? sA0 + (uint32_t)__builtin_ctzll(mA) : j0;
                    if (two) mA &= mA - 1u;
                    next();
                    pairs(NbEntry::pack(j0, 13), NbEntry::pack(j1, 13), two);
                }
            };
            // rows in force_row order, two at a time (the centre row alone)
            for (int k = 0; k < R; k += 2) {
                const int ra = force_row(DIM, k), rb = k + 1 < R ? force_row(DIM, k + 1) : -1;
                // both masks are loaded ahead of the walls (a walked row's is not used)
                const uint2 wa = masks[(size_t)ra * stride + i];
                const uint2 wb = rb >= 0 ? masks[(size_t)rb * stride + i] : make_uint2(0u, 0u);
                uint32_t s, e, sA0 = 0u, sB0 = 0u;
                uint64_t mA = 0u, mB = 0u;
                main_run(ra, s, e);
                const bool walk_a = e - s > (uint32_t)MASK_BITS;
                if (walk_a) {
                    walk_run_sub(sortedPos, s, e, pi.x, pi.y, pi.z, 13, cand);
                } else {
                    run_walls(ra);
                    mA = row_mask(wa, s, e, sA0);
                }
                last_walls(ra);
                if (rb >= 0) {
                    main_run(rb, s, e);
                    if (e - s > (uint32_t)MASK_BITS) {
                        // row a's pair terms come before the ones row b queues
                        if (!walk_a) {
                            if (nq.count != 0) nq.drain(CAP, pairs);
                            mask_pairs(mA, sA0, 0u, 0u);
                            mA = 0u;
                        }
                        walk_run_sub(sortedPos, s, e, pi.x, pi.y, pi.z, 13, cand);
                    } else {
                        run_walls(rb);
                        mB = row_mask(wb, s, e, sB0);
                    }
                    last_walls(rb);
                }
                if (__ballot((mA | mB) != 0u)) {
                    // neighbours queued by earlier walked rows come first
                    if (__ballot(nq.count != 0)) nq.drain(CAP, pairs);
                    mask_pairs(mA, sA0, mB, sB0);
                }
            }
        } else if (lp) {
            walk_lookup<DIM, false, true>(P, pi, gp, *lp, *lt, sortedPos, cand);
        } else {
            walk_stencil_pf<DIM, HASH, false, true>(P, pi, gp, cells, sortedPos, cand);
        }
    } else {
        walk_stencil_pf<DIM, HASH, SYS != CMAG_SYS_DAMBREAK>(P, pi, gp, cells, sortedPos, cand);
    }
    nq.drain(CAP, pairs);
    const uint32_t k = index[i];
    if (k < n) acc[k] = make_float4(F.fx + F.wx, F.fy + F.wy, F.fz + F.wz, 0.0f);
}

// ---- Peristalsis: pressure and viscous passes -------------------------------------
// The fused variant produces both from one neighbour walk (same per-pair
// arithmetic as the split kernels)
template <bool WANT_VISC, bool WANT_PRES>
struct PeriSum {
    const cmag_params& P;
    float4 pi, vi;
    float dens, inv_dens, pterm_i, inv_h, m, gc, visc_k;
    float vx = 0.f, vy = 0.f, vz = 0.f, px = 0.f, py = 0.f, pz = 0.f;
    __device__ __forceinline__ PeriSum(const cmag_params& P_, const float4& pi_, const float4& vi_, const float4& mi)
        : P(P_), pi(pi_), vi(vi_) {
        dens = mi.x;
        inv_dens = 1.0f / dens;
        pterm_i = mi.y / (dens * dens);
        inv_h = 1.0f / P.smoothing_radius;
        m = P.particle_mass;
        gc = P.grad_coeff;
        visc_k = m * (P.mu + P.mu);
    }
    // valid = false adds nothing
    __device__ __forceinline__ void pair(const float3& q, const float4& pj, const float4& mj, const float4& vj,
                                         bool valid = true) {
        const float rx = q.x - pj.x, ry = q.y - pj.y, rz = q.z - pj.z;
        const float r2 = norm2(rx, ry, rz);
        const float inv_d = frsq(r2);
        const float dist = r2 * inv_d;
        const float qq = dist * inv_h;
        const float t = fmaf(-0.5f, qq, 1.0f);
        const float t2 = t * t;
        const float temp = gc * fmaf(t2, t2, -(t2 * t) * fmaf(2.0f, qq, 1.0f));
        const float inv_dj = frcp(mj.x);
        if (WANT_PRES) {
            const float kk = valid ? -m * temp * fmaf(mj.y * inv_dj, inv_dj, pterm_i) * inv_d : 0.0f;
            px = fmaf(kk, rx, px); py = fmaf(kk, ry, py); pz = fmaf(kk, rz, pz);
        }
        if (WANT_VISC) {
            const float kv = valid ? visc_k * temp * inv_dens * inv_dj * inv_d : 0.0f;
            vx = fmaf(kv, vi.x - vj.x, vx); vy = fmaf(kv, vi.y - vj.y, vy); vz = fmaf(kv, vi.z - vj.z, vz);
        }
    }
};

template <int DIM, int HASH, bool WANT_VISC, bool WANT_PRES, typename Cells, int CAP = NB_CAP>
__device__ __forceinline__ void peristalsis_force_body(
    const cmag_params& P, uint32_t i, float4* __restrict__ viscF, float4* __restrict__ presF,
    const float4* __restrict__ measures, const float4* __restrict__ sortedPos,
    const float4* __restrict__ sortedVel, const uint32_t* __restrict__ index, const Cells& cells, uint32_t n,
    uint32_t (*nbl)[64]) {
    const int lane = threadIdx.x & 63;
    const float4 pi = sortedPos[i];
    PeriSum<WANT_VISC, WANT_PRES> F(P, pi, WANT_VISC ? sortedVel[i] : make_float4(0.f, 0.f, 0.f, 0.f),
                                    measures[i]);
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float r2max = 4.0f * P.smoothing_radius * P.smoothing_radius;
    NbQueue<CAP> nq(nbl, lane);
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    auto pairs = [&](uint32_t e0, uint32_t e1, bool second) {
        const uint32_t j0 = NbEntry::index(e0), j1 = NbEntry::index(e1);
        const float4 p0 = row4(sortedPos, j0), m0 = row4(measures, j0);
        const float4 v0 = WANT_VISC ? row4(sortedVel, j0) : zero4;
        const float4 p1 = row4(sortedPos, j1), m1 = row4(measures, j1);
        const float4 v1 = WANT_VISC ? row4(sortedVel, j1) : zero4;
        F.pair(image_target(P, pi, NbEntry::code(e0)), p0, m0, v0);
        F.pair(image_target(P, pi, NbEntry::code(e1)), p1, m1, v1, second);
    };

    walk_stencil_pf<DIM, HASH>(P, pi, gp, cells, sortedPos,
                               [&](uint32_t j, const float4& pj, bool valid, float qx, float qy, float qz, int code) {
        const float rx = qx - pj.x, ry = qy - pj.y, rz = qz - pj.z;
        const float r2 = norm2(rx, ry, rz);
        if (valid && j != i && r2 < r2max) nq.push(NbEntry::pack(j, code));
        if (__ballot(nq.full())) nq.drain(NB_DRAIN, pairs);
    });
    nq.drain(CAP, pairs);
    const uint32_t k = index[i];
    if (k >= n) return;
    if (WANT_VISC) viscF[k] = make_float4(F.vx, F.vy, F.vz, 0.0f);
    if (WANT_PRES) presF[k] = make_float4(F.px, F.py, F.pz, 0.0f);
}

}  // namespace cmag
