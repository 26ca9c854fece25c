// Fused, temporally blocked 2D TE time stepping for gfx950: the register-tile kernel.
//
// One launch advances the fields T full steps.  Per step n (etalon/fdtd/solver.py:
// 120-128): H update -> M injection (t = nΔt) -> E update -> J injection
// (t = (n+½)Δt); the monitored values of Hz (after M) and of Ex, Ey (after J) go to
// a trace buffer that a separate DFT kernel folds into the accumulators in step
// order.  Every node update performs the reference's operations in the reference's
// order (etalon/fdtd/backends/numpy_backend.py:41-61), so the results are bit-identical
// to T single steps whatever T or the tile shape.
//
// Decomposition.  One wave (64 lanes) owns an output tile of BR rows (x), from the
// row its descriptor names, by
// WO = 64*V - 2*T columns (y).  It loads a window of RW = BR + 2T rows by 64*V
// columns — lane l holds columns o0 - T + l*V .. + V-1 of every window row in
// registers (Ey one row more) — in a single burst of independent vector loads,
// performs the T steps entirely in registers, and stores its BR x WO output.  Row
// neighbours are registers of the same lane, column neighbours come from the
// adjacent lane (DPP wave shifts).  Values beyond the window are unknown: they corrupt
// one row and one column per half-step from each side, which the T-wide halo
// absorbs, so the owned cells are exact.  Physical boundaries need no halo: tangential
// E on the outer boundary (Ex at j = 0, ny; Ey at i = 0, nx) is never updated, so
// rows and columns outside the grid never reach inside it.  Reads come from buffer A
// and writes go to buffer B (ping-pong): tiles never see each other's writes.
//
// All register indices are compile-time constants (rows are unrolled), so the tile
// lives in VGPRs and the only waits are the one after the load burst.  Thousands of
// tiles per launch keep many loads in flight per SIMD.
//
// Variants.  Strip tiles (S != 0) touch a CPML strip: they load, carry and store the
// psi arrays of the strips they meet (S bit 0: the y strips, psi_hz_y and psi_ex_y;
// bit 1: the x strips, psi_hz_x and psi_ey_x) where the 1D profiles differ from
// (b, c, 1/kappa) = (1, 0, 1).  Elsewhere psi is +0 and the reference's term
// d*(1/kappa) + psi is d*1 + (+0), which equals d bit for bit unless d is -0; d is a
// difference of two field values times a positive constant, and a - b is -0 only
// when a is -0.  The plan checks at creation that no Ex, Ey or Hz value is -0, and
// the reference's updates never make one from values that are not (x - y and x + y
// are -0 only when x is; the psi terms they add are never -0), so the term is d.
// EXTRA tiles also inject sources and trace monitors.
#pragma once

#include "common.h"
#include "fused.h"

namespace etl {

// Rows of the window loaded ahead of the front that first needs them (time-skewed
// tiles), beyond the two fronts a row must lead by.
#ifndef ETL_LOAD_AHEAD
#define ETL_LOAD_AHEAD 3
#endif

// V consecutive REALs with the alignment of one REAL: Ex rows have an odd pitch
// (ny + 1) and gfx950 global loads/stores need only dword alignment.
template <typename R, int V>
using Vec = R __attribute__((ext_vector_type(V), aligned(sizeof(R))));

// Pointers the kernel reads from the argument array are generic; casting them to
// the global address space lets the compiler issue global (not flat) loads and
// stores, which do not tie up the LDS counter.
template <typename T>
using global_t = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ global_t<T>* to_global(T* p) {
    return (global_t<T>*)p;
}
// int4 descriptors read through global pointers (HIP's int4 class has no
// address-space-qualified copy constructor; the builtin vector type needs none).
using int4v = int __attribute__((ext_vector_type(4)));
__device__ __forceinline__ const global_t<const int4v>* to_global_v(const int4* p) {
    return (const global_t<const int4v>*)p;
}

// dst[k] = p[row * pitch + min(c, limit - 1) + k]: always one full vector load.
// Arrays come from etl_malloc, which keeps kGuard valid bytes before and after each
// one, and c >= -T, so the overhang past a row (or before the first) stays in
// memory; elements outside the row are garbage the caller never uses.
template <typename R, int V>
__device__ __forceinline__ void ldv(R (&dst)[V], const global_t<R>* __restrict__ base, int row,
                                    int pitch, int c, int limit) {
    const global_t<R>* p = base + int64_t(row) * pitch + min(c, limit - 1);
    if constexpr (V == 1) {
        dst[0] = p[0];
    } else {
        const Vec<R, V> x = *reinterpret_cast<const global_t<Vec<R, V>>*>(p);
#pragma unroll
        for (int k = 0; k < V; ++k) dst[k] = x[k];
    }
}

// p[c + k] = src[k] for the elements with lo <= c + k < hi.
template <typename R, int V>
__device__ __forceinline__ void store_row(global_t<R>* __restrict__ p, const R (&src)[V], int c,
                                          int lo, int hi) {
    if (c >= lo && c + V <= hi) {
        if constexpr (V == 1) {
            p[c] = src[0];
        } else {
            Vec<R, V> x;
#pragma unroll
            for (int k = 0; k < V; ++k) x[k] = src[k];
            *reinterpret_cast<global_t<Vec<R, V>>*>(p + c) = x;
        }
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k)
            if (c + k >= lo && c + k < hi) p[c + k] = src[k];
    }
}

template <typename R, int V>
__device__ __forceinline__ void store_masked(global_t<R>* __restrict__ p, const R (&src)[V], int c,
                                             const bool (&m)[V]) {
#pragma unroll
    for (int k = 0; k < V; ++k)
        if (m[k]) p[c + k] = src[k];
}

// Row accesses as a wave-uniform pointer `rowcs` to element cs of a row (the window's
// first column) plus a per-lane byte offset from it that is the same for every row.
// (Keeping the pointer whole in SGPRs, so that loads take the saddr form, was measured
// slower: the scalar 64-bit adds per row cost more than the vector adds they replace.)
template <typename T>
__device__ __forceinline__ const global_t<T>* at_byte(const global_t<T>* p, uint32_t b) {
    return reinterpret_cast<const global_t<T>*>(reinterpret_cast<const global_t<char>*>(p) + b);
}
template <typename T>
__device__ __forceinline__ global_t<T>* at_byte(global_t<T>* p, uint32_t b) {
    return reinterpret_cast<global_t<T>*>(reinterpret_cast<global_t<char>*>(p) + b);
}

// ldv with the row pointer and offset split: `boff` = (min(c, limit - 1) - cs) *
// sizeof(R) >= 0, so the load is the one ldv issues.
template <typename R, int V>
__device__ __forceinline__ void ldr(R (&dst)[V], const global_t<R>* rowcs, uint32_t boff) {
    const global_t<R>* p = at_byte(rowcs, boff);
    if constexpr (V == 1) {
        dst[0] = p[0];
    } else {
        const Vec<R, V> x = *reinterpret_cast<const global_t<Vec<R, V>>*>(p);
#pragma unroll
        for (int k = 0; k < V; ++k) dst[k] = x[k];
    }
}

// store_row with the row pointer and offset split: `soff` = (c - cs) * sizeof(R), the
// lane's first element; the predicates on c are store_row's.
template <typename R, int V>
__device__ __forceinline__ void str(global_t<R>* rowcs, uint32_t soff, const R (&src)[V], int c,
                                    int lo, int hi) {
    global_t<R>* p = at_byte(rowcs, soff);
    if (c >= lo && c + V <= hi) {
        if constexpr (V == 1) {
            p[0] = src[0];
        } else {
            Vec<R, V> x;
#pragma unroll
            for (int k = 0; k < V; ++k) x[k] = src[k];
            *reinterpret_cast<global_t<Vec<R, V>>*>(p) = x;
        }
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k)
            if (c + k >= lo && c + k < hi) p[k] = src[k];
    }
}

template <typename R, int V>
__device__ __forceinline__ void str_masked(global_t<R>* rowcs, uint32_t soff, const R (&src)[V],
                                           const bool (&m)[V]) {
    global_t<R>* p = at_byte(rowcs, soff);
#pragma unroll
    for (int k = 0; k < V; ++k)
        if (m[k]) p[k] = src[k];
}

// v of lane + 1 (wave_shl:1) and of lane - 1 (wave_shr:1): DPP moves inside the VALU,
// no LDS round trip.  Lanes 63 and 0 receive a value outside the window (0: bound_ctrl),
// which nothing owned depends on; with bound_ctrl the move folds into the VALU op that
// consumes it.
template <int CTRL, typename R>
__device__ __forceinline__ R lane_shift(R v) {
    if constexpr (sizeof(R) == 4) {
        return __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp(0, int(b & 0xffffffffLL), CTRL, 0xf, 0xf, true);
        const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xf, 0xf, true);
        return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
    }
}
template <typename R>
__device__ __forceinline__ R from_next_lane(R v) {
    return lane_shift<0x130>(v);
}
template <typename R>
__device__ __forceinline__ R from_prev_lane(R v) {
    return lane_shift<0x138>(v);
}

// Broadcast lane `l` of v to the whole wave (v_readlane; l is wave-uniform).
template <typename R>
__device__ __forceinline__ R lane_value(R v, int l) {
    if constexpr (sizeof(R) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffffLL), l);
        const int hi = __builtin_amdgcn_readlane(int(b >> 32), l);
        return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
    }
}

// Point sources of a tile: lane s holds source s (comp, row, column) and its
// deltas for the T steps of the launch; `rows[comp]` has bit l set when window row
// l carries a source of that component.
template <typename R, int T>
struct TileSources {
    int n = 0;
    int comp = 0, row = 0, col = 0;
    R delta[T];
    unsigned long long rows[3] = {0, 0, 0};
};

// Subtract this step's deltas of the sources of component `comp` on row i, in
// source order.  x - (+0) == x bit for bit, so the untouched elements are
// unchanged; the select keeps v[] in registers (no dynamic index).
template <typename R, int V, int T>
__device__ __forceinline__ void inject_row(const TileSources<R, T>& src, int t, int comp, int i,
                                           int c, R (&v)[V]) {
    for (int s = 0; s < src.n; ++s) {
        if (__builtin_amdgcn_readlane(src.comp, s) != comp ||
            __builtin_amdgcn_readlane(src.row, s) != i)
            continue;
        const int js = __builtin_amdgcn_readlane(src.col, s) - c;
        const R dl = lane_value(src.delta[t], s);
#pragma unroll
        for (int k = 0; k < V; ++k) v[k] = v[k] - ((k == js) ? dl : R(0));
    }
}

// Monitored values of one component's owned rows into the trace row: the rows go
// to LDS, then the lanes copy the tile's entries (row, column, slot).
template <typename R, int V, int RW, int T, int BR>
__device__ __forceinline__ void trace_tile(const global_t<const int4v>* __restrict__ ent, int e0,
                                           int e1, R* __restrict__ lds, const R (&f)[RW][V],
                                           int i0, int cs, global_t<R>* __restrict__ trow) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int l = 0; l < BR; ++l)
#pragma unroll
        for (int k = 0; k < V; ++k) lds[(l * 64 + lane) * V + k] = f[T + l][k];
    __syncthreads();
    for (int q = e0 + lane; q < e1; q += 64) {
        const int4v en = ent[q];  // row, column, slot
        trow[en.z] = lds[(en.x - i0) * 64 * V + (en.y - cs)];
    }
    __syncthreads();
}

// The tile program of one wave: tile li = ((b & 7) * per_xcd + (b >> 3)) * waves + wave
// of `tiles`, for workgroup b of n_groups.  Workgroups are dealt to the XCDs
// round-robin, so consecutive workgroup indices are spread XCD-major: each XCD
// receives a contiguous range of tiles, whose neighbours share its L2.  EXTRA tiles
// synchronise the workgroup (one wave per workgroup) and stage rows through `lds`
// (BR * 64 * V REALs).  UCB tiles (interior only) have one cb_ex and one cb_ey value
// over their whole window, uvals[2 li] and uvals[2 li + 1]: they load no cb row and
// multiply by those values, the same bits the rows hold.
template <typename R, int V, int T, int BR, int S, bool EXTRA, bool UCB = false>
__device__ __forceinline__ void tile_body(const FusedArgs<R>* __restrict__ args,
                                          const int4* __restrict__ tiles, int n_tiles, int b,
                                          int n_groups, int waves, int wave,
                                          const int* __restrict__ counter, int kstep0, int parity,
                                          R* __restrict__ lds,
                                          const R* __restrict__ uvals = nullptr) {
    static_assert(!UCB || (S == 0 && !EXTRA), "uniform-material tiles are interior tiles");
    constexpr int RW = BR + 2 * T;  // window rows (Ey: one more)
    // S: which CPML strips the window meets — bit 0 the y strips (columns), bit 1 the x
    // strips (rows, and the grid's first and last rows).  Only those psi arrays are
    // loaded, updated and stored, and only those boundary predicates evaluated.
    constexpr bool SY = (S & 1) != 0, SX = (S & 2) != 0;
    const int per_xcd = (n_groups + 7) >> 3;
    const int li = ((b & 7) * per_xcd + (b >> 3)) * waves + wave;
    if (li >= n_tiles) return;
    const int4 td = tiles[li];  // first row, tile column, extra-tile index, simulation
    // The simulation's arguments are read once, here, into registers: after the
    // kernel's first store the compiler could no longer assume them unchanged.
    const FusedArgs<R>& a = args[td.w];
    const int xh_lo = a.xh_lo, xh_hi = a.xh_hi, xe_lo = a.xe_lo, xe_hi = a.xe_hi;
    const int yh_lo = a.yh_lo, yh_hi = a.yh_hi, ye_lo = a.ye_lo, ye_hi = a.ye_hi;
    const int wo = a.wo;
    const int tj = td.y;
    const int lane = threadIdx.x & 63;
    const int nx = a.nx, ny = a.ny, px = ny + 1;
    const int o0 = tj * wo;  // first output column
    const int cs = o0 - T;     // first window column
    const int c = cs + lane * V;
    const int i0 = td.x;
    const int g0 = i0 - T;  // global row of window row 0

    using G = global_t<R>;
    const G* __restrict__ exA = to_global(a.f[0][parity]);
    const G* __restrict__ eyA = to_global(a.f[1][parity]);
    const G* __restrict__ hzA = to_global(a.f[2][parity]);
    const G* __restrict__ pexA = to_global(a.f[3][parity]);  // psi_ex_y
    const G* __restrict__ peyA = to_global(a.f[4][parity]);  // psi_ey_x
    const G* __restrict__ phxA = to_global(a.f[5][parity]);  // psi_hz_x
    const G* __restrict__ phyA = to_global(a.f[6][parity]);  // psi_hz_y
    G* __restrict__ exB = to_global(const_cast<R*>(a.f[0][parity ^ 1]));
    G* __restrict__ eyB = to_global(const_cast<R*>(a.f[1][parity ^ 1]));
    G* __restrict__ hzB = to_global(const_cast<R*>(a.f[2][parity ^ 1]));
    G* __restrict__ pexB = to_global(const_cast<R*>(a.f[3][parity ^ 1]));
    G* __restrict__ peyB = to_global(const_cast<R*>(a.f[4][parity ^ 1]));
    G* __restrict__ phxB = to_global(const_cast<R*>(a.f[5][parity ^ 1]));
    G* __restrict__ phyB = to_global(const_cast<R*>(a.f[6][parity ^ 1]));
    const G* __restrict__ cbx = to_global(a.cb_ex);
    const G* __restrict__ cby = to_global(a.cb_ey);

    // Per-lane byte offsets from element cs of a row, the same for every row: of a
    // clamped window element (loads; arrays of pitch px and of pitch ny) and of the
    // lane's own element (stores).
    const uint32_t offx = uint32_t(min(c, px - 1) - cs) * uint32_t(sizeof(R));
    const uint32_t offy = uint32_t(min(c, ny - 1) - cs) * uint32_t(sizeof(R));
    const uint32_t soff = uint32_t(lane * V) * uint32_t(sizeof(R));
    // element cs of row r of an array of pitch px / ny (wave-uniform)
    auto rx = [&](auto* base, int r) __attribute__((always_inline)) {
        return base + (int64_t(r) * px + cs);
    };
    auto ry = [&](auto* base, int r) __attribute__((always_inline)) {
        return base + (int64_t(r) * ny + cs);
    };

    // ---- loads: the window's rows (all up front for source/monitor tiles; in the
    // time-skewed order a few fronts ahead of their first use, so that rows stored at
    // the top and rows not yet loaded at the bottom hold no registers) ----
    constexpr int PY = SY ? RW : 1;  // psi rows held (none in interior tiles)
    constexpr int PX = SX ? RW : 1;
    R ex[RW][V], hz[RW][V], ey[RW + 1][V], cx[RW][V], cy[RW][V];
    R phy[PY][V], phx[PX][V], pey[PY][V], pex[PX][V];
    auto load_row = [&](int l) __attribute__((always_inline)) {
        if (l == RW) {  // Ey has one row more
            ldr<R, V>(ey[RW], ry(eyA, min(max(g0 + RW, 0), nx)), offy);
            return;
        }
        const int g = g0 + l;
        const int gx = min(max(g, 0), nx - 1);  // rows of Ex, Hz, cb_ex, psi_hz, psi_ex
        const int ge = min(max(g, 0), nx);      // rows of Ey, cb_ey, psi_ey
        ldr<R, V>(ex[l], rx(exA, gx), offx);
        ldr<R, V>(hz[l], ry(hzA, gx), offy);
        ldr<R, V>(ey[l], ry(eyA, ge), offy);
        if constexpr (!UCB) {
            ldr<R, V>(cx[l], rx(cbx, gx), offx);
            ldr<R, V>(cy[l], ry(cby, ge), offy);
        }
        if constexpr (SY) {
            ldr<R, V>(phy[SY ? l : 0], ry(phyA, gx), offy);
            ldr<R, V>(pey[SY ? l : 0], rx(pexA, gx), offx);
        }
        if constexpr (SX) {
            ldr<R, V>(phx[SX ? l : 0], ry(phxA, gx), offy);
            ldr<R, V>(pex[SX ? l : 0], ry(peyA, ge), offy);
        }
    };
    // rows loaded before front f of the skewed order: [0, f + 2 + kLoadAhead]
    constexpr int kLoadAhead = ETL_LOAD_AHEAD;
    constexpr int first_rows = EXTRA ? RW + 1 : (RW + 1 < 2 + kLoadAhead ? RW + 1 : 2 + kLoadAhead);
#pragma unroll
    for (int l = 0; l < first_rows; ++l) load_row(l);

    const R db = a.db, idx = a.inv_dx, idy = a.inv_dy;
    R ucx = R(0), ucy = R(0);  // the window's cb_ex and cb_ey (UCB tiles)
    if constexpr (UCB) {
        ucx = to_global(uvals)[2 * li];
        ucy = to_global(uvals)[2 * li + 1];
    }

    // y-strip membership and profiles of this lane's columns
    bool mH[V], ysh[V], yse[V];
    R byh[V], cyh[V], ikyh[V], bye[V], cye[V], ikye[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int j = c + k;
        mH[k] = (j >= 0 && j < ny);
        ysh[k] = SY && mH[k] && (j < yh_lo || j >= yh_hi);
        yse[k] = SY && j >= 1 && j <= ny - 1 && (j < ye_lo || j >= ye_hi);
        byh[k] = ysh[k] ? to_global(a.by_h)[j] : R(1);
        cyh[k] = ysh[k] ? to_global(a.cy_h)[j] : R(0);
        ikyh[k] = ysh[k] ? to_global(a.iky_h)[j] : R(1);
        bye[k] = yse[k] ? to_global(a.by_e)[j] : R(1);
        cye[k] = yse[k] ? to_global(a.cy_e)[j] : R(0);
        ikye[k] = yse[k] ? to_global(a.iky_e)[j] : R(1);
    }

    // Row profiles of the x strips: lane l holds those of window row l (RW <= 64),
    // read back with v_readlane where a row needs them.
    R pbxh = R(1), pcxh = R(0), pikxh = R(1), pbxe = R(1), pcxe = R(0), pikxe = R(1);
    if constexpr (SX) {
        const int gl = g0 + min(lane, RW - 1);
        const int gh = min(max(gl, 0), nx - 1), ge = min(max(gl, 0), nx);
        pbxh = to_global(a.bx_h)[gh];
        pcxh = to_global(a.cx_h)[gh];
        pikxh = to_global(a.ikx_h)[gh];
        pbxe = to_global(a.bx_e)[ge];
        pcxe = to_global(a.cx_e)[ge];
        pikxe = to_global(a.ikx_e)[ge];
    }

    // Sources and monitor entries of an extra tile.
    TileSources<R, T> src;
    int mon_lo[3] = {0, 0, 0}, mon_hi[3] = {0, 0, 0};
    G* trace_base = nullptr;
    int n_slots = 0;
    const global_t<const int4v>* ent[3] = {nullptr, nullptr, nullptr};
    if constexpr (EXTRA) {
        const int nbase = *counter + kstep0;  // table row of this launch's first step
        src.n = a.n_src;
        if (lane < a.n_src) {
            const int4v sd = to_global_v(a.src)[lane];
            src.comp = sd.x;
            src.row = sd.y;
            src.col = sd.z;
#pragma unroll
            for (int t = 0; t < T; ++t)
                src.delta[t] = to_global(a.delta)[int64_t(nbase + t) * a.n_src + lane];
        }
        for (int s = 0; s < src.n; ++s) {
            const int l = __builtin_amdgcn_readlane(src.row, s) - g0;
            if (l < 0 || l >= RW) continue;
            // constant indices keep src.rows in registers
            const int comp = __builtin_amdgcn_readlane(src.comp, s);
            const unsigned long long bit = 1ull << l;
            if (comp == 0) src.rows[0] |= bit;
            else if (comp == 1) src.rows[1] |= bit;
            else src.rows[2] |= bit;
        }
        if (a.has_mon) {
            trace_base = to_global(a.trace);
            n_slots = a.n_slots;
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) ent[comp] = to_global_v(a.tile_ent[comp]);
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) {
                if (a.tile_mon[comp] == nullptr) continue;
                mon_lo[comp] = to_global(a.tile_mon[comp])[td.z];
                mon_hi[comp] = to_global(a.tile_mon[comp])[td.z + 1];
            }
        }
    }
    const int lo = max(o0, 0), hi = o0 + wo;
    // The H half-step of step t on window row l (rows [t, RW - t)), in place.
    auto h_row = [&](int t, int l) __attribute__((always_inline)) {
        const int g = g0 + l;
        bool xh = false;
        R bx = R(1), cxp = R(0), ikx = R(1);
        if constexpr (SX) {
            xh = g < xh_lo || g >= xh_hi;
            if (xh) {
                bx = lane_value(pbxh, l);
                cxp = lane_value(pcxh, l);
                ikx = lane_value(pikxh, l);
            }
        }
        const R exr = from_next_lane(ex[l][0]);  // lane 63: outside the window
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const R exn = (k + 1 < V) ? ex[l][k + 1] : exr;
            const R d = (exn - ex[l][k]) * idy;
            const R e = (ey[l + 1][k] - ey[l][k]) * idx;
            R ty, tx;
            if (SY && ysh[k]) {
                const R py = byh[k] * phy[SY ? l : 0][k] + cyh[k] * d;
                phy[SY ? l : 0][k] = py;
                ty = d * ikyh[k] + py;
            } else {
                ty = d;
            }
            if (SX && xh) {
                const R pxv = bx * phx[SX ? l : 0][k] + cxp * e;
                phx[SX ? l : 0][k] = pxv;
                tx = e * ikx + pxv;
            } else {
                tx = e;
            }
            hz[l][k] = hz[l][k] + db * (ty - tx);
        }
        if constexpr (EXTRA) {
            if ((src.rows[2] >> l) & 1) inject_row<R, V, T>(src, t, 2, g, c, hz[l]);
        }
    };
    // The E half-step of step t on window row l: Ex rows [t, RW - t), Ey rows
    // [t + 1, RW - t), in place.
    auto e_row = [&](int t, int l) __attribute__((always_inline)) {
        const int g = g0 + l;
        const R hl = from_prev_lane(hz[l][V - 1]);  // lane 0: outside the window
        // Without SX the window lies inside rows [1, nx - 1], without SY inside
        // columns [1, ny - 1].
        const bool ex_row = !SX || (g >= 0 && g <= nx - 1);
        const bool ey_row = (l >= t + 1) && (!SX || (g >= 1 && g <= nx - 1));
        bool xe = false;
        R bxe = R(1), cxe = R(0), ikxe = R(1);
        if constexpr (SX) {
            xe = ey_row && (g < xe_lo || g >= xe_hi);
            if (xe) {
                bxe = lane_value(pbxe, l);
                cxe = lane_value(pcxe, l);
                ikxe = lane_value(pikxe, l);
            }
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int j = c + k;
            // Ex(g, j); j = 0, ny stay unchanged (a select, not a branch)
            if (ex_row) {
                const bool upd = !SY || (j >= 1 && j <= ny - 1);
                const R hleft = (k == 0) ? hl : hz[l][k - 1];
                const R d = (hz[l][k] - hleft) * idy;
                R tt;
                if (SY && yse[k]) {
                    const R p = bye[k] * pey[SY ? l : 0][k] + cye[k] * d;
                    pey[SY ? l : 0][k] = p;
                    tt = d * ikye[k] + p;
                } else {
                    tt = d;
                }
                const R exn = ex[l][k] + (UCB ? ucx : cx[l][k]) * tt;
                ex[l][k] = upd ? exn : ex[l][k];
            }
            if (ey_row) {  // Ey(g, j), g in [1, nx - 1]; rows 0 and nx stay unchanged
                const R d = (hz[l][k] - hz[l > 0 ? l - 1 : 0][k]) * idx;
                R tt;
                if (SX && xe) {
                    const R p = bxe * pex[SX ? l : 0][k] + cxe * d;
                    pex[SX ? l : 0][k] = p;
                    tt = d * ikxe + p;
                } else {
                    tt = d;
                }
                ey[l][k] = ey[l][k] - (UCB ? ucy : cy[l][k]) * tt;
            }
        }
        if constexpr (EXTRA) {
            if ((src.rows[0] >> l) & 1) inject_row<R, V, T>(src, t, 0, g, c, ex[l]);
            if ((src.rows[1] >> l) & 1) inject_row<R, V, T>(src, t, 1, g, c, ey[l]);
        }
    };
    // Store owned row l (window rows [T, T + BR), columns [o0, o0 + WO)).
    auto store_owned = [&](int l) __attribute__((always_inline)) {
        const int g = g0 + l;
        if (g >= nx) return;
        str<R, V>(ry(hzB, g), soff, hz[l], c, lo, min(hi, ny));
        str<R, V>(rx(exB, g), soff, ex[l], c, lo, min(hi, px));
        str<R, V>(ry(eyB, g), soff, ey[l], c, lo, min(hi, ny));
        if constexpr (SY) {
            bool moy[V], moe[V];
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const bool own = (c + k >= lo && c + k < hi);
                moy[k] = own && ysh[k];
                moe[k] = own && yse[k];
            }
            str_masked<R, V>(ry(phyB, g), soff, phy[SY ? l : 0], moy);
            str_masked<R, V>(rx(pexB, g), soff, pey[SY ? l : 0], moe);
        }
        if constexpr (SX) {
            bool mo[V];
#pragma unroll
            for (int k = 0; k < V; ++k) mo[k] = (c + k >= lo && c + k < hi) && mH[k];
            if (g < xh_lo || g >= xh_hi)
                str_masked<R, V>(ry(phxB, g), soff, phx[SX ? l : 0], mo);
            if (g >= 1 && (g < xe_lo || g >= xe_hi))
                str_masked<R, V>(ry(peyB, g), soff, pex[SX ? l : 0], mo);
        }
    };
    if constexpr (EXTRA) {
        // Step-major: the monitors' trace rows are written per step.
#pragma unroll
        for (int t = 0; t < T; ++t) {
#pragma unroll
            for (int l = t; l < RW - t; ++l) h_row(t, l);
#pragma unroll
            for (int l = t; l < RW - t; ++l) e_row(t, l);
            G* trow = trace_base + int64_t(kstep0 + t) * n_slots;
            if (mon_lo[2] < mon_hi[2])
                trace_tile<R, V, RW, T, BR>(ent[2], mon_lo[2], mon_hi[2], lds, hz, i0, cs, trow);
            if (mon_lo[0] < mon_hi[0])
                trace_tile<R, V, RW, T, BR>(ent[0], mon_lo[0], mon_hi[0], lds, ex, i0, cs, trow);
            if (mon_lo[1] < mon_hi[1])
                trace_tile<R, V, RW + 1, T, BR>(ent[1], mon_lo[1], mon_hi[1], lds, ey, i0, cs,
                                                trow);
        }
#pragma unroll
        for (int l = T; l < T + BR; ++l) store_owned(l);
    } else {
        // Time-skewed: at front f, step t advances window row f - t.  Row l + 1 is at
        // step t - 1 when step t reads its Ey, row l - 1 already at step t when the E
        // half-step reads its Hz, and row l's own values at step t - 1 — the values the
        // step-major order reads.  Front f needs window rows up to f + 1 only, so the
        // first rows compute while the later ones are still loading, and each owned row
        // is stored as soon as its last step is done.
#pragma unroll
        for (int f = 0; f < RW; ++f) {
            if (f + 2 + kLoadAhead <= RW) load_row(f + 2 + kLoadAhead);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int l = f - t;
                if (l < t || l >= RW - t) continue;
                h_row(t, l);
                e_row(t, l);
                if (t == T - 1 && l >= T && l < T + BR) store_owned(l);
            }
        }
    }

    // the last band copies Ey row nx (outer tangential E)
    if (i0 + BR >= nx) {
        R eyl[V];
        ldr<R, V>(eyl, ry(eyA, nx), offy);
        str<R, V>(ry(eyB, nx), soff, eyl, c, lo, min(hi, ny));
    }
}

// One wave per tile, `blockDim.x / 64` waves per workgroup (always one for EXTRA).
template <typename R, int V, int T, int BR, int S, bool EXTRA>
__global__ __launch_bounds__(256) void tile_kernel(const FusedArgs<R>* __restrict__ args,
                                                   const int4* __restrict__ tiles, int n_tiles,
                                                   const int* __restrict__ counter, int kstep0,
                                                   int parity) {
    __shared__ R lds[EXTRA ? BR * 64 * V : 1];
    const int waves = blockDim.x >> 6;
    // the wave index is uniform across the wave: readfirstlane puts it (and all that
    // derives from it — the descriptor, the simulation's arguments) in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    tile_body<R, V, T, BR, S, EXTRA>(args, tiles, n_tiles, blockIdx.x,
                                         (n_tiles + waves - 1) / waves, waves, wave, counter,
                                         kstep0, parity, lds);
}

}  // namespace etl
