// Kernel arguments of the fused, temporally blocked time step (see te_fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace etl {

constexpr int kMaxSources = 16;  // point-source nodes per simulation (bilinear: <= 4 each)

template <typename R>
struct FusedArgs {
    // f[k][p]: field k (ex, ey, hz, psi_ex_y, psi_ey_x, psi_hz_x, psi_hz_y), buffer p.
    const R* f[7][2];
    const R* cb_ex;
    const R* cb_ey;
    const R *bx_e, *cx_e, *ikx_e, *bx_h, *cx_h, *ikx_h;
    const R *by_e, *cy_e, *iky_e, *by_h, *cy_h, *iky_h;
    R db, inv_dx, inv_dy;
    int nx, ny;
    int wo;  // output columns per tile (64 * V - 2 * deepest T)
    // CPML strips: index k of an axis is in the strip iff k < lo or k >= hi.
    int xh_lo, xh_hi, xe_lo, xe_hi, yh_lo, yh_hi, ye_lo, ye_hi;
    // Point sources: n_src nodes (comp, i, j, 0) on the device; delta table rows of
    // n_src REALs, one row per step.
    int n_src;
    const int4* src;
    const R* delta;
    // Monitor trace: per component a CSR over tiles of entries (row, column, slot,
    // 0) of the tile's owned nodes; the trace holds one row of n_slots REALs per step
    // of a chunk.
    int has_mon;
    const int* tile_mon[3];
    const int4* tile_ent[3];
    R* trace;
    int n_slots;
};

// Strip of an axis: every index whose profile differs from (1, 0, 1) lies in
// [0, lo) or [hi, n).
template <typename R>
void strip_bounds(const std::vector<R>& b, const std::vector<R>& c, const std::vector<R>& ik,
                  int& lo, int& hi) {
    const int n = int(b.size());
    const int half = (n + 1) / 2;
    lo = 0;
    hi = n;
    for (int k = 0; k < n; ++k) {
        const bool nontrivial = (b[k] != R(1)) || (c[k] != R(0)) || (ik[k] != R(1));
        if (!nontrivial) continue;
        if (k < half)
            lo = k + 1;
        else if (hi == n)
            hi = k;
    }
}

// Output columns of a tile for vector width v and temporal depth t.
inline int tile_out_cols(int v, int t) { return 64 * v - 2 * t; }

// Tile descriptors (first row, tile column, extra-tile index, simulation) split by
// kernel variant: interior tiles carry no psi and are `band` rows high; strip tiles
// touch a CPML strip and carry psi; extra tiles also inject sources or trace
// monitors.  Strip and extra tiles are kSubBand rows high: they are few, so
// splitting them shortens the longest wave of a launch.  The simulation index
// selects the FusedArgs of the tile in the plan's argument array: one launch
// advances every simulation of a batch.
constexpr int kSubBand = 4;
struct TileList {
    const int4* interior;
    int n_interior;
    const int4* strip;  // the y-strip tiles, then the x-strip tiles, then the corners
    int n_strip;
    const int4* extra;  // extra tiles off the CPML strips (no psi)
    int n_extra;
    const int4* extra_strip;  // extra tiles that meet a strip (rare)
    int n_extra_strip;
    int n_strip_y;  // strip tiles meeting only y strips (columns)
    int n_strip_x;  // then those meeting only x strips (rows)
    int n_uniform;  // the last n_uniform interior tiles have one cb_ex and one cb_ey value
                    // over their window (the plan's uvals, two REALs each, in tile order)
};

// The tile list of descriptors stored in the order interior, strip (y, x, corner),
// extra, extra-strip.
inline TileList make_tile_list(const int4* t, int n_interior, int n_strip, int n_strip_y,
                               int n_strip_x, int n_extra, int n_extra_strip,
                               int n_uniform = 0) {
    return TileList{t,
                    n_interior,
                    t + n_interior,
                    n_strip,
                    t + n_interior + n_strip,
                    n_extra,
                    t + n_interior + n_strip + n_extra,
                    n_extra_strip,
                    n_strip_y,
                    n_strip_x,
                    n_uniform};
}

// Which strips a sub-tile's window (rows [r0 - T, r1 + T), columns [cs, cs + W))
// meets: bit 0 the y strips (columns outside [y_lo, y_hi)), bit 1 the x strips (rows
// outside [x_lo, x_hi), and the first row, where Ey stays fixed).  A strip sub-tile
// that meets neither (the sub-tiles of a band that does) is run as an x-strip tile.
inline int strip_kind(int r0, int r1, int cs, int T, int W, int x_lo, int x_hi, int y_lo,
                      int y_hi) {
    const bool rows = (r0 - T < (x_lo > 1 ? x_lo : 1)) || (r1 + T > x_hi);
    const bool cols = (cs < y_lo) || (cs + W > y_hi);
    return (cols ? 1 : 0) | (rows || !cols ? 2 : 0);
}

// Register-tile shapes compiled into the library: (precision, vec, depth, band rows).
// Every shape of depth T > 1 comes with its depth-1 sibling for step-count tails.
struct TileShape {
    int precision, vec, depth, band;
};
// Depths 6 and 8 (vec 1) are compiled for the one-launch joint kernels only (8: float
// forward plans with 16-row tiles).
constexpr TileShape kTileShapes[] = {
    {0, 1, 1, 8},  {0, 1, 2, 8},  {0, 1, 4, 8},  {0, 1, 6, 8},  {0, 1, 1, 16}, {0, 1, 2, 16},
    {0, 1, 4, 16}, {0, 1, 6, 16}, {0, 2, 1, 8},  {0, 2, 2, 8},  {0, 2, 4, 8},  {1, 1, 1, 8},
    {1, 1, 2, 8},  {1, 1, 4, 8},  {1, 1, 6, 8},  {1, 1, 1, 16}, {1, 1, 2, 16},
    {1, 1, 4, 16}, {1, 1, 6, 16}, {0, 1, 8, 16},
};
inline bool tile_shape_supported(int precision, int vec, int depth, int band) {
    for (const TileShape& s : kTileShapes)
        if (s.precision == precision && s.vec == vec && s.depth == depth && s.band == band)
            return true;
    return false;
}

// Whether plans over these grids count as large: at least 8 tiles of 16 x 56 per CU
// of the device.  Large plans of one launch per step block default to temporal depth
// 6 (a launch moves about the same bytes as at depth 4 for 6 steps instead of 4) and,
// in float, 16-row tiles; smaller ones to 8-row tiles and depth 4, which fill the
// device with more, shorter waves.  Float adjoint plans use the same rule, so that
// the two plans of a co-plan over the same grids agree on the depth.
template <typename Desc>
inline bool large_grids(const Desc* descs, int n_sims, int cu_count) {
    int64_t tiles = 0;
    for (int k = 0; k < n_sims; ++k)
        tiles += int64_t((descs[k].nx + 15) / 16) * ((descs[k].ny + 55) / 56);
    return tiles >= int64_t(8) * (cu_count > 0 ? cu_count : 1);
}

// Depth of the next launch with r >= 1 steps left at plan depth T: T, or for a tail
// of fewer than T steps the largest compiled smaller depth that fits (6, 4, 2, then 1)
// — a launch moves the bytes of a whole window whatever its depth, so a tail of
// depth-1 launches would cost nearly a full block per step.  Forward and adjoint
// plans of one depth decompose a step count alike, so a co-plan pairs their tails.
inline int next_depth(int r, int T, int precision, int vec, int band) {
    if (r >= T) return T;
    for (int d : {6, 4, 2})
        if (d < T && d <= r && tile_shape_supported(precision, vec, d, band)) return d;
    return 1;
}

// Launch one fused step of depth T over all tiles: the interior variant on s_int, the
// strip variant on s_strip and the extra variant on s_extra (the caller orders the
// streams).  Interior and strip tiles run `waves` waves (tiles) per workgroup; extra
// tiles one (they synchronise the workgroup).
template <typename R>
int launch_fused(const FusedArgs<R>* args, const TileList& tl, int vec, int depth, int band,
                 int waves, const int* counter, int kstep0, int parity, hipStream_t s_int,
                 hipStream_t s_strip, hipStream_t s_extra);

// Monitor trace of one simulation, for the DFT fold.
template <typename R>
struct TraceSim {
    const R* trace;  // (rows, n_slots)
    const R* ph;     // (table rows, nq) complex
    int n_slots, nq;
};
// pairs[p] = (slot, phasor column, accumulator, complex element), pair_sim[p] its
// simulation.
template <typename R>
int launch_trace_dft(const TraceSim<R>* sims, int rows, int* counter, const int4* pairs,
                     const int* pair_sim, int n_pairs, R* const* accs, hipStream_t stream);
int launch_bump(int* counter, int by, hipStream_t stream);
// A time-domain probe's copy: its n trace slots from slot_base of a simulation's trace
// (n_slots REALs per row) to rows of `out` (n REALs per row).
struct ProbeCopy {
    const void* trace;
    void* out;
    int n_slots, slot_base, n, pad;
};
// For every probe d of tab[0 .. n_probes): out[(counter[0] - counter[2] + r) * n + k] =
// trace[r * n_slots + slot_base + k], r < rows — one launch for all of them (max_n: the
// largest n).
template <typename R>
int launch_probe_copies(const ProbeCopy* tab, int n_probes, int max_n, int rows,
                        const int* counter, hipStream_t stream);

}  // namespace etl
