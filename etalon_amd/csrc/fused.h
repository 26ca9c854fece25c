// Kernel arguments of the fused, temporally blocked time step (see te_fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace etl {

constexpr int kMaxSources = 16;  // point-source nodes per plan (a bilinear source uses <= 4)

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
    int band;           // rows per tile
    int wo;             // output columns per tile (64 * V - 2 * deepest T)
    int n_bands, n_tj;  // tiles along x and along y
    int tiles_per_xcd;
    // CPML strips: index k of an axis is in the strip iff k < lo or k >= hi.
    int xh_lo, xh_hi, xe_lo, xe_hi, yh_lo, yh_hi, ye_lo, ye_hi;
    // Point sources: n_src nodes (comp, i, j, 0) on the device; delta table rows of
    // n_src REALs, one row per step.
    int n_src;
    const int4* src;
    const R* delta;
    // Per tile: bit 0 = monitored nodes in its rows, bit 1 = a source in its window.
    const unsigned char* tile_flags;
    // Monitor trace: per component a CSR over (row * n_tj + tile column) of entries
    // (column, slot); the trace holds one row of n_slots REALs per step of a chunk.
    int has_mon;
    const int* mon_ptr[3];
    const int2* mon_ent[3];
    R* trace;
    int n_slots;
};

// Output columns of a tile for vector width v and temporal depth t.
inline int tile_out_cols(int v, int t) { return 64 * v - 2 * t; }

template <typename R>
int launch_fused(const FusedArgs<R>& a, int vec, int depth, const int* counter, int kstep0,
                 int parity, hipStream_t stream);
template <typename R>
int launch_trace_dft(const R* trace, int n_slots, int rows, const R* ph, int nq,
                     const int* counter, const int4* pairs, int n_pairs, R* const* accs,
                     hipStream_t stream);
int launch_bump(int* counter, int by, hipStream_t stream);

}  // namespace etl
