// One kernel launch per step block: every tile kind of a forward plan (interior,
// strip, extra, extra-strip tiles; tile_body in te_tile.h) and/or of an adjoint plan
// (adj_tile_body in adj_tile.h) in a single grid, one wave per workgroup, the kinds
// in consecutive workgroup ranges.  Against one launch per kind on three streams
// joined by events, this drops the fork/join latency that dominates small grids,
// where each launch is bound by one wave's run through its tile, and lets a forward
// recompute and the reverse steps of another segment share one launch.  No
// reference counterpart (the reference steps with one launch per half-step).
#pragma once

#include "adj_tile.h"
#include "fused.h"

namespace etl {

// Tiles of a forward step block: kinds 0 interior (8 rows), 1 strip, 2 extra,
// 3 extra-strip (kSubBand rows); `counter`, `kstep0`, `parity` as tile_kernel takes.
template <typename R>
struct FwdSet {
    const FusedArgs<R>* args;
    const int4* tiles[4];
    int n[4];
    const int* counter;
    int kstep0;
    int parity;
};

// Tiles of an adjoint step block, the same kinds; `step`, `rtop`, `parity` as
// adj_tile_kernel takes.
template <typename R>
struct AdjSet {
    const AdjTileArgs<R>* args;
    const int4* tiles[4];
    int n[4];
    const int* step;
    int rtop;
    int parity;
};

// Plans of at most this many tiles (summed over their simulations, at the full band)
// default to one launch per step block.
constexpr int64_t kSingleLaunchTiles = 4096;

// Workgroups of a kind with n tiles: a multiple of 8, so that every range starts on
// XCD 0 and the XCD-major order of tile_body holds within it.
__host__ __device__ inline int joint_groups(int n) { return 8 * ((n + 7) / 8); }

template <typename R>
inline FwdSet<R> fwd_set(const FusedArgs<R>* args, const TileList& tl, const int* counter,
                         int kstep0, int parity) {
    return FwdSet<R>{args,
                     {tl.interior, tl.strip, tl.extra, tl.extra_strip},
                     {tl.n_interior, tl.n_strip, tl.n_extra, tl.n_extra_strip},
                     counter,
                     kstep0,
                     parity};
}

template <typename R>
inline AdjSet<R> adj_set(const AdjTileArgs<R>* args, const TileList& tl, const int* step,
                         int rtop, int parity) {
    return AdjSet<R>{args,
                     {tl.interior, tl.strip, tl.extra, tl.extra_strip},
                     {tl.n_interior, tl.n_strip, tl.n_extra, tl.n_extra_strip},
                     step,
                     rtop,
                     parity};
}

// Depths launch_joint accepts for (forward, adjoint): (T, 0), (0, T), (T, T), T in 1, 2, 4.
inline bool joint_depths_supported(int tf, int ta) {
    auto ok = [](int t) { return t == 1 || t == 2 || t == 4; };
    return (ok(tf) && ta == 0) || (tf == 0 && ok(ta)) || (ok(tf) && tf == ta);
}

// Instantiations, compiled in separate files: forward only, adjoint only, both.
template <typename R>
int launch_joint_fwd(const FwdSet<R>* f, int tf, hipStream_t s);
template <typename R>
int launch_joint_adj(const AdjSet<R>* a, int ta, hipStream_t s);
template <typename R>
int launch_joint_pair(const FwdSet<R>* f, const AdjSet<R>* a, int t, hipStream_t s);

// One step block: forward depth tf (0: no forward tiles), adjoint depth ta (0: none).
template <typename R>
int launch_joint(const FwdSet<R>* f, int tf, const AdjSet<R>* a, int ta, hipStream_t s) {
    if (!joint_depths_supported(tf, ta)) {
        set_error("joint step: no compiled depths (%d, %d)", tf, ta);
        return ETL_EARG;
    }
    if (ta == 0) return launch_joint_fwd<R>(f, tf, s);
    if (tf == 0) return launch_joint_adj<R>(a, ta, s);
    return launch_joint_pair<R>(f, a, tf, s);
}

template <typename R, int TF, int TA>
__global__ __launch_bounds__(64) void joint_kernel(FwdSet<R> f, AdjSet<R> a) {
    __shared__ R lds[kSubBand * 64];
    int b = blockIdx.x;
    if constexpr (TF > 0) {
        int g = joint_groups(f.n[0]);
        if (b < g) {
            tile_body<R, 1, TF, 8, false, false>(f.args, f.tiles[0], f.n[0], b, f.n[0], 1, 0,
                                                 f.counter, f.kstep0, f.parity, lds);
            return;
        }
        b -= g;
        g = joint_groups(f.n[1]);
        if (b < g) {
            tile_body<R, 1, TF, kSubBand, true, false>(f.args, f.tiles[1], f.n[1], b, f.n[1], 1,
                                                       0, f.counter, f.kstep0, f.parity, lds);
            return;
        }
        b -= g;
        g = joint_groups(f.n[2]);
        if (b < g) {
            tile_body<R, 1, TF, kSubBand, false, true>(f.args, f.tiles[2], f.n[2], b, f.n[2], 1,
                                                       0, f.counter, f.kstep0, f.parity, lds);
            return;
        }
        b -= g;
        g = joint_groups(f.n[3]);
        if (b < g) {
            tile_body<R, 1, TF, kSubBand, true, true>(f.args, f.tiles[3], f.n[3], b, f.n[3], 1,
                                                      0, f.counter, f.kstep0, f.parity, lds);
            return;
        }
        b -= g;
    }
    if constexpr (TA > 0) {
        int g = joint_groups(a.n[0]);
        if (b < g) {
            adj_tile_body<R, 1, TA, 8, false, false>(a.args, a.tiles[0], a.n[0], b, a.n[0], 1, 0,
                                                     a.step, a.rtop, a.parity);
            return;
        }
        b -= g;
        g = joint_groups(a.n[1]);
        if (b < g) {
            adj_tile_body<R, 1, TA, kSubBand, true, false>(a.args, a.tiles[1], a.n[1], b, a.n[1],
                                                           1, 0, a.step, a.rtop, a.parity);
            return;
        }
        b -= g;
        g = joint_groups(a.n[2]);
        if (b < g) {
            adj_tile_body<R, 1, TA, kSubBand, false, true>(a.args, a.tiles[2], a.n[2], b, a.n[2],
                                                           1, 0, a.step, a.rtop, a.parity);
            return;
        }
        b -= g;
        g = joint_groups(a.n[3]);
        if (b < g)
            adj_tile_body<R, 1, TA, kSubBand, true, true>(a.args, a.tiles[3], a.n[3], b, a.n[3],
                                                          1, 0, a.step, a.rtop, a.parity);
    }
}

template <typename R, int TF, int TA>
int launch_joint_t(const FwdSet<R>* f, const AdjSet<R>* a, hipStream_t s) {
    int groups = 0;
    FwdSet<R> fz{};
    AdjSet<R> az{};
    for (int k = 0; k < 4; ++k) {
        if (TF > 0) groups += joint_groups(f->n[k]);
        if (TA > 0) groups += joint_groups(a->n[k]);
    }
    if (groups == 0) return ETL_OK;
    hipLaunchKernelGGL((joint_kernel<R, TF, TA>), dim3(unsigned(groups)), dim3(64), 0, s,
                       TF > 0 ? *f : fz, TA > 0 ? *a : az);
    ETL_LAUNCHED();
    return ETL_OK;
}

}  // namespace etl
