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

// Tiles of a step block by kind: 0 interior (8 or 16 rows), then kSubBand-row
// sub-tiles: 1 y-strip, 2 x-strip, 3 corner strip, 4 extra, 5 extra-strip; 6 interior
// tiles of uniform material (forward plans: the last n_uniform interior tiles).
constexpr int kJointKinds = 7;

// A forward step block; `counter`, `kstep0`, `parity` as tile_kernel takes.
template <typename R>
struct FwdSet {
    const FusedArgs<R>* args;
    const int4* tiles[kJointKinds];
    int n[kJointKinds];
    const int* counter;
    int kstep0;
    int parity;
    int band;  // rows of an interior tile (8, or 16 for float forward-only launches)
    const void* uvals;  // kind 6: cb_ex, cb_ey of each tile (REAL pairs)
};

// An adjoint step block, the same kinds; `step`, `rtop`, `parity` as adj_tile_kernel
// takes.
template <typename R>
struct AdjSet {
    const AdjTileArgs<R>* args;
    const int4* tiles[kJointKinds];
    int n[kJointKinds];
    const int* step;
    int rtop;
    int parity;
};

// Workgroups of a kind with n tiles: a multiple of 8, so that every range starts on
// XCD 0 and the XCD-major order of tile_body holds within it.
__host__ __device__ inline int joint_groups(int n) { return 8 * ((n + 7) / 8); }

// The kinds of a tile list: pointers and counts.
inline void split_kinds(const TileList& tl, const int4* (&t)[kJointKinds], int (&n)[kJointKinds]) {
    const int ny_ = tl.n_strip_y, nx_ = tl.n_strip_x, nxy = tl.n_strip - ny_ - nx_;
    t[0] = tl.interior;
    n[0] = tl.n_interior - tl.n_uniform;
    t[1] = tl.strip;
    n[1] = ny_;
    t[2] = tl.strip + ny_;
    n[2] = nx_;
    t[3] = tl.strip + ny_ + nx_;
    n[3] = nxy;
    t[4] = tl.extra;
    n[4] = tl.n_extra;
    t[5] = tl.extra_strip;
    n[5] = tl.n_extra_strip;
    t[6] = tl.interior + n[0];
    n[6] = tl.n_uniform;
}

template <typename R>
inline FwdSet<R> fwd_set(const FusedArgs<R>* args, const TileList& tl, const int* counter,
                         int kstep0, int parity, int band = 8, const void* uvals = nullptr) {
    FwdSet<R> f{};
    f.args = args;
    split_kinds(tl, f.tiles, f.n);
    f.counter = counter;
    f.kstep0 = kstep0;
    f.parity = parity;
    f.band = band;
    f.uvals = uvals;
    return f;
}

template <typename R>
inline AdjSet<R> adj_set(const AdjTileArgs<R>* args, const TileList& tl, const int* step,
                         int rtop, int parity) {
    AdjSet<R> a{};
    a.args = args;
    split_kinds(tl, a.tiles, a.n);
    a.step = step;
    a.rtop = rtop;
    a.parity = parity;
    return a;
}

// Depths launch_joint accepts for (forward, adjoint): (T, 0), (0, T), (T, T), T in 1, 2, 4,
// and 6 in float, and forward-only 8 (float, 16-row tiles; the dispatchers refuse what is
// not compiled).
inline bool joint_depths_supported(int tf, int ta) {
    auto ok = [](int t) { return t == 1 || t == 2 || t == 4 || t == 6; };
    return ((ok(tf) || tf == 8) && ta == 0) || (tf == 0 && ok(ta)) || (ok(tf) && tf == ta);
}

// Instantiations, compiled in separate files: forward only, adjoint only, both.
template <typename R>
int launch_joint_fwd(const FwdSet<R>* f, int tf, int part, hipStream_t s);
template <typename R>
int launch_joint_adj(const AdjSet<R>* a, int ta, int part, hipStream_t s);
template <typename R>
int launch_joint_pair(const FwdSet<R>* f, const AdjSet<R>* a, int t, int part, hipStream_t s);

// One step block: forward depth tf (0: no forward tiles), adjoint depth ta (0: none),
// the tile kinds of `part` (a JointPart).
template <typename R>
int launch_joint(const FwdSet<R>* f, int tf, const AdjSet<R>* a, int ta, hipStream_t s,
                 int part = 0) {
    if (!joint_depths_supported(tf, ta)) {
        set_error("joint step: no compiled depths (%d, %d)", tf, ta);
        return ETL_EARG;
    }
    if (ta == 0) return launch_joint_fwd<R>(f, tf, part, s);
    if (tf == 0) return launch_joint_adj<R>(a, ta, part, s);
    return launch_joint_pair<R>(f, a, tf, part, s);
}

// Which tile kinds a launch covers: every kind, the interior tiles only (light: few
// registers, bandwidth-bound), or all but the interior tiles (the CPML strips and the
// source/monitor tiles: heavy, latency-bound, few).
enum JointPart { kAllKinds = 0, kInteriorOnly = 1, kEdgeOnly = 2 };

__host__ __device__ constexpr bool part_has(int part, int kind) {
    const bool interior = kind == 0 || kind == 6;
    return part == kAllKinds || (part == kInteriorOnly ? interior : !interior);
}

// One kind K of a forward (adjoint) step block: when workgroup b falls in its range
// it runs its tile and returns true, else b moves past the range.
template <typename R, int T, int K, int BR, int S, bool EXTRA>
__device__ __forceinline__ bool fwd_kind(const FwdSet<R>& f, int& b, R* lds) {
    const int g = joint_groups(f.n[K]);
    if (b >= g) {
        b -= g;
        return false;
    }
    tile_body<R, 1, T, BR, S, EXTRA, K == 6>(f.args, f.tiles[K], f.n[K], b, f.n[K], 1, 0,
                                                 f.counter, f.kstep0, f.parity, lds,
                                                 static_cast<const R*>(f.uvals));
    return true;
}

template <typename R, int T, int K, int BR, int S, bool EXTRA>
__device__ __forceinline__ bool adj_kind(const AdjSet<R>& a, int& b) {
    const int g = joint_groups(a.n[K]);
    if (b >= g) {
        b -= g;
        return false;
    }
    adj_tile_body<R, 1, T, BR, S, EXTRA>(a.args, a.tiles[K], a.n[K], b, a.n[K], 1, 0, a.step,
                                             a.rtop, a.parity);
    return true;
}

// Workgroup b of a joint launch, in the order: the forward plan's strip, extra and
// extra-strip tiles, the adjoint plan's, then the forward and the adjoint interior
// tiles.  The latency-bound CPML and source/monitor tiles come first, so they are
// dispatched before the bulk of interior tiles fills the GPU and finish under it.
template <typename R, int TF, int TA, int PART, int JB>
__device__ __forceinline__ void joint_body(const FwdSet<R>& f, const AdjSet<R>& a) {
    __shared__ R lds[part_has(PART, 4) ? kSubBand * 64 : 1];  // the extra tiles' staging
    int b = blockIdx.x;
#ifndef ETL_JOINT_EDGE_FIRST
#define ETL_JOINT_EDGE_FIRST 1
#endif
    if constexpr (!ETL_JOINT_EDGE_FIRST && part_has(PART, 0)) {
        if constexpr (TF > 0) {
            if (fwd_kind<R, TF, 0, JB, 0, false>(f, b, lds)) return;
            if (fwd_kind<R, TF, 6, JB, 0, false>(f, b, lds)) return;
        }
        if constexpr (TA > 0) {
            if (adj_kind<R, TA, 0, 8, 0, false>(a, b)) return;
        }
    }
    if constexpr (part_has(PART, 1)) {
        if constexpr (TF > 0) {
            if (fwd_kind<R, TF, 1, kSubBand, 1, false>(f, b, lds) ||
                fwd_kind<R, TF, 2, kSubBand, 2, false>(f, b, lds) ||
                fwd_kind<R, TF, 3, kSubBand, 3, false>(f, b, lds) ||
                fwd_kind<R, TF, 4, kSubBand, 0, true>(f, b, lds) ||
                fwd_kind<R, TF, 5, kSubBand, 3, true>(f, b, lds))
                return;
        }
        if constexpr (TA > 0) {
            if (adj_kind<R, TA, 1, kSubBand, 1, false>(a, b) ||
                adj_kind<R, TA, 2, kSubBand, 2, false>(a, b) ||
                adj_kind<R, TA, 3, kSubBand, 3, false>(a, b) ||
                adj_kind<R, TA, 4, kSubBand, 0, true>(a, b) ||
                adj_kind<R, TA, 5, kSubBand, 3, true>(a, b))
                return;
        }
    }
    if constexpr (ETL_JOINT_EDGE_FIRST && part_has(PART, 0)) {
        if constexpr (TF > 0) {
            if (fwd_kind<R, TF, 0, JB, 0, false>(f, b, lds)) return;
            if (fwd_kind<R, TF, 6, JB, 0, false>(f, b, lds)) return;
        }
        if constexpr (TA > 0) adj_kind<R, TA, 0, 8, 0, false>(a, b);
    }
}

// kJointWaves, kJointWavesF64: the waves per SIMD the register allocation of a float
// (double) joint kernel must allow.  Its heaviest tiles (strip, source/monitor) need
// more registers than the interior tiles; capped, they spill a little, and the
// interior tiles, which are most of the work and bandwidth-bound, keep enough waves in
// flight.  Measured at 2024²: float 3 (the best of 2, 3, 4); double 2 (265 Gcell/s,
// against 217 uncapped — one wave per SIMD — and 179 at 3, where every tile spills).
#ifndef ETL_JOINT_WAVES
#define ETL_JOINT_WAVES 3
#endif
#ifndef ETL_JOINT_WAVES_F64
#define ETL_JOINT_WAVES_F64 2
#endif
constexpr int kJointWaves = ETL_JOINT_WAVES;
constexpr int kJointWavesF64 = ETL_JOINT_WAVES_F64;

template <typename R, int TF, int TA, int PART, int JB>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? kJointWaves : kJointWavesF64)))
void joint_kernel(FwdSet<R> f, AdjSet<R> a) {
    joint_body<R, TF, TA, PART, JB>(f, a);
}

template <typename R, int TF, int TA, int PART, int JB = 8>
int launch_joint_t(const FwdSet<R>* f, const AdjSet<R>* a, hipStream_t s) {
    int groups = 0;
    FwdSet<R> fz{};
    AdjSet<R> az{};
    for (int k = 0; k < kJointKinds; ++k) {
        if (!part_has(PART, k)) continue;
        if (TF > 0) groups += joint_groups(f->n[k]);
        if (TA > 0) groups += joint_groups(a->n[k]);
    }
    if (groups == 0) return ETL_OK;
    hipLaunchKernelGGL((joint_kernel<R, TF, TA, PART, JB>), dim3(unsigned(groups)), dim3(64), 0,
                       s, TF > 0 ? *f : fz, TA > 0 ? *a : az);
    ETL_LAUNCHED();
    return ETL_OK;
}

// A dispatch of one joint kernel variant over `groups` workgroups and no tiles: every
// workgroup finds no tile of its range and returns.  A queue's scratch is sized for a
// kernel's private segment (register spills) at the first dispatch that needs more than
// the queue holds, and that dispatch waits for the allocation; warm_scratch runs these
// at set-up so that the wait falls outside any timed or captured work.
template <typename R, int TF, int TA, int PART, int JB = 8>
int launch_joint_empty(int groups, hipStream_t s) {
    FwdSet<R> fz{};
    AdjSet<R> az{};
    hipLaunchKernelGGL((joint_kernel<R, TF, TA, PART, JB>), dim3(unsigned(groups)), dim3(64), 0,
                       s, fz, az);
    ETL_LAUNCHED();
    return ETL_OK;
}

// Empty dispatches (launch_joint_empty), on stream s, of the variants with the largest
// private segments: the double forward kernel of depth 6 (joint_fwd.hip, 580 bytes per
// lane) and the float forward + adjoint pair of depth 6 (joint_pair_f32.hip, 572), over
// enough workgroups to fill every CU.
// An empty one-workgroup dispatch of the forward variant of depth tf (every tile kind,
// interior tiles of `band` rows): a plan runs these at creation so that no kernel is
// dispatched for the first time inside its steps.
template <typename R>
int launch_joint_fwd_empty(int tf, int band, hipStream_t s);

constexpr int kWarmGroups = 8192;  // 32 one-wave workgroups per CU of 256
int warm_scratch_fwd(hipStream_t s);
int warm_scratch_pair(hipStream_t s);

// launch_joint_t at a run-time part.
template <typename R, int TF, int TA, int JB = 8>
int launch_joint_part(const FwdSet<R>* f, const AdjSet<R>* a, int part, hipStream_t s) {
    switch (part) {
        case kAllKinds: return launch_joint_t<R, TF, TA, kAllKinds, JB>(f, a, s);
        case kInteriorOnly: return launch_joint_t<R, TF, TA, kInteriorOnly, JB>(f, a, s);
        case kEdgeOnly: return launch_joint_t<R, TF, TA, kEdgeOnly, JB>(f, a, s);
    }
    set_error("joint step: unknown part %d", part);
    return ETL_EARG;
}

}  // namespace etl
