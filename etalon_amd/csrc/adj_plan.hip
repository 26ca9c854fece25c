// Adjoint plans: the reverse steps of one or many simulations with the register-tile
// adjoint kernels (adj_tile.h), per T steps one launch of each tile kernel over the
// tiles of every simulation, replayed from a graph per segment length.  No reference
// counterpart (the reference has no adjoint).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "adj_tile.h"
#include "coplan.h"
#include "joint.h"
#include "common.h"

namespace etl {
namespace {

__global__ void set_adj_step_kernel(int* step, int value, int hist_row0) {
    if (threadIdx.x == 0) {
        step[0] = value;
        step[1] = hist_row0;
    }
}

}  // namespace
}  // namespace etl

using namespace etl;

struct EtlAdjPlan {
    int precision = 0;
    EtlTune tune{};
    int parity = 0;
    unsigned layout_gen = 0;  // bumped by every layout rebuild (graphs hold its pointers)
    int lane = 0;  // stream lane of its work
    struct Sim {
        EtlAdjointDesc desc{};
        int nx = 0, ny = 0;
        void* lamB[7] = {nullptr};
        int strips[8] = {0};
        // sources, sorted by (component, flat index); perm[k] = caller's column
        std::vector<int> comp;
        std::vector<int64_t> flat;
        std::vector<int> perm;
        void* src = nullptr;  // (src_rows, n_src) REAL, columns in sorted order
        int src_rows = 0;
        int* tile_runs = nullptr;
        int4* runs = nullptr;
    };
    std::vector<Sim> sims;
    bool dirty = true;
    void* args_dev = nullptr;
    int4* tiles_dev = nullptr;
    int n_interior = 0, n_strip = 0, n_extra = 0, n_extra_strip = 0;
    int n_strip_y = 0, n_strip_x = 0;  // the strip tiles: y only, then x only, then both
    int* step = nullptr;
    hipStream_t s_strip = nullptr, s_extra = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    int graph_count = 0, graph_parity = 0, graph_launches = 0;  // what the graph replays

    int elem() const { return precision == ETL_F32 ? 4 : 8; }
    int wo() const { return tile_out_cols(1, tune.temporal); }
};

namespace {

constexpr int kAdjBand = 8;

void free_adj_graph(EtlAdjPlan* p) {
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->graph_src) (void)hipGraphDestroy(p->graph_src);
    p->graph = nullptr;
    p->graph_src = nullptr;
    p->graph_count = 0;
}

void free_adj_layout(EtlAdjPlan* p) {
    for (auto& s : p->sims) {
        etl_free(s.tile_runs);
        etl_free(s.runs);
        s.tile_runs = nullptr;
        s.runs = nullptr;
    }
    etl_free(p->tiles_dev);
    p->tiles_dev = nullptr;
}

int upload_bytes(void** dst, const void* src, size_t bytes) {
    int rc = etl_malloc(dst, bytes);
    if (rc) return rc;
    return etl_memcpy_h2d(*dst, src, bytes);
}

// Tiles as in the forward plan: a full tile (8 rows x WO columns) is interior when
// its window meets no CPML strip, no adjoint source and it owns no design-box node;
// otherwise it splits into kSubBand-row sub-tiles, extra when its window holds a
// source or it owns box nodes, else strip.  Extra sub-tiles get the runs of sources
// in their window: one run per (component, row) and consecutive columns.
int build_adj_layout(EtlAdjPlan* p) {
    if (int rc_ = sync_lanes()) return rc_;
    ++p->layout_gen;
    free_adj_layout(p);
    free_adj_graph(p);
    const int T = p->tune.temporal, wo = p->wo(), W = 64;
    std::vector<int4> interior, strip_y, strip_x, strip, extra, extra_strip;
    for (int si = 0; si < int(p->sims.size()); ++si) {
        auto& sim = p->sims[si];
        const int nx = sim.nx, ny = sim.ny;
        const EtlAdjointDesc& d = sim.desc;
        const bool box = d.box_ni > 0 && d.box_nj > 0;
        struct Node {
            int comp, row, col, idx;
        };
        std::vector<Node> nodes;
        for (size_t k = 0; k < sim.comp.size(); ++k) {
            const int pitch = sim.comp[k] == 0 ? ny + 1 : ny;
            const int row = int(sim.flat[k] / pitch);
            nodes.push_back({sim.comp[k], row, int(sim.flat[k] - int64_t(row) * pitch), int(k)});
        }
        const int* st = sim.strips;
        const int x_lo = std::max(st[0], st[2]), x_hi = std::min(st[1], st[3]);
        const int y_lo = std::max(st[4], st[6]), y_hi = std::min(st[5], st[7]);
        auto sees_source = [&](int r0, int r1, int tj) {
            const int cs = tj * wo - T;
            for (const Node& n : nodes)
                if (n.row >= r0 - T && n.row <= r1 + T && n.col >= cs && n.col < cs + W) return true;
            return false;
        };
        auto owns_box = [&](int r0, int r1, int tj) {
            if (!box) return false;
            const int o0 = tj * wo;
            return r0 < d.box_i0 + d.box_ni && r1 > d.box_i0 && o0 < d.box_j0 + d.box_nj &&
                   o0 + wo > d.box_j0;
        };
        const int ntj = (ny + 1 + wo - 1) / wo;
        const int nb = (nx + kAdjBand - 1) / kAdjBand;
        std::vector<int4> sim_extra;
        auto sub_touches = [&](int r0, int r1, int tj) {
            const int cs = tj * wo - T;
            return (r0 - T < std::max(x_lo, 1)) || (r1 + T > x_hi) || (cs < y_lo) ||
                   (cs + W > y_hi);
        };
        for (int tb = 0; tb < nb; ++tb) {
            const int i0 = tb * kAdjBand, i1 = std::min(i0 + kAdjBand, nx);
            const bool rows_touch = (i0 - T < std::max(x_lo, 1)) || (i1 + T > x_hi);
            for (int tj = 0; tj < ntj; ++tj) {
                const int cs = tj * wo - T;
                const bool cols_touch = (cs < y_lo) || (cs + W > y_hi);
                if (!rows_touch && !cols_touch && !sees_source(i0, i1, tj) && !owns_box(i0, i1, tj)) {
                    interior.push_back(make_int4(i0, tj, 0, si));
                    continue;
                }
                for (int r0 = i0; r0 < i1; r0 += kSubBand) {
                    const int r1 = std::min(r0 + kSubBand, nx);
                    if (sees_source(r0, r1, tj) || owns_box(r0, r1, tj)) {
                        const int4 d = make_int4(r0, tj, int(sim_extra.size()), si);
                        sim_extra.push_back(d);
                        (sub_touches(r0, r1, tj) ? extra_strip : extra).push_back(d);
                    } else {  // by the strips the sub-tile meets; corners go to `strip`
                        const int kind =
                            strip_kind(r0, r1, tj * wo - T, T, W, x_lo, x_hi, y_lo, y_hi);
                        (kind == 1 ? strip_y : kind == 2 ? strip_x : strip)
                            .push_back(make_int4(r0, tj, 0, si));
                    }
                }
            }
        }
        // runs per extra sub-tile of this simulation
        const int n_ex = int(sim_extra.size());
        std::vector<int> ptr(n_ex + 1, 0);
        std::vector<int4> runs;
        for (int e = 0; e < n_ex; ++e) {
            const int4 x = sim_extra[e];
            const int r0 = x.x, cs = x.y * wo - T;
            // nodes are sorted by (comp, row, col): consecutive columns of one row with
            // consecutive indices form a run
            for (size_t k = 0; k < nodes.size();) {
                const Node& n = nodes[k];
                const bool in = n.row >= r0 - T && n.row <= r0 + kSubBand + T && n.col >= cs &&
                                n.col < cs + W;
                if (!in) {
                    ++k;
                    continue;
                }
                size_t m = k + 1;
                while (m < nodes.size() && nodes[m].comp == n.comp && nodes[m].row == n.row &&
                       nodes[m].col == nodes[m - 1].col + 1 && nodes[m].col < cs + W)
                    ++m;
                runs.push_back(make_int4(n.row | (n.comp << 28), n.col, nodes[m - 1].col + 1, n.idx));
                k = m;
            }
            ptr[e + 1] = int(runs.size());
        }
        int rc;
        if (!runs.empty()) {
            if ((rc = upload_bytes((void**)&sim.tile_runs, ptr.data(), ptr.size() * sizeof(int))) ||
                (rc = upload_bytes((void**)&sim.runs, runs.data(), runs.size() * sizeof(int4))))
                return rc;
        }
    }
    std::vector<int4> tiles(interior);
    tiles.insert(tiles.end(), strip_y.begin(), strip_y.end());
    tiles.insert(tiles.end(), strip_x.begin(), strip_x.end());
    tiles.insert(tiles.end(), strip.begin(), strip.end());
    tiles.insert(tiles.end(), extra.begin(), extra.end());
    tiles.insert(tiles.end(), extra_strip.begin(), extra_strip.end());
    int rc;
    if (!tiles.empty() &&
        (rc = upload_bytes((void**)&p->tiles_dev, tiles.data(), tiles.size() * sizeof(int4))))
        return rc;
    p->n_interior = int(interior.size());
    p->n_strip_y = int(strip_y.size());
    p->n_strip_x = int(strip_x.size());
    p->n_strip = int(strip_y.size() + strip_x.size() + strip.size());
    p->n_extra = int(extra.size());
    p->n_extra_strip = int(extra_strip.size());
    p->dirty = false;
    return ETL_OK;
}

template <typename R>
int upload_adj_args_t(EtlAdjPlan* p) {
    std::vector<AdjTileArgs<R>> args;
    for (const auto& s : p->sims) {
        const EtlAdjointDesc& d = s.desc;
        AdjTileArgs<R> a{};
        for (int k = 0; k < 7; ++k) {
            a.lam[k][0] = (const R*)d.lam[k];
            a.lam[k][1] = (const R*)s.lamB[k];
        }
        a.cb_ex = (const R*)d.cb_ex;
        a.cb_ey = (const R*)d.cb_ey;
        a.bx_e = (const R*)d.bx_e;
        a.cx_e = (const R*)d.cx_e;
        a.ikx_e = (const R*)d.ikx_e;
        a.bx_h = (const R*)d.bx_h;
        a.cx_h = (const R*)d.cx_h;
        a.ikx_h = (const R*)d.ikx_h;
        a.by_e = (const R*)d.by_e;
        a.cy_e = (const R*)d.cy_e;
        a.iky_e = (const R*)d.iky_e;
        a.by_h = (const R*)d.by_h;
        a.cy_h = (const R*)d.cy_h;
        a.iky_h = (const R*)d.iky_h;
        a.db = R(d.db);
        a.inv_dx = R(d.inv_dx);
        a.inv_dy = R(d.inv_dy);
        a.nx = s.nx;
        a.ny = s.ny;
        a.wo = p->wo();
        a.xh_lo = s.strips[0];
        a.xh_hi = s.strips[1];
        a.xe_lo = s.strips[2];
        a.xe_hi = s.strips[3];
        a.yh_lo = s.strips[4];
        a.yh_hi = s.strips[5];
        a.ye_lo = s.strips[6];
        a.ye_hi = s.strips[7];
        a.tile_runs = s.tile_runs;
        a.runs = s.runs;
        a.src = (const R*)s.src;
        a.n_src = int(s.comp.size());
        const bool box = d.box_ni > 0 && d.box_nj > 0;
        a.bi0 = d.box_i0;
        a.bj0 = d.box_j0;
        a.bni = box ? d.box_ni : 0;
        a.bnj = box ? d.box_nj : 0;
        a.gx = d.grad_ex;
        a.gy = d.grad_ey;
        a.hist = (const R*)d.hist;
        args.push_back(a);
    }
    return etl_memcpy_h2d(p->args_dev, args.data(), args.size() * sizeof(AdjTileArgs<R>));
}

// Reverse steps of history rows count-1 .. 0: launches of depth T, then depth 1.
template <typename R>
int enqueue_adj(EtlAdjPlan* p, int count, int parity, int* launches) {
    const AdjTileArgs<R>* args = static_cast<const AdjTileArgs<R>*>(p->args_dev);
    const int4* t = p->tiles_dev;
    const TileList tl = make_tile_list(t, p->n_interior, p->n_strip, p->n_strip_y,
                                       p->n_strip_x, p->n_extra, p->n_extra_strip);
    hipStream_t s = lane_stream(p->lane);
    const int T = p->tune.temporal;
    int rtop = count - 1, n = 0;
    while (rtop >= 0) {
        const int depth = next_depth(rtop + 1, T, p->precision, 1, p->tune.band_rows);
        if (p->tune.launch == 1) {  // every tile kind in one kernel
            const AdjSet<R> as = adj_set<R>(args, tl, p->step, rtop, parity ^ (n & 1));
            if (int rc = launch_joint<R>(nullptr, 0, &as, depth, s)) return rc;
            rtop -= depth;
            ++n;
            continue;
        }
        // strip tiles on the main stream (dispatched first), interior tiles on a side one
        ETL_HIP(hipEventRecord(p->ev_fork, s));
        ETL_HIP(hipStreamWaitEvent(p->s_strip, p->ev_fork, 0));
        ETL_HIP(hipStreamWaitEvent(p->s_extra, p->ev_fork, 0));
        int rc = launch_adj<R>(args, tl, depth, p->tune.waves, p->step, rtop, parity ^ (n & 1),
                               p->s_strip, s, p->s_extra);
        if (rc) return rc;
        ETL_HIP(hipEventRecord(p->ev_join, p->s_strip));
        ETL_HIP(hipEventRecord(p->ev_join2, p->s_extra));
        ETL_HIP(hipStreamWaitEvent(s, p->ev_join, 0));
        ETL_HIP(hipStreamWaitEvent(s, p->ev_join2, 0));
        rtop -= depth;
        ++n;
    }
    *launches = n;
    return ETL_OK;
}

int enqueue_adj_any(EtlAdjPlan* p, int count, int parity, int* launches) {
    return p->precision == ETL_F32 ? enqueue_adj<float>(p, count, parity, launches)
                                   : enqueue_adj<double>(p, count, parity, launches);
}

void destroy_adj_plan(EtlAdjPlan* p) {
    free_adj_graph(p);
    free_adj_layout(p);
    for (auto& s : p->sims) {
        for (void* b : s.lamB) etl_free(b);
        etl_free(s.src);
    }
    etl_free(p->args_dev);
    etl_free(p->step);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->ev_join2) (void)hipEventDestroy(p->ev_join2);
    for (hipStream_t side : {p->s_strip, p->s_extra})
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
    delete p;
}

template <typename R>
int adj_strips(const EtlAdjointDesc& d, int* out) {
    auto fetch = [&](const void* src, int n, std::vector<R>& v) -> int {
        v.resize(n);
        return etl_memcpy_d2h(v.data(), src, size_t(n) * sizeof(R));
    };
    std::vector<R> b, c, ik;
    const void* prof[4][3] = {{d.bx_h, d.cx_h, d.ikx_h},
                              {d.bx_e, d.cx_e, d.ikx_e},
                              {d.by_h, d.cy_h, d.iky_h},
                              {d.by_e, d.cy_e, d.iky_e}};
    const int len[4] = {d.nx, d.nx + 1, d.ny, d.ny + 1};
    for (int k = 0; k < 4; ++k) {
        int rc;
        if ((rc = fetch(prof[k][0], len[k], b)) || (rc = fetch(prof[k][1], len[k], c)) ||
            (rc = fetch(prof[k][2], len[k], ik)))
            return rc;
        strip_bounds<R>(b, c, ik, out[2 * k], out[2 * k + 1]);
    }
    return ETL_OK;
}

}  // namespace

namespace {

// `step_out` non-null: the two step-record values are returned there instead of set.
int adj_prepare_run(EtlAdjPlan* p, int64_t n_first, int count, int hist_row0,
                    int* step_out = nullptr) {
    ETL_REQUIRE(count >= 0 && n_first >= 0 && n_first < (int64_t(1) << 31),
                "etl_adjplan_run: bad step range");
    for (size_t k = 0; k < p->sims.size(); ++k) {
        const auto& s = p->sims[k];
        const bool box = s.desc.box_ni > 0 && s.desc.box_nj > 0;
        ETL_REQUIRE(!box || (hist_row0 >= 0 && hist_row0 + count <= s.desc.hist_rows),
                    "etl_adjplan_run: %d steps from history row %d, simulation %d's history "
                    "holds %d", count, hist_row0, int(k), s.desc.hist_rows);
        ETL_REQUIRE(s.comp.empty() || n_first + count <= s.src_rows,
                    "etl_adjplan_run: simulation %d's source table covers steps [0, %d)", int(k),
                    s.src_rows);
    }
    int rc;
    if (p->dirty) {
        if ((rc = build_adj_layout(p))) return rc;
        if (int rc_ = sync_lanes()) return rc_;
        if ((rc = p->precision == ETL_F32 ? upload_adj_args_t<float>(p) : upload_adj_args_t<double>(p)))
            return rc;
    }
    if (step_out) {
        step_out[0] = int(n_first);
        step_out[1] = hist_row0;
        return ETL_OK;
    }
    hipStream_t s = lane_stream(p->lane);
    hipLaunchKernelGGL(set_adj_step_kernel, dim3(1), dim3(64), 0, s, p->step, int(n_first),
                       hist_row0);
    ETL_LAUNCHED();
    return ETL_OK;
}

}  // namespace

namespace etl {

int adj_prepare(EtlAdjPlan* p, int64_t n_first, int count, int hist_row0, int* step_out) {
    return adj_prepare_run(p, n_first, count, hist_row0, step_out);
}

void adj_info(const EtlAdjPlan* p, AdjInfo* out) {
    const int4* t = p->tiles_dev;
    out->precision = p->precision;
    out->temporal = p->tune.temporal;
    out->launch = p->tune.launch;
    out->band = p->tune.band_rows;
    out->lane = p->lane;
    out->args = p->args_dev;
    out->tl = make_tile_list(t, p->n_interior, p->n_strip, p->n_strip_y, p->n_strip_x, p->n_extra,
                                  p->n_extra_strip);
    out->step = p->step;
    out->parity = p->parity;
    out->gen = p->layout_gen;
}

void adj_set_parity(EtlAdjPlan* p, int parity) { p->parity = parity; }

}  // namespace etl

extern "C" {

int etl_adjplan_create(const EtlAdjointDesc* descs, int n_sims, const EtlTune* tune,
                       etl_adjplan_t* out) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(descs && out && n_sims >= 1, "etl_adjplan_create: bad arguments");
    const int precision = descs[0].precision;
    ETL_REQUIRE(precision == ETL_F32 || precision == ETL_F64,
                "etl_adjplan_create: unknown precision %d", precision);
    EtlTune t{};
    if (tune) t = *tune;
    if (t.launch == 0) t.launch = 1;  // every tile kind in one kernel
    // the forward plans' rule (fused.h: large_grids), seeing the forward band the
    // caller asks for, so that a co-plan's recompute and reverse plans agree
    if (t.temporal == 0)
        t.temporal = precision == ETL_F32 && t.launch == 1 &&
                             (t.band_rows == 0 || t.band_rows == 16) &&
                             large_grids(descs, n_sims, ctx().cu_count)
                         ? 6
                         : 4;
    if (t.waves == 0) t.waves = 1;
    t.vec = 1;
    t.band_rows = kAdjBand;
    ETL_REQUIRE(t.temporal == 1 || t.temporal == 2 || t.temporal == 4 ||
                    (t.temporal == 6 && precision == ETL_F32 && t.launch != 3),
                "etl_adjplan_create: temporal depth %d (1, 2 or 4; 6 in float with one launch "
                "per step block)", t.temporal);
    ETL_REQUIRE(t.waves == 1 || t.waves == 2 || t.waves == 4,
                "etl_adjplan_create: waves per workgroup must be 1, 2 or 4");
    ETL_REQUIRE(t.launch == 0 || t.launch == 1 || t.launch == 3,
                "etl_adjplan_create: launch must be 0, 1 or 3 (got %d)", t.launch);
    EtlAdjPlan* p = new EtlAdjPlan();
    p->precision = precision;
    p->tune = t;
    p->sims.resize(n_sims);
    int rc = ETL_OK;
    for (int k = 0; k < n_sims && rc == ETL_OK; ++k) {
        const EtlAdjointDesc& d = descs[k];
        auto& s = p->sims[k];
        if (d.precision != precision || d.nx < 2 || d.ny < 2) {
            set_error("etl_adjplan_create: simulation %d: precision or grid", k);
            rc = ETL_EARG;
            break;
        }
        s.desc = d;
        s.nx = d.nx;
        s.ny = d.ny;
        const size_t el = p->elem();
        const size_t sz_ex = size_t(d.nx) * (d.ny + 1) * el, sz_ey = size_t(d.nx + 1) * d.ny * el,
                     sz_hz = size_t(d.nx) * d.ny * el;
        const size_t need[7] = {sz_ex, sz_ey, sz_hz, sz_ex, sz_ey, sz_hz, sz_hz};
        bool ok = guarded_bytes(d.cb_ex) >= sz_ex && guarded_bytes(d.cb_ey) >= sz_ey;
        for (int f = 0; f < 7; ++f) ok = ok && guarded_bytes(d.lam[f]) >= need[f];
        const bool box = d.box_ni > 0 && d.box_nj > 0;
        if (box)
            ok = ok && d.box_i0 >= 1 && d.box_j0 >= 1 && d.box_i0 + d.box_ni <= d.nx - 1 &&
                 d.box_j0 + d.box_nj <= d.ny - 1 && d.grad_ex && d.grad_ey && d.hist &&
                 d.hist_rows >= 1;
        if (!ok) {
            set_error("etl_adjplan_create: simulation %d: lambda and cb must be etl_malloc "
                      "allocations of their shapes, the design box inside the grid with its "
                      "arrays", k);
            rc = ETL_EARG;
            break;
        }
        for (int f = 0; f < 7 && rc == ETL_OK; ++f) {
            rc = etl_malloc(&s.lamB[f], need[f]);
            if (rc == ETL_OK) rc = zero_async(s.lamB[f], need[f]);
        }
        if (rc == ETL_OK)
            rc = precision == ETL_F32 ? adj_strips<float>(d, s.strips) : adj_strips<double>(d, s.strips);
    }
    // side streams and their events: only launch mode 3 (one kernel per tile kind)
    // forks onto them
    if (rc == ETL_OK && p->tune.launch == 3) {
        if (hipStreamCreateWithFlags(&p->s_strip, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&p->s_extra, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join2, hipEventDisableTiming) != hipSuccess) {
            set_error("etl_adjplan_create: stream/event creation failed");
            rc = ETL_EHIP;
        }
    }
    const size_t args_bytes = size_t(n_sims) *
        (precision == ETL_F32 ? sizeof(AdjTileArgs<float>) : sizeof(AdjTileArgs<double>));
    if (rc == ETL_OK) rc = etl_malloc((void**)&p->step, 256);
    if (rc == ETL_OK) rc = etl_malloc(&p->args_dev, args_bytes);
    if (rc != ETL_OK) {
        destroy_adj_plan(p);
        return rc;
    }
    *out = p;
    return ETL_OK;
}

int etl_adjplan_tune(etl_adjplan_t p, EtlTune* out) {
    ETL_REQUIRE(p && out, "etl_adjplan_tune: null argument");
    *out = p->tune;
    return ETL_OK;
}

int etl_adjplan_destroy(etl_adjplan_t p) {
    if (p == nullptr) return ETL_OK;
    ETL_CHECK_INIT();
    if (int rc_ = sync_lanes()) return rc_;
    destroy_adj_plan(p);
    return ETL_OK;
}

int etl_adjplan_sources(etl_adjplan_t p, int sim, int n, const int* comp, const int64_t* flat,
                        int rows, const void* table) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p && sim >= 0 && sim < int(p->sims.size()), "etl_adjplan_sources: no simulation %d",
                sim);
    ETL_REQUIRE(n >= 0 && rows >= 0 && (n == 0 || (comp && flat && table)),
                "etl_adjplan_sources: bad arguments");
    auto& s = p->sims[sim];
    const int nx = s.nx, ny = s.ny;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    for (int k = 0; k < n; ++k) {
        ETL_REQUIRE(comp[k] >= 0 && comp[k] <= 2, "etl_adjplan_sources: component %d", comp[k]);
        const int64_t size = comp[k] == 0 ? int64_t(nx) * (ny + 1)
                                          : (comp[k] == 1 ? int64_t(nx + 1) * ny : int64_t(nx) * ny);
        ETL_REQUIRE(flat[k] >= 0 && flat[k] < size, "etl_adjplan_sources: node %lld outside "
                    "component %d", (long long)flat[k], comp[k]);
    }
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return comp[a] != comp[b] ? comp[a] < comp[b] : flat[a] < flat[b];
    });
    for (int k = 1; k < n; ++k)
        ETL_REQUIRE(!(comp[order[k]] == comp[order[k - 1]] && flat[order[k]] == flat[order[k - 1]]),
                    "etl_adjplan_sources: node listed twice (sum its columns)");
    const size_t el = p->elem();
    bool in_order = true;  // the caller's columns already in node order (the usual case)
    for (int k = 0; k < n; ++k) in_order = in_order && order[k] == k;
    // The table's columns in node order: the caller's table itself when they are, else
    // a copy with the columns gathered row by row.
    std::vector<unsigned char> gathered;
    auto sorted_table = [&]() -> const void* {
        if (in_order) return table;
        gathered.resize(size_t(rows) * n * el);
        const size_t w = el / 4;  // 32-bit words per REAL
        const uint32_t* tb = static_cast<const uint32_t*>(table);
        uint32_t* out = reinterpret_cast<uint32_t*>(gathered.data());
        for (int r = 0; r < rows; ++r)
            for (int k = 0; k < n; ++k)
                for (size_t b = 0; b < w; ++b)
                    out[(size_t(r) * n + k) * w + b] = tb[(size_t(r) * n + order[k]) * w + b];
        return gathered.data();
    };
    const size_t table_bytes = size_t(rows) * n * el;
    // The same nodes and rows as before (the next gradient of a design loop): only the
    // table changes, in place; the layout and the graphs stay.
    bool same = s.src != nullptr && n > 0 && rows == s.src_rows && size_t(n) == s.comp.size();
    for (int k = 0; same && k < n; ++k)
        same = s.comp[k] == comp[order[k]] && s.flat[k] == flat[order[k]];
    if (same) return etl_memcpy_h2d(s.src, sorted_table(), table_bytes);
    if (int rc_ = sync_lanes()) return rc_;
    s.comp.resize(n);
    s.flat.resize(n);
    s.perm = order;
    for (int k = 0; k < n; ++k) {
        s.comp[k] = comp[order[k]];
        s.flat[k] = flat[order[k]];
    }
    etl_free(s.src);
    s.src = nullptr;
    s.src_rows = rows;
    if (n && rows) {
        int rc = upload_bytes(&s.src, sorted_table(), table_bytes);
        if (rc) return rc;
    }
    p->dirty = true;
    free_adj_graph(p);
    return ETL_OK;
}

int etl_adjplan_run(etl_adjplan_t p, int64_t n_first, int count) {
    return etl_adjplan_run_rows(p, n_first, count, 0);
}

int etl_adjplan_set_lane(etl_adjplan_t p, int lane) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr && lane >= 0 && lane < kLanes, "etl_adjplan_set_lane: bad arguments");
    if (lane != p->lane) {
        if (int rc = sync_lanes()) return rc;
        p->lane = lane;
    }
    return ETL_OK;
}

int etl_adjplan_run_rows(etl_adjplan_t p, int64_t n_first, int count, int hist_row0) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_adjplan_run: null plan");
    ETL_REQUIRE(count >= 0 && n_first >= 0 && n_first < (int64_t(1) << 31),
                "etl_adjplan_run: bad step range");
    if (count == 0) return ETL_OK;
    int rc;
    if ((rc = adj_prepare_run(p, n_first, count, hist_row0))) return rc;
    hipStream_t s = lane_stream(p->lane);
    if (p->graph == nullptr || p->graph_count != count || p->graph_parity != p->parity) {
        // the previous graph's launches may still be queued: it is destroyed only once
        // its work is done
        if (p->graph && (rc = sync_lanes())) return rc;
        free_adj_graph(p);
        ETL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int launches = 0;
        rc = enqueue_adj_any(p, count, p->parity, &launches);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        ETL_HIP(e);
        p->graph_src = g;
        ETL_HIP(hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0));
        p->graph_count = count;
        p->graph_parity = p->parity;
        p->graph_launches = launches;
    }
    ETL_HIP(hipGraphLaunch(p->graph, s));
    p->parity ^= p->graph_launches & 1;
    return ETL_OK;
}

int etl_adjplan_reset(etl_adjplan_t p) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_adjplan_reset: null plan");
    static const int comp_of[7] = {0, 1, 2, 0, 1, 2, 2};
    for (auto& s : p->sims)
        for (int f = 0; f < 7; ++f) {
            const int c = comp_of[f];
            const size_t bytes = size_t(c == 1 ? s.nx + 1 : s.nx) * (c == 0 ? s.ny + 1 : s.ny) *
                                 p->elem();
            int rc = zero_async(s.desc.lam[f], bytes);
            if (rc == ETL_OK) rc = zero_async(s.lamB[f], bytes);
            if (rc) return rc;
        }
    p->parity = 0;
    return ETL_OK;
}

int etl_adjplan_lambda(etl_adjplan_t p, int sim, int field, void** ptr) {
    ETL_REQUIRE(p && ptr && sim >= 0 && sim < int(p->sims.size()) && field >= 0 && field < 7,
                "etl_adjplan_lambda: bad arguments");
    const auto& s = p->sims[sim];
    *ptr = p->parity ? s.lamB[field] : s.desc.lam[field];
    return ETL_OK;
}

}  // extern "C"
