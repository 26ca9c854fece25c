// Fused time-stepping plans: the C ABI around te_fused.hip.
//
// A plan replaces the host loop of FDTDSolver.step()/run() (etalon/fdtd/solver.py:
// 83-128,155-160).  The per-step host work of the reference — evaluating the source
// pulse and the DFT phasors in float64 — is done by the caller ahead of time into
// tables (same NumPy expressions, so the REAL values are identical), and the plan
// replays the steps on the device: one fused kernel per T steps, one DFT kernel per
// trace chunk and a counter bump, captured once into a hipGraph per buffer parity.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "fused.h"

namespace etl {
namespace {

__global__ void set_counter_kernel(int* counter, int value) {
    if (threadIdx.x == 0) {
        counter[0] = value;
        counter[1] = 0;  // block tally of the DFT kernel
    }
}

// flag[0] = 1 if psi is non-zero anywhere the fused kernel assumes it stays +0:
// outside the CPML strips, and on the outer tangential-E rows/columns of psi_E.
template <typename R>
__global__ void psi_check_kernel(const R* __restrict__ psi, int rows, int cols, int lo_r, int hi_r,
                                 int lo_c, int hi_c, int edge_rows, int edge_cols,
                                 int* __restrict__ flag) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= int64_t(rows) * cols) return;
    const int i = int(k / cols);
    const int j = int(k - int64_t(i) * cols);
    const bool in_strip = (i < lo_r || i >= hi_r || j < lo_c || j >= hi_c);
    const bool on_edge =
        (edge_rows && (i == 0 || i == rows - 1)) || (edge_cols && (j == 0 || j == cols - 1));
    if ((!in_strip || on_edge) && psi[k] != R(0)) flag[0] = 1;
}

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

}  // namespace
}  // namespace etl

using namespace etl;

struct EtlPlan {
    int precision = 0;
    int nx = 0, ny = 0;
    EtlTune tune{};
    EtlTeDesc desc{};
    int parity = 0;
    int strips[8] = {0};  // xh_lo, xh_hi, xe_lo, xe_hi, yh_lo, yh_hi, ye_lo, ye_hi

    // sources
    int n_src = 0;
    std::vector<int4> src;  // comp, i, j, 0
    int4* src_dev = nullptr;
    void* delta = nullptr;
    int64_t delta_n0 = 0;
    int delta_rows = 0;
    size_t delta_bytes = 0;
    // Tables are staged on the host and uploaded in one contiguous copy each by the
    // next advance (stale = staged but not yet on the device).
    std::vector<unsigned char> delta_host, ph_host;  // ph_host: (ph_host_rows, nq) complex
    int ph_host_rows = 0;
    bool delta_stale = false, ph_stale = false;

    // monitors
    struct Mon {
        int comp, n_nodes, n_f;
        void* acc;
        std::vector<int64_t> flat;
        int slot_base, q_base;
        int64_t n0 = 0;
        int rows = -1;
    };
    std::vector<Mon> mons;
    bool dirty = true;  // monitor CSR / tile flags need a rebuild
    int n_slots = 0, nq = 0, n_pairs = 0;
    int* tile_mon[3] = {nullptr, nullptr, nullptr};   // per component: CSR over tiles
    int4* tile_ent[3] = {nullptr, nullptr, nullptr};  // (row, column, slot, 0)
    int4* tiles_dev = nullptr;  // interior, then strip, then extra tile descriptors
    int n_interior = 0, n_strip = 0, n_extra = 0;
    hipStream_t s_strip = nullptr, s_extra = nullptr;  // beside the interior tiles
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
    void* trace = nullptr;
    int4* pairs = nullptr;
    void** accs = nullptr;
    void* ph = nullptr;  // (ph_rows, nq) complex
    int ph_rows_cap = 0;

    int* counter = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipGraph_t graph_src[2] = {nullptr, nullptr};

    int elem() const { return precision == ETL_F32 ? 4 : 8; }
    int wo() const { return tile_out_cols(tune.vec, tune.temporal); }
    int n_tj() const { return (ny + 1 + wo() - 1) / wo(); }
    int n_bands() const { return (nx + tune.band_rows - 1) / tune.band_rows; }
    int rows_of(int comp) const { return comp == 1 ? nx + 1 : nx; }
    int pitch_of(int comp) const { return comp == 0 ? ny + 1 : ny; }
};

namespace {

void free_graphs(EtlPlan* p) {
    for (int k = 0; k < 2; ++k) {
        if (p->graph[k]) (void)hipGraphExecDestroy(p->graph[k]);
        if (p->graph_src[k]) (void)hipGraphDestroy(p->graph_src[k]);
        p->graph[k] = nullptr;
        p->graph_src[k] = nullptr;
    }
}

void free_layout(EtlPlan* p) {
    for (int c = 0; c < 3; ++c) {
        etl_free(p->tile_mon[c]);
        etl_free(p->tile_ent[c]);
        p->tile_mon[c] = nullptr;
        p->tile_ent[c] = nullptr;
    }
    etl_free(p->trace);
    etl_free(p->pairs);
    etl_free(p->accs);
    etl_free(p->tiles_dev);
    p->tiles_dev = nullptr;
    p->trace = nullptr;
    p->pairs = nullptr;
    p->accs = nullptr;
}

int upload(void** dst, const void* src, size_t bytes) {
    int rc = etl_malloc(dst, bytes);
    if (rc) return rc;
    return etl_memcpy_h2d(*dst, src, bytes);
}

// Tile descriptors of the three kernel variants, the monitor CSR over extra tiles,
// DFT pairs and the trace buffer; rebuilt whenever sources, monitors or tiling change.
//
// A full tile is `band` rows by WO columns.  Its window — rows [i0 - T, i0 + band + T]
// and columns [o0 - T, o0 - T + W) — decides its variant: interior when it meets no
// CPML strip, no monitored node it owns and no source; otherwise it is split into
// kSubBand-row sub-tiles, each extra when it owns a monitored node or its own window
// holds a source, else strip.
int build_layout(EtlPlan* p) {
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    free_layout(p);
    free_graphs(p);
    const int ntj = p->n_tj();
    const int nb = p->n_bands();
    const int wo = p->wo();
    const int T = p->tune.temporal;
    const int W = 64 * p->tune.vec;
    const int band = p->tune.band_rows;
    const int nx = p->nx;

    // monitored nodes: slots, DFT pairs
    struct Node {
        int comp, row, col, slot;
    };
    std::vector<Node> nodes;
    std::vector<int4> pairs;
    std::vector<void*> accs;
    int slot = 0, q = 0;
    for (size_t m = 0; m < p->mons.size(); ++m) {
        auto& mon = p->mons[m];
        mon.slot_base = slot;
        mon.q_base = q;
        const int pitch = p->pitch_of(mon.comp);
        for (int k = 0; k < mon.n_nodes; ++k) {
            const int64_t fl = mon.flat[k];
            const int row = int(fl / pitch);
            nodes.push_back({mon.comp, row, int(fl - int64_t(row) * pitch), slot + k});
        }
        for (int f = 0; f < mon.n_f; ++f)
            for (int k = 0; k < mon.n_nodes; ++k)
                pairs.push_back(make_int4(slot + k, q + f, int(m), f * mon.n_nodes + k));
        accs.push_back(mon.acc);
        slot += mon.n_nodes;
        q += mon.n_f;
    }

    const int* st = p->strips;
    const int x_lo = std::max(st[0], st[2]), x_hi = std::min(st[1], st[3]);
    const int y_lo = std::max(st[4], st[6]), y_hi = std::min(st[5], st[7]);
    auto owns_node = [&](int r0, int r1, int tj) {
        for (const Node& n : nodes)
            if (n.row >= r0 && n.row < r1 && n.col / wo == tj) return true;
        return false;
    };
    auto sees_source = [&](int r0, int r1, int tj) {
        const int cs = tj * wo - T;
        for (const int4& s : p->src)
            if (s.y >= r0 - T && s.y <= r1 + T && s.z >= cs && s.z < cs + W) return true;
        return false;
    };
    std::vector<int4> interior, strip, extra;
    for (int tb = 0; tb < nb; ++tb) {
        const int i0 = tb * band;
        const int i1 = std::min(i0 + band, nx);
        const bool rows_touch = (i0 - T < x_lo) || (i1 + T > x_hi);
        for (int tj = 0; tj < ntj; ++tj) {
            const int cs = tj * wo - T;
            const bool cols_touch = (cs < y_lo) || (cs + W > y_hi);
            if (!rows_touch && !cols_touch && !owns_node(i0, i1, tj) && !sees_source(i0, i1, tj)) {
                interior.push_back(make_int4(i0, tj, 0, 0));
                continue;
            }
            for (int r0 = i0; r0 < i1; r0 += kSubBand) {
                const int r1 = std::min(r0 + kSubBand, nx);
                if (owns_node(r0, r1, tj) || sees_source(r0, r1, tj))
                    extra.push_back(make_int4(r0, tj, int(extra.size()), 0));
                else
                    strip.push_back(make_int4(r0, tj, 0, 0));
            }
        }
    }
    // monitor entries per extra sub-tile (row, column, slot)
    std::vector<std::vector<std::pair<int, int4>>> ent(3);
    for (const Node& n : nodes) {
        const int tj = n.col / wo;
        for (const int4& e : extra)
            if (e.y == tj && n.row >= e.x && n.row < e.x + kSubBand) {
                ent[n.comp].push_back({e.z, make_int4(n.row, n.col, n.slot, 0)});
                break;
            }
    }
    p->n_slots = slot;
    p->nq = q;
    p->n_pairs = int(pairs.size());
    int rc;
    for (int c = 0; c < 3; ++c) {
        if (ent[c].empty()) continue;
        std::stable_sort(ent[c].begin(), ent[c].end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        std::vector<int> ptr(extra.size() + 1, 0);
        for (auto& e : ent[c]) ptr[e.first + 1]++;
        for (size_t k = 0; k < extra.size(); ++k) ptr[k + 1] += ptr[k];
        std::vector<int4> flat(ent[c].size());
        for (size_t k = 0; k < ent[c].size(); ++k) flat[k] = ent[c][k].second;
        if ((rc = upload((void**)&p->tile_mon[c], ptr.data(), ptr.size() * sizeof(int)))) return rc;
        if ((rc = upload((void**)&p->tile_ent[c], flat.data(), flat.size() * sizeof(int4))))
            return rc;
    }
    std::vector<int4> tiles(interior);
    tiles.insert(tiles.end(), strip.begin(), strip.end());
    tiles.insert(tiles.end(), extra.begin(), extra.end());
    if ((rc = upload((void**)&p->tiles_dev, tiles.data(), tiles.size() * sizeof(int4)))) return rc;
    p->n_interior = int(interior.size());
    p->n_strip = int(strip.size());
    p->n_extra = int(extra.size());
    if (p->n_pairs) {
        if ((rc = upload((void**)&p->pairs, pairs.data(), pairs.size() * sizeof(int4)))) return rc;
        if ((rc = upload((void**)&p->accs, accs.data(), accs.size() * sizeof(void*)))) return rc;
        if ((rc = etl_malloc(&p->trace,
                             size_t(p->tune.trace_chunk) * std::max(1, p->n_slots) * p->elem())))
            return rc;
    }
    p->dirty = false;
    return ETL_OK;
}

template <typename R>
FusedArgs<R> make_args(const EtlPlan* p) {
    FusedArgs<R> a;
    std::memset(&a, 0, sizeof a);
    for (int k = 0; k < 7; ++k)
        for (int b = 0; b < 2; ++b) a.f[k][b] = static_cast<const R*>(p->desc.field[k][b]);
    const EtlTeDesc& d = p->desc;
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
    a.nx = p->nx;
    a.ny = p->ny;
    a.band = p->tune.band_rows;
    a.wo = p->wo();
    a.n_bands = p->n_bands();
    a.n_tj = p->n_tj();
    a.tiles_per_xcd = (a.n_bands * a.n_tj + 7) / 8;
    a.xh_lo = p->strips[0];
    a.xh_hi = p->strips[1];
    a.xe_lo = p->strips[2];
    a.xe_hi = p->strips[3];
    a.yh_lo = p->strips[4];
    a.yh_hi = p->strips[5];
    a.ye_lo = p->strips[6];
    a.ye_hi = p->strips[7];
    a.n_src = p->n_src;
    a.src = p->src_dev;
    a.delta = (const R*)p->delta;
    a.has_mon = p->n_slots > 0;
    for (int c = 0; c < 3; ++c) {
        a.tile_mon[c] = p->tile_mon[c];
        a.tile_ent[c] = p->tile_ent[c];
    }
    a.trace = (R*)p->trace;
    a.n_slots = p->n_slots;
    return a;
}

// Enqueue `count` (<= trace_chunk) steps starting at buffer parity `parity`: fused
// launches of depth T, then depth-1 launches for the remainder, the DFT of the trace
// rows and the counter bump.  Returns the number of fused launches in *launches.
// `events` (optional) receives four events per fused launch: fork, interior kernel
// start and end (on the main stream, where only that kernel runs), join.
template <typename R>
int enqueue_block(EtlPlan* p, int count, int parity, hipEvent_t* events, int* launches) {
    const FusedArgs<R> a = make_args<R>(p);
    hipStream_t s = ctx().stream;
    const TileList tl{p->tiles_dev,
                      p->n_interior,
                      p->tiles_dev + p->n_interior,
                      p->n_strip,
                      p->tiles_dev + p->n_interior + p->n_strip,
                      p->n_extra};
    const int T = p->tune.temporal;
    int k = 0, n = 0;
    while (k < count) {
        const int depth = (count - k >= T) ? T : 1;
        if (events) ETL_HIP(hipEventRecord(events[4 * n], s));
        // fork: strip and extra tiles run on their own streams, joined before the next step
        ETL_HIP(hipEventRecord(p->ev_fork, s));
        ETL_HIP(hipStreamWaitEvent(p->s_strip, p->ev_fork, 0));
        ETL_HIP(hipStreamWaitEvent(p->s_extra, p->ev_fork, 0));
        if (events) ETL_HIP(hipEventRecord(events[4 * n + 1], s));
        int rc = launch_fused<R>(a, tl, p->tune.vec, depth, p->tune.band_rows, p->counter, k,
                                 parity ^ (n & 1), s, p->s_strip, p->s_extra);
        if (rc) return rc;
        if (events) ETL_HIP(hipEventRecord(events[4 * n + 2], s));
        ETL_HIP(hipEventRecord(p->ev_join, p->s_strip));
        ETL_HIP(hipEventRecord(p->ev_join2, p->s_extra));
        ETL_HIP(hipStreamWaitEvent(s, p->ev_join, 0));
        ETL_HIP(hipStreamWaitEvent(s, p->ev_join2, 0));
        if (events) ETL_HIP(hipEventRecord(events[4 * n + 3], s));
        k += depth;
        ++n;
    }
    *launches = n;
    // the DFT kernel advances the step counter itself
    if (p->n_pairs && count)
        return launch_trace_dft<R>((const R*)p->trace, p->n_slots, count, (const R*)p->ph, p->nq,
                                   p->counter, p->pairs, p->n_pairs, (R* const*)p->accs, s);
    return launch_bump(p->counter, count, s);
}

int enqueue(EtlPlan* p, int count, int parity, hipEvent_t* events, int* launches) {
    return p->precision == ETL_F32 ? enqueue_block<float>(p, count, parity, events, launches)
                                   : enqueue_block<double>(p, count, parity, events, launches);
}

// Rows allocated for a table of `rows` rows: a power of two, at least 1024.
int table_capacity(int rows) {
    int cap = 1024;
    while (cap < rows && cap < (1 << 30)) cap <<= 1;
    return std::max(cap, rows);
}

// Upload the staged tables (after the steps in flight, which may read the old ones).
// A larger device table gets a new address, which invalidates the captured graphs.
int upload_tables(EtlPlan* p) {
    if (!p->delta_stale && !p->ph_stale) return ETL_OK;
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    int rc;
    if (p->delta_stale) {
        const size_t bytes = p->delta_host.size();
        if (bytes > p->delta_bytes) {
            const size_t cap = size_t(table_capacity(p->delta_rows)) * p->n_src * p->elem();
            etl_free(p->delta);
            p->delta = nullptr;
            if ((rc = etl_malloc(&p->delta, cap))) return rc;
            p->delta_bytes = cap;
            free_graphs(p);
        }
        if (bytes && (rc = etl_memcpy_h2d(p->delta, p->delta_host.data(), bytes))) return rc;
        p->delta_stale = false;
    }
    if (p->ph_stale) {
        const size_t row_bytes = size_t(p->nq) * 2 * p->elem();
        if (p->ph_host_rows > p->ph_rows_cap) {
            const int cap = table_capacity(p->ph_host_rows);
            etl_free(p->ph);
            p->ph = nullptr;
            if ((rc = etl_malloc(&p->ph, size_t(cap) * row_bytes))) return rc;
            p->ph_rows_cap = cap;
            free_graphs(p);
        }
        if (!p->ph_host.empty() &&
            (rc = etl_memcpy_h2d(p->ph, p->ph_host.data(), p->ph_host.size())))
            return rc;
        p->ph_stale = false;
    }
    return ETL_OK;
}

int ensure_graph(EtlPlan* p, int parity) {
    if (p->graph[parity]) return ETL_OK;
    hipStream_t s = ctx().stream;
    ETL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int launches = 0;
    int rc = enqueue(p, p->tune.trace_chunk, parity, nullptr, &launches);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    ETL_HIP(e);
    p->graph_src[parity] = g;
    ETL_HIP(hipGraphInstantiate(&p->graph[parity], g, nullptr, nullptr, 0));
    return ETL_OK;
}

int check_tables(EtlPlan* p, int64_t n, int64_t count) {
    if (p->n_src > 0) {
        ETL_REQUIRE(p->delta != nullptr, "etl_plan_advance: no source table");
        if (!(n >= p->delta_n0 && n + count <= p->delta_n0 + p->delta_rows)) {
            set_error("etl_plan_advance: source table covers steps [%lld, %lld), asked [%lld, %lld)",
                      (long long)p->delta_n0, (long long)(p->delta_n0 + p->delta_rows),
                      (long long)n, (long long)(n + count));
            return ETL_ESTATE;
        }
    }
    for (auto& m : p->mons) {
        if (m.rows < 0 || !(n >= m.n0 && n + count <= m.n0 + m.rows)) {
            set_error("etl_plan_advance: a phasor table does not cover steps [%lld, %lld)",
                      (long long)n, (long long)(n + count));
            return ETL_ESTATE;
        }
        if (p->n_src > 0 && m.n0 != p->delta_n0) {
            set_error("etl_plan_advance: source and phasor tables start at different steps");
            return ETL_ESTATE;
        }
    }
    for (size_t k = 1; k < p->mons.size(); ++k)
        if (p->mons[k].n0 != p->mons[0].n0) {
            set_error("etl_plan_advance: phasor tables start at different steps");
            return ETL_ESTATE;
        }
    return ETL_OK;
}

int run_steps(EtlPlan* p, int64_t n, int64_t count, hipEvent_t* events, int* n_launch) {
    int rc;
    if (p->dirty) {
        if ((rc = build_layout(p))) return rc;
    }
    if ((rc = upload_tables(p))) return rc;
    if ((rc = check_tables(p, n, count))) return rc;
    const int64_t n0 = p->n_src > 0 ? p->delta_n0 : (p->mons.empty() ? n : p->mons[0].n0);
    const int64_t start = n - n0;
    ETL_REQUIRE(start >= 0 && start + count < (int64_t(1) << 31), "etl_plan_advance: step range");
    hipStream_t s = ctx().stream;
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(64), 0, s, p->counter, int(start));
    ETL_LAUNCHED();
    const int K = p->tune.trace_chunk;
    int64_t done = 0;
    int total = 0;
    if (events == nullptr && p->tune.use_graph) {
        while (count - done >= K) {
            if ((rc = ensure_graph(p, p->parity))) return rc;
            ETL_HIP(hipGraphLaunch(p->graph[p->parity], s));
            done += K;  // an even number of launches: parity unchanged
            total += K / p->tune.temporal;
        }
    }
    while (done < count) {
        const int blk = int(std::min<int64_t>(K, count - done));
        int launches = 0;
        if ((rc = enqueue(p, blk, p->parity, events ? events + 4 * total : nullptr, &launches)))
            return rc;
        p->parity ^= (launches & 1);
        done += blk;
        total += launches;
    }
    if (n_launch) *n_launch = total;
    return ETL_OK;
}

}  // namespace

extern "C" {

int etl_plan_create(const EtlTeDesc* desc, const EtlTune* tune, etl_plan_t* out) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(desc && out, "etl_plan_create: null argument");
    ETL_REQUIRE(desc->precision == ETL_F32 || desc->precision == ETL_F64,
                "etl_plan_create: unknown precision %d", desc->precision);
    ETL_REQUIRE(desc->nx >= 1 && desc->ny >= 1 && desc->nx < (1 << 24) && desc->ny < (1 << 24),
                "etl_plan_create: bad grid %d x %d", desc->nx, desc->ny);
    for (int k = 0; k < 7; ++k)
        ETL_REQUIRE(desc->field[k][0] && desc->field[k][1] && desc->field[k][0] != desc->field[k][1],
                    "etl_plan_create: field %d needs two distinct buffers", k);
    ETL_REQUIRE(desc->cb_ex && desc->cb_ey && desc->bx_e && desc->cx_e && desc->ikx_e &&
                    desc->bx_h && desc->cx_h && desc->ikx_h && desc->by_e && desc->cy_e &&
                    desc->iky_e && desc->by_h && desc->cy_h && desc->iky_h,
                "etl_plan_create: null coefficient array");
    EtlTune t{};
    if (tune) t = *tune;
    const bool f32 = desc->precision == ETL_F32;
    if (t.vec == 0) t.vec = 1;
    if (t.temporal == 0) t.temporal = 4;
    if (t.band_rows == 0) t.band_rows = 8;
    if (t.trace_chunk == 0) t.trace_chunk = 64;
    if (tune == nullptr) t.use_graph = 1;
    ETL_REQUIRE(tile_shape_supported(desc->precision, t.vec, t.temporal, t.band_rows) &&
                    tile_shape_supported(desc->precision, t.vec, 1, t.band_rows),
                "etl_plan_create: no compiled tile shape vec=%d temporal=%d band_rows=%d",
                t.vec, t.temporal, t.band_rows);
    ETL_REQUIRE(t.trace_chunk >= 2 && t.trace_chunk <= 4096 &&
                    t.trace_chunk % (2 * t.temporal) == 0,
                "etl_plan_create: trace_chunk must be a multiple of 2 * temporal, <= 4096");

    EtlPlan* p = new EtlPlan();
    p->precision = desc->precision;
    p->nx = desc->nx;
    p->ny = desc->ny;
    p->tune = t;
    p->desc = *desc;
    const int nx = p->nx, ny = p->ny;

    // CPML strips from the 1D profiles.
    auto fetch = [&](const void* src, int n, auto& v) -> int {
        v.resize(n);
        return etl_memcpy_d2h(v.data(), src, size_t(n) * p->elem());
    };
    auto bounds = [&](auto tag, int* out) -> int {
        using R = decltype(tag);
        std::vector<R> b, c, ik;
        const void* prof[4][3] = {{desc->bx_h, desc->cx_h, desc->ikx_h},
                                  {desc->bx_e, desc->cx_e, desc->ikx_e},
                                  {desc->by_h, desc->cy_h, desc->iky_h},
                                  {desc->by_e, desc->cy_e, desc->iky_e}};
        const int len[4] = {nx, nx + 1, ny, ny + 1};
        for (int a = 0; a < 4; ++a) {
            int rc;
            if ((rc = fetch(prof[a][0], len[a], b)) || (rc = fetch(prof[a][1], len[a], c)) ||
                (rc = fetch(prof[a][2], len[a], ik)))
                return rc;
            strip_bounds<R>(b, c, ik, out[2 * a], out[2 * a + 1]);
        }
        return ETL_OK;
    };
    // The fused kernel loads whole vectors that overhang rows by up to 64*vec
    // elements... clamped to the row end, so by less than vec elements, and starts up
    // to `temporal` elements before a row: every array must carry etl_malloc's guard
    // bands and be at least its shape's size.
    {
        const size_t el = p->elem();
        const size_t sz_ex = size_t(nx) * (ny + 1) * el, sz_ey = size_t(nx + 1) * ny * el,
                     sz_hz = size_t(nx) * ny * el;
        const size_t need[7] = {sz_ex, sz_ey, sz_hz, sz_ex, sz_ey, sz_hz, sz_hz};
        bool ok = guarded_bytes(desc->cb_ex) >= sz_ex && guarded_bytes(desc->cb_ey) >= sz_ey;
        for (int k = 0; k < 7; ++k)
            for (int b = 0; b < 2; ++b) ok = ok && guarded_bytes(desc->field[k][b]) >= need[k];
        if (!ok) {
            delete p;
            set_error("etl_plan_create: field and cb arrays must be etl_malloc allocations of "
                      "their full shape (the fused kernel reads into their guard bands)");
            return ETL_EARG;
        }
    }
    int rc = f32 ? bounds(float(0), p->strips) : bounds(double(0), p->strips);
    if (rc == ETL_OK) {
        if (hipStreamCreateWithFlags(&p->s_strip, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&p->s_extra, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join2, hipEventDisableTiming) != hipSuccess) {
            set_error("etl_plan_create: stream/event creation failed");
            rc = ETL_EHIP;
        }
    }
    if (rc == ETL_OK) rc = etl_malloc((void**)&p->counter, 256);
    // psi must be +0 wherever the fused kernel does not touch it.
    if (rc == ETL_OK) {
        int* flag = nullptr;
        rc = etl_malloc((void**)&flag, 256);
        if (rc == ETL_OK) rc = zero_async(flag, 4);
        struct {
            int field, rows, cols, lo_r, hi_r, lo_c, hi_c, er, ec;
        } chk[4] = {
            {3, nx, ny + 1, 0, nx, p->strips[6], p->strips[7], 0, 1},  // psi_ex_y
            {4, nx + 1, ny, p->strips[2], p->strips[3], 0, ny, 1, 0},  // psi_ey_x
            {5, nx, ny, p->strips[0], p->strips[1], 0, ny, 0, 0},      // psi_hz_x
            {6, nx, ny, 0, nx, p->strips[4], p->strips[5], 0, 0},      // psi_hz_y
        };
        for (int k = 0; k < 4 && rc == ETL_OK; ++k) {
            const auto& q = chk[k];
            const int64_t n = int64_t(q.rows) * q.cols;
            const dim3 grid(unsigned((n + 255) / 256));
            if (f32)
                hipLaunchKernelGGL(psi_check_kernel<float>, grid, dim3(256), 0, ctx().stream,
                                   (const float*)desc->field[q.field][0], q.rows, q.cols, q.lo_r,
                                   q.hi_r, q.lo_c, q.hi_c, q.er, q.ec, flag);
            else
                hipLaunchKernelGGL(psi_check_kernel<double>, grid, dim3(256), 0, ctx().stream,
                                   (const double*)desc->field[q.field][0], q.rows, q.cols, q.lo_r,
                                   q.hi_r, q.lo_c, q.hi_c, q.er, q.ec, flag);
            if (hipGetLastError() != hipSuccess) rc = ETL_EHIP;
        }
        int host_flag = 0;
        if (rc == ETL_OK) rc = etl_memcpy_d2h(&host_flag, flag, 4);
        etl_free(flag);
        if (rc == ETL_OK && host_flag) {
            set_error("etl_plan_create: psi is non-zero outside the CPML strips; the fused "
                      "path needs psi = +0 there (use the per-half-step primitives)");
            rc = ETL_ESTATE;
        }
    }
    if (rc != ETL_OK) {
        etl_free(p->counter);
        if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
        if (p->ev_join) (void)hipEventDestroy(p->ev_join);
        if (p->ev_join2) (void)hipEventDestroy(p->ev_join2);
        if (p->s_strip) (void)hipStreamDestroy(p->s_strip);
        if (p->s_extra) (void)hipStreamDestroy(p->s_extra);
        delete p;
        return rc;
    }
    *out = p;
    return ETL_OK;
}

int etl_plan_destroy(etl_plan_t p) {
    if (p == nullptr) return ETL_OK;
    ETL_CHECK_INIT();
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    free_graphs(p);
    free_layout(p);
    etl_free(p->src_dev);
    etl_free(p->delta);
    etl_free(p->ph);
    etl_free(p->counter);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->ev_join2) (void)hipEventDestroy(p->ev_join2);
    for (hipStream_t side : {p->s_strip, p->s_extra})
        if (side) {
            ETL_HIP(hipStreamSynchronize(side));
            (void)hipStreamDestroy(side);
        }
    delete p;
    return ETL_OK;
}

int etl_plan_set_sources(etl_plan_t p, int n_nodes, const int* comp, const int64_t* i,
                         const int64_t* j) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_plan_set_sources: null plan");
    ETL_REQUIRE(n_nodes >= 0 && n_nodes <= kMaxSources,
                "etl_plan_set_sources: %d source nodes (at most %d)", n_nodes, kMaxSources);
    ETL_REQUIRE(n_nodes == 0 || (comp && i && j), "etl_plan_set_sources: null array");
    std::vector<int4> src;
    for (int s = 0; s < n_nodes; ++s) {
        ETL_REQUIRE(comp[s] >= 0 && comp[s] <= 2, "etl_plan_set_sources: component %d", comp[s]);
        ETL_REQUIRE(i[s] >= 0 && i[s] < p->rows_of(comp[s]) && j[s] >= 0 &&
                        j[s] < p->pitch_of(comp[s]),
                    "etl_plan_set_sources: node (%lld, %lld) outside component %d",
                    (long long)i[s], (long long)j[s], comp[s]);
        // Outer tangential E never changes in the update; a source there would be
        // the only writer, which the fused step does not model.
        const bool outer = (comp[s] == 0 && (j[s] == 0 || j[s] == p->ny)) ||
                           (comp[s] == 1 && (i[s] == 0 || i[s] == p->nx));
        ETL_REQUIRE(!outer, "etl_plan_set_sources: node on the outer tangential-E boundary");
        src.push_back(make_int4(comp[s], int(i[s]), int(j[s]), 0));
    }
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    etl_free(p->src_dev);
    p->src_dev = nullptr;
    if (n_nodes) {
        int rc = upload((void**)&p->src_dev, src.data(), src.size() * sizeof(int4));
        if (rc) return rc;
    }
    p->src = src;
    p->n_src = n_nodes;
    p->dirty = true;
    free_graphs(p);
    return ETL_OK;
}

int etl_plan_source_table(etl_plan_t p, int64_t n0, int rows, const void* delta) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_plan_source_table: null plan");
    ETL_REQUIRE(rows >= 0 && n0 >= 0, "etl_plan_source_table: bad rows");
    if (p->n_src == 0) return ETL_OK;
    ETL_REQUIRE(delta != nullptr, "etl_plan_source_table: null table");
    const size_t bytes = size_t(rows) * p->n_src * p->elem();
    const unsigned char* b = static_cast<const unsigned char*>(delta);
    p->delta_host.assign(b, b + bytes);
    p->delta_n0 = n0;
    p->delta_rows = rows;
    p->delta_stale = true;
    return ETL_OK;
}

int etl_plan_add_monitor(etl_plan_t p, int comp, int n_nodes, const int64_t* flat_index, int n_f,
                         void* acc, int* id) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p && id, "etl_plan_add_monitor: null argument");
    ETL_REQUIRE(comp >= 0 && comp <= 2, "etl_plan_add_monitor: component %d", comp);
    ETL_REQUIRE(n_nodes >= 1 && n_f >= 1 && flat_index && acc,
                "etl_plan_add_monitor: bad monitor (%d nodes, %d frequencies)", n_nodes, n_f);
    const int64_t size = int64_t(p->rows_of(comp)) * p->pitch_of(comp);
    EtlPlan::Mon m;
    m.comp = comp;
    m.n_nodes = n_nodes;
    m.n_f = n_f;
    m.acc = acc;
    m.flat.assign(flat_index, flat_index + n_nodes);
    const int pitch = p->pitch_of(comp);
    for (int64_t f : m.flat) {
        ETL_REQUIRE(f >= 0 && f < size, "etl_plan_add_monitor: index %lld outside component %d",
                    (long long)f, comp);
        const int64_t row = f / pitch;
        ETL_REQUIRE(!(comp == 1 && row == p->nx),
                    "etl_plan_add_monitor: node on the outer tangential-E boundary");
    }
    m.slot_base = m.q_base = 0;
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    p->mons.push_back(m);
    p->dirty = true;
    *id = int(p->mons.size()) - 1;
    // The combined phasor table changes shape: drop it.
    etl_free(p->ph);
    p->ph = nullptr;
    p->ph_rows_cap = 0;
    p->ph_host.clear();
    p->ph_host_rows = 0;
    p->ph_stale = false;
    for (auto& mm : p->mons) mm.rows = -1;
    free_graphs(p);
    return ETL_OK;
}

int etl_plan_phasor_table(etl_plan_t p, int id, int64_t n0, int rows, const void* ph) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_plan_phasor_table: null plan");
    ETL_REQUIRE(id >= 0 && id < int(p->mons.size()), "etl_plan_phasor_table: no monitor %d", id);
    ETL_REQUIRE(rows >= 0 && n0 >= 0 && (ph || rows == 0), "etl_plan_phasor_table: bad table");
    int rc;
    if (p->dirty) {
        if ((rc = build_layout(p))) return rc;
    }
    const size_t cplx = size_t(2) * p->elem();
    const size_t row_bytes = size_t(p->nq) * cplx;
    if (rows > p->ph_host_rows) {  // rows are appended: the other monitors' columns stay
        p->ph_host.resize(size_t(rows) * row_bytes);
        p->ph_host_rows = rows;
    }
    auto& m = p->mons[id];
    const unsigned char* src = static_cast<const unsigned char*>(ph);
    for (int r = 0; r < rows; ++r)
        std::memcpy(p->ph_host.data() + size_t(r) * row_bytes + size_t(m.q_base) * cplx,
                    src + size_t(r) * m.n_f * cplx, size_t(m.n_f) * cplx);
    m.n0 = n0;
    m.rows = rows;
    p->ph_stale = true;
    return ETL_OK;
}

int etl_plan_tune(etl_plan_t p, EtlTune* out) {
    ETL_REQUIRE(p && out, "etl_plan_tune: null argument");
    *out = p->tune;
    return ETL_OK;
}

int etl_plan_advance(etl_plan_t p, int64_t n, int64_t count, int* parity) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_plan_advance: null plan");
    ETL_REQUIRE(n >= 0 && count >= 0, "etl_plan_advance: bad step range");
    if (count > 0) {
        int rc = run_steps(p, n, count, nullptr, nullptr);
        if (rc) return rc;
    }
    if (parity) *parity = p->parity;
    return ETL_OK;
}

int etl_plan_time(etl_plan_t p, int64_t n, int64_t count, EtlTiming* out, int* parity) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr && out != nullptr, "etl_plan_time: null argument");
    ETL_REQUIRE(count >= 1 && count <= 100000, "etl_plan_time: count %lld outside [1, 100000]",
                (long long)count);
    std::vector<hipEvent_t> ev(4 * count);  // at most one launch per step
    for (auto& e : ev) ETL_HIP(hipEventCreate(&e));
    hipEvent_t t0, t1;
    ETL_HIP(hipEventCreate(&t0));
    ETL_HIP(hipEventCreate(&t1));
    hipStream_t s = ctx().stream;
    ETL_HIP(hipEventRecord(t0, s));
    int launches = 0;
    int rc = run_steps(p, n, count, ev.data(), &launches);
    if (rc == ETL_OK) {
        ETL_HIP(hipEventRecord(t1, s));
        ETL_HIP(hipEventSynchronize(t1));
        double all = 0.0, interior = 0.0;
        for (int k = 0; k < launches; ++k) {
            float a = 0.f, b = 0.f;
            ETL_HIP(hipEventElapsedTime(&a, ev[4 * k], ev[4 * k + 3]));
            ETL_HIP(hipEventElapsedTime(&b, ev[4 * k + 1], ev[4 * k + 2]));
            all += a;
            interior += b;
        }
        float tot = 0.f;
        ETL_HIP(hipEventElapsedTime(&tot, t0, t1));
        const int L = std::max(launches, 1);
        out->launch_ms = all / L;
        out->interior_ms = p->n_interior ? interior / L : 0.0;
        out->total_ms = tot;
        out->interior_cells = int64_t(p->n_interior) * p->tune.band_rows * p->wo();
        out->launches = launches;
        out->depth = p->tune.temporal;
        if (parity) *parity = p->parity;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return rc;
}

}  // extern "C"
