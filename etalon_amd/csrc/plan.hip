// Fused time-stepping plans: the C ABI around the register-tile kernels.
//
// A plan replaces the host loop of FDTDSolver.step()/run() (etalon/fdtd/solver.py:
// 83-128,155-160) for one simulation or for a batch of independent simulations of
// any sizes that advance together.  The per-step host work of the reference —
// evaluating the source pulse and the DFT phasors in float64 — is done by the
// caller ahead of time into tables (same NumPy expressions, so the REAL values are
// identical), and the plan replays the steps on the device: per T steps one launch
// of each tile kernel over the tiles of every simulation (a work-list of tile
// descriptors), per trace chunk one DFT fold over every monitor, captured once into
// a hipGraph per buffer parity.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "fused.h"
#include "coplan.h"
#include "joint.h"

namespace etl {
namespace {

__global__ void set_counter_kernel(int* counter, int value, int probe_row0) {
    if (threadIdx.x == 0) {
        counter[0] = value;
        counter[1] = 0;                   // block tally of the DFT kernel
        counter[2] = value - probe_row0;  // probe row of step counter[0]: counter[0] - counter[2]
    }
}

// Checkpoint copies: entry e (one field of one simulation) copies `words` 32-bit
// words between buffer `parity` of the field and slot `seg` of its checkpoint array.
// A window entry (rows > 0) has slots that hold a window of the field: `rows` rows of
// `row_words` words, stored one after another in the slot (of `words` words), taken
// from the field `off` words in, `pitch` words apart.
struct CkEntry {
    uint32_t* ck;
    uint32_t* f[2];
    int64_t words;
    int64_t rows, row_words, pitch, off;
};

// 16-byte moves where both ends are 16-byte aligned (the fields are: etl_malloc
// blocks are 256-byte aligned, and a checkpoint slot starts at seg * words words),
// word moves otherwise and for the tail.  Window entries: block x takes rows x,
// x + gridDim.x, ..., its threads the words of a row.
__global__ void ckpt_copy_kernel(const CkEntry* __restrict__ tab, int seg, int parity,
                                 int restore) {
    const CkEntry& e = tab[blockIdx.y];
    // a select, not e.f[parity]: a dynamically indexed local copy of the entry would
    // be promoted to LDS (32 KB per workgroup)
    uint32_t* f = parity ? e.f[1] : e.f[0];
    uint32_t* ck = e.ck + seg * e.words;
    if (e.rows > 0) {
        uint32_t* fw = f + e.off;
        for (int64_t r = blockIdx.x; r < e.rows; r += gridDim.x)
            for (int64_t q = threadIdx.x; q < e.row_words; q += blockDim.x) {
                if (restore) fw[r * e.pitch + q] = ck[r * e.row_words + q];
                else ck[r * e.row_words + q] = fw[r * e.pitch + q];
            }
        return;
    }
    const uint32_t* src = restore ? ck : f;
    uint32_t* dst = restore ? f : ck;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const int64_t t0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t done = 0;
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0) {
        const int64_t quads = e.words / 4;
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        for (int64_t i = t0; i < quads; i += stride) d4[i] = s4[i];
        done = quads * 4;
    }
    for (int64_t i = done + t0; i < e.words; i += stride) dst[i] = src[i];
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

// flag[1] = 1 if any of the n values is -0 (the tile kernels need fields without -0).
template <typename R>
__global__ void negzero_check_kernel(const R* __restrict__ f, int64_t n, int* __restrict__ flag) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const R v = f[k];
    if (v == R(0) && signbit(v)) flag[1] = 1;
}

}  // namespace
}  // namespace etl

using namespace etl;

struct EtlPlan {
    int precision = 0;
    EtlTune tune{};
    int parity = 0;
    unsigned layout_gen = 0;  // bumped by every layout rebuild (graphs hold its pointers)
    int lane = 0;  // stream lane of its work

    // One simulation of the plan.
    struct Mon {
        int comp, n_nodes, n_f;
        void* acc;
        std::vector<int64_t> flat;
        int slot_base = 0, q_base = 0;
        int64_t n0 = 0;
        int rows = -1;
        int owner = -1;  // >= 0: reads the phasor columns of that monitor
    };
    // A time-domain probe: the values of its nodes at every step of an advance go to
    // rows of `out` (n_nodes REALs per row, row 0 = the first step of the advance).
    struct Probe {
        int comp;
        std::vector<int64_t> flat;
        void* out;
        int out_rows;
        int slot_base = 0;
    };
    struct Sim {
        int nx = 0, ny = 0;
        EtlTeDesc desc{};
        int strips[8] = {0};  // xh_lo, xh_hi, xe_lo, xe_hi, yh_lo, yh_hi, ye_lo, ye_hi
        // sources
        int n_src = 0;
        std::vector<int4> src;  // comp, i, j, 0
        int4* src_dev = nullptr;
        void* delta = nullptr;
        int64_t delta_n0 = 0;
        int delta_rows = 0;
        size_t delta_bytes = 0;
        // Tables are staged on the host and uploaded in one contiguous copy each by
        // the next advance (stale = staged but not yet on the device).
        std::vector<unsigned char> delta_host, ph_host;  // ph_host: (ph_host_rows, nq)
        int ph_host_rows = 0;
        bool delta_stale = false, ph_stale = false;
        // monitors and probes
        std::vector<Mon> mons;
        std::vector<Probe> probes;
        int n_slots = 0, nq = 0;
        int* tile_mon[3] = {nullptr, nullptr, nullptr};   // per component: CSR over extra tiles
        int4* tile_ent[3] = {nullptr, nullptr, nullptr};  // (row, column, slot, 0)
        void* trace = nullptr;                            // (trace_chunk, n_slots)
        void* ph = nullptr;                               // (ph_rows_cap, nq) complex
        int ph_rows_cap = 0;

        // checkpoints: ck[k] = (ck_seg, size of field k) REALs; with ck_window, each
        // slot holds only the window of the fields with Hz nodes [ck_row0, + ck_nx) x
        // [ck_col0, + ck_ny) (Ex one column more, Ey one row more)
        void* ck[7] = {nullptr};
        int ck_seg = 0;
        bool ck_window = false;
        int ck_nx = 0, ck_ny = 0, ck_row0 = 0, ck_col0 = 0;

        int rows_of(int comp) const { return comp == 1 ? nx + 1 : nx; }
        int pitch_of(int comp) const { return comp == 0 ? ny + 1 : ny; }
    };
    std::vector<Sim> sims;

    bool dirty = true;        // tiles / monitor CSR / DFT pairs need a rebuild
    bool args_stale = true;   // the device copies of the per-simulation arguments
    void* args_dev = nullptr;        // FusedArgs<R>[n_sims]
    void* trace_sims_dev = nullptr;  // TraceSim<R>[n_sims]
    int4* tiles_dev = nullptr;       // interior, then strip, then extra tile descriptors
    int n_interior = 0, n_strip = 0, n_extra = 0, n_extra_strip = 0;
    int n_uniform = 0;               // the last interior tiles, of uniform material
    void* uvals = nullptr;           // their cb_ex, cb_ey: REAL pairs in tile order
    ProbeCopy* probe_tab = nullptr;  // every probe's copy, all simulations
    int n_probe_tab = 0, probe_max_n = 0;
    int n_strip_y = 0, n_strip_x = 0;  // the strip tiles: y only, then x only, then both
    int64_t interior_cells = 0;      // Hz cells the interior tiles own
    int4* pairs = nullptr;           // DFT pairs of every simulation
    int* pair_sim = nullptr;
    void** accs = nullptr;
    int n_pairs = 0;
    hipStream_t s_strip = nullptr, s_extra = nullptr;  // beside the interior tiles
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;

    CkEntry* ck_tab = nullptr;  // checkpoint copy entries of every simulation
    int ck_entries = 0;
    bool ck_stale = true;

    int* counter = nullptr;
    // What the device counter holds once the work enqueued so far has run, when the
    // host knows it: counter[0] == counter_at (-1: unknown), counter[1] == 0,
    // counter[2] == counter_probe.  An advance that starts there needs no counter set.
    int64_t counter_at = -1;
    int64_t counter_next = -1;  // counter[0] after the steps being enqueued
    int counter_probe = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipGraph_t graph_src[2] = {nullptr, nullptr};

    int elem() const { return precision == ETL_F32 ? 4 : 8; }
    int wo() const { return tile_out_cols(tune.vec, tune.temporal); }
    int n_tj(const Sim& s) const { return (s.ny + 1 + wo() - 1) / wo(); }
    int n_bands(const Sim& s) const { return (s.nx + tune.band_rows - 1) / tune.band_rows; }
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
    for (auto& s : p->sims) {
        for (int c = 0; c < 3; ++c) {
            etl_free(s.tile_mon[c]);
            etl_free(s.tile_ent[c]);
            s.tile_mon[c] = nullptr;
            s.tile_ent[c] = nullptr;
        }
        etl_free(s.trace);
        s.trace = nullptr;
    }
    etl_free(p->pairs);
    etl_free(p->pair_sim);
    etl_free(p->accs);
    etl_free(p->tiles_dev);
    etl_free(p->uvals);
    p->tiles_dev = nullptr;
    p->pairs = nullptr;
    p->pair_sim = nullptr;
    p->accs = nullptr;
    p->uvals = nullptr;
    p->n_uniform = 0;
    etl_free(p->probe_tab);
    p->probe_tab = nullptr;
    p->n_probe_tab = p->probe_max_n = 0;
}

int upload(void** dst, const void* src, size_t bytes) {
    int rc = etl_malloc(dst, bytes);
    if (rc) return rc;
    return etl_memcpy_h2d(*dst, src, bytes);
}

// Uniform-material interior tiles: those whose window (rows [i0 - T, i0 + band + T),
// columns [cs, cs + W), the cb rows an interior tile loads) holds one cb_ex and one
// cb_ey value, bit for bit.  `tiles` (interior tiles, any simulations) is reordered
// stably so that they come last; `vals` receives their two values each.  Reads every
// simulation's cb arrays back once (the layout is rebuilt when they change).
template <typename R>
int split_uniform(EtlPlan* p, std::vector<int4>& tiles, std::vector<R>& vals) {
    const int T = p->tune.temporal, band = p->tune.band_rows, W = 64 * p->tune.vec;
    const int wo = p->wo();
    std::vector<char> uniform(tiles.size(), 0);
    std::vector<R> pair(2 * tiles.size(), R(0));
    for (int si = 0; si < int(p->sims.size()); ++si) {
        const auto& sim = p->sims[si];
        bool any = false;
        for (const int4& t : tiles) any = any || t.w == si;
        if (!any) continue;
        const int nx = sim.nx, ny = sim.ny, px = ny + 1;
        std::vector<R> cx(size_t(nx) * px), cy(size_t(nx + 1) * ny);
        int rc;
        if ((rc = etl_memcpy_d2h(cx.data(), sim.desc.cb_ex, cx.size() * sizeof(R))) ||
            (rc = etl_memcpy_d2h(cy.data(), sim.desc.cb_ey, cy.size() * sizeof(R))))
            return rc;
        // run[r * pitch + c]: how many elements from (r, c) on equal it, capped at W
        auto runs = [W](const std::vector<R>& a, int rows, int pitch) {
            std::vector<uint8_t> run(a.size());
            for (int r = 0; r < rows; ++r) {
                const R* row = a.data() + size_t(r) * pitch;
                uint8_t* out = run.data() + size_t(r) * pitch;
                out[pitch - 1] = 1;
                for (int c = pitch - 2; c >= 0; --c)
                    out[c] = std::memcmp(&row[c], &row[c + 1], sizeof(R)) == 0
                                 ? uint8_t(std::min(W, out[c + 1] + 1))
                                 : uint8_t(1);
            }
            return run;
        };
        const std::vector<uint8_t> rx = runs(cx, nx, px), ry = runs(cy, nx + 1, ny);
        for (size_t k = 0; k < tiles.size(); ++k) {
            const int4& t = tiles[k];
            if (t.w != si) continue;
            const int cs = t.y * wo - T;
            const int r0 = t.x - T, r1 = t.x + band + T;
            if (cs < 0 || cs + W > ny || r0 < 0 || r1 > nx) continue;  // never for interior
            const R vx = cx[size_t(r0) * px + cs], vy = cy[size_t(r0) * ny + cs];
            bool same = true;
            for (int r = r0; r < r1 && same; ++r)
                same = rx[size_t(r) * px + cs] >= W && ry[size_t(r) * ny + cs] >= W &&
                       std::memcmp(&cx[size_t(r) * px + cs], &vx, sizeof(R)) == 0 &&
                       std::memcmp(&cy[size_t(r) * ny + cs], &vy, sizeof(R)) == 0;
            uniform[k] = same;
            pair[2 * k] = vx;
            pair[2 * k + 1] = vy;
        }
    }
    std::vector<int4> mixed, flat;
    vals.clear();
    for (size_t k = 0; k < tiles.size(); ++k) {
        if (uniform[k]) {
            flat.push_back(tiles[k]);
            vals.push_back(pair[2 * k]);
            vals.push_back(pair[2 * k + 1]);
        } else {
            mixed.push_back(tiles[k]);
        }
    }
    p->n_uniform = int(flat.size());
    tiles.swap(mixed);
    tiles.insert(tiles.end(), flat.begin(), flat.end());
    return ETL_OK;
}

// Tile descriptors of the three kernel variants, the monitor CSR over extra tiles,
// DFT pairs and the trace buffers; rebuilt whenever sources, monitors or tiling change.
//
// A full tile of a simulation is `band` rows by WO columns.  Its window — rows
// [i0 - T, i0 + band + T] and columns [o0 - T, o0 - T + W) — decides its variant:
// interior when it meets no CPML strip, no monitored node it owns and no source;
// otherwise it is split into kSubBand-row sub-tiles, each extra when it owns a
// monitored node or its own window holds a source, else strip.
int build_layout(EtlPlan* p) {
    if (int rc_ = sync_lanes()) return rc_;
    ++p->layout_gen;
    free_layout(p);
    free_graphs(p);
    const int wo = p->wo();
    const int T = p->tune.temporal;
    const int W = 64 * p->tune.vec;
    const int band = p->tune.band_rows;

    std::vector<int4> interior, strip_y, strip_x, strip, extra, extra_strip, pairs;
    std::vector<int> pair_sim;
    std::vector<void*> accs;
    int64_t interior_cells = 0;
    for (int si = 0; si < int(p->sims.size()); ++si) {
        auto& sim = p->sims[si];
        const int ntj = p->n_tj(sim);
        const int nb = p->n_bands(sim);
        const int nx = sim.nx;

        // monitored nodes: slots, DFT pairs
        struct Node {
            int comp, row, col, slot;
        };
        std::vector<Node> nodes;
        int slot = 0, q = 0;
        for (auto& mon : sim.mons) {
            mon.slot_base = slot;
            mon.q_base = q;
            const int pitch = sim.pitch_of(mon.comp);
            for (int k = 0; k < mon.n_nodes; ++k) {
                const int64_t fl = mon.flat[k];
                const int row = int(fl / pitch);
                nodes.push_back({mon.comp, row, int(fl - int64_t(row) * pitch), slot + k});
            }
            if (mon.owner >= 0) mon.q_base = sim.mons[mon.owner].q_base;
            const int acc_index = int(accs.size());
            for (int f = 0; f < mon.n_f; ++f)
                for (int k = 0; k < mon.n_nodes; ++k) {
                    pairs.push_back(
                        make_int4(slot + k, mon.q_base + f, acc_index, f * mon.n_nodes + k));
                    pair_sim.push_back(si);
                }
            accs.push_back(mon.acc);
            slot += mon.n_nodes;
            if (mon.owner < 0) q += mon.n_f;
        }
        for (auto& pr : sim.probes) {
            pr.slot_base = slot;
            const int pitch = sim.pitch_of(pr.comp);
            for (size_t k = 0; k < pr.flat.size(); ++k) {
                const int row = int(pr.flat[k] / pitch);
                nodes.push_back({pr.comp, row, int(pr.flat[k] - int64_t(row) * pitch),
                                 slot + int(k)});
            }
            slot += int(pr.flat.size());
        }
        sim.n_slots = slot;
        sim.nq = q;

        const int* st = sim.strips;
        const int x_lo = std::max(st[0], st[2]), x_hi = std::min(st[1], st[3]);
        const int y_lo = std::max(st[4], st[6]), y_hi = std::min(st[5], st[7]);
        auto owns_node = [&](int r0, int r1, int tj) {
            for (const Node& n : nodes)
                if (n.row >= r0 && n.row < r1 && n.col / wo == tj) return true;
            return false;
        };
        auto sees_source = [&](int r0, int r1, int tj) {
            const int cs = tj * wo - T;
            for (const int4& s : sim.src)
                if (s.y >= r0 - T && s.y <= r1 + T && s.z >= cs && s.z < cs + W) return true;
            return false;
        };
        std::vector<int4> sim_extra;  // this simulation's extra sub-tiles, numbered in order
        auto sub_touches = [&](int r0, int r1, int tj) {
            const int cs = tj * wo - T;
            return (r0 - T < std::max(x_lo, 1)) || (r1 + T > x_hi) || (cs < y_lo) ||
                   (cs + W > y_hi);
        };
        for (int tb = 0; tb < nb; ++tb) {
            const int i0 = tb * band;
            const int i1 = std::min(i0 + band, nx);
            // the first row counts as a strip row: Ey row 0 stays fixed
            const bool rows_touch = (i0 - T < std::max(x_lo, 1)) || (i1 + T > x_hi);
            for (int tj = 0; tj < ntj; ++tj) {
                const int cs = tj * wo - T;
                const bool cols_touch = (cs < y_lo) || (cs + W > y_hi);
                if (!rows_touch && !cols_touch && !owns_node(i0, i1, tj) &&
                    !sees_source(i0, i1, tj)) {
                    interior.push_back(make_int4(i0, tj, 0, si));
                    interior_cells += int64_t(band) * wo;
                    continue;
                }
                for (int r0 = i0; r0 < i1; r0 += kSubBand) {
                    const int r1 = std::min(r0 + kSubBand, nx);
                    if (owns_node(r0, r1, tj) || sees_source(r0, r1, tj)) {
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
        // monitor entries per extra sub-tile of this simulation (row, column, slot)
        const int n_extra_sim = int(sim_extra.size());
        std::vector<std::vector<std::pair<int, int4>>> ent(3);
        for (const Node& n : nodes) {
            const int tj = n.col / wo;
            for (int e = 0; e < n_extra_sim; ++e) {
                const int4& x = sim_extra[e];
                if (x.y == tj && n.row >= x.x && n.row < x.x + kSubBand) {
                    ent[n.comp].push_back({x.z, make_int4(n.row, n.col, n.slot, 0)});
                    break;
                }
            }
        }
        int rc;
        for (int c = 0; c < 3; ++c) {
            if (ent[c].empty()) continue;
            std::stable_sort(ent[c].begin(), ent[c].end(),
                             [](const auto& a, const auto& b) { return a.first < b.first; });
            std::vector<int> ptr(n_extra_sim + 1, 0);
            for (auto& e : ent[c]) ptr[e.first + 1]++;
            for (int k = 0; k < n_extra_sim; ++k) ptr[k + 1] += ptr[k];
            std::vector<int4> flat(ent[c].size());
            for (size_t k = 0; k < ent[c].size(); ++k) flat[k] = ent[c][k].second;
            if ((rc = upload((void**)&sim.tile_mon[c], ptr.data(), ptr.size() * sizeof(int))))
                return rc;
            if ((rc = upload((void**)&sim.tile_ent[c], flat.data(), flat.size() * sizeof(int4))))
                return rc;
        }
        if (sim.n_slots &&
            (rc = etl_malloc(&sim.trace, size_t(p->tune.trace_chunk) * sim.n_slots * p->elem())))
            return rc;
    }
    // the probes' copies (the trace buffers are this layout's)
    {
        std::vector<ProbeCopy> pc;
        for (const auto& sim : p->sims)
            for (const auto& pr : sim.probes) {
                pc.push_back(ProbeCopy{sim.trace, pr.out, sim.n_slots, pr.slot_base,
                                       int(pr.flat.size()), 0});
                p->probe_max_n = std::max(p->probe_max_n, int(pr.flat.size()));
            }
        p->n_probe_tab = int(pc.size());
        int rc_p;
        if (!pc.empty() &&
            (rc_p = upload((void**)&p->probe_tab, pc.data(), pc.size() * sizeof(ProbeCopy))))
            return rc_p;
    }
    // uniform-material interior tiles last, with their two cb values
    if (p->tune.uniform_cb > 0 && !interior.empty()) {
        int rc_u;
        if (p->precision == ETL_F32) {
            std::vector<float> vals;
            rc_u = split_uniform<float>(p, interior, vals);
            if (rc_u == ETL_OK && !vals.empty())
                rc_u = upload(&p->uvals, vals.data(), vals.size() * sizeof(float));
        } else {
            std::vector<double> vals;
            rc_u = split_uniform<double>(p, interior, vals);
            if (rc_u == ETL_OK && !vals.empty())
                rc_u = upload(&p->uvals, vals.data(), vals.size() * sizeof(double));
        }
        if (rc_u) return rc_u;
    }
    std::vector<int4> tiles(interior);
    tiles.insert(tiles.end(), strip_y.begin(), strip_y.end());
    tiles.insert(tiles.end(), strip_x.begin(), strip_x.end());
    tiles.insert(tiles.end(), strip.begin(), strip.end());
    tiles.insert(tiles.end(), extra.begin(), extra.end());
    tiles.insert(tiles.end(), extra_strip.begin(), extra_strip.end());
    int rc;
    if (!tiles.empty() &&
        (rc = upload((void**)&p->tiles_dev, tiles.data(), tiles.size() * sizeof(int4))))
        return rc;
    p->n_interior = int(interior.size());
    p->n_strip_y = int(strip_y.size());
    p->n_strip_x = int(strip_x.size());
    p->n_strip = int(strip_y.size() + strip_x.size() + strip.size());
    p->n_extra = int(extra.size());
    p->n_extra_strip = int(extra_strip.size());
    p->interior_cells = interior_cells;
    p->n_pairs = int(pairs.size());
    if (p->n_pairs) {
        if ((rc = upload((void**)&p->pairs, pairs.data(), pairs.size() * sizeof(int4)))) return rc;
        if ((rc = upload((void**)&p->pair_sim, pair_sim.data(), pair_sim.size() * sizeof(int))))
            return rc;
        if ((rc = upload((void**)&p->accs, accs.data(), accs.size() * sizeof(void*)))) return rc;
    }
    p->dirty = false;
    p->args_stale = true;
    return ETL_OK;
}

template <typename R>
FusedArgs<R> make_args(const EtlPlan* p, const EtlPlan::Sim& s) {
    FusedArgs<R> a{};
    for (int k = 0; k < 7; ++k)
        for (int b = 0; b < 2; ++b) a.f[k][b] = (const R*)s.desc.field[k][b];
    const EtlTeDesc& d = s.desc;
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
    a.n_src = s.n_src;
    a.src = s.src_dev;
    a.delta = (const R*)s.delta;
    a.has_mon = s.n_slots > 0;
    for (int c = 0; c < 3; ++c) {
        a.tile_mon[c] = s.tile_mon[c];
        a.tile_ent[c] = s.tile_ent[c];
    }
    a.trace = (R*)s.trace;
    a.n_slots = s.n_slots;
    return a;
}

// Refresh the device copies of the per-simulation arguments (after the steps in
// flight, which read them).  Their addresses never change, so captured graphs stay
// valid.
template <typename R>
int upload_args_t(EtlPlan* p) {
    if (int rc_ = sync_lanes()) return rc_;
    std::vector<FusedArgs<R>> args;
    std::vector<TraceSim<R>> traces;
    for (const auto& s : p->sims) {
        args.push_back(make_args<R>(p, s));
        traces.push_back(TraceSim<R>{(const R*)s.trace, (const R*)s.ph, s.n_slots, s.nq});
    }
    int rc;
    if ((rc = etl_memcpy_h2d(p->args_dev, args.data(), args.size() * sizeof(FusedArgs<R>))))
        return rc;
    if ((rc = etl_memcpy_h2d(p->trace_sims_dev, traces.data(),
                             traces.size() * sizeof(TraceSim<R>))))
        return rc;
    p->args_stale = false;
    return ETL_OK;
}

int upload_args(EtlPlan* p) {
    return p->precision == ETL_F32 ? upload_args_t<float>(p) : upload_args_t<double>(p);
}

// After `count` steps of a trace chunk: probe copies, then the DFT fold of the trace
// rows (which advances the step counter) or a counter bump.
template <typename R>
int after_chunk(EtlPlan* p, int count, hipStream_t s) {
    if (int rc = launch_probe_copies<R>(p->probe_tab, p->n_probe_tab, p->probe_max_n, count,
                                        p->counter, s))
        return rc;
    if (p->n_pairs && count)
        return launch_trace_dft<R>(static_cast<const TraceSim<R>*>(p->trace_sims_dev), count,
                                   p->counter, p->pairs, p->pair_sim, p->n_pairs,
                                   (R* const*)p->accs, s);
    return launch_bump(p->counter, count, s);
}

// Enqueue `count` (<= trace_chunk) steps starting at buffer parity `parity`: fused
// launches of depth T, then depth-1 launches for the remainder, and the DFT of the
// trace rows (which advances the step counter) or a counter bump.  Returns the
// number of fused launches in *launches.  `events` (optional) receives four events
// per fused launch: fork, interior kernel start and end (on the side stream where
// only that kernel runs), join.
template <typename R>
int enqueue_block(EtlPlan* p, int count, int parity, hipEvent_t* events, int* launches) {
    const FusedArgs<R>* args = static_cast<const FusedArgs<R>*>(p->args_dev);
    hipStream_t s = lane_stream(p->lane);
    const int4* t = p->tiles_dev;
    const TileList tl = make_tile_list(t, p->n_interior, p->n_strip, p->n_strip_y,
                                       p->n_strip_x, p->n_extra, p->n_extra_strip,
                                       p->n_uniform);
    const int T = p->tune.temporal;
    int k = 0, n = 0;
    while (k < count) {
        const int depth =
            next_depth(count - k, T, p->precision, p->tune.vec, p->tune.band_rows);
        if (p->tune.launch == 1) {  // every tile kind in one kernel
            if (events) {
                ETL_HIP(hipEventRecord(events[4 * n], s));
                ETL_HIP(hipEventRecord(events[4 * n + 1], s));
            }
            const FwdSet<R> fs = fwd_set<R>(args, tl, p->counter, k, parity ^ (n & 1),
                                            p->tune.band_rows, p->uvals);
            if (int rc = launch_joint<R>(&fs, depth, nullptr, 0, s)) return rc;
            if (events) {
                ETL_HIP(hipEventRecord(events[4 * n + 2], s));
                ETL_HIP(hipEventRecord(events[4 * n + 3], s));
            }
            k += depth;
            ++n;
            continue;
        }
        if (events) ETL_HIP(hipEventRecord(events[4 * n], s));
        // fork: the strip tiles run on the main stream, which needs no cross-stream wait
        // and so is dispatched first — the latency-bound CPML tiles get their CUs before
        // the bulk of interior tiles fills the GPU; the interior tiles (side stream
        // s_strip) and the extra tiles (s_extra) wait for the fork; all join before the
        // next step.
        hipStream_t s_int = p->s_strip;
        ETL_HIP(hipEventRecord(p->ev_fork, s));
        ETL_HIP(hipStreamWaitEvent(s_int, p->ev_fork, 0));
        ETL_HIP(hipStreamWaitEvent(p->s_extra, p->ev_fork, 0));
        if (events) ETL_HIP(hipEventRecord(events[4 * n + 1], s_int));
        int rc = launch_fused<R>(args, tl, p->tune.vec, depth, p->tune.band_rows, p->tune.waves,
                                 p->counter, k, parity ^ (n & 1), s_int, s, p->s_extra);
        if (rc) return rc;
        if (events) ETL_HIP(hipEventRecord(events[4 * n + 2], s_int));
        ETL_HIP(hipEventRecord(p->ev_join, s_int));
        ETL_HIP(hipEventRecord(p->ev_join2, p->s_extra));
        ETL_HIP(hipStreamWaitEvent(s, p->ev_join, 0));
        ETL_HIP(hipStreamWaitEvent(s, p->ev_join2, 0));
        if (events) ETL_HIP(hipEventRecord(events[4 * n + 3], s));
        k += depth;
        ++n;
    }
    *launches = n;
    return after_chunk<R>(p, count, s);
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
// A larger table gets a new address, which the argument arrays carry.
int upload_tables(EtlPlan* p) {
    bool any = false;
    for (const auto& s : p->sims) any = any || s.delta_stale || s.ph_stale;
    if (!any) return ETL_OK;
    if (int rc_ = sync_lanes()) return rc_;
    int rc;
    for (auto& s : p->sims) {
        if (s.delta_stale) {
            const size_t bytes = s.delta_host.size();
            if (bytes > s.delta_bytes) {
                const size_t cap = size_t(table_capacity(s.delta_rows)) * s.n_src * p->elem();
                etl_free(s.delta);
                s.delta = nullptr;
                if ((rc = etl_malloc(&s.delta, cap))) return rc;
                s.delta_bytes = cap;
                p->args_stale = true;
            }
            if (bytes && (rc = etl_memcpy_h2d(s.delta, s.delta_host.data(), bytes))) return rc;
            s.delta_stale = false;
        }
        if (s.ph_stale) {
            const size_t row_bytes = size_t(s.nq) * 2 * p->elem();
            if (s.ph_host_rows > s.ph_rows_cap) {
                const int cap = table_capacity(s.ph_host_rows);
                etl_free(s.ph);
                s.ph = nullptr;
                if ((rc = etl_malloc(&s.ph, size_t(cap) * row_bytes))) return rc;
                s.ph_rows_cap = cap;
                p->args_stale = true;
            }
            if (!s.ph_host.empty() &&
                (rc = etl_memcpy_h2d(s.ph, s.ph_host.data(), s.ph_host.size())))
                return rc;
            s.ph_stale = false;
        }
    }
    return ETL_OK;
}

int ensure_graph(EtlPlan* p, int parity) {
    if (p->graph[parity]) return ETL_OK;
    hipStream_t s = lane_stream(p->lane);
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

// Every table of every simulation must cover [n, n + count) and all must start at the
// same step (one step counter indexes them all).  *n0 receives that step.
int check_tables(EtlPlan* p, int64_t n, int64_t count, int64_t* n0) {
    bool have = false;
    *n0 = n;
    auto same_start = [&](int64_t start) -> bool {
        if (!have) {
            *n0 = start;
            have = true;
            return true;
        }
        return start == *n0;
    };
    for (size_t si = 0; si < p->sims.size(); ++si) {
        const auto& s = p->sims[si];
        if (s.n_src > 0) {
            if (s.delta_rows <= 0 || !(n >= s.delta_n0 && n + count <= s.delta_n0 + s.delta_rows)) {
                set_error("etl_plan_advance: the source table of simulation %d covers steps "
                          "[%lld, %lld), asked [%lld, %lld)", int(si), (long long)s.delta_n0,
                          (long long)(s.delta_n0 + s.delta_rows), (long long)n,
                          (long long)(n + count));
                return ETL_ESTATE;
            }
            if (!same_start(s.delta_n0)) {
                set_error("etl_plan_advance: tables start at different steps");
                return ETL_ESTATE;
            }
        }
        for (const auto& mm : s.mons) {
            const auto& m = mm.owner >= 0 ? s.mons[mm.owner] : mm;
            if (m.rows < 0 || !(n >= m.n0 && n + count <= m.n0 + m.rows)) {
                set_error("etl_plan_advance: a phasor table of simulation %d does not cover "
                          "steps [%lld, %lld)", int(si), (long long)n, (long long)(n + count));
                return ETL_ESTATE;
            }
            if (!same_start(m.n0)) {
                set_error("etl_plan_advance: tables start at different steps");
                return ETL_ESTATE;
            }
        }
    }
    return ETL_OK;
}

// `counter_out` non-null: the three counter values are returned there for the caller
// to set (a co-plan sets both plans' counters in one launch) instead of launched here.
int prepare_steps(EtlPlan* p, int64_t n, int64_t count, int probe_row0,
                  int* counter_out = nullptr) {
    int rc;
    if (p->dirty) {
        if ((rc = build_layout(p))) return rc;
    }
    if ((rc = upload_tables(p))) return rc;
    int64_t n0 = n;
    if ((rc = check_tables(p, n, count, &n0))) return rc;
    if (p->args_stale) {
        if ((rc = upload_args(p))) return rc;
    }
    for (const auto& sim : p->sims)
        for (const auto& pr : sim.probes)
            ETL_REQUIRE(probe_row0 >= 0 && probe_row0 + count <= pr.out_rows,
                        "etl_plan_advance: %lld steps from probe row %d, a probe holds %d rows",
                        (long long)count, probe_row0, pr.out_rows);
    const int64_t start = n - n0;
    ETL_REQUIRE(start >= 0 && start + count < (int64_t(1) << 31), "etl_plan_advance: step range");
    const int64_t known = p->counter_at;
    p->counter_at = -1;  // until the steps are enqueued (run_steps)
    p->counter_next = start + count;
    if (counter_out) {   // the co-plan's kernels set and advance the counter
        counter_out[0] = int(start);
        counter_out[1] = 0;
        counter_out[2] = int(start) - probe_row0;
        return ETL_OK;
    }
    // The DFT (or bump) kernel that ends the previous advance leaves counter[0] at its
    // end step and counter[1] at 0: an advance that continues from there, with the same
    // probe offset or no probes, finds the counter set.
    bool probes = false;
    for (const auto& sim : p->sims) probes = probes || !sim.probes.empty();
    if (known == start && (!probes || p->counter_probe == int(start) - probe_row0))
        return ETL_OK;
    hipStream_t s = lane_stream(p->lane);
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(64), 0, s, p->counter, int(start),
                       probe_row0);
    ETL_LAUNCHED();
    p->counter_probe = int(start) - probe_row0;
    return ETL_OK;
}

int run_steps(EtlPlan* p, int64_t n, int64_t count, hipEvent_t* events, int* n_launch,
              int probe_row0 = 0) {
    int rc;
    if ((rc = prepare_steps(p, n, count, probe_row0))) return rc;
    hipStream_t s = lane_stream(p->lane);
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
    p->counter_at = p->counter_next;  // each block's DFT or bump kernel added its rows
    return ETL_OK;
}

// Elements of field k (ex ey hz psi_ex_y psi_ey_x psi_hz_x psi_hz_y) of a simulation.
int64_t field_elems(const EtlPlan::Sim& s, int k) {
    static const int comp_of[7] = {0, 1, 2, 0, 1, 2, 2};
    return int64_t(s.rows_of(comp_of[k])) * s.pitch_of(comp_of[k]);
}

// Component (0 Ex, 1 Ey, 2 Hz) of field k (ex ey hz psi_ex_y psi_ey_x psi_hz_x psi_hz_y).
int comp_of_field(int k) {
    static const int comp[7] = {0, 1, 2, 0, 1, 2, 2};
    return comp[k];
}

// (Re)upload the checkpoint copy entries of every simulation.
int ck_upload(EtlPlan* p) {
    std::vector<CkEntry> tab;
    const int64_t w = p->elem() / 4;  // words per REAL
    for (const auto& sim : p->sims)
        for (int k = 0; k < 7; ++k) {
            CkEntry e{static_cast<uint32_t*>(sim.ck[k]),
                      {static_cast<uint32_t*>(sim.desc.field[k][0]),
                       static_cast<uint32_t*>(sim.desc.field[k][1])},
                      field_elems(sim, k) * w, 0, 0, 0, 0};
            if (sim.ck_window) {
                const int c = comp_of_field(k);
                const int64_t rows = c == 1 ? sim.ck_nx + 1 : sim.ck_nx;
                const int64_t cols = c == 0 ? sim.ck_ny + 1 : sim.ck_ny;
                e.words = rows * cols * w;
                e.rows = rows;
                e.row_words = cols * w;
                e.pitch = int64_t(sim.pitch_of(c)) * w;
                e.off = (int64_t(sim.ck_row0) * sim.pitch_of(c) + sim.ck_col0) * w;
            }
            tab.push_back(e);
        }
    etl_free(p->ck_tab);
    p->ck_tab = nullptr;
    if (int rc = upload((void**)&p->ck_tab, tab.data(), tab.size() * sizeof(CkEntry))) return rc;
    p->ck_entries = int(tab.size());
    p->ck_stale = false;
    return ETL_OK;
}

// Save the state into (restore = 0) or load it from (restore = 1) checkpoint slot `seg`
// of every simulation, on the plan's lane, in the buffers of the current parity.
int ck_copy(EtlPlan* p, int seg, int restore) {
    for (size_t k = 0; k < p->sims.size(); ++k) {
        ETL_REQUIRE(p->sims[k].ck_seg > seg && seg >= 0,
                    "checkpoint %d: simulation %d has %d checkpoint slots", seg, int(k),
                    p->sims[k].ck_seg);
    }
    if (p->ck_stale)
        if (int rc = ck_upload(p)) return rc;
    // one 16-byte move per thread and block row for the largest field (one block per
    // row of the largest window), at most 2048 blocks
    int64_t blocks = 1;
    for (const auto& sim : p->sims)
        for (int k = 0; k < 7; ++k)
            blocks = std::max(blocks, sim.ck_window
                                          ? int64_t(comp_of_field(k) == 1 ? sim.ck_nx + 1 : sim.ck_nx)
                                          : (field_elems(sim, k) * p->elem() / 4 / 4 + 255) / 256);
    const int gx = int(std::min<int64_t>(2048, blocks));
    hipLaunchKernelGGL(ckpt_copy_kernel, dim3(unsigned(gx), unsigned(p->ck_entries)), dim3(256),
                       0, lane_stream(p->lane), p->ck_tab, seg, p->parity, restore);
    ETL_LAUNCHED();
    return ETL_OK;
}

// Validate one simulation's descriptor and find its CPML strips; checks that psi is +0
// wherever the fused kernel leaves it untouched.
int init_sim(EtlPlan* p, const EtlTeDesc* desc, EtlPlan::Sim& sim) {
    ETL_REQUIRE(desc->precision == p->precision,
                "etl_plan_create: every simulation of a plan needs the same precision");
    ETL_REQUIRE(desc->nx >= 1 && desc->ny >= 1 && desc->nx < (1 << 24) && desc->ny < (1 << 24),
                "etl_plan_create: bad grid %d x %d", desc->nx, desc->ny);
    for (int k = 0; k < 7; ++k)
        ETL_REQUIRE(desc->field[k][0] && desc->field[k][1] && desc->field[k][0] != desc->field[k][1],
                    "etl_plan_create: field %d needs two distinct buffers", k);
    ETL_REQUIRE(desc->cb_ex && desc->cb_ey && desc->bx_e && desc->cx_e && desc->ikx_e &&
                    desc->bx_h && desc->cx_h && desc->ikx_h && desc->by_e && desc->cy_e &&
                    desc->iky_e && desc->by_h && desc->cy_h && desc->iky_h,
                "etl_plan_create: null coefficient array");
    sim.nx = desc->nx;
    sim.ny = desc->ny;
    sim.desc = *desc;
    const int nx = sim.nx, ny = sim.ny;
    const bool f32 = p->precision == ETL_F32;
    // The fused kernel loads whole vectors that start up to `temporal` elements before
    // a row and may end past it: every array must carry etl_malloc's guard bands and be
    // at least its shape's size.
    {
        const size_t el = p->elem();
        const size_t sz_ex = size_t(nx) * (ny + 1) * el, sz_ey = size_t(nx + 1) * ny * el,
                     sz_hz = size_t(nx) * ny * el;
        const size_t need[7] = {sz_ex, sz_ey, sz_hz, sz_ex, sz_ey, sz_hz, sz_hz};
        bool ok = guarded_bytes(desc->cb_ex) >= sz_ex && guarded_bytes(desc->cb_ey) >= sz_ey;
        for (int k = 0; k < 7; ++k)
            for (int b = 0; b < 2; ++b) ok = ok && guarded_bytes(desc->field[k][b]) >= need[k];
        if (!ok) {
            set_error("etl_plan_create: field and cb arrays must be etl_malloc allocations of "
                      "their full shape (the fused kernel reads into their guard bands)");
            return ETL_EARG;
        }
    }
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
    int rc = f32 ? bounds(float(0), sim.strips) : bounds(double(0), sim.strips);
    if (rc) return rc;
    // psi must be +0 wherever the fused kernel does not touch it, and no field -0.
    int* flag = nullptr;
    if ((rc = etl_malloc((void**)&flag, 256))) return rc;
    rc = zero_async(flag, 8);
    struct {
        int field, rows, cols, lo_r, hi_r, lo_c, hi_c, er, ec;
    } chk[4] = {
        {3, nx, ny + 1, 0, nx, sim.strips[6], sim.strips[7], 0, 1},  // psi_ex_y
        {4, nx + 1, ny, sim.strips[2], sim.strips[3], 0, ny, 1, 0},  // psi_ey_x
        {5, nx, ny, sim.strips[0], sim.strips[1], 0, ny, 0, 0},      // psi_hz_x
        {6, nx, ny, 0, nx, sim.strips[4], sim.strips[5], 0, 0},      // psi_hz_y
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
    const int64_t n_field[3] = {int64_t(nx) * (ny + 1), int64_t(nx + 1) * ny, int64_t(nx) * ny};
    for (int k = 0; k < 3 && rc == ETL_OK; ++k) {
        const dim3 grid(unsigned((n_field[k] + 255) / 256));
        if (f32)
            hipLaunchKernelGGL(negzero_check_kernel<float>, grid, dim3(256), 0, ctx().stream,
                               (const float*)desc->field[k][0], n_field[k], flag);
        else
            hipLaunchKernelGGL(negzero_check_kernel<double>, grid, dim3(256), 0, ctx().stream,
                               (const double*)desc->field[k][0], n_field[k], flag);
        if (hipGetLastError() != hipSuccess) rc = ETL_EHIP;
    }
    int host_flag[2] = {0, 0};
    if (rc == ETL_OK) rc = etl_memcpy_d2h(host_flag, flag, 8);
    etl_free(flag);
    if (rc == ETL_OK && host_flag[0]) {
        set_error("etl_plan_create: psi is non-zero outside the CPML strips; the fused "
                  "path needs psi = +0 there (use the per-half-step primitives)");
        rc = ETL_ESTATE;
    }
    if (rc == ETL_OK && host_flag[1]) {
        set_error("etl_plan_create: a field holds -0; the fused path needs fields without "
                  "-0 (use the per-half-step primitives)");
        rc = ETL_ESTATE;
    }
    return rc;
}

void destroy_plan(EtlPlan* p) {
    free_graphs(p);
    free_layout(p);
    for (auto& s : p->sims) {
        etl_free(s.src_dev);
        etl_free(s.delta);
        etl_free(s.ph);
    }
    etl_free(p->args_dev);
    etl_free(p->trace_sims_dev);
    etl_free(p->counter);
    etl_free(p->ck_tab);
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

EtlPlan::Sim* sim_of(EtlPlan* p, int sim) {
    if (p == nullptr || sim < 0 || sim >= int(p->sims.size())) return nullptr;
    return &p->sims[sim];
}

}  // namespace

namespace etl {

int fwd_prepare(EtlPlan* p, int64_t n, int64_t count, int probe_row0, int* counter_out) {
    return prepare_steps(p, n, count, probe_row0, counter_out);
}

void fwd_info(const EtlPlan* p, FwdInfo* out) {
    const int4* t = p->tiles_dev;
    out->precision = p->precision;
    out->temporal = p->tune.temporal;
    out->trace_chunk = p->tune.trace_chunk;
    out->launch = p->tune.launch;
    out->lane = p->lane;
    out->vec = p->tune.vec;
    out->band = p->tune.band_rows;
    out->args = p->args_dev;
    out->tl = make_tile_list(t, p->n_interior, p->n_strip, p->n_strip_y, p->n_strip_x, p->n_extra,
                                  p->n_extra_strip);
    out->counter = p->counter;
    out->parity = p->parity;
    out->gen = p->layout_gen;
}

void fwd_set_parity(EtlPlan* p, int parity) { p->parity = parity; }

int fwd_checkpoint(EtlPlan* p, int seg, int restore) { return ck_copy(p, seg, restore); }

int fwd_after_chunk(EtlPlan* p, int rows, hipStream_t s) {
    return p->precision == ETL_F32 ? after_chunk<float>(p, rows, s)
                                   : after_chunk<double>(p, rows, s);
}

}  // namespace etl

extern "C" {

namespace {

// Large plans (the rule in fused.h): vec 1, not launch mode 3.
bool large_plan(const EtlTeDesc* descs, int n_sims, const EtlTune& t) {
    return (t.vec == 0 || t.vec == 1) && t.launch != 3 &&
           large_grids(descs, n_sims, ctx().cu_count);
}

}  // namespace

// Whether a float plan runs temporal depth 8 by default: the fields a step block
// streams (both buffers of Ex, Ey, Hz and cb_ex, cb_ey: 32 bytes per cell) exceed the
// 256 MiB Infinity Cache, so that every launch reads them from HBM, and every grid has
// at least 4 M cells, so that interior tiles dominate.  Depth 8 moves the same window
// bytes per launch for 8 steps instead of 6 (8024^2: 73 against 86 us per step); where
// the fields fit the cache it measured level with depth 6 (2024^2: 5.86 us per step
// both), and on batches of smaller grids, a third of whose tiles are CPML strip and
// source sub-tiles with the wider halo, slower (32 x 1024^2: 0.493 against 0.47 s for
// 8000 steps).
static bool deep_plan(const EtlTeDesc* descs, int n_sims) {
    int64_t cells = 0, smallest = INT64_MAX;
    for (int k = 0; k < n_sims; ++k) {
        const int64_t c = int64_t(descs[k].nx) * descs[k].ny;
        cells += c;
        smallest = std::min(smallest, c);
    }
    return cells * 32 > (int64_t(256) << 20) && smallest >= (int64_t(4) << 20);
}

int etl_plan_create_batch(const EtlTeDesc* descs, int n_sims, const EtlTune* tune,
                          etl_plan_t* out) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(descs && out, "etl_plan_create: null argument");
    ETL_REQUIRE(n_sims >= 1 && n_sims <= (1 << 20), "etl_plan_create: %d simulations", n_sims);
    const int precision = descs[0].precision;
    ETL_REQUIRE(precision == ETL_F32 || precision == ETL_F64,
                "etl_plan_create: unknown precision %d", precision);
    EtlTune t{};
    if (tune) t = *tune;
    if (t.vec == 0) t.vec = 1;
    const bool large = large_plan(descs, n_sims, t);
    const int large_band = 16;
    if (t.temporal == 0)
        t.temporal = large && (t.band_rows == 0 || t.band_rows == large_band)
                         ? (precision == ETL_F32 && deep_plan(descs, n_sims) ? 8 : 6)
                         : 4;
    if (t.band_rows == 0) t.band_rows = large ? large_band : 8;
    if (t.trace_chunk == 0) t.trace_chunk = t.temporal >= 6 ? 96 : 64;
    if (t.waves == 0) t.waves = 1;
    if (tune == nullptr) t.use_graph = 1;
    ETL_REQUIRE(t.uniform_cb >= -1 && t.uniform_cb <= 1, "etl_plan_create: uniform_cb %d",
                t.uniform_cb);
    ETL_REQUIRE(tile_shape_supported(precision, t.vec, t.temporal, t.band_rows) &&
                    tile_shape_supported(precision, t.vec, 1, t.band_rows),
                "etl_plan_create: no compiled tile shape vec=%d temporal=%d band_rows=%d",
                t.vec, t.temporal, t.band_rows);
    ETL_REQUIRE(t.trace_chunk >= 2 && t.trace_chunk <= 4096 &&
                    t.trace_chunk % (2 * t.temporal) == 0,
                "etl_plan_create: trace_chunk must be a multiple of 2 * temporal, <= 4096");
    ETL_REQUIRE(t.waves == 1 || t.waves == 2 || t.waves == 4,
                "etl_plan_create: waves per workgroup must be 1, 2 or 4 (got %d)", t.waves);
    ETL_REQUIRE(t.launch == 0 || t.launch == 1 || t.launch == 3,
                "etl_plan_create: launch must be 0, 1 or 3 (got %d)", t.launch);
    ETL_REQUIRE(t.temporal < 6 || t.launch != 3,
                "etl_plan_create: depth %d runs with one launch per step block (launch 1)",
                t.temporal);
    ETL_REQUIRE(t.launch != 1 || (t.vec == 1 && (t.band_rows == 8 || t.band_rows == 16)),
                "etl_plan_create: one launch per step block needs vec 1 and band 8 or 16");

    EtlPlan* p = new EtlPlan();
    p->precision = precision;
    p->tune = t;
    p->sims.resize(n_sims);
    int rc = ETL_OK;
    for (int k = 0; k < n_sims && rc == ETL_OK; ++k) rc = init_sim(p, &descs[k], p->sims[k]);
    // one launch per step block wherever the joint kernel is compiled for the shape
    if (rc == ETL_OK && p->tune.launch == 0)
        p->tune.launch =
            (t.vec == 1 && (t.band_rows == 8 || t.band_rows == 16)) ? 1 : 3;
    // uniform-material tiles: the joint kernel's kind 6, so launch mode 1 only
    if (p->tune.uniform_cb == 0) p->tune.uniform_cb = 1;
    if (p->tune.launch != 1) p->tune.uniform_cb = -1;
    // side streams and their events: only launch mode 3 (one kernel per tile kind)
    // forks onto them
    if (rc == ETL_OK && p->tune.launch == 3) {
        if (hipStreamCreateWithFlags(&p->s_strip, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&p->s_extra, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join2, hipEventDisableTiming) != hipSuccess) {
            set_error("etl_plan_create: stream/event creation failed");
            rc = ETL_EHIP;
        }
    }
    const size_t args_bytes =
        size_t(n_sims) * (precision == ETL_F32 ? sizeof(FusedArgs<float>) : sizeof(FusedArgs<double>));
    const size_t trace_bytes =
        size_t(n_sims) * (precision == ETL_F32 ? sizeof(TraceSim<float>) : sizeof(TraceSim<double>));
    if (rc == ETL_OK) rc = etl_malloc((void**)&p->counter, 256);
    if (rc == ETL_OK) rc = etl_malloc(&p->args_dev, args_bytes);
    if (rc == ETL_OK) rc = etl_malloc(&p->trace_sims_dev, trace_bytes);
    // the depths its step blocks and tails launch, dispatched once empty (joint.h)
    if (rc == ETL_OK && p->tune.launch == 1)
        for (int d : {8, 6, 4, 2, 1})  // the plan's depth and its tails' (next_depth)
            if (rc == ETL_OK && d <= p->tune.temporal &&
                tile_shape_supported(precision, 1, d, p->tune.band_rows))
                rc = precision == ETL_F32
                         ? launch_joint_fwd_empty<float>(d, p->tune.band_rows, lane_stream(p->lane))
                         : launch_joint_fwd_empty<double>(d, p->tune.band_rows,
                                                          lane_stream(p->lane));
    if (rc != ETL_OK) {
        destroy_plan(p);
        return rc;
    }
    *out = p;
    return ETL_OK;
}

int etl_plan_create(const EtlTeDesc* desc, const EtlTune* tune, etl_plan_t* out) {
    return etl_plan_create_batch(desc, 1, tune, out);
}

int etl_plan_destroy(etl_plan_t p) {
    if (p == nullptr) return ETL_OK;
    ETL_CHECK_INIT();
    if (int rc_ = sync_lanes()) return rc_;
    destroy_plan(p);
    return ETL_OK;
}

int etl_plan_touch(etl_plan_t p) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_plan_touch: null plan");
    // only the uniform-material classification depends on the cb values: a plan
    // without it reads the cb rows everywhere and keeps its layout and graphs
    if (p->tune.uniform_cb == 1) p->dirty = true;
    return ETL_OK;
}

int etl_plan_set_sources_sim(etl_plan_t p, int sim, int n_nodes, const int* comp, const int64_t* i,
                             const int64_t* j) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr, "etl_plan_set_sources: no simulation %d", sim);
    ETL_REQUIRE(n_nodes >= 0 && n_nodes <= kMaxSources,
                "etl_plan_set_sources: %d source nodes (at most %d)", n_nodes, kMaxSources);
    ETL_REQUIRE(n_nodes == 0 || (comp && i && j), "etl_plan_set_sources: null array");
    std::vector<int4> src;
    for (int k = 0; k < n_nodes; ++k) {
        ETL_REQUIRE(comp[k] >= 0 && comp[k] <= 2, "etl_plan_set_sources: component %d", comp[k]);
        ETL_REQUIRE(i[k] >= 0 && i[k] < s->rows_of(comp[k]) && j[k] >= 0 &&
                        j[k] < s->pitch_of(comp[k]),
                    "etl_plan_set_sources: node (%lld, %lld) outside component %d",
                    (long long)i[k], (long long)j[k], comp[k]);
        // Outer tangential E never changes in the update; a source there would be
        // the only writer, which the fused step does not model.
        const bool outer = (comp[k] == 0 && (j[k] == 0 || j[k] == s->ny)) ||
                           (comp[k] == 1 && (i[k] == 0 || i[k] == s->nx));
        ETL_REQUIRE(!outer, "etl_plan_set_sources: node on the outer tangential-E boundary");
        src.push_back(make_int4(comp[k], int(i[k]), int(j[k]), 0));
    }
    if (int rc_ = sync_lanes()) return rc_;
    etl_free(s->src_dev);
    s->src_dev = nullptr;
    if (n_nodes) {
        int rc = upload((void**)&s->src_dev, src.data(), src.size() * sizeof(int4));
        if (rc) return rc;
    }
    s->src = src;
    s->n_src = n_nodes;
    s->delta_rows = 0;  // a table of the old source set no longer applies
    p->dirty = true;
    p->args_stale = true;
    free_graphs(p);
    return ETL_OK;
}

int etl_plan_set_sources(etl_plan_t p, int n_nodes, const int* comp, const int64_t* i,
                         const int64_t* j) {
    return etl_plan_set_sources_sim(p, 0, n_nodes, comp, i, j);
}

int etl_plan_source_table_sim(etl_plan_t p, int sim, int64_t n0, int rows, const void* delta) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr, "etl_plan_source_table: no simulation %d", sim);
    ETL_REQUIRE(rows >= 0 && n0 >= 0, "etl_plan_source_table: bad rows");
    if (s->n_src == 0) return ETL_OK;
    ETL_REQUIRE(delta != nullptr, "etl_plan_source_table: null table");
    const size_t bytes = size_t(rows) * s->n_src * p->elem();
    const unsigned char* b = static_cast<const unsigned char*>(delta);
    s->delta_host.assign(b, b + bytes);
    s->delta_n0 = n0;
    s->delta_rows = rows;
    s->delta_stale = true;
    return ETL_OK;
}

int etl_plan_source_table(etl_plan_t p, int64_t n0, int rows, const void* delta) {
    return etl_plan_source_table_sim(p, 0, n0, rows, delta);
}

int etl_plan_add_monitor_sim(etl_plan_t p, int sim, int comp, int n_nodes,
                             const int64_t* flat_index, int n_f, void* acc, int* id) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr && id, "etl_plan_add_monitor: no simulation %d or null id", sim);
    ETL_REQUIRE(comp >= 0 && comp <= 2, "etl_plan_add_monitor: component %d", comp);
    ETL_REQUIRE(n_nodes >= 1 && n_f >= 1 && flat_index && acc,
                "etl_plan_add_monitor: bad monitor (%d nodes, %d frequencies)", n_nodes, n_f);
    const int64_t size = int64_t(s->rows_of(comp)) * s->pitch_of(comp);
    EtlPlan::Mon m;
    m.comp = comp;
    m.n_nodes = n_nodes;
    m.n_f = n_f;
    m.acc = acc;
    m.flat.assign(flat_index, flat_index + n_nodes);
    const int pitch = s->pitch_of(comp);
    for (int64_t f : m.flat) {
        ETL_REQUIRE(f >= 0 && f < size, "etl_plan_add_monitor: index %lld outside component %d",
                    (long long)f, comp);
        const int64_t row = f / pitch;
        ETL_REQUIRE(!(comp == 1 && row == s->nx),
                    "etl_plan_add_monitor: node on the outer tangential-E boundary");
    }
    if (int rc_ = sync_lanes()) return rc_;
    s->mons.push_back(m);
    p->dirty = true;
    *id = int(s->mons.size()) - 1;
    // This simulation's combined phasor table changes shape: drop it.
    etl_free(s->ph);
    s->ph = nullptr;
    s->ph_rows_cap = 0;
    s->ph_host.clear();
    s->ph_host_rows = 0;
    s->ph_stale = false;
    for (auto& mm : s->mons) mm.rows = -1;
    p->args_stale = true;
    free_graphs(p);
    return ETL_OK;
}

int etl_plan_add_probe_sim(etl_plan_t p, int sim, int comp, int n_nodes,
                           const int64_t* flat_index, void* out, int out_rows) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr, "etl_plan_add_probe: no simulation %d", sim);
    ETL_REQUIRE(comp >= 0 && comp <= 2, "etl_plan_add_probe: component %d", comp);
    ETL_REQUIRE(n_nodes >= 1 && flat_index && out && out_rows >= 1,
                "etl_plan_add_probe: bad probe (%d nodes, %d rows)", n_nodes, out_rows);
    const int64_t size = int64_t(s->rows_of(comp)) * s->pitch_of(comp);
    const int pitch = s->pitch_of(comp);
    EtlPlan::Probe pr;
    pr.comp = comp;
    pr.flat.assign(flat_index, flat_index + n_nodes);
    for (int64_t f : pr.flat) {
        ETL_REQUIRE(f >= 0 && f < size, "etl_plan_add_probe: index %lld outside component %d",
                    (long long)f, comp);
        ETL_REQUIRE(!(comp == 1 && f / pitch == s->nx),
                    "etl_plan_add_probe: node on the outer tangential-E boundary");
    }
    ETL_REQUIRE(guarded_bytes(out) >= size_t(out_rows) * n_nodes * p->elem(),
                "etl_plan_add_probe: out must be an etl_malloc allocation of out_rows x n_nodes");
    pr.out = out;
    pr.out_rows = out_rows;
    if (int rc_ = sync_lanes()) return rc_;
    s->probes.push_back(std::move(pr));
    p->dirty = true;
    p->args_stale = true;
    free_graphs(p);
    return ETL_OK;
}

int etl_plan_add_monitor(etl_plan_t p, int comp, int n_nodes, const int64_t* flat_index, int n_f,
                         void* acc, int* id) {
    return etl_plan_add_monitor_sim(p, 0, comp, n_nodes, flat_index, n_f, acc, id);
}

int etl_plan_phasor_table_sim(etl_plan_t p, int sim, int id, int64_t n0, int rows,
                              const void* ph) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr, "etl_plan_phasor_table: no simulation %d", sim);
    ETL_REQUIRE(id >= 0 && id < int(s->mons.size()), "etl_plan_phasor_table: no monitor %d", id);
    ETL_REQUIRE(rows >= 0 && n0 >= 0 && (ph || rows == 0), "etl_plan_phasor_table: bad table");
    int rc;
    if (p->dirty) {
        if ((rc = build_layout(p))) return rc;  // assigns the monitors' phasor columns
    }
    const size_t cplx = size_t(2) * p->elem();
    const size_t row_bytes = size_t(s->nq) * cplx;
    if (rows > s->ph_host_rows) {  // rows are appended: the other monitors' columns stay
        s->ph_host.resize(size_t(rows) * row_bytes);
        s->ph_host_rows = rows;
    }
    auto& m = s->mons[id];
    ETL_REQUIRE(m.owner < 0, "etl_plan_phasor_table: monitor %d reads the table of monitor %d",
                id, m.owner);
    const unsigned char* src = static_cast<const unsigned char*>(ph);
    for (int r = 0; r < rows; ++r)
        std::memcpy(s->ph_host.data() + size_t(r) * row_bytes + size_t(m.q_base) * cplx,
                    src + size_t(r) * m.n_f * cplx, size_t(m.n_f) * cplx);
    m.n0 = n0;
    m.rows = rows;
    s->ph_stale = true;
    return ETL_OK;
}

int etl_plan_share_phasors(etl_plan_t p, int sim, int id, int owner) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr, "etl_plan_share_phasors: no simulation %d", sim);
    ETL_REQUIRE(owner >= 0 && owner < id && id < int(s->mons.size()),
                "etl_plan_share_phasors: monitor %d cannot share the table of monitor %d", id,
                owner);
    ETL_REQUIRE(s->mons[owner].owner < 0 && s->mons[owner].n_f == s->mons[id].n_f,
                "etl_plan_share_phasors: monitor %d must own a table of %d frequencies", owner,
                s->mons[id].n_f);
    if (int rc_ = sync_lanes()) return rc_;
    s->mons[id].owner = owner;
    // the combined table loses monitor id's columns: drop it
    etl_free(s->ph);
    s->ph = nullptr;
    s->ph_rows_cap = 0;
    s->ph_host.clear();
    s->ph_host_rows = 0;
    s->ph_stale = false;
    for (auto& mm : s->mons) mm.rows = -1;
    p->dirty = true;
    p->args_stale = true;
    free_graphs(p);
    return ETL_OK;
}

int etl_plan_phasor_table(etl_plan_t p, int id, int64_t n0, int rows, const void* ph) {
    return etl_plan_phasor_table_sim(p, 0, id, n0, rows, ph);
}

int etl_plan_tune(etl_plan_t p, EtlTune* out) {
    ETL_REQUIRE(p && out, "etl_plan_tune: null argument");
    *out = p->tune;
    return ETL_OK;
}

int etl_plan_prepare(etl_plan_t p) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_plan_prepare: null plan");
    return p->dirty ? build_layout(p) : ETL_OK;
}

int etl_plan_advance(etl_plan_t p, int64_t n, int64_t count, int* parity) {
    return etl_plan_advance_rows(p, n, count, 0, parity);
}

int etl_plan_advance_rows(etl_plan_t p, int64_t n, int64_t count, int probe_row0, int* parity) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_plan_advance: null plan");
    ETL_REQUIRE(n >= 0 && count >= 0, "etl_plan_advance: bad step range");
    if (count > 0) {
        int rc = run_steps(p, n, count, nullptr, nullptr, probe_row0);
        if (rc) return rc;
    }
    if (parity) *parity = p->parity;
    return ETL_OK;
}

int etl_plan_set_checkpoints_sim(etl_plan_t p, int sim, void* const* ptr, int n_seg) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr && ptr != nullptr && n_seg >= 1,
                "etl_plan_set_checkpoints: bad arguments (simulation %d, %d slots)", sim, n_seg);
    for (int k = 0; k < 7; ++k)
        ETL_REQUIRE(ptr[k] != nullptr && guarded_bytes(ptr[k]) >=
                        size_t(n_seg) * size_t(field_elems(*s, k)) * p->elem(),
                    "etl_plan_set_checkpoints: field %d needs an etl_malloc array of %d x %lld "
                    "REALs", k, n_seg, (long long)field_elems(*s, k));
    for (int k = 0; k < 7; ++k) s->ck[k] = ptr[k];
    s->ck_seg = n_seg;
    s->ck_window = false;
    p->ck_stale = true;
    return ETL_OK;
}

int etl_plan_set_checkpoint_window_sim(etl_plan_t p, int sim, void* const* ptr, int n_seg,
                                       int win_nx, int win_ny, int row0, int col0) {
    ETL_CHECK_INIT();
    EtlPlan::Sim* s = sim_of(p, sim);
    ETL_REQUIRE(s != nullptr && ptr != nullptr && n_seg >= 1,
                "etl_plan_set_checkpoint_window: bad arguments (simulation %d, %d slots)", sim,
                n_seg);
    ETL_REQUIRE(win_nx >= 1 && win_ny >= 1 && row0 >= 0 && col0 >= 0 &&
                    row0 + win_nx <= s->nx && col0 + win_ny <= s->ny,
                "etl_plan_set_checkpoint_window: a %d x %d window from (%d, %d) does not fit "
                "the %d x %d grid", win_nx, win_ny, row0, col0, s->nx, s->ny);
    for (int k = 0; k < 7; ++k) {
        const int c = comp_of_field(k);
        const int64_t win = int64_t(c == 1 ? win_nx + 1 : win_nx) * (c == 0 ? win_ny + 1 : win_ny);
        ETL_REQUIRE(ptr[k] != nullptr && guarded_bytes(ptr[k]) >= size_t(n_seg) * size_t(win) * p->elem(),
                    "etl_plan_set_checkpoint_window: field %d needs an etl_malloc array of %d x "
                    "%lld REALs", k, n_seg, (long long)win);
    }
    for (int k = 0; k < 7; ++k) s->ck[k] = ptr[k];
    s->ck_seg = n_seg;
    s->ck_window = true;
    s->ck_nx = win_nx;
    s->ck_ny = win_ny;
    s->ck_row0 = row0;
    s->ck_col0 = col0;
    p->ck_stale = true;
    return ETL_OK;
}

int etl_plan_checkpoint(etl_plan_t p, int seg, int restore) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr && (restore == 0 || restore == 1),
                "etl_plan_checkpoint: bad arguments");
    return ck_copy(p, seg, restore);
}

int etl_plan_advance_checkpointed(etl_plan_t p, int64_t n, int64_t count, int K, int* parity) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr && K >= 1 && n >= 0 && count >= 0 && n % K == 0,
                "etl_plan_advance_checkpointed: bad arguments (n %lld, K %d)", (long long)n, K);
    for (int64_t c = n; c < n + count; c += K) {
        if (int rc = ck_copy(p, int(c / K), 0)) return rc;
        if (int rc = run_steps(p, c, std::min<int64_t>(K, n + count - c), nullptr, nullptr))
            return rc;
    }
    if (parity) *parity = p->parity;
    return ETL_OK;
}

int etl_plan_set_lane(etl_plan_t p, int lane) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr && lane >= 0 && lane < kLanes, "etl_plan_set_lane: bad arguments");
    if (lane != p->lane) {
        if (int rc = sync_lanes()) return rc;
        p->lane = lane;
    }
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
    hipStream_t s = lane_stream(p->lane);
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
        out->kernel_ms = (p->n_interior || p->tune.launch == 1) ? interior / L : 0.0;
        out->total_ms = tot;
        int64_t all_cells = 0;
        for (const auto& sim : p->sims) all_cells += int64_t(sim.nx) * sim.ny;
        out->kernel_cells = p->tune.launch == 1 ? all_cells : p->interior_cells;
        out->launches = launches;
        out->depth = p->tune.temporal;
        out->launch = p->tune.launch;
        out->uniform_cells = int64_t(p->n_uniform) * p->tune.band_rows * p->wo();
        if (parity) *parity = p->parity;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return rc;
}

}  // extern "C"
