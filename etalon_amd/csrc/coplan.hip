// Co-plans: a forward plan and an adjoint plan advanced together, step block i of
// each in one launch of the joint kernel (joint.h) — e.g. the recompute of checkpoint
// segment k - 1 beside the reverse steps of segment k, whose tiles are independent.
// Replayed from a graph per (counts, parities, layouts).  No reference counterpart
// (the reference has no adjoint).
#include <vector>

#include "coplan.h"
#include "joint.h"

using namespace etl;

struct EtlCoPlan {
    EtlPlan* f = nullptr;
    EtlAdjPlan* a = nullptr;
    struct Entry {
        int count_f, count_a, fpar, apar;
        unsigned fgen, agen;
        hipGraph_t src;
        hipGraphExec_t exec;
        int f_launches, a_launches;
    };
    std::vector<Entry> graphs;
};

namespace {

// The step counter of the forward plan and the step record of the adjoint plan of a
// co-plan run, in one launch.
__global__ void co_set_kernel(int* counter, int c0, int c1, int c2, int* step, int s0,
                              int s1) {
    if (threadIdx.x == 0) {
        if (counter) {
            counter[0] = c0;
            counter[1] = c1;
            counter[2] = c2;
        }
        if (step) {
            step[0] = s0;
            step[1] = s1;
        }
    }
}

void free_co_graphs(EtlCoPlan* c) {
    for (auto& e : c->graphs) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.src);
    }
    c->graphs.clear();
}

// Enqueue count_f forward steps (trace chunks with their probe copies and DFT fold)
// and count_a reverse steps: blocks of equal depth pair up in one launch, a block
// without a partner of its depth runs alone.
template <typename R>
int co_enqueue(EtlCoPlan* c, const FwdInfo& fi, int count_f, const AdjInfo& ai, int count_a,
               hipStream_t s, int* f_launches, int* a_launches) {
    struct FBlock {
        int depth, kstep0, chunk_rows;  // chunk_rows > 0: the chunk ends with this block
    };
    std::vector<FBlock> fb;
    for (int done = 0; done < count_f;) {
        const int blk = std::min(fi.trace_chunk, count_f - done);
        for (int k = 0; k < blk;) {
            const int d = next_depth(blk - k, fi.temporal, fi.precision, fi.vec, fi.band);
            fb.push_back({d, k, 0});
            k += d;
        }
        fb.back().chunk_rows = blk;
        done += blk;
    }
    std::vector<int2> ab;  // (depth, rtop)
    for (int rtop = count_a - 1; rtop >= 0;) {
        const int d = next_depth(rtop + 1, ai.temporal, ai.precision, 1, ai.band);
        ab.push_back(make_int2(d, rtop));
        rtop -= d;
    }
    const bool split = fi.launch != 1 || ai.launch != 1;
    size_t i = 0, j = 0;
    int nf = 0, na = 0;
    while (i < fb.size() || j < ab.size()) {
        const int df = i < fb.size() ? fb[i].depth : 0;
        const int da = j < ab.size() ? ab[j].x : 0;
        const bool do_f = df > 0 && df >= da;
        const bool do_a = da > 0 && da >= df;
        FwdSet<R> fs{};
        AdjSet<R> as{};
        if (do_f)
            fs = fwd_set<R>(static_cast<const FusedArgs<R>*>(fi.args), fi.tl, fi.counter,
                            fb[i].kstep0, fi.parity ^ (nf & 1), fi.band);
        if (do_a)
            as = adj_set<R>(static_cast<const AdjTileArgs<R>*>(ai.args), ai.tl, ai.step, ab[j].y,
                            ai.parity ^ (na & 1));
        // Small plans (launch mode 1): one launch over every tile kind.  Large ones: the
        // interior tiles of both plans in one launch (few registers, many waves per CU
        // for the bandwidth-bound bulk), then the strip and source/monitor tiles.
        if (split) {
            for (int part : {int(kInteriorOnly), int(kEdgeOnly)})
                if (int rc = launch_joint<R>(do_f ? &fs : nullptr, do_f ? df : 0,
                                             do_a ? &as : nullptr, do_a ? da : 0, s, part))
                    return rc;
        } else if (int rc = launch_joint<R>(do_f ? &fs : nullptr, do_f ? df : 0,
                                            do_a ? &as : nullptr, do_a ? da : 0, s)) {
            return rc;
        }
        if (do_f) {
            if (fb[i].chunk_rows)
                if (int rc = fwd_after_chunk(c->f, fb[i].chunk_rows, s)) return rc;
            ++i;
            ++nf;
        }
        if (do_a) {
            ++j;
            ++na;
        }
    }
    *f_launches = nf;
    *a_launches = na;
    return ETL_OK;
}

int co_run(EtlCoPlan* c, int64_t n, int count_f, int probe_row0, int64_t n_first, int count_a,
           int hist_row0) {
    ETL_REQUIRE(count_f >= 0 && count_a >= 0 && n >= 0 && n_first >= 0,
                "etl_coplan_run: bad step ranges");
    if (count_f == 0 && count_a == 0) return ETL_OK;
    FwdInfo fi;
    AdjInfo ai;
    fwd_info(c->f, &fi);
    adj_info(c->a, &ai);
    ETL_REQUIRE(fi.lane == ai.lane, "etl_coplan_run: the plans are on different lanes");
    int rc;
    int cv[3] = {0, 0, 0}, sv[2] = {0, 0};
    if (count_f > 0 && (rc = fwd_prepare(c->f, n, count_f, probe_row0, cv))) return rc;
    if (count_a > 0 && (rc = adj_prepare(c->a, n_first, count_a, hist_row0, sv))) return rc;
    fwd_info(c->f, &fi);  // after the layouts are up to date
    adj_info(c->a, &ai);
    hipStream_t s = lane_stream(fi.lane);
    hipLaunchKernelGGL(co_set_kernel, dim3(1), dim3(64), 0, s, count_f > 0 ? fi.counter : nullptr,
                       cv[0], cv[1], cv[2], count_a > 0 ? ai.step : nullptr, sv[0], sv[1]);
    ETL_LAUNCHED();
    EtlCoPlan::Entry* hit = nullptr;
    for (auto& e : c->graphs)
        if (e.count_f == count_f && e.count_a == count_a && e.fpar == fi.parity &&
            e.apar == ai.parity && e.fgen == fi.gen && e.agen == ai.gen)
            hit = &e;
    if (hit == nullptr) {
        if (c->graphs.size() >= 8) {
            // the cached graphs' launches may still be queued: a graph is destroyed
            // only once its work is done
            if ((rc = sync_lanes())) return rc;
            free_co_graphs(c);
        }
        EtlCoPlan::Entry e{count_f, count_a, fi.parity, ai.parity, fi.gen, ai.gen,
                           nullptr, nullptr, 0, 0};
        ETL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = fi.precision == ETL_F32
                 ? co_enqueue<float>(c, fi, count_f, ai, count_a, s, &e.f_launches, &e.a_launches)
                 : co_enqueue<double>(c, fi, count_f, ai, count_a, s, &e.f_launches,
                                      &e.a_launches);
        const hipError_t err = hipStreamEndCapture(s, &e.src);
        if (rc) {
            if (e.src) (void)hipGraphDestroy(e.src);
            return rc;
        }
        ETL_HIP(err);
        if (hipGraphInstantiate(&e.exec, e.src, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(e.src);
            set_error("etl_coplan_run: graph instantiation failed");
            return ETL_EHIP;
        }
        c->graphs.push_back(e);
        hit = &c->graphs.back();
    }
    ETL_HIP(hipGraphLaunch(hit->exec, s));
    fwd_set_parity(c->f, fi.parity ^ (hit->f_launches & 1));
    adj_set_parity(c->a, ai.parity ^ (hit->a_launches & 1));
    return ETL_OK;
}

}  // namespace

extern "C" {

int etl_coplan_create(etl_plan_t f, etl_adjplan_t a, etl_coplan_t* out) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(f && a && out, "etl_coplan_create: null argument");
    FwdInfo fi;
    AdjInfo ai;
    fwd_info(f, &fi);
    adj_info(a, &ai);
    ETL_REQUIRE(fi.precision == ai.precision, "etl_coplan_create: precisions differ");
    ETL_REQUIRE(fi.vec == 1 && (fi.band == 8 || (fi.band == 16 && fi.precision == ETL_F32)),
                "etl_coplan_create: the forward plan needs vec 1 and band 8 (or 16 in float; "
                "got %d, %d)", fi.vec, fi.band);
    EtlCoPlan* c = new EtlCoPlan();
    c->f = f;
    c->a = a;
    *out = c;
    return ETL_OK;
}

int etl_coplan_destroy(etl_coplan_t c) {
    if (c == nullptr) return ETL_OK;
    ETL_CHECK_INIT();
    int rc = sync_lanes();
    free_co_graphs(c);
    delete c;
    return rc;
}

int etl_coplan_run(etl_coplan_t c, int64_t n, int count_f, int probe_row0, int64_t n_first,
                   int count_a, int hist_row0) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(c != nullptr, "etl_coplan_run: null co-plan");
    return co_run(c, n, count_f, probe_row0, n_first, count_a, hist_row0);
}

int etl_coplan_sweep(etl_coplan_t c, int64_t n_steps, int K) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(c != nullptr && n_steps >= 1 && K >= 1 && (n_steps + K - 1) / K < (1 << 30),
                "etl_coplan_sweep: bad arguments");
    const int last = int((n_steps - 1) / K);
    auto count = [&](int k) { return int(std::min<int64_t>(K, n_steps - int64_t(k) * K)); };
    auto half = [&](int k) { return (k & 1) * K; };
    int rc;
    if ((rc = fwd_checkpoint(c->f, last, 1)) ||
        (rc = co_run(c, int64_t(last) * K, count(last), half(last), 0, 0, 0)))
        return rc;
    for (int k = last; k >= 1; --k) {
        if ((rc = fwd_checkpoint(c->f, k - 1, 1)) ||
            (rc = co_run(c, int64_t(k - 1) * K, count(k - 1), half(k - 1), int64_t(k) * K,
                         count(k), half(k))))
            return rc;
    }
    return co_run(c, 0, 0, 0, 0, count(0), half(0));
}

}  // extern "C"

