// Discrete adjoint of the 2D TE step, for permittivity gradients of DFT-monitor
// objectives.  No reference counterpart (the reference has no adjoint).
//
// One forward step is linear in the state x = (Ex, Ey, Hz, psi):
//   H:  d = Dy Ex, e = Dx Ey, psi_y' = by psi_y + cy d, psi_x' = bx psi_x + cx e,
//       Hz' = Hz + db (d iky + psi_y' - e ikx - psi_x')
//   E:  a = Dy Hz', psi' = by psi + cy a, u = a iky + psi', Ex' = Ex + cb_ex u
//       b = Dx Hz', psi' = bx psi + cx b, w = b ikx + psi', Ey' = Ey - cb_ey w
// The reverse step n applies the transposes in reverse order to the adjoint state
// lambda (same shapes as x):
//   sources    lambda of the step's outputs += dJ/d(monitored values of step n)
//   adj_h      per Hz node: lambda_Hz' gathers the E transpose of its four E
//              neighbours (each recomputed from lambda_E', lambda_psiE), then the H
//              transpose: lambda_psiH updates, and L_d, L_e for the E neighbours
//   adj_e      per E node: lambda_psiE update, dJ/dcb += lambda_E' u (design box),
//              lambda_E += the H transpose gathered from L_d / L_e
// so that after the reverse step lambda holds dJ/dx of step n's inputs.  The
// gradient needs the forward increments u, w of each step on the design box only;
// they are recomputed from Hz' of that step (a history the forward recompute
// writes), since inside the box the CPML is absent (u = a, w = b).
#include <vector>

#include "common.h"

namespace etl {
namespace {

template <typename R>
struct AdjArgs {
    R *lex, *ley, *lhz, *lpey, *lpex, *lphx, *lphy;  // lambda: ex ey hz psi_ex_y psi_ey_x psi_hz_x psi_hz_y
    const R *cbx, *cby;
    const R *bx_e, *cx_e, *ikx_e, *bx_h, *cx_h, *ikx_h;
    const R *by_e, *cy_e, *iky_e, *by_h, *cy_h, *iky_h;
    R db, idx, idy;
    int nx, ny;
    R *ld, *le;  // H-transpose results, Hz-shaped
    int bi0, bj0, bni, bnj;
    double *gx, *gy;  // dJ/dcb_ex, dJ/dcb_ey on the box, (bni, bnj)
    const R* hist;    // (rows, (bni + 1) * (bnj + 1)): Hz' on rows bi0-1.., columns bj0-1..
    int n_src;
    const int* src_comp;
    const int64_t* src_flat;
    const R* src;          // (steps, n_src)
    const int* step;       // first step of the current run (device)
};

// E-transpose coupling of Ex node (i, j) into Hz (zero where Ex is never updated).
template <typename R>
__device__ __forceinline__ R la_at(const AdjArgs<R>& a, int i, int j) {
    if (j < 1 || j > a.ny - 1) return R(0);
    const int64_t k = int64_t(i) * (a.ny + 1) + j;
    const R lu = a.cbx[k] * a.lex[k];
    const R lp = a.lpey[k] + lu;
    return a.iky_e[j] * lu + a.cy_e[j] * lp;
}

// ... of Ey node (i, j).
template <typename R>
__device__ __forceinline__ R lb_at(const AdjArgs<R>& a, int i, int j) {
    if (i < 1 || i > a.nx - 1) return R(0);
    const int64_t k = int64_t(i) * a.ny + j;
    const R lw = -a.cby[k] * a.ley[k];
    const R lq = a.lpex[k] + lw;
    return a.ikx_e[i] * lw + a.cx_e[i] * lq;
}

// lambda_field.flat[k] += dJ/dv_k of step `row` (distinct nodes per component).
template <typename R>
__global__ void adj_src_kernel(const AdjArgs<R> a, int row) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n_src) return;
    const int c = a.src_comp[s];
    R* f = c == 0 ? a.lex : (c == 1 ? a.ley : a.lhz);
    const int64_t r = int64_t(*a.step) + row;
    f[a.src_flat[s]] += a.src[r * a.n_src + s];
}

template <typename R>
__global__ __launch_bounds__(256) void adj_h_kernel(const AdjArgs<R> a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= a.ny) return;
    const int64_t k = int64_t(i) * a.ny + j;
    const R lh = a.lhz[k] + a.idy * (la_at(a, i, j) - la_at(a, i, j + 1)) +
                 a.idx * (lb_at(a, i, j) - lb_at(a, i + 1, j));
    a.lhz[k] = lh;
    const R lpy = a.lphy[k] + a.db * lh;
    a.ld[k] = a.db * a.iky_h[j] * lh + a.cy_h[j] * lpy;
    a.lphy[k] = a.by_h[j] * lpy;
    const R lpx = a.lphx[k] - a.db * lh;
    a.le[k] = -a.db * a.ikx_h[i] * lh + a.cx_h[i] * lpx;
    a.lphx[k] = a.bx_h[i] * lpx;
}

template <typename R>
__global__ __launch_bounds__(256) void adj_e_kernel(const AdjArgs<R> a, int row) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;  // 0..ny
    const int i = blockIdx.y;                              // 0..nx
    const int nx = a.nx, ny = a.ny;
    if (j > ny) return;
    const bool in_box = i >= a.bi0 && i < a.bi0 + a.bni && j >= a.bj0 && j < a.bj0 + a.bnj;
    const int hw = a.bnj + 1;
    const R* h = a.hist + int64_t(row) * (a.bni + 1) * hw;
    auto hz = [&](int ii, int jj) { return h[(ii - a.bi0 + 1) * hw + (jj - a.bj0 + 1)]; };
    if (i < nx) {  // Ex node (i, j)
        const int64_t k = int64_t(i) * (ny + 1) + j;
        const R lold = a.lex[k];
        if (j >= 1 && j <= ny - 1) {
            const R lu = a.cbx[k] * lold;
            const R lp = a.lpey[k] + lu;
            a.lpey[k] = a.by_e[j] * lp;
            if (in_box) {
                const double u = (double(hz(i, j)) - double(hz(i, j - 1))) * double(a.idy);
                a.gx[(i - a.bi0) * a.bnj + (j - a.bj0)] += double(lold) * u;
            }
        }
        R add = R(0);
        if (j >= 1) add += a.ld[int64_t(i) * ny + j - 1];
        if (j < ny) add -= a.ld[int64_t(i) * ny + j];
        a.lex[k] = lold + a.idy * add;
    }
    if (j < ny) {  // Ey node (i, j)
        const int64_t k = int64_t(i) * ny + j;
        const R lold = a.ley[k];
        if (i >= 1 && i <= nx - 1) {
            const R lw = -a.cby[k] * lold;
            const R lq = a.lpex[k] + lw;
            a.lpex[k] = a.bx_e[i] * lq;
            if (in_box) {
                const double w = (double(hz(i, j)) - double(hz(i - 1, j))) * double(a.idx);
                a.gy[(i - a.bi0) * a.bnj + (j - a.bj0)] -= double(lold) * w;
            }
        }
        R add = R(0);
        if (i >= 1) add += a.le[int64_t(i - 1) * ny + j];
        if (i < nx) add -= a.le[int64_t(i) * ny + j];
        a.ley[k] = lold + a.idx * add;
    }
}

__global__ void set_step_kernel(int* step, int value) {
    if (threadIdx.x == 0) step[0] = value;
}

}  // namespace
}  // namespace etl

using namespace etl;

struct EtlAdjoint {
    EtlAdjointDesc desc{};
    void* ld = nullptr;
    void* le = nullptr;
    int* step = nullptr;
    int n_src = 0;
    int* src_comp = nullptr;
    int64_t* src_flat = nullptr;
    void* src = nullptr;
    int src_rows = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    int graph_count = 0;
    int elem() const { return desc.precision == ETL_F32 ? 4 : 8; }
};

namespace {

template <typename R>
AdjArgs<R> adj_args(const EtlAdjoint* p) {
    const EtlAdjointDesc& d = p->desc;
    AdjArgs<R> a{};
    a.lex = (R*)d.lam[0];
    a.ley = (R*)d.lam[1];
    a.lhz = (R*)d.lam[2];
    a.lpey = (R*)d.lam[3];
    a.lpex = (R*)d.lam[4];
    a.lphx = (R*)d.lam[5];
    a.lphy = (R*)d.lam[6];
    a.cbx = (const R*)d.cb_ex;
    a.cby = (const R*)d.cb_ey;
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
    a.idx = R(d.inv_dx);
    a.idy = R(d.inv_dy);
    a.nx = d.nx;
    a.ny = d.ny;
    a.ld = (R*)p->ld;
    a.le = (R*)p->le;
    a.bi0 = d.box_i0;
    a.bj0 = d.box_j0;
    a.bni = d.box_ni;
    a.bnj = d.box_nj;
    a.gx = d.grad_ex;
    a.gy = d.grad_ey;
    a.hist = (const R*)d.hist;
    a.n_src = p->n_src;
    a.src_comp = p->src_comp;
    a.src_flat = p->src_flat;
    a.src = (const R*)p->src;
    a.step = p->step;
    return a;
}

// Reverse steps of history rows count-1 .. 0.
template <typename R>
int enqueue_reverse(EtlAdjoint* p, int count, hipStream_t s) {
    const AdjArgs<R> a = adj_args<R>(p);
    const dim3 block(256);
    const dim3 grid_h(unsigned((a.ny + 255) / 256), unsigned(a.nx));
    const dim3 grid_e(unsigned((a.ny + 1 + 255) / 256), unsigned(a.nx + 1));
    for (int r = count - 1; r >= 0; --r) {
        if (a.n_src) {
            hipLaunchKernelGGL(adj_src_kernel<R>, dim3(unsigned((a.n_src + 255) / 256)), block, 0,
                               s, a, r);
            ETL_LAUNCHED();
        }
        hipLaunchKernelGGL(adj_h_kernel<R>, grid_h, block, 0, s, a);
        ETL_LAUNCHED();
        hipLaunchKernelGGL(adj_e_kernel<R>, grid_e, block, 0, s, a, r);
        ETL_LAUNCHED();
    }
    return ETL_OK;
}

int enqueue_reverse_any(EtlAdjoint* p, int count, hipStream_t s) {
    return p->desc.precision == ETL_F32 ? enqueue_reverse<float>(p, count, s)
                                        : enqueue_reverse<double>(p, count, s);
}

void free_graph(EtlAdjoint* p) {
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->graph_src) (void)hipGraphDestroy(p->graph_src);
    p->graph = nullptr;
    p->graph_src = nullptr;
    p->graph_count = 0;
}

}  // namespace

extern "C" {

int etl_adjoint_create(const EtlAdjointDesc* d, etl_adjoint_t* out) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(d && out, "etl_adjoint_create: null argument");
    ETL_REQUIRE(d->precision == ETL_F32 || d->precision == ETL_F64,
                "etl_adjoint_create: unknown precision %d", d->precision);
    ETL_REQUIRE(d->nx >= 2 && d->ny >= 2, "etl_adjoint_create: bad grid %d x %d", d->nx, d->ny);
    for (int k = 0; k < 7; ++k) ETL_REQUIRE(d->lam[k], "etl_adjoint_create: null adjoint field %d", k);
    ETL_REQUIRE(d->cb_ex && d->cb_ey && d->bx_e && d->cx_e && d->ikx_e && d->bx_h && d->cx_h &&
                    d->ikx_h && d->by_e && d->cy_e && d->iky_e && d->by_h && d->cy_h && d->iky_h,
                "etl_adjoint_create: null coefficient array");
    ETL_REQUIRE(d->box_ni >= 0 && d->box_nj >= 0, "etl_adjoint_create: bad design box");
    if (d->box_ni > 0 && d->box_nj > 0) {
        ETL_REQUIRE(d->box_i0 >= 1 && d->box_j0 >= 1 && d->box_i0 + d->box_ni <= d->nx - 1 &&
                        d->box_j0 + d->box_nj <= d->ny - 1,
                    "etl_adjoint_create: design box must lie inside the grid, off its faces");
        ETL_REQUIRE(d->grad_ex && d->grad_ey && d->hist && d->hist_rows >= 1,
                    "etl_adjoint_create: a design box needs gradient and history arrays");
    }
    EtlAdjoint* p = new EtlAdjoint();
    p->desc = *d;
    const size_t hz = size_t(d->nx) * d->ny * p->elem();
    int rc = etl_malloc(&p->ld, hz);
    if (rc == ETL_OK) rc = etl_malloc(&p->le, hz);
    if (rc == ETL_OK) rc = etl_malloc((void**)&p->step, 256);
    if (rc != ETL_OK) {
        etl_free(p->ld);
        etl_free(p->le);
        delete p;
        return rc;
    }
    *out = p;
    return ETL_OK;
}

int etl_adjoint_destroy(etl_adjoint_t p) {
    if (p == nullptr) return ETL_OK;
    ETL_CHECK_INIT();
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    free_graph(p);
    etl_free(p->ld);
    etl_free(p->le);
    etl_free(p->step);
    etl_free(p->src_comp);
    etl_free(p->src_flat);
    etl_free(p->src);
    delete p;
    return ETL_OK;
}

int etl_adjoint_sources(etl_adjoint_t p, int n, const int* comp, const int64_t* flat, int rows,
                        const void* table) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_adjoint_sources: null handle");
    ETL_REQUIRE(n >= 0 && rows >= 0 && (n == 0 || (comp && flat && table)),
                "etl_adjoint_sources: bad arguments");
    const int nx = p->desc.nx, ny = p->desc.ny;
    std::vector<std::pair<int, int64_t>> seen;
    for (int s = 0; s < n; ++s) {
        ETL_REQUIRE(comp[s] >= 0 && comp[s] <= 2, "etl_adjoint_sources: component %d", comp[s]);
        const int64_t size = comp[s] == 0 ? int64_t(nx) * (ny + 1)
                                          : (comp[s] == 1 ? int64_t(nx + 1) * ny : int64_t(nx) * ny);
        ETL_REQUIRE(flat[s] >= 0 && flat[s] < size, "etl_adjoint_sources: node %lld outside "
                    "component %d", (long long)flat[s], comp[s]);
        for (const auto& q : seen)
            ETL_REQUIRE(!(q.first == comp[s] && q.second == flat[s]),
                        "etl_adjoint_sources: node listed twice (sum its columns)");
        seen.push_back({comp[s], flat[s]});
    }
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    free_graph(p);
    etl_free(p->src_comp);
    etl_free(p->src_flat);
    etl_free(p->src);
    p->src_comp = nullptr;
    p->src_flat = nullptr;
    p->src = nullptr;
    p->n_src = n;
    p->src_rows = rows;
    if (n == 0) return ETL_OK;
    int rc;
    if ((rc = etl_malloc((void**)&p->src_comp, size_t(n) * sizeof(int))) ||
        (rc = etl_memcpy_h2d(p->src_comp, comp, size_t(n) * sizeof(int))) ||
        (rc = etl_malloc((void**)&p->src_flat, size_t(n) * sizeof(int64_t))) ||
        (rc = etl_memcpy_h2d(p->src_flat, flat, size_t(n) * sizeof(int64_t))))
        return rc;
    const size_t bytes = size_t(rows) * n * p->elem();
    if (bytes && ((rc = etl_malloc(&p->src, bytes)) || (rc = etl_memcpy_h2d(p->src, table, bytes))))
        return rc;
    return ETL_OK;
}

int etl_adjoint_run(etl_adjoint_t p, int64_t n_first, int count) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(p != nullptr, "etl_adjoint_run: null handle");
    ETL_REQUIRE(count >= 0 && n_first >= 0, "etl_adjoint_run: bad step range");
    if (count == 0) return ETL_OK;
    const bool box = p->desc.box_ni > 0 && p->desc.box_nj > 0;
    ETL_REQUIRE(!box || count <= p->desc.hist_rows,
                "etl_adjoint_run: %d steps, the history holds %d", count, p->desc.hist_rows);
    ETL_REQUIRE(p->n_src == 0 || n_first + count <= p->src_rows,
                "etl_adjoint_run: the source table covers steps [0, %d)", p->src_rows);
    ETL_REQUIRE(n_first < (int64_t(1) << 31), "etl_adjoint_run: step range");
    hipStream_t s = ctx().stream;
    hipLaunchKernelGGL(set_step_kernel, dim3(1), dim3(64), 0, s, p->step, int(n_first));
    ETL_LAUNCHED();
    // The reverse steps of a full segment are replayed from a graph (captured once per
    // segment length); the step index comes from the device.
    if (p->graph_count != count) {
        free_graph(p);
        ETL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_reverse_any(p, count, s);
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
    }
    ETL_HIP(hipGraphLaunch(p->graph, s));
    return ETL_OK;
}

}  // extern "C"
