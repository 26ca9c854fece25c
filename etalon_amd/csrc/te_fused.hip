// Fused, temporally blocked 2D TE time stepping for gfx950.
//
// One launch advances the fields T full steps.  Per step n (etalon/fdtd/solver.py:
// 120-128): H update -> M injection (t = nΔt) -> E update -> J injection
// (t = (n+½)Δt); the monitored values of Hz (after M) and of Ex, Ey (after J) go to
// a trace buffer that a separate DFT kernel folds into the accumulators in step
// order.  Every node update performs the reference's operations in the reference's
// order (etalon/fdtd/backends/numpy_backend.py:41-61), so the results are bit-identical
// to T single steps whatever T, the tile shape or the launch order.
//
// Decomposition.  One wave (64 lanes) owns an output tile of `band` rows (x) by
// WO <= 64*V - 2*T columns (y).  Lane l holds the V consecutive columns
// cs + l*V .. + V-1 of a window that extends T columns past the tile on each side
// (cs = first output column - T).  The wave marches down rows r = i0 - T ..
// i1 + T - 2 and, at row r, performs step t on row r - t (a skew of one row per
// step): step t needs E^t of rows r-t and r-t+1 — the latter produced moments
// before by step t-1 — and H^(t+½) of row r-t-1, kept from the previous row.  Each
// row's state between steps (fields, psi, cb) stays in registers, so every array is
// read from memory once and Ex, Ey, Hz written once per T steps.  Row r + 1 is
// loaded while row r is computed, and no other load sits between them: gfx950
// retires vector loads in order, so any load waited on inside the step loop would
// also drain the prefetch.  Values outside the window are unknown; they
// corrupt one column per half-step from each side and one row per step from the
// top, which the T-wide halos absorb; the physical boundaries (PEC tangential E,
// copied unchanged) need no halo.  Reads come from buffer A, writes go to buffer B,
// so tiles never see each other's results within a launch.
//
// CPML.  psi lives only in the border strips, where the 1D profiles differ from
// (b, c, 1/kappa) = (1, 0, 1); there it is carried in registers between steps like
// the fields.  Elsewhere psi is +0 and the update reduces to d*1 + (+0), exactly the
// reference's arithmetic (the +0 fixes the sign of a zero difference).
//
// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs; block b takes
// tile (b % 8) * tiles_per_xcd + b / 8, so each XCD sweeps a contiguous slab of
// rows and the halo rows shared by vertically adjacent tiles stay in its L2.
#include "common.h"
#include "fused.h"

namespace etl {
namespace {

// V consecutive REALs with the alignment of one REAL: Ex rows have an odd pitch
// (ny + 1) and gfx950 global loads/stores need only dword alignment.
template <typename R, int V>
using Vec = R __attribute__((ext_vector_type(V), aligned(sizeof(R))));

// dst[k] = p[c + k] for the elements with 0 <= c + k < limit (others 0).
template <typename R, int V>
__device__ __forceinline__ void load_row(R (&dst)[V], const R* __restrict__ p, int c, int limit) {
    if (c >= 0 && c + V <= limit) {
        if constexpr (V == 1) {
            dst[0] = p[c];
        } else {
            const Vec<R, V> x = *reinterpret_cast<const Vec<R, V>*>(p + c);
#pragma unroll
            for (int k = 0; k < V; ++k) dst[k] = x[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k) dst[k] = (c + k >= 0 && c + k < limit) ? p[c + k] : R(0);
    }
}

// p[c + k] = src[k] for the elements with lo <= c + k < hi.
template <typename R, int V>
__device__ __forceinline__ void store_row(R* __restrict__ p, const R (&src)[V], int c, int lo,
                                          int hi) {
    if (c >= lo && c + V <= hi) {
        if constexpr (V == 1) {
            p[c] = src[0];
        } else {
            Vec<R, V> x;
#pragma unroll
            for (int k = 0; k < V; ++k) x[k] = src[k];
            *reinterpret_cast<Vec<R, V>*>(p + c) = x;
        }
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k)
            if (c + k >= lo && c + k < hi) p[c + k] = src[k];
    }
}

// Elements whose mask is false read as 0 (no memory access).
template <typename R, int V>
__device__ __forceinline__ void load_masked(R (&dst)[V], const R* __restrict__ p, int c,
                                            const bool (&m)[V]) {
#pragma unroll
    for (int k = 0; k < V; ++k) dst[k] = m[k] ? p[c + k] : R(0);
}

template <typename R, int V>
__device__ __forceinline__ void store_masked(R* __restrict__ p, const R (&src)[V], int c,
                                             const bool (&m)[V]) {
#pragma unroll
    for (int k = 0; k < V; ++k)
        if (m[k]) p[c + k] = src[k];
}

// Subtract this step's deltas of the sources of component `comp` lying on row i,
// in source order.  x - (+0) == x bit for bit, so the untouched elements are
// unchanged; the select keeps v[] in registers (no dynamic index).
template <typename R, int V>
__device__ __forceinline__ void inject_row(const FusedArgs<R>& a, const R* __restrict__ drow,
                                           int comp, int i, int c, R (&v)[V]) {
    for (int s = 0; s < a.n_src; ++s) {
        const int4 sd = a.src[s];  // comp, i, j
        if (sd.x == comp && sd.y == i) {
            const int js = sd.z - c;
            if (js >= 0 && js < V) {
                const R dl = drow[s];
#pragma unroll
                for (int k = 0; k < V; ++k) v[k] = v[k] - ((k == js) ? dl : R(0));
            }
        }
    }
}

// Write the monitored values of row i of component `comp` into trace row `trow`.
template <typename R, int V>
__device__ __forceinline__ void trace_row(const FusedArgs<R>& a, R* __restrict__ lds, int comp,
                                          int i, int tj, int cs, const R (&v)[V],
                                          R* __restrict__ trow) {
    const int* ptr = a.mon_ptr[comp];
    if (ptr == nullptr) return;
    const int64_t cell = int64_t(i) * a.n_tj + tj;
    const int s = ptr[cell];
    const int e = ptr[cell + 1];
    if (s >= e) return;  // wave-uniform
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < V; ++k) lds[lane * V + k] = v[k];
    __syncthreads();
    const int2* ent = a.mon_ent[comp];
    for (int q = s + lane; q < e; q += 64) {
        const int2 en = ent[q];  // column, slot
        trow[en.y] = lds[en.x - cs];
    }
    __syncthreads();
}

template <typename R, int V, int T>
__global__ __launch_bounds__(64) void fused_kernel(const FusedArgs<R> a,
                                                   const int* __restrict__ counter, int kstep0,
                                                   int parity) {
    constexpr int W = 64 * V;
    const int WO = a.wo;  // the same tiles for every depth T; WO + 2T <= W
    const int b = blockIdx.x;
    const int tile = (b & 7) * a.tiles_per_xcd + (b >> 3);
    if (tile >= a.n_bands * a.n_tj) return;
    const int tb = tile / a.n_tj;
    const int tj = tile - tb * a.n_tj;
    const int lane = threadIdx.x;
    const int nx = a.nx, ny = a.ny, px = ny + 1;
    const int o0 = tj * WO;  // first output column of the tile
    const int cs = o0 - T;   // first column of the window
    const int c = cs + lane * V;
    const int i0 = tb * a.band;
    const int i1 = min(i0 + a.band, nx);
    const int rs = max(i0 - T, 0);

    __shared__ R lds[W];

    const R* __restrict__ exA = a.f[0][parity];
    const R* __restrict__ eyA = a.f[1][parity];
    const R* __restrict__ hzA = a.f[2][parity];
    const R* __restrict__ pexA = a.f[3][parity];  // psi_ex_y
    const R* __restrict__ peyA = a.f[4][parity];  // psi_ey_x
    const R* __restrict__ phxA = a.f[5][parity];  // psi_hz_x
    const R* __restrict__ phyA = a.f[6][parity];  // psi_hz_y
    R* __restrict__ exB = const_cast<R*>(a.f[0][parity ^ 1]);
    R* __restrict__ eyB = const_cast<R*>(a.f[1][parity ^ 1]);
    R* __restrict__ hzB = const_cast<R*>(a.f[2][parity ^ 1]);
    R* __restrict__ pexB = const_cast<R*>(a.f[3][parity ^ 1]);
    R* __restrict__ peyB = const_cast<R*>(a.f[4][parity ^ 1]);
    R* __restrict__ phxB = const_cast<R*>(a.f[5][parity ^ 1]);
    R* __restrict__ phyB = const_cast<R*>(a.f[6][parity ^ 1]);

    const int nbase = *counter + kstep0;  // table row of this launch's first step
    const int flags = a.tile_flags ? int(a.tile_flags[tile]) : 3;
    const bool has_src = (flags & 2) && a.n_src > 0;
    const bool has_mon = (flags & 1) && a.has_mon;

    const R db = a.db, idx = a.inv_dx, idy = a.inv_dy;
    const R zero = R(0);

    // ---- per-lane column masks and y-strip profiles ----
    const int own_lo = o0, own_hi = o0 + WO;
    bool mH[V], ysh[V], yse[V];
    R byh[V], cyh[V], ikyh[V], bye[V], cye[V], ikye[V];
    const bool tile_ys = (cs < max(a.yh_lo, a.ye_lo)) || (cs + W > min(a.yh_hi, a.ye_hi));
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int j = c + k;
        mH[k] = (j >= 0 && j < ny);  // Hz / Ey column
        ysh[k] = tile_ys && mH[k] && (j < a.yh_lo || j >= a.yh_hi);
        yse[k] = tile_ys && j >= 1 && j <= ny - 1 && (j < a.ye_lo || j >= a.ye_hi);
        byh[k] = ysh[k] ? a.by_h[j] : R(1);
        cyh[k] = ysh[k] ? a.cy_h[j] : R(0);
        ikyh[k] = ysh[k] ? a.iky_h[j] : R(1);
        bye[k] = yse[k] ? a.by_e[j] : R(1);
        cye[k] = yse[k] ? a.cy_e[j] : R(0);
        ikye[k] = yse[k] ? a.iky_e[j] : R(1);
    }
    auto xs_h = [&](int i) { return i < a.xh_lo || i >= a.xh_hi; };
    auto xs_e = [&](int i) { return i >= 1 && i <= nx - 1 && (i < a.xe_lo || i >= a.xe_hi); };

    // ---- state of step t on its current row r - t (slot t) ----
    R sex[T][V], sey[T][V], shz[T][V];                 // E^t, H^(t-½): step t's inputs
    R hpv[T][V];                                       // H^(t+½) of the row above
    R scbx[T][V], scby[T][V];                          // cb_ex, cb_ey of step t's row
    R spyh[T][V], spxh[T][V], spye[T][V], spxe[T][V];  // psi inputs of step t
    R eyn0[V];                                         // Ey^0 of row r + 1
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int k = 0; k < V; ++k) {
            sex[t][k] = sey[t][k] = shz[t][k] = hpv[t][k] = scbx[t][k] = scby[t][k] = R(0);
            spyh[t][k] = spxh[t][k] = spye[t][k] = spxe[t][k] = R(0);
        }

    // row rs into slot 0
    load_row<R, V>(sex[0], exA + int64_t(rs) * px, c, px);
    load_row<R, V>(shz[0], hzA + int64_t(rs) * ny, c, ny);
    load_row<R, V>(sey[0], eyA + int64_t(rs) * ny, c, ny);
    load_row<R, V>(eyn0, eyA + int64_t(rs + 1) * ny, c, ny);
    load_row<R, V>(scbx[0], a.cb_ex + int64_t(rs) * px, c, px);
    load_row<R, V>(scby[0], a.cb_ey + int64_t(rs) * ny, c, ny);
    if (tile_ys) {
        load_masked<R, V>(spyh[0], phyA + int64_t(rs) * ny, c, ysh);
        load_masked<R, V>(spye[0], pexA + int64_t(rs) * px, c, yse);
    }
    if (xs_h(rs)) load_row<R, V>(spxh[0], phxA + int64_t(rs) * ny, c, ny);
    if (xs_e(rs)) load_row<R, V>(spxe[0], peyA + int64_t(rs) * ny, c, ny);

    const int r_end = i1 + T - 1;
    for (int r = rs; r < r_end; ++r) {
        // ---- prefetch row r + 1 for step 0 of the next row ----
        const int rn = r + 1;
        R pex[V], phz[V], pey[V], ppyh[V], ppxh[V], ppye[V], ppxe[V], pcbx[V], pcby[V];
#pragma unroll
        for (int k = 0; k < V; ++k)
            pex[k] = phz[k] = pey[k] = ppyh[k] = ppxh[k] = ppye[k] = ppxe[k] = pcbx[k] =
                pcby[k] = R(0);
        if (rn < nx) {
            load_row<R, V>(pex, exA + int64_t(rn) * px, c, px);
            load_row<R, V>(phz, hzA + int64_t(rn) * ny, c, ny);
            load_row<R, V>(pcbx, a.cb_ex + int64_t(rn) * px, c, px);
            load_row<R, V>(pcby, a.cb_ey + int64_t(rn) * ny, c, ny);
            if (tile_ys) {
                load_masked<R, V>(ppyh, phyA + int64_t(rn) * ny, c, ysh);
                load_masked<R, V>(ppye, pexA + int64_t(rn) * px, c, yse);
            }
            if (xs_h(rn)) load_row<R, V>(ppxh, phxA + int64_t(rn) * ny, c, ny);
            if (xs_e(rn)) load_row<R, V>(ppxe, peyA + int64_t(rn) * ny, c, ny);
        }
        if (rn + 1 <= nx) load_row<R, V>(pey, eyA + int64_t(rn + 1) * ny, c, ny);

#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int rho = r - t;
            if (rho < rs || rho >= nx) continue;  // wave-uniform
            const int64_t rx = int64_t(rho) * px;
            const int64_t rh = int64_t(rho) * ny;
            // E^t of row rho + 1: from memory for step 0, else step t-1's output of this
            // row; Ey row nx is outer tangential E and never changes.
            R eyN[V];
            if (t == 0) {
#pragma unroll
                for (int k = 0; k < V; ++k) eyN[k] = eyn0[k];
            } else if (rho + 1 == nx) {
                load_row<R, V>(eyN, eyA + int64_t(nx) * ny, c, ny);
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) eyN[k] = sey[t > 0 ? t - 1 : 0][k];
            }

            // ---- H half-step on row rho ----
            const bool xh = xs_h(rho);
            R bx = R(1), cx = R(0), ikx = R(1);
            if (xh) {
                bx = a.bx_h[rho];
                cx = a.cx_h[rho];
                ikx = a.ikx_h[rho];
            }
            const R exr = __shfl_down(sex[t][0], 1, 64);  // lane 63: outside the window
            R hn[V];
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const R exn = (k + 1 < V) ? sex[t][k + 1] : exr;
                const R d = (exn - sex[t][k]) * idy;
                const R e = (eyN[k] - sey[t][k]) * idx;
                R ty, tx;
                if (ysh[k]) {
                    const R py = byh[k] * spyh[t][k] + cyh[k] * d;
                    spyh[t][k] = py;
                    ty = d * ikyh[k] + py;
                } else {
                    ty = d + zero;
                }
                if (xh) {
                    const R pxv = bx * spxh[t][k] + cx * e;
                    spxh[t][k] = pxv;
                    tx = e * ikx + pxv;
                } else {
                    tx = e + zero;
                }
                hn[k] = shz[t][k] + db * (ty - tx);
            }
            const R* drow = a.delta + int64_t(nbase + t) * a.n_src;
            if (has_src) inject_row<R, V>(a, drow, 2, rho, c, hn);

            // ---- E half-step on row rho ----
            const R hl = __shfl_up(hn[V - 1], 1, 64);  // lane 0: outside the window
            const bool xe = xs_e(rho);
            R bxe = R(1), cxe = R(0), ikxe = R(1);
            if (xe) {
                bxe = a.bx_e[rho];
                cxe = a.cx_e[rho];
                ikxe = a.ikx_e[rho];
            }
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int j = c + k;
                if (j >= 1 && j <= ny - 1) {  // Ex(rho, j); j = 0, ny stay unchanged
                    const R hleft = (k == 0) ? hl : hn[k - 1];
                    const R d = (hn[k] - hleft) * idy;
                    R tt;
                    if (yse[k]) {
                        const R p = bye[k] * spye[t][k] + cye[k] * d;
                        spye[t][k] = p;
                        tt = d * ikye[k] + p;
                    } else {
                        tt = d + zero;
                    }
                    sex[t][k] = sex[t][k] + scbx[t][k] * tt;
                }
                if (rho >= 1) {  // Ey(rho, j), rho <= nx - 1; row 0 stays unchanged
                    const R d = (hn[k] - hpv[t][k]) * idx;
                    R tt;
                    if (xe) {
                        const R p = bxe * spxe[t][k] + cxe * d;
                        spxe[t][k] = p;
                        tt = d * ikxe + p;
                    } else {
                        tt = d + zero;
                    }
                    sey[t][k] = sey[t][k] - scby[t][k] * tt;
                }
            }
            if (has_src) {
                inject_row<R, V>(a, drow, 0, rho, c, sex[t]);
                inject_row<R, V>(a, drow, 1, rho, c, sey[t]);
            }
            const bool owned_row = rho >= i0 && rho < i1;
            if (has_mon && owned_row) {
                R* trow = a.trace + int64_t(kstep0 + t) * a.n_slots;
                trace_row<R, V>(a, lds, 2, rho, tj, cs, hn, trow);
                trace_row<R, V>(a, lds, 0, rho, tj, cs, sex[t], trow);
                trace_row<R, V>(a, lds, 1, rho, tj, cs, sey[t], trow);
            }
            if (t == T - 1 && owned_row) {
                store_row<R, V>(hzB + rh, hn, c, max(own_lo, 0), min(own_hi, ny));
                store_row<R, V>(exB + rx, sex[t], c, max(own_lo, 0), min(own_hi, px));
                store_row<R, V>(eyB + rh, sey[t], c, max(own_lo, 0), min(own_hi, ny));
                bool mo[V], moy[V], moe[V];
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const bool own = (c + k >= own_lo && c + k < own_hi);
                    mo[k] = own && mH[k];
                    moy[k] = own && ysh[k];
                    moe[k] = own && yse[k];
                }
                if (tile_ys) {
                    store_masked<R, V>(phyB + rh, spyh[t], c, moy);
                    store_masked<R, V>(pexB + rx, spye[t], c, moe);
                }
                if (xh) store_masked<R, V>(phxB + rh, spxh[t], c, mo);
                if (xe) store_masked<R, V>(peyB + rh, spxe[t], c, mo);
            }
#pragma unroll
            for (int k = 0; k < V; ++k) {
                shz[t][k] = hn[k];  // step t+1's H input after the shift below
                hpv[t][k] = hn[k];
            }
        }

        // ---- shift: step t's outputs become step t+1's inputs; row r+1 enters slot 0 ----
#pragma unroll
        for (int t = T - 1; t >= 1; --t)
#pragma unroll
            for (int k = 0; k < V; ++k) {
                sex[t][k] = sex[t - 1][k];
                sey[t][k] = sey[t - 1][k];
                shz[t][k] = shz[t - 1][k];
                scbx[t][k] = scbx[t - 1][k];
                scby[t][k] = scby[t - 1][k];
                spyh[t][k] = spyh[t - 1][k];
                spxh[t][k] = spxh[t - 1][k];
                spye[t][k] = spye[t - 1][k];
                spxe[t][k] = spxe[t - 1][k];
            }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            sex[0][k] = pex[k];
            shz[0][k] = phz[k];
            sey[0][k] = eyn0[k];
            eyn0[k] = pey[k];
            scbx[0][k] = pcbx[k];
            scby[0][k] = pcby[k];
            spyh[0][k] = ppyh[k];
            spxh[0][k] = ppxh[k];
            spye[0][k] = ppye[k];
            spxe[0][k] = ppxe[k];
        }
    }

    // ---- last band: Ey row nx (outer tangential E) is copied ----
    if (i1 == nx) {
        R eyl[V];
        load_row<R, V>(eyl, eyA + int64_t(nx) * ny, c, ny);
        store_row<R, V>(eyB + int64_t(nx) * ny, eyl, c, max(own_lo, 0), min(own_hi, ny));
    }
}

// acc[pair] += sum over the chunk's rows, in step order, of trace * phasor.
template <typename R>
__global__ __launch_bounds__(256) void trace_dft_kernel(const R* __restrict__ trace, int n_slots,
                                                        int rows, const R* __restrict__ ph, int nq,
                                                        const int* __restrict__ counter,
                                                        const int4* __restrict__ pairs,
                                                        int n_pairs, R* const* __restrict__ accs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int4 pr = pairs[p];  // slot, q, monitor, complex element index
    R* acc = accs[pr.z] + 2 * int64_t(pr.w);
    R re = acc[0], im = acc[1];
    const int base = *counter;
    const R* __restrict__ php = ph + 2 * (int64_t(base) * nq + pr.y);
    for (int r = 0; r < rows; ++r) {
        const R v = trace[int64_t(r) * n_slots + pr.x];
        const R pre = php[2 * int64_t(r) * nq];
        const R pim = php[2 * int64_t(r) * nq + 1];
        re = re + v * pre;
        im = im + v * pim;
    }
    acc[0] = re;
    acc[1] = im;
}

__global__ void bump_counter_kernel(int* counter, int by) {
    if (threadIdx.x == 0) counter[0] = counter[0] + by;
}

template <typename R, int V>
int launch_v(const FusedArgs<R>& a, int depth, const int* counter, int kstep0, int parity,
             hipStream_t stream) {
    const dim3 grid(unsigned(8 * a.tiles_per_xcd)), block(64);
    switch (depth) {
        case 1:
            hipLaunchKernelGGL((fused_kernel<R, V, 1>), grid, block, 0, stream, a, counter, kstep0,
                               parity);
            break;
        case 2:
            hipLaunchKernelGGL((fused_kernel<R, V, 2>), grid, block, 0, stream, a, counter, kstep0,
                               parity);
            break;
        case 4:
            hipLaunchKernelGGL((fused_kernel<R, V, 4>), grid, block, 0, stream, a, counter, kstep0,
                               parity);
            break;
        default:
            set_error("fused step: unsupported temporal depth %d", depth);
            return ETL_EARG;
    }
    ETL_LAUNCHED();
    return ETL_OK;
}

}  // namespace

template <typename R>
int launch_fused(const FusedArgs<R>& a, int vec, int depth, const int* counter, int kstep0,
                 int parity, hipStream_t stream) {
    switch (vec) {
        case 1:
            return launch_v<R, 1>(a, depth, counter, kstep0, parity, stream);
        case 2:
            return launch_v<R, 2>(a, depth, counter, kstep0, parity, stream);
        case 4:
            return launch_v<R, 4>(a, depth, counter, kstep0, parity, stream);
        default:
            set_error("fused step: unsupported vector width %d", vec);
            return ETL_EARG;
    }
}

template <typename R>
int launch_trace_dft(const R* trace, int n_slots, int rows, const R* ph, int nq,
                     const int* counter, const int4* pairs, int n_pairs, R* const* accs,
                     hipStream_t stream) {
    if (n_pairs == 0 || rows == 0) return ETL_OK;
    hipLaunchKernelGGL(trace_dft_kernel<R>, dim3((n_pairs + 255) / 256), dim3(256), 0, stream,
                       trace, n_slots, rows, ph, nq, counter, pairs, n_pairs, accs);
    ETL_LAUNCHED();
    return ETL_OK;
}

int launch_bump(int* counter, int by, hipStream_t stream) {
    hipLaunchKernelGGL(bump_counter_kernel, dim3(1), dim3(64), 0, stream, counter, by);
    ETL_LAUNCHED();
    return ETL_OK;
}

template int launch_fused<float>(const FusedArgs<float>&, int, int, const int*, int, int,
                                 hipStream_t);
template int launch_fused<double>(const FusedArgs<double>&, int, int, const int*, int, int,
                                  hipStream_t);
template int launch_trace_dft<float>(const float*, int, int, const float*, int, const int*,
                                     const int4*, int, float* const*, hipStream_t);
template int launch_trace_dft<double>(const double*, int, int, const double*, int, const int*,
                                      const int4*, int, double* const*, hipStream_t);

}  // namespace etl
