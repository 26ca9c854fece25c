// Fused 2D TE time step for gfx950: one launch advances Ex, Ey, Hz by one full step.
//
// Per step n (etalon/fdtd/solver.py:120-128):  H update -> M injection (t = nΔt) ->
// E update -> J injection (t = (n+½)Δt); the monitored node values of Hz (after M)
// and of Ex, Ey (after J) are written to a trace buffer, and a separate DFT kernel
// folds a chunk of trace rows into the accumulators in step order.
//
// Work decomposition.  One wave (64 lanes) owns a tile of `band` rows (x) by 64*V
// columns (y).  Lane l holds V consecutive columns c = c0 + l*V .. c+V-1 and marches
// down the rows; Ey of the next row and Hz of the previous row stay in registers,
// so each field value is loaded once.  Neighbours along y come from the adjacent
// lane (ds_bpermute); the wave's first lane recomputes the Hz value of the column to
// the left of the tile (halo), and before its first row the wave recomputes the Hz
// row above the band.  Reads come from buffer A and writes go to buffer B
// (ping-pong), so the halo values recomputed here are the old ones whatever order
// the tiles run in.  Outer tangential E (Ex at j = 0, ny; Ey at i = 0, nx) is copied.
//
// CPML.  psi is read and written only in the border strips where the 1D profiles
// differ from (b, c, 1/kappa) = (1, 0, 1); elsewhere psi is +0 and the update
// reduces to  d*1 + (+0)  exactly as in the reference (the +0 keeps the sign of zero).
//
// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs; block b is
// mapped to tile (b % 8) * tiles_per_xcd + b / 8, so each XCD sweeps one contiguous
// slab of rows and the halo rows shared by vertically adjacent bands stay in that
// XCD's L2.
#include "common.h"
#include "fused.h"

namespace etl {
namespace {

// V consecutive REALs with only the alignment of one REAL: rows of Ex have an odd
// pitch (ny + 1), and gfx950 global loads/stores need only dword alignment.
template <typename R, int V>
using Vec = R __attribute__((ext_vector_type(V), aligned(sizeof(R))));

// Load V consecutive elements p[c..c+V-1], those with index >= limit read as 0.
template <typename R, int V>
__device__ __forceinline__ void load_row(R (&dst)[V], const R* __restrict__ p, int c, int limit) {
    if (c + V <= limit) {
        if constexpr (V == 1) {
            dst[0] = p[c];
        } else {
            const Vec<R, V> x = *reinterpret_cast<const Vec<R, V>*>(p + c);
#pragma unroll
            for (int k = 0; k < V; ++k) dst[k] = x[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k) dst[k] = (c + k < limit) ? p[c + k] : R(0);
    }
}

template <typename R, int V>
__device__ __forceinline__ void store_row(R* __restrict__ p, const R (&src)[V], int c, int limit) {
    if (c + V <= limit) {
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
            if (c + k < limit) p[c + k] = src[k];
    }
}

template <typename R>
__device__ __forceinline__ R shfl_from_next(R x) {  // lane l receives lane l+1's x
    return __shfl_down(x, 1, 64);
}
template <typename R>
__device__ __forceinline__ R shfl_from_prev(R x) {  // lane l receives lane l-1's x
    return __shfl_up(x, 1, 64);
}

// Subtract this step's deltas of the sources of component `comp` lying on row i.
template <typename R, int V>
__device__ __forceinline__ void inject_row(const FusedArgs<R>& a, const R* __restrict__ drow,
                                           int comp, int i, int c, R (&v)[V]) {
#pragma unroll
    for (int s = 0; s < kMaxSources; ++s) {
        if (s < a.n_src && a.src_comp[s] == comp && a.src_i[s] == i) {
            const int js = a.src_j[s] - c;
            if (js >= 0 && js < V) {
                const R dl = drow[s];
                // x - (+0) == x bit for bit, so the untouched elements are unchanged;
                // a select per element keeps v[] in registers (no dynamic index).
#pragma unroll
                for (int k = 0; k < V; ++k) v[k] = v[k] - ((k == js) ? dl : R(0));
            }
        }
    }
}

template <typename R>
__device__ __forceinline__ R inject_one(const FusedArgs<R>& a, const R* __restrict__ drow, int comp,
                                        int i, int j, R v) {
#pragma unroll
    for (int s = 0; s < kMaxSources; ++s)
        if (s < a.n_src && a.src_comp[s] == comp && a.src_i[s] == i && a.src_j[s] == j)
            v = v - drow[s];
    return v;
}

// Write the monitored values of row i of component `comp` into the trace row.
template <typename R, int V>
__device__ __forceinline__ void trace_row(const FusedArgs<R>& a, R* __restrict__ lds, int comp,
                                          int i, int tj, int c0, const R (&v)[V], R* trow) {
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
        const int2 en = ent[q];
        trow[en.y] = lds[en.x - c0];
    }
    __syncthreads();
}

template <typename R, int V>
__global__ __launch_bounds__(64) void fused_step_kernel(FusedArgs<R> a,
                                                        const int* __restrict__ counter,
                                                        int kstep, int parity) {
    const int b = blockIdx.x;
    const int tile = (b & 7) * a.tiles_per_xcd + (b >> 3);
    if (tile >= a.n_bands * a.n_tj) return;
    const int tb = tile / a.n_tj;
    const int tj = tile - tb * a.n_tj;
    const int lane = threadIdx.x;
    const int c0 = tj * 64 * V;
    const int c = c0 + lane * V;
    const int nx = a.nx, ny = a.ny;
    const int px = ny + 1;  // Ex pitch
    const int i0 = tb * a.band;
    const int i1 = min(i0 + a.band, nx);

    __shared__ R lds[64 * V];

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

    const int trow_index = *counter + kstep;  // row of the source / phasor tables
    const R* __restrict__ drow = a.delta + int64_t(trow_index) * a.n_src;
    R* __restrict__ trow = a.trace + int64_t(kstep) * a.n_slots;

    const R db = a.db, idx = a.inv_dx, idy = a.inv_dy;
    const R zero = R(0);

    // Column profiles of this lane's columns (only read inside the y strips).
    const bool tile_ys_h = (c0 - 1 < a.yh_lo) || (c0 + 64 * V > a.yh_hi);
    const bool tile_ys_e = (c0 < a.ye_lo) || (c0 + 64 * V > a.ye_hi);
    bool ysh[V], yse[V];
    R byh[V], cyh[V], ikyh[V], bye[V], cye[V], ikye[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int j = c + k;
        ysh[k] = tile_ys_h && (j < a.yh_lo || j >= a.yh_hi) && j < ny;
        yse[k] = tile_ys_e && (j < a.ye_lo || j >= a.ye_hi) && j >= 1 && j <= ny - 1;
        byh[k] = ysh[k] ? a.by_h[j] : R(1);
        cyh[k] = ysh[k] ? a.cy_h[j] : R(0);
        ikyh[k] = ysh[k] ? a.iky_h[j] : R(1);
        bye[k] = yse[k] ? a.by_e[j] : R(1);
        cye[k] = yse[k] ? a.cy_e[j] : R(0);
        ikye[k] = yse[k] ? a.iky_e[j] : R(1);
    }
    // Left halo column of the tile (first lane only).
    const int cw = c0 - 1;
    const bool has_left = (tj > 0);
    const bool ysw = has_left && (cw < a.yh_lo || cw >= a.yh_hi);
    R byw = R(1), cyw = R(0), ikyw = R(1);
    if (lane == 0 && ysw) {
        byw = a.by_h[cw];
        cyw = a.cy_h[cw];
        ikyw = a.iky_h[cw];
    }

    // Hz update of one node; psi is read from A and, when `store`, written to B.
    auto h_node = [&](R hz, R d, R e, bool ys, R by, R cy, R iky, bool xs, R bx, R cx, R ikx,
                      int64_t k_hz, bool store) -> R {
        R ty, tx;
        if (ys) {
            const R py = by * phyA[k_hz] + cy * d;
            if (store) phyB[k_hz] = py;
            ty = d * iky + py;
        } else {
            ty = d + zero;
        }
        if (xs) {
            const R pxv = bx * phxA[k_hz] + cx * e;
            if (store) phxB[k_hz] = pxv;
            tx = e * ikx + pxv;
        } else {
            tx = e + zero;
        }
        return hz + db * (ty - tx);
    };

    R hprev[V];  // Hz^(n+1/2) of the previous row
    R eyC[V];    // Ey^n of the current row
    R eyCw = R(0);

    // ---- prologue: Ey of the first row, and the Hz row above the band (halo) ----
    load_row<R, V>(eyC, eyA + int64_t(i0) * ny, c, ny);
    if (lane == 0 && has_left) eyCw = eyA[int64_t(i0) * ny + cw];
    if (i0 > 0) {
        const int ih = i0 - 1;
        R exh[V], eyh[V], hzh[V];
        load_row<R, V>(exh, exA + int64_t(ih) * px, c, px);
        load_row<R, V>(eyh, eyA + int64_t(ih) * ny, c, ny);
        load_row<R, V>(hzh, hzA + int64_t(ih) * ny, c, ny);
        R exr = shfl_from_next(exh[0]);
        if (lane == 63 && c + V <= ny) exr = exA[int64_t(ih) * px + c + V];
        const bool xs = (ih < a.xh_lo || ih >= a.xh_hi);
        R bx = R(1), cx = R(0), ikx = R(1);
        if (xs) {
            bx = a.bx_h[ih];
            cx = a.cx_h[ih];
            ikx = a.ikx_h[ih];
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const R exn = (k + 1 < V) ? exh[k + 1] : exr;
            const R d = (exn - exh[k]) * idy;
            const R e = (eyC[k] - eyh[k]) * idx;
            hprev[k] = (c + k < ny)
                           ? h_node(hzh[k], d, e, ysh[k], byh[k], cyh[k], ikyh[k], xs, bx, cx,
                                    ikx, int64_t(ih) * ny + c + k, false)
                           : R(0);
        }
        if (a.n_src) inject_row<R, V>(a, drow, 2, ih, c, hprev);
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k) hprev[k] = R(0);
    }

    // ---- march down the band ----
    for (int i = i0; i < i1; ++i) {
        R ex[V], hz[V], eyN[V], cbx[V], cby[V] = {};
        const int64_t rx = int64_t(i) * px;
        const int64_t rh = int64_t(i) * ny;
        load_row<R, V>(ex, exA + rx, c, px);
        load_row<R, V>(hz, hzA + rh, c, ny);
        load_row<R, V>(eyN, eyA + rh + ny, c, ny);
        load_row<R, V>(cbx, a.cb_ex + rx, c, px);
        if (i >= 1) load_row<R, V>(cby, a.cb_ey + rh, c, ny);
        R exr = shfl_from_next(ex[0]);
        if (lane == 63 && c + V <= ny) exr = exA[rx + c + V];

        const bool xsh = (i < a.xh_lo || i >= a.xh_hi);
        R bx = R(1), cx = R(0), ikx = R(1);
        if (xsh) {
            bx = a.bx_h[i];
            cx = a.cx_h[i];
            ikx = a.ikx_h[i];
        }

        // H half-step
        R hn[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const R exn = (k + 1 < V) ? ex[k + 1] : exr;
            const R d = (exn - ex[k]) * idy;
            const R e = (eyN[k] - eyC[k]) * idx;
            hn[k] = (c + k < ny) ? h_node(hz[k], d, e, ysh[k], byh[k], cyh[k], ikyh[k], xsh, bx,
                                          cx, ikx, rh + c + k, true)
                                 : R(0);
        }
        if (a.n_src) inject_row<R, V>(a, drow, 2, i, c, hn);

        // Hz^(n+1/2) at column c - 1
        R hl = shfl_from_prev(hn[V - 1]);
        R eyNw = R(0);
        if (lane == 0 && has_left) {
            const R exw = exA[rx + cw];
            const R hzw = hzA[rh + cw];
            eyNw = eyA[rh + ny + cw];
            const R d = (ex[0] - exw) * idy;
            const R e = (eyNw - eyCw) * idx;
            hl = h_node(hzw, d, e, ysw, byw, cyw, ikyw, xsh, bx, cx, ikx, rh + cw, false);
            if (a.n_src) hl = inject_one(a, drow, 2, i, cw, hl);
        }

        // E half-step
        const bool xse = (i < a.xe_lo || i >= a.xe_hi);
        R bxe = R(1), cxe = R(0), ikxe = R(1);
        if (xse) {
            bxe = a.bx_e[i];
            cxe = a.cx_e[i];
            ikxe = a.ikx_e[i];
        }
        R exo[V], eyo[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int j = c + k;
            // Ex(i, j), j in [1, ny - 1]
            if (j >= 1 && j <= ny - 1) {
                const R hleft = (k == 0) ? hl : hn[k - 1];
                const R d = (hn[k] - hleft) * idy;
                R t;
                if (yse[k]) {
                    const R p = bye[k] * pexA[rx + j] + cye[k] * d;
                    pexB[rx + j] = p;
                    t = d * ikye[k] + p;
                } else {
                    t = d + zero;
                }
                exo[k] = ex[k] + cbx[k] * t;
            } else {
                exo[k] = ex[k];
            }
            // Ey(i, j), i in [1, nx - 1]
            if (i >= 1 && j < ny) {
                const R d = (hn[k] - hprev[k]) * idx;
                R t;
                if (xse) {
                    const R p = bxe * peyA[rh + j] + cxe * d;
                    peyB[rh + j] = p;
                    t = d * ikxe + p;
                } else {
                    t = d + zero;
                }
                eyo[k] = eyC[k] - cby[k] * t;
            } else {
                eyo[k] = eyC[k];
            }
        }
        if (a.n_src) {
            inject_row<R, V>(a, drow, 0, i, c, exo);
            inject_row<R, V>(a, drow, 1, i, c, eyo);
        }

        store_row<R, V>(hzB + rh, hn, c, ny);
        store_row<R, V>(exB + rx, exo, c, px);
        store_row<R, V>(eyB + rh, eyo, c, ny);

        if (a.has_mon) {
            trace_row<R, V>(a, lds, 2, i, tj, c0, hn, trow);
            trace_row<R, V>(a, lds, 0, i, tj, c0, exo, trow);
            trace_row<R, V>(a, lds, 1, i, tj, c0, eyo, trow);
        }

#pragma unroll
        for (int k = 0; k < V; ++k) {
            hprev[k] = hn[k];
            eyC[k] = eyN[k];
        }
        eyCw = eyNw;
    }

    // ---- last band: Ey row nx is outer tangential E (copied) ----
    if (i1 == nx) {
        if (a.n_src) inject_row<R, V>(a, drow, 1, nx, c, eyC);
        store_row<R, V>(eyB + int64_t(nx) * ny, eyC, c, ny);
        if (a.has_mon) trace_row<R, V>(a, lds, 1, nx, tj, c0, eyC, trow);
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

}  // namespace

template <typename R>
int launch_fused_step(const FusedArgs<R>& a, int vec, const int* counter, int kstep, int parity,
                      hipStream_t stream) {
    const int n_tiles = a.n_bands * a.n_tj;
    const int grid = 8 * a.tiles_per_xcd;
    (void)n_tiles;
    switch (vec) {
        case 1:
            hipLaunchKernelGGL((fused_step_kernel<R, 1>), dim3(grid), dim3(64), 0, stream, a,
                               counter, kstep, parity);
            break;
        case 2:
            hipLaunchKernelGGL((fused_step_kernel<R, 2>), dim3(grid), dim3(64), 0, stream, a,
                               counter, kstep, parity);
            break;
        case 4:
            hipLaunchKernelGGL((fused_step_kernel<R, 4>), dim3(grid), dim3(64), 0, stream, a,
                               counter, kstep, parity);
            break;
        default:
            set_error("fused step: unsupported vector width %d", vec);
            return ETL_EARG;
    }
    ETL_LAUNCHED();
    return ETL_OK;
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

template int launch_fused_step<float>(const FusedArgs<float>&, int, const int*, int, int,
                                      hipStream_t);
template int launch_fused_step<double>(const FusedArgs<double>&, int, const int*, int, int,
                                       hipStream_t);
template int launch_trace_dft<float>(const float*, int, int, const float*, int, const int*,
                                     const int4*, int, float* const*, hipStream_t);
template int launch_trace_dft<double>(const double*, int, int, const double*, int, const int*,
                                      const int4*, int, double* const*, hipStream_t);

}  // namespace etl
