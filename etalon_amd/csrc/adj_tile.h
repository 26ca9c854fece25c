// Temporally blocked reverse steps of the discrete adjoint (the register-tile
// counterpart of the reverse-step kernels in adjoint.hip).
//
// Reverse step n, per node (lambda = adjoint state, same shapes as the forward state):
//   sources  lambda_F += dJ/dF of step n at the objective's monitored nodes
//   A (Hz)   La = [1 <= j <= ny-1] (iky_e Lu + cy_e (lambda_psiEy + Lu)),  Lu = cb_ex lambda_Ex
//            Lb = [1 <= i <= nx-1] (ikx_e Lw + cx_e (lambda_psiEx + Lw)),  Lw = -cb_ey lambda_Ey
//            lambda_Hz += idy (La(i,j) - La(i,j+1)) + idx (Lb(i,j) - Lb(i+1,j))
//            Ld = db iky_h lambda_Hz + cy_h Lpy, Lpy = lambda_psiHy + db lambda_Hz,
//            lambda_psiHy = by_h Lpy   (and Le, lambda_psiHx with -db, x profiles)
//   B (E)    lambda_psiEy = by_e (lambda_psiEy + Lu), dJ/dcb_ex += lambda_Ex u
//            lambda_Ex += idy (Ld(i,j-1) - Ld(i,j));  Ey alike with Le and rows
// A depends on the next column and the next row, B on the previous ones: the same
// dependency pattern as the forward H and E half-steps, so the forward kernel's
// decomposition carries over unchanged — a wave loads a window of BR + 2T rows by
// 64 V columns into registers, runs T reverse steps there, stores its BR x WO
// owned nodes into the other buffer of each pair.
//
// Interior tiles hold no lambda_psi: where c = 0 it never feeds back (La, Ld reduce
// to Lu, db lambda_Hz), so it is left untouched.  Strip tiles apply the full
// transposed CPML.  Extra tiles also inject the adjoint sources (runs of
// consecutive columns of one row) and accumulate dJ/dcb on the design box in
// double, from the forward Hz of each step (the history array).
#pragma once

#include "te_tile.h"

namespace etl {

template <typename R>
struct AdjTileArgs {
    const R* lam[7][2];  // ex ey hz psi_ex_y psi_ey_x psi_hz_x psi_hz_y, buffer p
    const R* cb_ex;
    const R* cb_ey;
    const R *bx_e, *cx_e, *ikx_e, *bx_h, *cx_h, *ikx_h;
    const R *by_e, *cy_e, *iky_e, *by_h, *cy_h, *iky_h;
    R db, inv_dx, inv_dy;
    int nx, ny, wo;
    int xh_lo, xh_hi, xe_lo, xe_hi, yh_lo, yh_hi, ye_lo, ye_hi;
    // sources: per extra tile a CSR of runs (row | comp << 28, first column, end
    // column, source index of the first column); table (steps, n_src)
    const int* tile_runs;
    const int4* runs;
    const R* src;
    int n_src;
    // design box and gradient
    int bi0, bj0, bni, bnj;
    double *gx, *gy;
    const R* hist;  // (rows, (bni + 1) * (bnj + 1))
};

template <typename R, int V, int RWR>
__device__ __forceinline__ void add_run(R (&f)[RWR][V], int lr, int c, int c_lo, int c_hi,
                                        const global_t<const R>* __restrict__ vals) {
#pragma unroll
    for (int l = 0; l < RWR; ++l) {
        if (l != lr) continue;  // uniform
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int j = c + k;
            const bool in = j >= c_lo && j < c_hi;
            f[l][k] = f[l][k] + (in ? vals[in ? j - c_lo : 0] : R(0));
        }
    }
}

// add_run with the values already in registers (v[k] = the run's value at column c + k,
// +0 outside it): the same sums.
template <typename R, int V, int RWR>
__device__ __forceinline__ void add_vals(R (&f)[RWR][V], int lr, const R (&v)[V]) {
#pragma unroll
    for (int l = 0; l < RWR; ++l) {
        if (l != lr) continue;  // uniform
#pragma unroll
        for (int k = 0; k < V; ++k) f[l][k] = f[l][k] + v[k];
    }
}

// Source runs of an extra tile whose values are loaded before the step loop.
constexpr int kRunRegs = 2;

// The reverse-step tile program of one wave: tile li = ((b & 7) * per_xcd + (b >> 3)) *
// waves + wave of `tiles`, for workgroup b of n_groups (XCD-major, as the forward tiles).
template <typename R, int V, int T, int BR, int S, bool EXTRA>
__device__ __forceinline__ void adj_tile_body(const AdjTileArgs<R>* __restrict__ args,
                                              const int4* __restrict__ tiles, int n_tiles, int b,
                                              int n_groups, int waves, int wave,
                                              const int* __restrict__ step, int rtop,
                                              int parity) {
    constexpr int RW = BR + 2 * T;
    // S as in tile_body: bit 0 the y strips (lambda psi_hz_y, psi_ex_y), bit 1 the x
    // strips (lambda psi_hz_x, psi_ey_x)
    constexpr bool SY = (S & 1) != 0, SX = (S & 2) != 0;
    const int per_xcd = (n_groups + 7) >> 3;
    const int li = ((b & 7) * per_xcd + (b >> 3)) * waves + wave;
    if (li >= n_tiles) return;
    const int4 td = tiles[li];  // first row, tile column, extra-tile index, simulation
    const AdjTileArgs<R>& a = args[td.w];
    const int xh_lo = a.xh_lo, xh_hi = a.xh_hi, xe_lo = a.xe_lo, xe_hi = a.xe_hi;
    const int yh_lo = a.yh_lo, yh_hi = a.yh_hi, ye_lo = a.ye_lo, ye_hi = a.ye_hi;
    const int wo = a.wo;
    const int lane = threadIdx.x & 63;
    const int nx = a.nx, ny = a.ny, px = ny + 1;
    const int o0 = td.y * wo;
    const int cs = o0 - T;
    const int c = cs + lane * V;
    const int i0 = td.x;
    const int g0 = i0 - T;

    using G = global_t<R>;
    const G* __restrict__ exA = to_global(a.lam[0][parity]);
    const G* __restrict__ eyA = to_global(a.lam[1][parity]);
    const G* __restrict__ hzA = to_global(a.lam[2][parity]);
    const G* __restrict__ pexA = to_global(a.lam[3][parity]);  // lambda psi_ex_y (Ex nodes)
    const G* __restrict__ peyA = to_global(a.lam[4][parity]);  // lambda psi_ey_x (Ey nodes)
    const G* __restrict__ phxA = to_global(a.lam[5][parity]);
    const G* __restrict__ phyA = to_global(a.lam[6][parity]);
    G* __restrict__ exB = to_global(const_cast<R*>(a.lam[0][parity ^ 1]));
    G* __restrict__ eyB = to_global(const_cast<R*>(a.lam[1][parity ^ 1]));
    G* __restrict__ hzB = to_global(const_cast<R*>(a.lam[2][parity ^ 1]));
    G* __restrict__ pexB = to_global(const_cast<R*>(a.lam[3][parity ^ 1]));
    G* __restrict__ peyB = to_global(const_cast<R*>(a.lam[4][parity ^ 1]));
    G* __restrict__ phxB = to_global(const_cast<R*>(a.lam[5][parity ^ 1]));
    G* __restrict__ phyB = to_global(const_cast<R*>(a.lam[6][parity ^ 1]));
    const G* __restrict__ cbx = to_global(a.cb_ex);
    const G* __restrict__ cby = to_global(a.cb_ey);

    // ---- the window: lambda_Ex, lambda_Hz, cb_ex rows 0..RW-1; lambda_Ey, cb_ey 0..RW
    constexpr int PY = SY ? RW : 1;
    constexpr int PX = SX ? RW : 1;
    constexpr int PE = SX ? RW + 1 : 1;
    R ex[RW][V], hz[RW][V], ey[RW + 1][V], cx[RW][V], cy[RW + 1][V];
    R phy[PY][V], phx[PX][V], pey[PY][V], pex[PE][V];
#pragma unroll
    for (int l = 0; l <= RW; ++l) {
        const int g = g0 + l;
        const int gx = min(max(g, 0), nx - 1);
        const int ge = min(max(g, 0), nx);
        if (l < RW) {
            ldv<R, V>(ex[l], exA, gx, px, c, px);
            ldv<R, V>(hz[l], hzA, gx, ny, c, ny);
            ldv<R, V>(cx[l], cbx, gx, px, c, px);
            if constexpr (SY) {
                ldv<R, V>(phy[SY ? l : 0], phyA, gx, ny, c, ny);
                ldv<R, V>(pey[SY ? l : 0], pexA, gx, px, c, px);
            }
            if constexpr (SX) ldv<R, V>(phx[SX ? l : 0], phxA, gx, ny, c, ny);
        }
        ldv<R, V>(ey[l], eyA, ge, ny, c, ny);
        ldv<R, V>(cy[l], cby, ge, ny, c, ny);
        if constexpr (SX) ldv<R, V>(pex[SX ? l : 0], peyA, ge, ny, c, ny);
    }

    const R db = a.db, idx = a.inv_dx, idy = a.inv_dy;
    bool mEx[V], ysh[V], yse[V];
    R byh[V], cyh[V], ikyh[V], bye[V], cye[V], ikye[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int j = c + k;
        mEx[k] = j >= 1 && j <= ny - 1;  // Ex nodes the forward updates
        ysh[k] = SY && j >= 0 && j < ny && (j < yh_lo || j >= yh_hi);
        yse[k] = SY && mEx[k] && (j < ye_lo || j >= ye_hi);
        byh[k] = ysh[k] ? to_global(a.by_h)[j] : R(1);
        cyh[k] = ysh[k] ? to_global(a.cy_h)[j] : R(0);
        ikyh[k] = ysh[k] ? to_global(a.iky_h)[j] : R(1);
        bye[k] = yse[k] ? to_global(a.by_e)[j] : R(1);
        cye[k] = yse[k] ? to_global(a.cy_e)[j] : R(0);
        ikye[k] = yse[k] ? to_global(a.iky_e)[j] : R(1);
    }
    // x-strip row profiles, one window row per lane (RW + 1 <= 64)
    R pbxh = R(1), pcxh = R(0), pikxh = R(1), pbxe = R(1), pcxe = R(0), pikxe = R(1);
    if constexpr (SX) {
        const int gl = g0 + min(lane, RW);
        const int gh = min(max(gl, 0), nx - 1), ge = min(max(gl, 0), nx);
        pbxh = to_global(a.bx_h)[gh];
        pcxh = to_global(a.cx_h)[gh];
        pikxh = to_global(a.ikx_h)[gh];
        pbxe = to_global(a.bx_e)[ge];
        pcxe = to_global(a.cx_e)[ge];
        pikxe = to_global(a.ikx_e)[ge];
    }

    // extra tiles: source runs and the design box
    int run_lo = 0, run_hi = 0;
    bool boxed = false;
    double gxa[EXTRA ? BR : 1][V], gya[EXTRA ? BR : 1][V];
    if constexpr (EXTRA) {
        if (a.tile_runs) {
            run_lo = to_global(a.tile_runs)[td.z];
            run_hi = to_global(a.tile_runs)[td.z + 1];
        }
        boxed = a.bni > 0 && i0 < a.bi0 + a.bni && i0 + BR > a.bi0 && o0 < a.bj0 + a.bnj &&
                o0 + wo > a.bj0;
#pragma unroll
        for (int l = 0; l < BR; ++l)
#pragma unroll
            for (int k = 0; k < V; ++k) gxa[l][k] = gya[l][k] = 0.0;
    }
    const int n_first = step[0];
    const int hrow0 = step[1];  // history row of step n_first

    // Loaded up front, so that no step waits on memory: the first kRunRegs source runs
    // (descriptor and the T steps' values) and the forward Hz history of the box rows
    // i0 - 1 .. i0 + BR - 1 at this lane's columns for the T steps.
    int n_cached = 0;
    int4v rd[kRunRegs];
    R rv[EXTRA ? kRunRegs : 1][T][V];
    R hv[EXTRA ? T : 1][EXTRA ? BR + 1 : 1][V];
    const int hw = a.bnj + 1;
    if constexpr (EXTRA) {
        n_cached = min(run_hi - run_lo, kRunRegs);
#pragma unroll
        for (int qq = 0; qq < kRunRegs; ++qq) {
            rd[qq] = int4v{0, 0, 0, 0};
            if (qq < n_cached) rd[qq] = to_global_v(a.runs)[run_lo + qq];
#pragma unroll
            for (int t = 0; t < T; ++t)
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const int j = c + k;
                    const bool in = qq < n_cached && j >= rd[qq].y && j < rd[qq].z;
                    rv[qq][t][k] = in ? to_global(a.src)[int64_t(n_first + rtop - t) * a.n_src +
                                                         rd[qq].w + (j - rd[qq].y)]
                                      : R(0);
                }
        }
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int m = 0; m <= BR; ++m)
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const int hr = i0 - 1 + m - a.bi0 + 1, hc = c + k - a.bj0 + 1;
                    const bool ok = boxed && hr >= 0 && hr <= a.bni && hc >= 0 && hc <= a.bnj;
                    hv[t][m][k] = ok ? to_global(a.hist)[int64_t(hrow0 + rtop - t) *
                                                             (a.bni + 1) * hw + hr * hw + hc]
                                     : R(0);
                }
    }

    // Per step: ld, le of the rows A has done for B (shared by the steps: a row's
    // values are read by B of the same step before A of the next step rewrites them),
    // and Lb of the row A does next.
    R ld[RW][V], le[RW][V], lbc[T][V];
    // Lb of window Ey row l.
    auto lb_row = [&](int l, R (&out)[V]) __attribute__((always_inline)) {
        const int g = g0 + l;
        const bool row_ok = g >= 1 && g <= nx - 1;
        bool xe = false;
        R ikx = R(1), cxv = R(0);
        if constexpr (SX) {
            xe = row_ok && (g < xe_lo || g >= xe_hi);
            if (xe) {
                ikx = lane_value(pikxe, l);
                cxv = lane_value(pcxe, l);
            }
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const R lw = -cy[l][k] * ey[l][k];
            R v = lw;
            if (SX && xe) v = ikx * lw + cxv * (pex[SX ? l : 0][k] + lw);
            out[k] = row_ok ? v : R(0);
        }
    };
    // The adjoint sources of step t (EXTRA tiles).
    auto add_sources = [&](int t) __attribute__((always_inline)) {
        const int r = rtop - t;  // table row n_first + r, history row hrow0 + r of this step
        if constexpr (EXTRA) {
#pragma unroll
            for (int qq = 0; qq < kRunRegs; ++qq) {
                if (qq >= n_cached) break;  // uniform
                const int comp = rd[qq].x >> 28, lr = (rd[qq].x & 0x0fffffff) - g0;
                if (comp == 0) add_vals<R, V, RW>(ex, lr, rv[qq][t]);
                else if (comp == 1) add_vals<R, V, RW + 1>(ey, lr, rv[qq][t]);
                else add_vals<R, V, RW>(hz, lr, rv[qq][t]);
            }
            for (int q = run_lo + n_cached; q < run_hi; ++q) {
                const int4v ru = to_global_v(a.runs)[q];
                const int comp = ru.x >> 28, g = ru.x & 0x0fffffff;
                const global_t<const R>* vals =
                    to_global(a.src) + int64_t(n_first + r) * a.n_src + ru.w;
                const int lr = g - g0;
                if (comp == 0) add_run<R, V, RW>(ex, lr, c, ru.y, ru.z, vals);
                else if (comp == 1) add_run<R, V, RW + 1>(ey, lr, c, ru.y, ru.z, vals);
                else add_run<R, V, RW>(hz, lr, c, ru.y, ru.z, vals);
            }
        }
    };
    // A of step t on window row l (Hz rows [t, RW - t)): lambda_Hz, and Ld, Le for B.
    auto a_row = [&](int t, int l) __attribute__((always_inline)) {
        R la[V], lb_next[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const R lu = cx[l][k] * ex[l][k];
            R v = lu;
            if (SY && yse[k]) v = ikye[k] * lu + cye[k] * (pey[SY ? l : 0][k] + lu);
            la[k] = mEx[k] ? v : R(0);
        }
        lb_row(l + 1, lb_next);
        const R la_r = from_next_lane(la[0]);
        const int g = g0 + l;
        bool xh = false;
        R bx = R(1), cxh = R(0), ikx = R(1);
        if constexpr (SX) {
            xh = g < xh_lo || g >= xh_hi;
            if (xh) {
                bx = lane_value(pbxh, l);
                cxh = lane_value(pcxh, l);
                ikx = lane_value(pikxh, l);
            }
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const R la_n = (k + 1 < V) ? la[k + 1] : la_r;
            const R lh = hz[l][k] + idy * (la[k] - la_n) + idx * (lbc[t][k] - lb_next[k]);
            hz[l][k] = lh;
            if (SY && ysh[k]) {
                const R lpy = phy[SY ? l : 0][k] + db * lh;
                ld[l][k] = db * ikyh[k] * lh + cyh[k] * lpy;
                phy[SY ? l : 0][k] = byh[k] * lpy;
            } else {
                ld[l][k] = db * lh;
            }
            if (SX && xh) {
                const R lpx = phx[SX ? l : 0][k] - db * lh;
                le[l][k] = -db * ikx * lh + cxh * lpx;
                phx[SX ? l : 0][k] = bx * lpx;
            } else {
                le[l][k] = -db * lh;
            }
            lbc[t][k] = lb_next[k];
        }
    };
    // B of step t on window row l: Ex rows [t, RW - t), Ey rows [t + 1, RW - t).
    auto b_row = [&](int t, int l) __attribute__((always_inline)) {
        const int g = g0 + l;
        const R ld_l = from_prev_lane(ld[l][V - 1]);
        const bool ey_row = l >= t + 1;
        const bool xe = SX && ey_row && g >= 1 && g <= nx - 1 && (g < xe_lo || g >= xe_hi);
        R bxe = R(1);
        if (SX && xe) bxe = lane_value(pbxe, l);
        const bool owned = EXTRA && l >= T && l < T + BR;
        // history of this row (m) and the one above (m - 1), and column j - 1
        const int m = owned ? l - T + 1 : 1;
        R h_left = R(0);
        if (owned && boxed) h_left = from_prev_lane(hv[EXTRA ? t : 0][EXTRA ? m : 0][V - 1]);
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int j = c + k;
            const R lex = ex[l][k];
            if (SY && yse[k]) pey[SY ? l : 0][k] = bye[k] * (pey[SY ? l : 0][k] + cx[l][k] * lex);
            if (owned && boxed && mEx[k] && g >= a.bi0 && g < a.bi0 + a.bni && j >= a.bj0 &&
                j < a.bj0 + a.bnj && j >= o0 && j < o0 + wo) {
                const R h0 = (k > 0) ? hv[EXTRA ? t : 0][EXTRA ? m : 0][k > 0 ? k - 1 : 0]
                                     : h_left;
                const double u =
                    (double(hv[EXTRA ? t : 0][EXTRA ? m : 0][k]) - double(h0)) * double(idy);
                gxa[owned ? l - T : 0][k] += double(lex) * u;
            }
            const R ldp = (k > 0) ? ld[l][k - 1] : ld_l;
            ex[l][k] = lex + idy * (ldp - ld[l][k]);
            if (ey_row) {
                const R ley = ey[l][k];
                if (SX && xe) pex[SX ? l : 0][k] = bxe * (pex[SX ? l : 0][k] - cy[l][k] * ley);
                if (owned && boxed && g >= a.bi0 && g < a.bi0 + a.bni && g >= 1 && g <= nx - 1 &&
                    j >= a.bj0 && j < a.bj0 + a.bnj && j >= o0 && j < o0 + wo) {
                    const double w = (double(hv[EXTRA ? t : 0][EXTRA ? m : 0][k]) -
                                      double(hv[EXTRA ? t : 0][EXTRA ? m - 1 : 0][k])) *
                                     double(idx);
                    gya[owned ? l - T : 0][k] -= double(ley) * w;
                }
                ey[l][k] = ley + idx * (le[l - 1][k] - le[l][k]);
            }
        }
    };
    // ---- store the owned rows [i0, i0 + BR) and columns [o0, o0 + WO)
    const int lo = max(o0, 0), hi = o0 + wo;
    // Store owned row l (window rows [T, T + BR)).
    auto store_owned = [&](int l) __attribute__((always_inline)) {
        const int g = g0 + l;
        if (g >= nx) return;
        store_row<R, V>(hzB + int64_t(g) * ny, hz[l], c, lo, min(hi, ny));
        store_row<R, V>(exB + int64_t(g) * px, ex[l], c, lo, min(hi, px));
        store_row<R, V>(eyB + int64_t(g) * ny, ey[l], c, lo, min(hi, ny));
        if constexpr (SY) {
            bool moy[V], moe[V];
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const bool own = (c + k >= lo && c + k < hi);
                moy[k] = own && ysh[k];
                moe[k] = own && yse[k];
            }
            store_masked<R, V>(phyB + int64_t(g) * ny, phy[SY ? l : 0], c, moy);
            store_masked<R, V>(pexB + int64_t(g) * px, pey[SY ? l : 0], c, moe);
        }
        if constexpr (SX) {
            bool mh[V];
#pragma unroll
            for (int k = 0; k < V; ++k) mh[k] = (c + k >= lo && c + k < hi) && c + k < ny;
            if (g < xh_lo || g >= xh_hi)
                store_masked<R, V>(phxB + int64_t(g) * ny, phx[SX ? l : 0], c, mh);
            if (g >= 1 && (g < xe_lo || g >= xe_hi))
                store_masked<R, V>(peyB + int64_t(g) * ny, pex[SX ? l : 0], c, mh);
        }
        if constexpr (EXTRA) {
            if (boxed) {
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const int j = c + k;
                    if (g >= a.bi0 && g < a.bi0 + a.bni && j >= a.bj0 && j < a.bj0 + a.bnj &&
                        j >= o0 && j < o0 + wo) {
                        const int64_t q = int64_t(g - a.bi0) * a.bnj + (j - a.bj0);
                        to_global(a.gx)[q] += gxa[l - T][k];
                        to_global(a.gy)[q] += gya[l - T][k];
                    }
                }
            }
        }
    };
    if constexpr (EXTRA) {
        // Step-major: sources and the forward history enter per step.
#pragma unroll
        for (int t = 0; t < T; ++t) {
            add_sources(t);
            lb_row(t, lbc[t]);
#pragma unroll
            for (int l = t; l < RW - t; ++l) a_row(t, l);
#pragma unroll
            for (int l = t; l < RW - t; ++l) b_row(t, l);
        }
#pragma unroll
        for (int l = T; l < T + BR; ++l) store_owned(l);
    } else {
        // Time-skewed, as the forward tiles: at front f, step t does row f - t (A then
        // B), each owned row stored when its last step is done.
#pragma unroll
        for (int f = 0; f < RW; ++f) {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int l = f - t;
                if (l < t || l >= RW - t) continue;
                if (l == t) lb_row(t, lbc[t]);
                a_row(t, l);
                b_row(t, l);
                if (t == T - 1 && l >= T && l < T + BR) store_owned(l);
            }
        }
    }
}

template <typename R, int V, int T, int BR, int S, bool EXTRA>
__global__ __launch_bounds__(256) void adj_tile_kernel(const AdjTileArgs<R>* __restrict__ args,
                                                       const int4* __restrict__ tiles, int n_tiles,
                                                       const int* __restrict__ step, int rtop,
                                                       int parity) {
    const int waves = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    adj_tile_body<R, V, T, BR, S, EXTRA>(args, tiles, n_tiles, blockIdx.x,
                                             (n_tiles + waves - 1) / waves, waves, wave, step,
                                             rtop, parity);
}

// One adjoint launch of depth T over all tiles (vec 1, 8-row interior tiles).
template <typename R>
int launch_adj(const AdjTileArgs<R>* args, const TileList& tl, int depth, int waves,
               const int* step, int rtop, int parity, hipStream_t s_int, hipStream_t s_strip,
               hipStream_t s_extra);

}  // namespace etl
