// Adjoint register-tile kernels of precision float: vec 1, band 8, depth 1, 2, 4.
#include "adj_tile.h"

namespace etl {
namespace {

template <typename R, int T>
int launch_depth(const AdjTileArgs<R>* args, const TileList& tl, int waves, const int* step,
                 int rtop, int parity, hipStream_t s_int, hipStream_t s_strip,
                 hipStream_t s_extra) {
    auto grid = [](int n, int w) { return dim3(unsigned(8 * (((n + w - 1) / w + 7) / 8))); };
    if (tl.n_interior) {
        hipLaunchKernelGGL((adj_tile_kernel<R, 1, T, 8, 0, false>), grid(tl.n_interior, waves),
                           dim3(64 * waves), 0, s_int, args, tl.interior, tl.n_interior, step, rtop,
                           parity);
        ETL_LAUNCHED();
    }
    if (tl.n_strip) {
        hipLaunchKernelGGL((adj_tile_kernel<R, 1, T, kSubBand, 3, false>),
                           grid(tl.n_strip, waves), dim3(64 * waves), 0, s_strip, args, tl.strip,
                           tl.n_strip, step, rtop, parity);
        ETL_LAUNCHED();
    }
    if (tl.n_extra_strip) {  // after the strip tiles, on their stream
        hipLaunchKernelGGL((adj_tile_kernel<R, 1, T, kSubBand, 3, true>),
                           grid(tl.n_extra_strip, 1), dim3(64), 0, s_strip, args, tl.extra_strip,
                           tl.n_extra_strip, step, rtop, parity);
        ETL_LAUNCHED();
    }
    if (tl.n_extra) {
        hipLaunchKernelGGL((adj_tile_kernel<R, 1, T, kSubBand, 0, true>), grid(tl.n_extra, 1),
                           dim3(64), 0, s_extra, args, tl.extra, tl.n_extra, step, rtop, parity);
        ETL_LAUNCHED();
    }
    return ETL_OK;
}

}  // namespace

template <>
int launch_adj<float>(const AdjTileArgs<float>* args, const TileList& tl, int depth, int waves,
                      const int* step, int rtop, int parity, hipStream_t s_int,
                      hipStream_t s_strip, hipStream_t s_extra) {
    switch (depth) {
        case 1:
            return launch_depth<float, 1>(args, tl, waves, step, rtop, parity, s_int, s_strip,
                                          s_extra);
        case 2:
            return launch_depth<float, 2>(args, tl, waves, step, rtop, parity, s_int, s_strip,
                                          s_extra);
        case 4:
            return launch_depth<float, 4>(args, tl, waves, step, rtop, parity, s_int, s_strip,
                                          s_extra);
        default:
            set_error("adjoint step: no compiled depth %d", depth);
            return ETL_EARG;
    }
}

}  // namespace etl
