// Register-tile kernels of precision double: the shapes listed in fused.h (kTileShapes).
#include "te_tile.h"

namespace etl {
namespace {

template <typename R, int V, int T, int BR>
int launch_shape(const FusedArgs<R>* args, const TileList& tl, int waves, const int* counter,
                 int kstep0, int parity, hipStream_t s_int, hipStream_t s_strip,
                 hipStream_t s_extra) {
    auto grid = [](int n, int w) { return dim3(unsigned(8 * (((n + w - 1) / w + 7) / 8))); };
    if (tl.n_interior) {
        hipLaunchKernelGGL((tile_kernel<R, V, T, BR, 0, false>), grid(tl.n_interior, waves),
                           dim3(64 * waves), 0, s_int, args, tl.interior, tl.n_interior, counter,
                           kstep0, parity);
        ETL_LAUNCHED();
    }
    if (tl.n_strip) {
        hipLaunchKernelGGL((tile_kernel<R, V, T, kSubBand, 3, false>), grid(tl.n_strip, waves),
                           dim3(64 * waves), 0, s_strip, args, tl.strip, tl.n_strip, counter,
                           kstep0, parity);
        ETL_LAUNCHED();
    }
    if (tl.n_extra_strip) {  // after the strip tiles, on their stream
        hipLaunchKernelGGL((tile_kernel<R, V, T, kSubBand, 3, true>), grid(tl.n_extra_strip, 1),
                           dim3(64), 0, s_strip, args, tl.extra_strip, tl.n_extra_strip, counter,
                           kstep0, parity);
        ETL_LAUNCHED();
    }
    if (tl.n_extra) {
        hipLaunchKernelGGL((tile_kernel<R, V, T, kSubBand, 0, true>), grid(tl.n_extra, 1),
                           dim3(64), 0, s_extra, args, tl.extra, tl.n_extra, counter, kstep0,
                           parity);
        ETL_LAUNCHED();
    }
    return ETL_OK;
}

}  // namespace

template <>
int launch_fused<double>(const FusedArgs<double>* args, const TileList& tl, int vec, int depth,
                        int band, int waves, const int* counter, int kstep0, int parity,
                        hipStream_t s_int, hipStream_t s_strip, hipStream_t s_extra) {
    if (vec == 1 && depth == 1 && band == 8)
        return launch_shape<double, 1, 1, 8>(args, tl, waves, counter, kstep0, parity, s_int,
                                                s_strip, s_extra);
    if (vec == 1 && depth == 2 && band == 8)
        return launch_shape<double, 1, 2, 8>(args, tl, waves, counter, kstep0, parity, s_int,
                                                s_strip, s_extra);
    if (vec == 1 && depth == 4 && band == 8)
        return launch_shape<double, 1, 4, 8>(args, tl, waves, counter, kstep0, parity, s_int,
                                                s_strip, s_extra);
    if (vec == 1 && depth == 1 && band == 16)
        return launch_shape<double, 1, 1, 16>(args, tl, waves, counter, kstep0, parity, s_int,
                                                s_strip, s_extra);
    if (vec == 1 && depth == 2 && band == 16)
        return launch_shape<double, 1, 2, 16>(args, tl, waves, counter, kstep0, parity, s_int,
                                                s_strip, s_extra);
    if (vec == 1 && depth == 4 && band == 16)
        return launch_shape<double, 1, 4, 16>(args, tl, waves, counter, kstep0, parity, s_int,
                                                s_strip, s_extra);
    set_error("fused step: no compiled tile shape vec=%d depth=%d band=%d", vec, depth, band);
    return ETL_EARG;
}

}  // namespace etl
