// Protocol primitives: one launch per half-step, in place, node for node the
// operations of the NumPy reference backend.
//
//   H:  d  = (Ex[i][j+1] - Ex[i][j]) * inv_dy        e = (Ey[i+1][j] - Ey[i][j]) * inv_dx
//       py = by_h[j]*psi_y + cy_h[j]*d                px = bx_h[i]*psi_x + cx_h[i]*e
//       Hz = Hz + db*((d*iky_h[j] + py) - (e*ikx_h[i] + px))
//       (etalon/fdtd/backends/numpy_backend.py:41-49)
//   E:  Ex (j in 1..ny-1) += cb_ex*(dhz_dy*iky_e[j] + psi),  psi = by_e*psi + cy_e*dhz_dy
//       Ey (i in 1..nx-1) -= cb_ey*(dhz_dx*ikx_e[i] + psi),  psi = bx_e*psi + cx_e*dhz_dx
//       (numpy_backend.py:51-61)
//   DFT: acc[f][k] += v*ph[f] as (re + v*ph_re, im + v*ph_im)   (numpy_backend.py:63-67)
//
// Threads: 64 lanes of a wave run along y (the contiguous axis), 4 waves along x.
#include "common.h"

namespace {

constexpr int kBX = 64;  // along y
constexpr int kBY = 4;   // along x

template <typename R>
__global__ __launch_bounds__(kBX* kBY) void plain_update_h(
    R* __restrict__ hz, const R* __restrict__ ex, const R* __restrict__ ey,
    R* __restrict__ psi_hz_x, R* __restrict__ psi_hz_y,
    const R* __restrict__ bx_h, const R* __restrict__ cx_h, const R* __restrict__ ikx_h,
    const R* __restrict__ by_h, const R* __restrict__ cy_h, const R* __restrict__ iky_h,
    R db, R inv_dx, R inv_dy, int nx, int ny) {
    const int j = blockIdx.x * kBX + threadIdx.x;
    const int i = blockIdx.y * kBY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const int64_t k = int64_t(i) * ny + j;
    const int64_t kx = int64_t(i) * (ny + 1) + j;
    const R d = (ex[kx + 1] - ex[kx]) * inv_dy;
    const R e = (ey[k + ny] - ey[k]) * inv_dx;
    const R py = by_h[j] * psi_hz_y[k] + cy_h[j] * d;
    const R px = bx_h[i] * psi_hz_x[k] + cx_h[i] * e;
    psi_hz_y[k] = py;
    psi_hz_x[k] = px;
    hz[k] = hz[k] + db * ((d * iky_h[j] + py) - (e * ikx_h[i] + px));
}

template <typename R>
__global__ __launch_bounds__(kBX* kBY) void plain_update_e(
    R* __restrict__ ex, R* __restrict__ ey, const R* __restrict__ hz,
    R* __restrict__ psi_ex_y, R* __restrict__ psi_ey_x,
    const R* __restrict__ cb_ex, const R* __restrict__ cb_ey,
    const R* __restrict__ bx_e, const R* __restrict__ cx_e, const R* __restrict__ ikx_e,
    const R* __restrict__ by_e, const R* __restrict__ cy_e, const R* __restrict__ iky_e,
    R inv_dx, R inv_dy, int nx, int ny) {
    const int j = blockIdx.x * kBX + threadIdx.x;  // grid covers (nx + 1) x (ny + 1)
    const int i = blockIdx.y * kBY + threadIdx.y;
    if (i < nx && j >= 1 && j <= ny - 1) {
        const int64_t k = int64_t(i) * (ny + 1) + j;
        const int64_t kh = int64_t(i) * ny + j;
        const R d = (hz[kh] - hz[kh - 1]) * inv_dy;
        const R p = by_e[j] * psi_ex_y[k] + cy_e[j] * d;
        psi_ex_y[k] = p;
        ex[k] = ex[k] + cb_ex[k] * (d * iky_e[j] + p);
    }
    if (i >= 1 && i <= nx - 1 && j < ny) {
        const int64_t k = int64_t(i) * ny + j;
        const R d = (hz[k] - hz[k - ny]) * inv_dx;
        const R p = bx_e[i] * psi_ey_x[k] + cx_e[i] * d;
        psi_ey_x[k] = p;
        ey[k] = ey[k] - cb_ey[k] * (d * ikx_e[i] + p);
    }
}

template <typename R>
__global__ void plain_dft(R* __restrict__ acc, const R* __restrict__ field,
                          const int64_t* __restrict__ index, const R* __restrict__ ph, int n_f,
                          int n_nodes) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (k >= n_nodes || f >= n_f) return;
    const R v = field[index[k]];
    const int64_t a = 2 * (int64_t(f) * n_nodes + k);
    acc[a] = acc[a] + v * ph[2 * f];
    acc[a + 1] = acc[a + 1] + v * ph[2 * f + 1];
}

template <typename R>
__global__ void scatter_sub_kernel(R* __restrict__ field, int64_t pitch,
                                   const int64_t* __restrict__ ii, const int64_t* __restrict__ jj,
                                   const R* __restrict__ delta, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t at = ii[k] * pitch + jj[k];
    field[at] = field[at] - delta[k];
}

template <typename R>
__global__ void gather_kernel(R* __restrict__ out, const R* __restrict__ field, int64_t pitch,
                              const int64_t* __restrict__ ii, const int64_t* __restrict__ jj,
                              int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = field[ii[k] * pitch + jj[k]];
}

template <typename R>
int launch_h(void* hz, const void* ex, const void* ey, void* pxh, void* pyh, const void* bx,
             const void* cx, const void* ikx, const void* by, const void* cy, const void* iky,
             double db, double idx, double idy, int nx, int ny) {
    dim3 grid((ny + kBX - 1) / kBX, (nx + kBY - 1) / kBY);
    hipLaunchKernelGGL(plain_update_h<R>, grid, dim3(kBX, kBY), 0, etl::ctx().stream,
                       (R*)hz, (const R*)ex, (const R*)ey, (R*)pxh, (R*)pyh, (const R*)bx,
                       (const R*)cx, (const R*)ikx, (const R*)by, (const R*)cy, (const R*)iky,
                       R(db), R(idx), R(idy), nx, ny);
    ETL_LAUNCHED();
    return ETL_OK;
}

template <typename R>
int launch_e(void* ex, void* ey, const void* hz, void* pey, void* pex, const void* cbx,
             const void* cby, const void* bx, const void* cx, const void* ikx, const void* by,
             const void* cy, const void* iky, double idx, double idy, int nx, int ny) {
    dim3 grid((ny + 1 + kBX - 1) / kBX, (nx + 1 + kBY - 1) / kBY);
    hipLaunchKernelGGL(plain_update_e<R>, grid, dim3(kBX, kBY), 0, etl::ctx().stream, (R*)ex,
                       (R*)ey, (const R*)hz, (R*)pey, (R*)pex, (const R*)cbx, (const R*)cby,
                       (const R*)bx, (const R*)cx, (const R*)ikx, (const R*)by, (const R*)cy,
                       (const R*)iky, R(idx), R(idy), nx, ny);
    ETL_LAUNCHED();
    return ETL_OK;
}

}  // namespace

extern "C" {

int etl_update_h_te(int precision, void* hz, const void* ex, const void* ey, void* psi_hz_x,
                    void* psi_hz_y, const void* bx_h, const void* cx_h, const void* ikx_h,
                    const void* by_h, const void* cy_h, const void* iky_h, double db,
                    double inv_dx, double inv_dy, int nx, int ny) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(nx > 0 && ny > 0, "etl_update_h_te: bad grid %d x %d", nx, ny);
    ETL_REQUIRE(hz && ex && ey && psi_hz_x && psi_hz_y && bx_h && cx_h && ikx_h && by_h && cy_h &&
                    iky_h,
                "etl_update_h_te: null array");
    ETL_REQUIRE(nx <= 65535 * kBY, "etl_update_h_te: nx = %d too large", nx);
    if (precision == ETL_F32)
        return launch_h<float>(hz, ex, ey, psi_hz_x, psi_hz_y, bx_h, cx_h, ikx_h, by_h, cy_h,
                               iky_h, db, inv_dx, inv_dy, nx, ny);
    if (precision == ETL_F64)
        return launch_h<double>(hz, ex, ey, psi_hz_x, psi_hz_y, bx_h, cx_h, ikx_h, by_h, cy_h,
                                iky_h, db, inv_dx, inv_dy, nx, ny);
    ETL_REQUIRE(false, "etl_update_h_te: unknown precision %d", precision);
}

int etl_update_e_te(int precision, void* ex, void* ey, const void* hz, void* psi_ex_y,
                    void* psi_ey_x, const void* cb_ex, const void* cb_ey, const void* bx_e,
                    const void* cx_e, const void* ikx_e, const void* by_e, const void* cy_e,
                    const void* iky_e, double inv_dx, double inv_dy, int nx, int ny) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(nx > 0 && ny > 0, "etl_update_e_te: bad grid %d x %d", nx, ny);
    ETL_REQUIRE(ex && ey && hz && psi_ex_y && psi_ey_x && cb_ex && cb_ey && bx_e && cx_e &&
                    ikx_e && by_e && cy_e && iky_e,
                "etl_update_e_te: null array");
    ETL_REQUIRE(nx + 1 <= 65535 * kBY, "etl_update_e_te: nx = %d too large", nx);
    if (precision == ETL_F32)
        return launch_e<float>(ex, ey, hz, psi_ex_y, psi_ey_x, cb_ex, cb_ey, bx_e, cx_e, ikx_e,
                               by_e, cy_e, iky_e, inv_dx, inv_dy, nx, ny);
    if (precision == ETL_F64)
        return launch_e<double>(ex, ey, hz, psi_ex_y, psi_ey_x, cb_ex, cb_ey, bx_e, cx_e, ikx_e,
                                by_e, cy_e, iky_e, inv_dx, inv_dy, nx, ny);
    ETL_REQUIRE(false, "etl_update_e_te: unknown precision %d", precision);
}

int etl_accumulate_dft(int precision, void* acc, const void* field, const int64_t* index,
                       const void* phasor, int n_f, int n_nodes) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(n_f >= 0 && n_nodes >= 0 && n_f <= 65535, "etl_accumulate_dft: bad sizes %d, %d",
                n_f, n_nodes);
    if (n_f == 0 || n_nodes == 0) return ETL_OK;
    ETL_REQUIRE(acc && field && index && phasor, "etl_accumulate_dft: null array");
    dim3 grid((n_nodes + 127) / 128, n_f);
    if (precision == ETL_F32) {
        hipLaunchKernelGGL(plain_dft<float>, grid, dim3(128), 0, etl::ctx().stream, (float*)acc,
                           (const float*)field, index, (const float*)phasor, n_f, n_nodes);
    } else if (precision == ETL_F64) {
        hipLaunchKernelGGL(plain_dft<double>, grid, dim3(128), 0, etl::ctx().stream, (double*)acc,
                           (const double*)field, index, (const double*)phasor, n_f, n_nodes);
    } else {
        ETL_REQUIRE(false, "etl_accumulate_dft: unknown precision %d", precision);
    }
    ETL_LAUNCHED();
    return ETL_OK;
}

int etl_scatter_sub(int precision, void* field, int64_t pitch, const int64_t* i, const int64_t* j,
                    const void* delta, int n) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(n >= 0 && pitch > 0, "etl_scatter_sub: bad sizes n=%d pitch=%lld", n,
                (long long)pitch);
    if (n == 0) return ETL_OK;
    ETL_REQUIRE(field && i && j && delta, "etl_scatter_sub: null array");
    dim3 grid((n + 63) / 64);
    if (precision == ETL_F32) {
        hipLaunchKernelGGL(scatter_sub_kernel<float>, grid, dim3(64), 0, etl::ctx().stream,
                           (float*)field, pitch, i, j, (const float*)delta, n);
    } else if (precision == ETL_F64) {
        hipLaunchKernelGGL(scatter_sub_kernel<double>, grid, dim3(64), 0, etl::ctx().stream,
                           (double*)field, pitch, i, j, (const double*)delta, n);
    } else {
        ETL_REQUIRE(false, "etl_scatter_sub: unknown precision %d", precision);
    }
    ETL_LAUNCHED();
    return ETL_OK;
}

int etl_gather(int precision, void* out, const void* field, int64_t pitch, const int64_t* i,
               const int64_t* j, int n) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(n >= 0 && pitch > 0, "etl_gather: bad sizes");
    if (n == 0) return ETL_OK;
    ETL_REQUIRE(out && field && i && j, "etl_gather: null array");
    dim3 grid((n + 63) / 64);
    if (precision == ETL_F32) {
        hipLaunchKernelGGL(gather_kernel<float>, grid, dim3(64), 0, etl::ctx().stream,
                           (float*)out, (const float*)field, pitch, i, j, n);
    } else if (precision == ETL_F64) {
        hipLaunchKernelGGL(gather_kernel<double>, grid, dim3(64), 0, etl::ctx().stream,
                           (double*)out, (const double*)field, pitch, i, j, n);
    } else {
        ETL_REQUIRE(false, "etl_gather: unknown precision %d", precision);
    }
    ETL_LAUNCHED();
    return ETL_OK;
}

}  // extern "C"
