// Monitor DFT from the trace buffer, and the step counter of fused plans.
//
// The fused kernels write each monitored node's value of every step into a trace
// row; this kernel folds a chunk of rows into the complex accumulators, one thread
// per (node, frequency), in step order: acc = acc + v*ph, re and im separately — the
// operations of NumpyBackend.accumulate_dft (etalon/fdtd/backends/numpy_backend.py:
// 63-67) applied step after step, so the accumulators are bit-identical.
#include "common.h"
#include "fused.h"

namespace etl {
namespace {

// acc[pair] += sum over the chunk's rows, in step order, of trace * phasor.  The
// rows are read in batches of B independent loads before the sequential fold, so
// the memory latency is paid once per batch; the last batch reads rows past the end
// as the last row again (in bounds, folded in by no one), so a chunk of up to B rows
// waits for memory once.  counter[0] is the step of the chunk's first row; the last
// block to finish advances it by `rows` (every block has read it by then) and clears
// the block tally counter[1] for the next chunk.
template <typename R>
__global__ __launch_bounds__(256) void trace_dft_kernel(const TraceSim<R>* __restrict__ sims,
                                                        int rows, int* __restrict__ counter,
                                                        const int4* __restrict__ pairs,
                                                        const int* __restrict__ pair_sim,
                                                        int n_pairs, R* const* __restrict__ accs) {
    constexpr int B = 32;  // two loads per row (trace value, phasor pair): 64 in flight
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int base = __atomic_load_n(counter, __ATOMIC_RELAXED);
    if (p < n_pairs) {
        const int4 pr = pairs[p];  // slot, q, accumulator, complex element index
        const TraceSim<R> sim = sims[pair_sim[p]];
        const int n_slots = sim.n_slots, nq = sim.nq;
        // global-address-space views: the pointers come from memory, where the compiler
        // cannot see their address space, and flat loads would also wait on the LDS and
        // scalar counters
        using GW = __attribute__((address_space(1))) R;
        using GR = __attribute__((address_space(1))) const R;
        GW* acc = (GW*)(accs[pr.z] + 2 * int64_t(pr.w));
        GR* __restrict__ php = (GR*)(sim.ph + 2 * (int64_t(base) * nq + pr.y));
        GR* __restrict__ tp = (GR*)(sim.trace + pr.x);
        R re = acc[0], im = acc[1];
        int r0 = 0;
        for (; r0 + B <= rows; r0 += B) {
            R v[B], pre[B], pim[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                v[b] = tp[int64_t(r0 + b) * n_slots];
                pre[b] = php[2 * int64_t(r0 + b) * nq];
                pim[b] = php[2 * int64_t(r0 + b) * nq + 1];
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                re = re + v[b] * pre[b];
                im = im + v[b] * pim[b];
            }
        }
        if (r0 < rows) {  // the last rows, fewer than B
            R v[B], pre[B], pim[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int64_t r = min(r0 + b, rows - 1);
                v[b] = tp[r * n_slots];
                pre[b] = php[2 * r * nq];
                pim[b] = php[2 * r * nq + 1];
            }
            const int m = rows - r0;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (b < m) {
                    re = re + v[b] * pre[b];
                    im = im + v[b] * pim[b];
                }
            }
        }
        acc[0] = re;
        acc[1] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned done = atomicAdd(reinterpret_cast<unsigned*>(counter + 1), 1u);
        if (done + 1 == gridDim.x) {
            counter[1] = 0;
            __atomic_store_n(counter, base + rows, __ATOMIC_RELAXED);
        }
    }
}

// One probe per blockIdx.y; its rows x n values spread over blockIdx.x.
template <typename R>
__global__ void probe_copy_kernel(const ProbeCopy* __restrict__ tab, int rows,
                                  const int* __restrict__ counter) {
    const ProbeCopy d = tab[blockIdx.y];
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(rows) * d.n) return;
    const int r = int(t / d.n);
    const int k = int(t - int64_t(r) * d.n);
    const int64_t base = int64_t(counter[0]) - counter[2];
    static_cast<R*>(d.out)[(base + r) * d.n + k] =
        static_cast<const R*>(d.trace)[int64_t(r) * d.n_slots + d.slot_base + k];
}

__global__ void bump_counter_kernel(int* counter, int by) {
    if (threadIdx.x == 0) counter[0] = counter[0] + by;
}

}  // namespace

template <typename R>
int launch_trace_dft(const TraceSim<R>* sims, int rows, int* counter, const int4* pairs,
                     const int* pair_sim, int n_pairs, R* const* accs, hipStream_t stream) {
    if (n_pairs == 0 || rows == 0) return ETL_OK;
    hipLaunchKernelGGL(trace_dft_kernel<R>, dim3((n_pairs + 255) / 256), dim3(256), 0, stream,
                       sims, rows, counter, pairs, pair_sim, n_pairs, accs);
    ETL_LAUNCHED();
    return ETL_OK;
}

template <typename R>
int launch_probe_copies(const ProbeCopy* tab, int n_probes, int max_n, int rows,
                        const int* counter, hipStream_t stream) {
    const int64_t total = int64_t(rows) * max_n;
    if (total == 0 || n_probes == 0) return ETL_OK;
    ETL_REQUIRE(n_probes <= 65535, "probe copy: %d probes in one plan (at most 65535)", n_probes);
    hipLaunchKernelGGL(probe_copy_kernel<R>, dim3(unsigned((total + 255) / 256), unsigned(n_probes)),
                       dim3(256), 0, stream, tab, rows, counter);
    ETL_LAUNCHED();
    return ETL_OK;
}

int launch_bump(int* counter, int by, hipStream_t stream) {
    hipLaunchKernelGGL(bump_counter_kernel, dim3(1), dim3(64), 0, stream, counter, by);
    ETL_LAUNCHED();
    return ETL_OK;
}

template int launch_probe_copies<float>(const ProbeCopy*, int, int, int, const int*, hipStream_t);
template int launch_probe_copies<double>(const ProbeCopy*, int, int, int, const int*, hipStream_t);
template int launch_trace_dft<float>(const TraceSim<float>*, int, int*, const int4*, const int*,
                                     int, float* const*, hipStream_t);
template int launch_trace_dft<double>(const TraceSim<double>*, int, int*, const int4*,
                                      const int*, int, double* const*, hipStream_t);

}  // namespace etl
