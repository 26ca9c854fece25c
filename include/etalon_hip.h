/*
 * etalon_hip.h — C ABI of libetalon_hip.so, the MI355X (gfx950) backend of the
 * etalon 2D TE FDTD solver.
 *
 * Every function returns 0 on success and a negative ETL_E* code on failure; the
 * message of the last failure on the calling thread is read with etl_last_error().
 * All "device pointers" are addresses on the library's device (etl_init), such as
 * those returned by etl_malloc().  No function takes or returns a torch type.
 * Calls are not thread-safe on one handle; the library issues its work on HIP
 * streams of default priority: lane 0 carries every call, except plans moved to
 * lane 1 (etl_plan_set_lane, etl_adjplan_set_lane) so that two plans run at once;
 * etl_lane_wait orders the lanes and etl_synchronize() waits for both.
 *
 * Precision codes: ETL_F32 (REAL = float, complex = 2 x float interleaved) and
 * ETL_F64 (REAL = double).  Arrays are C-ordered (x, y) with y fastest:
 *   hz (nx, ny)   ex (nx, ny + 1)   ey (nx + 1, ny)
 * Complex accumulators are (n_f, n_nodes) complex, (re, im) interleaved.
 *
 * Reference interfaces each entry point replaces are cited as
 * etalon/<file>:<line> in the reference tree.
 */
#ifndef ETALON_HIP_H
#define ETALON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ETL_F32 0
#define ETL_F64 1

#define ETL_OK 0
#define ETL_EARG (-1)     /* bad argument (size, null pointer, precision) */
#define ETL_EHIP (-2)     /* a HIP runtime call failed */
#define ETL_ESTATE (-3)   /* call out of order (no init, tables missing, ...) */
#define ETL_ENOMEM (-4)   /* device allocation failed */

/* ------------------------------------------------------------------ runtime */

/* Message of the last failure on this thread (NUL-terminated, truncated to n). */
int etl_last_error(char* buf, size_t n);
/* Select HIP device `device` and create the library stream. Idempotent.
 * Replaces the device probing of etalon/fdtd/backends/cupy_backend.py:36-59. */
int etl_init(int device);
/* "<device name> gfx<arch> CUs=<n> mem=<bytes>" of the initialised device. */
int etl_device_info(char* buf, size_t n);
/* Number of HIP devices visible (0 when there is no GPU); needs no etl_init. */
int etl_device_count(int* count);

/* ------------------------------------------------------------------ memory
 * Backend.to_device / to_host / synchronize of the backend protocol,
 * etalon/fdtd/backends/cupy_backend.py:61-68 (cp.asarray / cp.asnumpy). */
int etl_malloc(void** ptr, size_t bytes);
int etl_free(void* ptr);
int etl_memcpy_h2d(void* dst, const void* src, size_t bytes);  /* synchronous */
int etl_memcpy_d2h(void* dst, const void* src, size_t bytes);  /* synchronous */
int etl_memcpy_d2d(void* dst, const void* src, size_t bytes);  /* stream-ordered */
int etl_zero(void* dst, size_t bytes);                         /* fill kernel, stream-ordered */
/* Measurement only, no reference counterpart: copies `bytes` (a multiple of 16) between
 * two fresh device buffers with a 16-byte streaming kernel, once untimed and `reps`
 * times timed with HIP events; *ms_per_copy = the average.  The achievable-bandwidth
 * figure that bench.py reports beside the nominal HBM peak.  Synchronous. */
int etl_copy_rate(size_t bytes, int reps, float* ms_per_copy);
int etl_synchronize(void);
/* Lanes (see above): work issued on lane `waiter` after this call starts once the
 * work issued on lane `signaler` before it has finished.  No reference counterpart. */
int etl_lane_wait(int waiter, int signaler);
int etl_memcpy_d2d_lane(void* dst, const void* src, size_t bytes, int lane);

/* ------------------------------------------------------------------ protocol primitives
 * One launch each, in place, exactly the operations of the NumPy reference backend. */

/* Hz^(n-1/2) -> Hz^(n+1/2) with the CPML psi_Hz,x / psi_Hz,y recurrences.
 * Replaces NumpyBackend.update_h_te etalon/fdtd/backends/numpy_backend.py:41-49 and
 * kernel update_h_te etalon/fdtd/backends/cuda/update_te.cu:14-39 (same argument order). */
int etl_update_h_te(int precision, void* hz, const void* ex, const void* ey,
                    void* psi_hz_x, void* psi_hz_y,
                    const void* bx_h, const void* cx_h, const void* ikx_h,
                    const void* by_h, const void* cy_h, const void* iky_h,
                    double db, double inv_dx, double inv_dy, int nx, int ny);

/* Ex (j = 1..ny-1), Ey (i = 1..nx-1) from Hz^(n+1/2); outer tangential E untouched.
 * Replaces NumpyBackend.update_e_te numpy_backend.py:51-61 and kernel update_e_te
 * cuda/update_te.cu:41-75. */
int etl_update_e_te(int precision, void* ex, void* ey, const void* hz,
                    void* psi_ex_y, void* psi_ey_x,
                    const void* cb_ex, const void* cb_ey,
                    const void* bx_e, const void* cx_e, const void* ikx_e,
                    const void* by_e, const void* cy_e, const void* iky_e,
                    double inv_dx, double inv_dy, int nx, int ny);

/* acc[f, k] += field.flat[index[k]] * phasor[f]  (complex acc and phasor).
 * Replaces NumpyBackend.accumulate_dft numpy_backend.py:63-67, kernel cuda/dft.cu:8-23. */
int etl_accumulate_dft(int precision, void* acc, const void* field, const int64_t* index,
                       const void* phasor, int n_f, int n_nodes);

/* field[i[k] * pitch + j[k]] -= delta[k] for k < n (distinct nodes).
 * Replaces the scatter `field[i, j] -= delta` of FDTDSolver._inject etalon/fdtd/solver.py:83-89. */
int etl_scatter_sub(int precision, void* field, int64_t pitch, const int64_t* i,
                    const int64_t* j, const void* delta, int n);
/* out[k] = field[i[k] * pitch + j[k]] (materialises a fancy-indexed read). */
int etl_gather(int precision, void* out, const void* field, int64_t pitch, const int64_t* i,
               const int64_t* j, int n);

/* ------------------------------------------------------------------ run observers
 * On-device reductions for progress reports, replacing the device-side xp.sum and
 * block mean of etalon/studio/runner.py:258-311 (field_energy, downsample). */

/* out (device, 3 doubles) = { sum ex^2 / cb_ex, sum ey^2 / cb_ey, sum hz^2 } over every
 * node, each term (double(v) * double(v)) / double(cb) as field_energy forms it
 * (etalon/studio/runner.py:258-280); float64 sums in a fixed order (two passes of
 * 4096-element blocks and 256-wide binary trees), equal to NumPy's to float64
 * rounding.  Partial sums in a buffer from etl_malloc.  Stream-ordered (lane 0). */
int etl_field_energy(int precision, const void* ex, const void* ey, const void* hz,
                     const void* cb_ex, const void* cb_ey, int nx, int ny, double* out);

/* etl_field_energy without any allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); no reference
 * counterpart.  The partial sums live in a fixed device array of the library (1.5 MiB
 * per device, held until the process ends), so each of ex, ey, hz may have at most
 * 268435456 nodes (ETL_EARG beyond).  Same kernels and reduction order as
 * etl_field_energy: the same bits.  Every call on a device shares that array and
 * leaves its partials there; calls are issued on lane 0 and run in its order, and
 * calls that overlap on one GPU (the same work replayed from a captured graph on
 * another stream while one runs) race. */
int etl_field_energy_noalloc(int precision, const void* ex, const void* ey, const void* hz,
                             const void* cb_ex, const void* cb_ey, int nx, int ny, double* out);

/* Block means: out (device, (nx / sx) x (ny / sy) floats) [a, b] = the mean of
 * field[a sx + p, b sy + q] over p < sx, q < sy, summed in float64 in (p, q) row-major
 * order, divided by sx sy in float64 and rounded to float once (the snapshot block
 * mean of downsample, etalon/studio/runner.py:283-291).  Stream-ordered (lane 0). */
int etl_downsample(int precision, const void* field, int nx, int ny, int sx, int sy, float* out);

/* ------------------------------------------------------------------ fused time stepping
 * A plan advances the state many steps per call with one fused kernel per step
 * (H half-step, M injection, E half-step, J injection, monitor trace) and a DFT
 * kernel per trace chunk, replayed from a hipGraph.  It replaces the host loop of
 * FDTDSolver.run / step, etalon/fdtd/solver.py:120-128,155-160: the result after
 * etl_plan_advance(n) is bit-identical to n calls of step(). */

typedef struct EtlPlan* etl_plan_t;

/* Every field and cb array of a plan must come from etl_malloc: the fused kernel
 * issues whole vector loads that may start a few elements before a row and end a
 * few past it, which etl_malloc's guard bands make safe.  At creation the state must
 * hold psi = +0 outside the CPML strips and no -0 in Ex, Ey, Hz (the reference's
 * updates keep both true, so a state from zeros, from step() or from a plan always
 * qualifies); otherwise etl_plan_create returns ETL_ESTATE and the caller steps with
 * the per-half-step primitives. */
typedef struct EtlTeDesc {
    int precision;
    int nx, ny;
    double db, inv_dx, inv_dy;
    /* Ping-pong field pairs: [0] holds the current state, [1] is scratch of the
     * same shape (zero-filled by the caller).  The seven arrays are, in order,
     * ex, ey, hz, psi_ex_y, psi_ey_x, psi_hz_x, psi_hz_y. */
    void* field[7][2];
    const void* cb_ex;
    const void* cb_ey;
    /* CPML profiles, etalon/fdtd/state.py:15-49 naming. */
    const void *bx_e, *cx_e, *ikx_e, *bx_h, *cx_h, *ikx_h;
    const void *by_e, *cy_e, *iky_e, *by_h, *cy_h, *iky_h;
} EtlTeDesc;

/* Shape of the fused kernel (0 = library default: vec 1, band 8 or 16, temporal 4, 6
 * or 8, graph on, 64- or 96-step trace chunks, one wave per workgroup).  (vec,
 * temporal, band_rows) must be a compiled register-tile shape: float32 vec 1 with band
 * 8 or 16 at temporal depth 1, 2, 4 or 6, and with band 16 at depth 8; float32 vec 2
 * with band 8 at depth 1, 2 or 4; float64 vec 1 with band 8 or 16 at depth 1, 2, 4 or
 * 6.  Depths 6 and 8 run with launch 1 only. */
typedef struct EtlTune {
    int band_rows;     /* output rows of one wave's tile; 0: 16 for large float plans
                        * (vec 1, not launch 3, at least 8 tiles of 16 x 56 per CU),
                        * else 8 (large: the rule of the temporal default too) */
    int vec;           /* REAL elements per lane along y */
    int use_graph;     /* replay steps from a hipGraph (1) or launch directly (0) */
    int trace_chunk;   /* steps per DFT flush: a multiple of 2 * temporal (0: 64, or 96
                        * at depths 6 and 8) */
    int temporal;      /* steps fused per launch: 1 (none), 2, 4, or 6 with launch 1
                        * (band 8 or 16), or 8 in float with launch 1 and band 16; 0: for large plans 8 in float when the fields exceed
                        * the 256 MiB Infinity Cache (32 bytes per cell) and every grid
                        * has at least 4 M cells, else 6; 4 for the others */
    int waves;         /* tiles (waves) per workgroup of the interior/strip kernels: 1, 2, 4 */
    int launch;        /* 1: one kernel per step block covering every tile kind (vec 1,
                        * band 8 or 16; one wave per workgroup); 3: one kernel
                        * per tile kind on three streams; 0: 1 where the shape allows
                        * it, else 3 */
    int uniform_cb;    /* launch 1: interior tiles whose window holds one cb_ex and one
                        * cb_ey value (bit for bit; found when the plan lays out its tiles)
                        * read those two values instead of the cb rows.  0: on (1) with
                        * launch 1, -1: off.  After writing cb_ex or cb_ey of a plan's
                        * simulation, call etl_plan_touch. */
} EtlTune;

/* Components: 0 = Ex, 1 = Ey, 2 = Hz. */
int etl_plan_create(const EtlTeDesc* desc, const EtlTune* tune, etl_plan_t* out);
/* A plan over n_sims independent simulations of any grid sizes (one precision),
 * advanced together: every fused launch covers the tiles of all of them (one tile
 * work-list), and one DFT fold covers all their monitors.  The functions without
 * _sim address simulation 0; etl_plan_create is the batch of one.  All the
 * simulations' tables must start at the same step.  No reference counterpart: the
 * reference runs one simulation per FDTDSolver (etalon/fdtd/solver.py:49-160). */
int etl_plan_create_batch(const EtlTeDesc* descs, int n_sims, const EtlTune* tune,
                          etl_plan_t* out);
int etl_plan_destroy(etl_plan_t plan);
/* The plan's cb_ex or cb_ey arrays were written: lay out the tiles again (the uniform-
 * material tiles, EtlTune.uniform_cb) before the next step; nothing to do for a plan
 * without them.  No reference counterpart (the reference has no plan). */
int etl_plan_touch(etl_plan_t plan);
/* Point sources: node (comp, i, j) receives field -= delta[n][k] in step n; a node
 * listed more than once is updated once per listing, in list order. */
int etl_plan_set_sources(etl_plan_t plan, int n_nodes, const int* comp, const int64_t* i,
                         const int64_t* j);
/* Host table delta[(n - n0) * n_nodes + k] of REAL for steps [n0, n0 + rows). */
int etl_plan_source_table(etl_plan_t plan, int64_t n0, int rows, const void* delta);
/* A monitor component: flat C-order node indices into its component array, and its
 * device accumulator acc (n_f, n_nodes) complex.  Returns the monitor id in *id. */
int etl_plan_add_monitor(etl_plan_t plan, int comp, int n_nodes, const int64_t* flat_index,
                         int n_f, void* acc, int* id);
/* Host phasor rows (complex REAL, (rows, n_f)) for steps [n0, n0 + rows):
 * ph[n][f] is the factor of the H sample of step n when comp = Hz and of the E
 * sample otherwise (etalon/fdtd/solver.py:91-110). */
int etl_plan_phasor_table(etl_plan_t plan, int id, int64_t n0, int rows, const void* ph);
/* The same, for simulation `sim` of a batch plan. */
int etl_plan_set_sources_sim(etl_plan_t plan, int sim, int n_nodes, const int* comp,
                             const int64_t* i, const int64_t* j);
int etl_plan_source_table_sim(etl_plan_t plan, int sim, int64_t n0, int rows, const void* delta);
int etl_plan_add_monitor_sim(etl_plan_t plan, int sim, int comp, int n_nodes,
                             const int64_t* flat_index, int n_f, void* acc, int* id);
int etl_plan_phasor_table_sim(etl_plan_t plan, int sim, int id, int64_t n0, int rows,
                              const void* ph);
/* A time-domain probe on simulation `sim`: during etl_plan_advance(n, count), row r
 * of `out` (an etl_malloc buffer of out_rows x n_nodes REALs, count <= out_rows)
 * receives the values of the nodes of step n + r, sampled where monitors are (Hz
 * after the H update and M injection, E after the E update and J injection).  No
 * reference counterpart; the reference records such traces with a Python loop of
 * step() and single-node reads (etalon/validation/m0.py:153-158). */
int etl_plan_add_probe_sim(etl_plan_t plan, int sim, int comp, int n_nodes,
                           const int64_t* flat_index, void* out, int out_rows);
/* Monitor `id` of simulation `sim` reads the phasor columns of monitor `owner` (< id,
 * same n_f) instead of its own table: for monitors whose tables are equal (same
 * frequencies, both H or both E), which then is computed and uploaded once.  Tables
 * of the owner only; the shared monitor takes none. */
int etl_plan_share_phasors(etl_plan_t plan, int sim, int id, int owner);
/* The tuning the plan resolved (library defaults filled in). */
int etl_plan_tune(etl_plan_t plan, EtlTune* out);
/* Build the plan's layout now (tile work-list, monitor and probe slots, device
 * tables) instead of at the first table or advance call; a no-op when it is up to
 * date.  No reference counterpart (the reference has no plan). */
int etl_plan_prepare(etl_plan_t plan);
/* Advance steps [n, n + count); the tables must cover them.  On return *parity
 * tells which buffer of each pair holds the state (0: field[k][0]). */
int etl_plan_advance(etl_plan_t plan, int64_t n, int64_t count, int* parity);
/* The same, with probe rows starting at row probe_row0 of each probe's buffer (step
 * n + r to row probe_row0 + r; probe_row0 + count <= out_rows), so that one buffer can
 * hold several segments (double buffering against a consumer on the other lane). */
int etl_plan_advance_rows(etl_plan_t plan, int64_t n, int64_t count, int probe_row0,
                          int* parity);
/* Issue this plan's work on lane 0 or 1 (default 0). */
int etl_plan_set_lane(etl_plan_t plan, int lane);
/* Checkpoint storage of simulation `sim`: ptr[k] (etl_malloc) holds n_seg slots of
 * field k (ex ey hz psi_ex_y psi_ey_x psi_hz_x psi_hz_y order, each slot the field's
 * full array).  For the checkpointed adjoint; no reference counterpart. */
int etl_plan_set_checkpoints_sim(etl_plan_t plan, int sim, void* const* ptr, int n_seg);
/* Checkpoint storage of simulation `sim` that holds only a window of its fields: ptr[k]
 * (etl_malloc) holds n_seg slots, each the block of field k with Hz nodes [row0, row0 +
 * win_nx) x [col0, col0 + win_ny) (Ex one column more, Ey one row more) stored row after
 * row; saves and restores copy that block.  A plan over the window itself (a grid of
 * win_nx x win_ny) takes the same arrays with etl_plan_set_checkpoints_sim.  For the
 * checkpointed adjoint, whose recompute needs only the region a design box depends on
 * over one segment; no reference counterpart. */
int etl_plan_set_checkpoint_window_sim(etl_plan_t plan, int sim, void* const* ptr, int n_seg,
                                       int win_nx, int win_ny, int row0, int col0);
/* Save (restore = 0) or load (restore = 1) checkpoint slot `seg` of every simulation,
 * in the buffers that hold its state now, on the plan's lane.  Stream-ordered. */
int etl_plan_checkpoint(etl_plan_t plan, int seg, int restore);
/* Advance [n, n + count) (n a multiple of K), saving the state of every simulation
 * into checkpoint slot c / K before each segment of K steps starting at step c. */
int etl_plan_advance_checkpointed(etl_plan_t plan, int64_t n, int64_t count, int K,
                                  int* parity);
/* Event timing of `count` steps from step n (launched directly, not from the graph). */
typedef struct EtlTiming {
    double launch_ms;        /* average fused launch, fork to join (every tile kind) */
    double kernel_ms;        /* average duration of the kernel that advances the interior
                              * tiles: with launch 1 the one kernel of every tile kind,
                              * with launch 3 the interior-tile kernel alone */
    double total_ms;         /* the whole call */
    int64_t kernel_cells;    /* Hz cells that kernel advances per step (all simulations) */
    int launches;            /* fused launches timed */
    int depth;               /* steps per launch (the plan's temporal depth) */
    int launch;              /* the plan's launch mode (1 or 3) */
    int64_t uniform_cells;   /* of kernel_cells, those the uniform-material interior tiles
                              * own (EtlTune.uniform_cb): they read no cb array */
} EtlTiming;
int etl_plan_time(etl_plan_t plan, int64_t n, int64_t count, EtlTiming* out, int* parity);

/* ------------------------------------------------------------------ adjoint
 * The discrete adjoint of the step for permittivity gradients of DFT-monitor
 * objectives (no reference counterpart: the reference has no adjoint).  The adjoint
 * state lambda has the forward state's seven arrays and shapes; a reverse step n
 * adds dJ/d(monitored values of step n) to lambda, then applies the transposed E
 * and H half-steps (CPML recurrences transposed), and accumulates dJ/dcb_ex and
 * dJ/dcb_ey on a design box from the forward Hz of that step (a history the
 * caller provides, e.g. from a probe of a forward recompute). */
typedef struct EtlAdjoint* etl_adjoint_t;
typedef struct EtlAdjointDesc {
    int precision;
    int nx, ny;
    double db, inv_dx, inv_dy;
    void* lam[7];   /* lambda: ex, ey, hz, psi_ex_y, psi_ey_x, psi_hz_x, psi_hz_y (zeroed) */
    const void* cb_ex;
    const void* cb_ey;
    const void *bx_e, *cx_e, *ikx_e, *bx_h, *cx_h, *ikx_h;
    const void *by_e, *cy_e, *iky_e, *by_h, *cy_h, *iky_h;
    /* Design box: Ex and Ey nodes (i, j), box_i0 <= i < box_i0 + box_ni, box_j0 <= j <
     * box_j0 + box_nj, off the CPML (its profiles must be (1, 0, 1) there). */
    int box_i0, box_j0, box_ni, box_nj;
    double* grad_ex;    /* (box_ni, box_nj) dJ/dcb_ex, accumulated */
    double* grad_ey;    /* (box_ni, box_nj) dJ/dcb_ey, accumulated */
    /* (hist_rows, (box_ni + 1) * (box_nj + 1)) REAL: Hz after the H update of each step
     * of a run, on rows box_i0 - 1 .. box_i0 + box_ni - 1 and columns box_j0 - 1 ..
     * box_j0 + box_nj - 1; row r is step n_first + r of etl_adjoint_run. */
    const void* hist;
    int hist_rows;
} EtlAdjointDesc;
int etl_adjoint_create(const EtlAdjointDesc* desc, etl_adjoint_t* out);
int etl_adjoint_destroy(etl_adjoint_t adj);
/* Adjoint sources: distinct nodes (comp, flat index) and a host table (rows, n) REAL
 * of dJ/dv for steps 0 .. rows - 1. */
int etl_adjoint_sources(etl_adjoint_t adj, int n, const int* comp, const int64_t* flat, int rows,
                        const void* table);
/* Reverse steps n_first + count - 1 down to n_first (stream-ordered). */
int etl_adjoint_run(etl_adjoint_t adj, int64_t n_first, int count);

/* The same reverse steps with the temporally blocked register-tile kernels, for one
 * or many simulations at once (a work-list over all their tiles, like
 * etl_plan_create_batch).  desc.lam are the A buffers of lambda (etl_malloc, zeroed);
 * the plan allocates the B buffers and etl_adjplan_lambda names the one holding
 * lambda after a run.  tune: temporal 1, 2 or 4 and waves (vec 1, 8-row tiles). */
typedef struct EtlAdjPlan* etl_adjplan_t;
int etl_adjplan_create(const EtlAdjointDesc* descs, int n_sims, const EtlTune* tune,
                       etl_adjplan_t* out);
/* The tuning an adjoint plan resolved (its temporal depth, launch mode, band), e.g. for
 * a recompute plan to run step blocks of the same depth beside it in a co-plan. */
int etl_adjplan_tune(etl_adjplan_t plan, EtlTune* out);
int etl_adjplan_destroy(etl_adjplan_t plan);
int etl_adjplan_sources(etl_adjplan_t plan, int sim, int n, const int* comp, const int64_t* flat,
                        int rows, const void* table);
/* Reverse steps n_first + count - 1 .. n_first of every simulation (history row r =
 * step n_first + r). */
int etl_adjplan_run(etl_adjplan_t plan, int64_t n_first, int count);
/* The same, with history row hist_row0 + r = step n_first + r (hist_row0 + count <=
 * hist_rows). */
int etl_adjplan_run_rows(etl_adjplan_t plan, int64_t n_first, int count, int hist_row0);
/* Issue this plan's work on lane 0 or 1 (default 0). */
int etl_adjplan_set_lane(etl_adjplan_t plan, int lane);
int etl_adjplan_lambda(etl_adjplan_t plan, int sim, int field, void** ptr);
/* Zero lambda (both buffers of every simulation) for the next gradient; the layout,
 * graphs and sources stay.  etl_adjplan_sources with the same nodes and rows as before
 * only replaces the table. */
int etl_adjplan_reset(etl_adjplan_t plan);

/* A forward plan and an adjoint plan advanced together: forward steps [n, n +
 * count_f) (probe rows from probe_row0) and reverse steps n_first + count_a - 1 ..
 * n_first (history rows from hist_row0), step block i of each in one kernel launch,
 * e.g. the recompute of checkpoint segment k - 1 beside the reverse steps of segment
 * k.  Bit-identical to etl_plan_advance_rows then etl_adjplan_run_rows.  The forward
 * plan needs vec 1 and band 8 (or 16 in float); both plans on one lane; the plans
 * outlive the co-plan. */
typedef struct EtlCoPlan* etl_coplan_t;
int etl_coplan_create(etl_plan_t fwd, etl_adjplan_t adj, etl_coplan_t* out);
int etl_coplan_destroy(etl_coplan_t co);
int etl_coplan_run(etl_coplan_t co, int64_t n, int count_f, int probe_row0, int64_t n_first,
                   int count_a, int hist_row0);
/* The backward sweep of a checkpointed adjoint over steps [0, n_steps) in segments
 * of K: restore checkpoint slot k - 1 into the forward plan (etl_plan_set_checkpoints),
 * recompute segment k - 1 (probe rows (k - 1 & 1) K ..) beside the reverse steps of
 * segment k (history rows (k & 1) K ..), from the last segment down; the probes and
 * histories hold 2K rows. */
int etl_coplan_sweep(etl_coplan_t co, int64_t n_steps, int K);

#ifdef __cplusplus
}
#endif

#endif /* ETALON_HIP_H */
