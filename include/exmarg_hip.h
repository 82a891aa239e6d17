/*
 * exmarg_hip — C ABI of the MI355X (gfx950) implementation of the exmarg
 * segment analysis (package exmarg.search of the reference).
 *
 * Plain pointers and sizes only.  Host arrays are owned by the caller and
 * are not retained after a call returns.  Device memory is owned by the
 * library behind opaque 64-bit handles and released with exm_free.
 * Every call returns EXM_OK (0) or a negative code and then sets a
 * thread-local message readable with exm_last_error.  Calls are
 * synchronous: results written to host pointers are complete on return.
 *
 * Each entry point names the reference interface it replaces
 * (path:line in the reference repository).
 */
#ifndef EXMARG_HIP_H
#define EXMARG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t exm_handle;

enum {
    EXM_OK = 0,
    EXM_ERR_ARG = -1,     /* bad argument: the Python layer raises ValueError */
    EXM_ERR_HIP = -2,     /* HIP or hipFFT runtime failure: RuntimeError */
    EXM_ERR_NOMEM = -3,   /* device allocation refused or over budget: MemoryError */
    EXM_ERR_HANDLE = -4,  /* unknown or wrong-kind handle: ValueError */
    EXM_ERR_STATE = -5    /* library not initialised / no device: RuntimeError */
};

/* ---------------------------------------------------------------- library */

/* Select device `device` and create the library's stream.  Idempotent. */
int exm_init(int device);

/* Copy the last error message of this thread into buf (NUL-terminated). */
int exm_last_error(char* buf, size_t n);

/* Device name, HBM bytes and compute-unit count of the initialised device. */
int exm_device_info(char* name, size_t n, int64_t* total_mem, int* num_cu);

/* Release any object.  Replaces src/exmarg/search/store.py:123-127 (free) for tables
 * and the implicit cupy frees of trajectories (src/exmarg/search/pe.py:198-199). */
int exm_free(exm_handle h);

/* Bytes of device memory the library currently holds. */
int exm_mem_in_use(int64_t* bytes);

/* Free and total device memory (hipMemGetInfo); replaces the cupy memGetInfo of
 * src/exmarg/search/pe.py:43 that sizes segment chunks. */
int exm_mem_info(int64_t* free_bytes, int64_t* total_bytes);

/* ---------------------------------------------------------------- tables */

/* A device stack of ntab tables of shape (ntau, nphif, nb, nd) and their grids.
 * Replaces src/exmarg/search/store.py:110-120 (upload): the library allocates the stack;
 * tables are then set one by one.  beta0, dbeta, delta0, ddelta are beta[0], beta[1]-beta[0],
 * delta[0], delta[1]-delta[0] of the store grids (src/exmarg/search/gather.py:110-111). */
int exm_store_create(int ntab, int ntau, int nphif, int nb, int nd, const int64_t* tau_idx,
                     double beta0, double dbeta, double delta0, double ddelta, exm_handle* out);

/* Copy table k (host float32, reference layout (ntau, nphif, nb, nd), C order) to the device.
 * Replaces one copy of src/exmarg/search/store.py:117-118. */
int exm_store_set_table(exm_handle store, int k, const float* table);

/* Tables read straight from the store files: for table k, the file paths[k] holds the float32 C-order
 * (ntau, nphif, nb, nd) array at byte offset offsets[k] (the uncompressed T.npy member of a bd_*.npz)
 * and goes to device table tables[k].  `readers` threads read chunks into pinned buffers, copies run on
 * their own stream and the layout transform of one table overlaps the reading of the next; *seconds
 * receives the wall time.  Replaces the read and upload of src/exmarg/search/store.py:55-87 and
 * pe.py:146-154 (readers=6 there); the tables never pass through host arrays. */
int exm_store_load_files(exm_handle store, int n, const char* const* paths, const int64_t* offsets,
                         const int* tables, int readers, double* seconds);

/* Fill table k on the device with synthetic store values (seed, node, kind 0 smooth / 1 rough),
 * bit-identical to the host writer of the synthetic store files.  No reference counterpart:
 * it stands in for reading 0.75 GB files per node (src/exmarg/search/store.py:55-67). */
int exm_store_fill_synthetic(exm_handle store, int k, uint64_t seed, int64_t node, int kind);

/* Copy table k back to the host in the reference layout (for checks). */
int exm_store_read_table(exm_handle store, int k, float* table);

/* ---------------------------------------------------------------- bank and matched filter */

/* Template factors W (n_tmpl, n_freq) complex64 as interleaved float pairs, first bin lo,
 * filter length n_time.  Replaces the device copy of src/exmarg/search/bank.py:54 (MassBank.W). */
int exm_bank_create(const float* W, int n_tmpl, int n_freq, int lo, int64_t n_time, exm_handle* out);

/* Template factors computed on the device from TaylorF2 (detector-frame total mass M, mass ratio
 * 0 < q <= 1, common aligned spin chi per template) on the bins lo .. lo + n_freq - 1 of a
 * segment of T seconds: W = (a / D0) e^{-i psi} (D0 = d0_mpc) and u = sum a^2 / (D0^2 sig2_H)
 * (float64, written to u).  The device form of the bank construction of
 * src/exmarg/search/bank.py:35-54 with TaylorF2 in place of IMRPhenomXAS (no device
 * lalsimulation exists); the host model exmarg_amd.synth.taylorf2 computes the same factors. */
int exm_bank_taylorf2(const double* M, const double* q, const double* chi, int n_tmpl, double T, int lo, int n_freq,
                      int64_t n_time, const double* sig2_H, double d0_mpc, double* u, exm_handle* out);

/* Copy W (n_tmpl, n_freq) complex64 of a bank back to the host. */
int exm_bank_read(exm_handle bank, float* W);

/* Matched filter of one segment against every template of the bank: trajectories lqH, phH,
 * lqL, phL, each (n_tmpl, n_time) float32.  dH, dL are (n_freq) complex128 as interleaved
 * double pairs, s2H, s2L (n_freq) float64.  Weighting and transform in float64, one rounding
 * to float32 at the output.  Replaces src/exmarg/search/bank.py:56-78 (MassBank.trajectories). */
int exm_filter(exm_handle bank, const double* dH, const double* dL, const double* s2H, const double* s2L,
               exm_handle* traj_out);

/* Trajectories from host arrays (n_tmpl, n_time) float32, and back. */
int exm_traj_create(int n_tmpl, int64_t n_time, const float* lqH, const float* phH, const float* lqL,
                    const float* phL, exm_handle* out);
int exm_traj_read(exm_handle traj, float* lqH, float* phH, float* lqL, float* phL);

/* ---------------------------------------------------------------- arrival times */

/* Arrival times of every template: the K loudest by e^{2 lqL} + e^{2 max_{lag_lo..lag_hi} lqH}
 * (float64 score, ties to the smaller time) plus S drawn ones, pick (n_tmpl, S) int64 indexes
 * into the remaining times in time order, log weight lw_rand; all sorted by time.
 * Replaces src/exmarg/search/gather.py:374-420 (select_times). */
int exm_select_times(exm_handle traj, int lag_lo, int lag_hi, int K, int S, const int64_t* pick, float lw_rand,
                     exm_handle* times_out);

/* As exm_select_times, with the random draw made on the device from the generator state of
 * numpy.random.default_rng(seed + row): rng_state = (state >> 64, state & (2^64 - 1), inc >> 64,
 * inc & (2^64 - 1)) of its PCG64 bit generator.  The draw is numpy's integers(0, n_time - K,
 * (n_tmpl, S)) exactly (PCG64 XSL-RR words, low half first, Lemire's bounded method), the one
 * src/exmarg/search/gather.py:397 makes on the host. */
int exm_select_times_seeded(exm_handle traj, int lag_lo, int lag_hi, int K, int S, const uint64_t* rng_state,
                            float lw_rand, exm_handle* times_out);

/* The same draw copied to out (n values in [0, high)): numpy default_rng integers from rng_state. */
int exm_draw_integers(const uint64_t* rng_state, int64_t n, int64_t high, int64_t* out);

/* sel (n_tmpl, K+S) int32 and lw (n_tmpl, K+S) float32 of a times object. */
int exm_times_read(exm_handle times, int32_t* sel, float* lw);

/* ---------------------------------------------------------------- table read */

/* ln Lambda of (template, table) pairs: out (ns, nsel) float64 on the host.
 * mass_ids (nsel) int64, lna and tab_ids (nsel, ns); times 0 walks every arrival time.
 * Replaces src/exmarg/search/gather.py:340-371 (gather_gpu) and :114-140 (gather_cpu),
 * including _lse_lags (:103-107). */
int exm_gather(exm_handle traj, exm_handle store, const int64_t* mass_ids, int64_t nsel, const double* lna,
               const int64_t* tab_ids, int ns, exm_handle times, double* out);

/* A reusable list of njobs (template, table, ln a) reads for exm_plan_gather.  Built once per
 * bank, store and distance bound from the node geometry (src/exmarg/search/pe.py:58-84,
 * 181-185). */
int exm_plan_create(exm_handle store, int n_tmpl, int64_t njobs, const int32_t* job_m, const int32_t* job_tab,
                    const double* job_lna, exm_handle* out);

/* ln Lambda of every job of a plan for one segment, out (njobs) float64: every node group and
 * the boundary reads of src/exmarg/search/pe.py:178-194 in one launch. */
int exm_plan_gather(exm_handle plan, exm_handle traj, exm_handle store, exm_handle times, double* out);

/* The distance-prior weights of every template over the points it reaches, float64 on the device:
 * yN (nrow, nm) the points y = sqrt(a) of each template (y <= 0: not reached; nrow <= 64, every
 * template at least two points); lnW (nrow, nm) receives ln w (-inf where not reached) with
 * sum_j w_j Lambda(y_j) = integral p(y) Lambda(y) dy for Lambda piecewise linear in ln y.
 * prior 0: p = 3 ymin^3 y^-4 integrated exactly per interval; prior 1: the density p(ln y) tabulated
 * on lnx (ntab, ascending; zero outside), trapezoid rule on 401 points per interval.  Replaces
 * src/exmarg/search/pe.py:260-286 (distance_weights) as pe_intrinsic applies it per template
 * (pe.py:321-327). */
int exm_distance_weights(const double* yN, int nrow, int nm, int prior, double ymin, const double* lnx,
                         const double* px, int ntab, double* lnW);

/* The node weights of a plan's reads: lnw (njobs) ln w of each read in the plan's order (the distance
 * weight of its node, -inf allowed), cellw (n_tmpl) the template average weights (summing to one).
 * Kept with the plan for exm_plan_gather_lse and for batches admitting it. */
int exm_plan_set_weights(exm_handle plan, const double* lnw, const double* cellw);

/* exm_plan_gather with the result assembled on the device in float64: ln_lambda (n_tmpl) = LSE over
 * each template's reads of ln w + ln Lambda_read (src/exmarg/search/pe.py:333-336, in node order) and
 * ln_b (1) = ln sum_m c_m e^{ln_lambda_m} over the finite entries (pe.py:135-142, _log_mean). */
int exm_plan_gather_lse(exm_handle plan, exm_handle traj, exm_handle store, exm_handle times, double* ln_lambda,
                        double* ln_b);

/* Bytes of the no-allocation scratch region (one per device, held from the library's load until the
 * process ends). */
#define EXM_NOALLOC_SCRATCH_BYTES (128u << 20)

/* exm_plan_gather without any device allocation: the same kernels and results, every scratch buffer
 * taken from the library's static per-device region of EXM_NOALLOC_SCRATCH_BYTES.  No counterpart in
 * the reference: for callers that cannot allocate device memory (tight serving loops, harnesses that
 * forbid it).  A plan, trajectories and times whose gather needs more scratch than the region holds
 * are refused with EXM_ERR_NOMEM before anything runs (at the reference geometry, 87 lags and 2048
 * walked times, about 2.3 MB per template that reaches 46 nodes: about 55 such templates).  Calls on one
 * device share the region: calls that overlap race. */
int exm_plan_gather_noalloc(exm_handle plan, exm_handle traj, exm_handle store, exm_handle times, double* out);

/* ---------------------------------------------------------------- batches
 * Many segments analysed together (no reference counterpart: the reference loops over segments in
 * src/exmarg/search/pe.py:186-194 with one launch and one device-to-host copy per node group and
 * segment).  A batch holds the walk records (8 bytes per template, lag and walked time) of every
 * admitted segment; exm_batch_run reads every (segment, template, node) in one gather launch. */

/* A batch over store for at most max_templates templates in total, each walking n_walk times. */
int exm_batch_create(exm_handle store, int64_t max_templates, int n_walk, exm_handle* out);

/* Admit one segment: matched filter with bank, arrival-time selection (as exm_select_times), walk
 * records into the batch; the reads of plan (built for this bank) join the batch.  *slot receives
 * the segment's index in the batch.  EXM_ERR_NOMEM when the batch is full. */
int exm_batch_add(exm_handle batch, exm_handle bank, exm_handle plan, const double* dH, const double* dL,
                  const double* s2H, const double* s2L, int lag_lo, int lag_hi, int K, int S, const int64_t* pick,
                  float lw_rand, int* slot);
/* As exm_batch_add, with the draw made on the device from the generator state (see
 * exm_select_times_seeded). */
int exm_batch_add_seeded(exm_handle batch, exm_handle bank, exm_handle plan, const double* dH, const double* dL,
                         const double* s2H, const double* s2L, int lag_lo, int lag_hi, int K, int S,
                         const uint64_t* rng_state, float lw_rand, int* slot);
/* n segments of one bank sharing one noise variance (s2H, s2L) admitted together: dH, dL (n, n_freq),
 * rng_states (n, 4), one generator state per segment; each segment is filtered apart, the selection and
 * the walk run once over the n x n_tmpl templates.  The segments take slots *first_slot ..
 * *first_slot + n - 1.  The same results as n calls of exm_batch_add_seeded; the reference's chunk of
 * segments per call (src/exmarg/search/pe.py:173-194). */
int exm_batch_add_seeded_many(exm_handle batch, exm_handle bank, exm_handle plan, int n, const double* dH,
                              const double* dL, const double* s2H, const double* s2L, int lag_lo, int lag_hi, int K,
                              int S, const uint64_t* rng_states, float lw_rand, int* first_slot);

/* ln Lambda of every read of the batch: out holds, segment after segment, the reads of each segment's
 * plan in plan order (float64). */
int exm_batch_run(exm_handle batch, double* out);

/* Asynchronous form of exm_batch_run for continuous batching: exm_batch_launch enqueues the gather of
 * every admitted segment on the batch's own stream and returns; exm_batch_poll reports whether it
 * has finished; exm_batch_collect waits for it and writes the results as exm_batch_run does.  While a
 * batch is in flight, segments can be admitted to another batch (filter, selection and walk run on
 * the library stream and overlap the gather on the device).  An in-flight batch refuses
 * exm_batch_add and exm_batch_reset.  No reference counterpart (the reference analyses one segment
 * group at a time, src/exmarg/search/pe.py:173-194). */
int exm_batch_launch(exm_handle batch);
/* Grow the library's work memory to what admitting a segment of n_tmpl templates, n_time samples,
 * n_freq bins and n_walk walked times needs, so that later admissions allocate nothing (a growth
 * while another batch is in flight would wait for it).  No reference counterpart. */
int exm_reserve(int n_tmpl, int64_t n_time, int n_freq, int n_walk);
int exm_batch_poll(exm_handle batch, int* done);
int exm_batch_collect(exm_handle batch, double* out);
/* As exm_batch_collect, with every admitted segment's result assembled on the device as
 * exm_plan_gather_lse does (every segment's plan must carry node weights): ln_lambda (templates
 * used, admission order) and ln_b (one per segment). */
int exm_batch_collect_lse(exm_handle batch, double* ln_lambda, double* ln_b);

/* Empty the batch (keeps its memory). */
int exm_batch_reset(exm_handle batch);

int exm_batch_info(exm_handle batch, int64_t* templates_used, int64_t* reads, int* segments);

/* Choose the gather kernel: 0 automatic (the band kernel in its scheduled form, walks longer than 8192
 * times in slices), 1 the reference-order kernel, 2 the band kernel's piece form (pieces dealt over the
 * whole workgroup), 3 its wave form (each wave deals its own reads' pieces), 4 its scheduled form (the
 * pieces listed once for every lag, dealt to the waves in chunks, each read's sum kept in LDS against a
 * fixed reference).  No reference counterpart (the reference has one kernel per device). */
int exm_set_gather_kernel(int which);

/* Walked times per piece of the band kernel: 4, 6 or 8 in its scheduled and wave forms, 4, 8 or 16 in its
 * piece form (0 restores the default: 8 in the scheduled form, 6 in the wave form, 8 in the piece form).  Any
 * length gives the same results to rounding; the choice is a tuning knob.  No reference counterpart. */
int exm_set_gather_piece(int n);

/* Arithmetic of the band kernel's table interpolation: 0 float64 (default), 1 float32.  Index and
 * fraction arithmetic and every log-sum-exp stay float64 in both; the float32 form is the precision
 * form of the batched analysis (the reference's CUDA kernel interpolates in float32 throughout,
 * src/exmarg/search/gather.py:143-306).  Applies to the band kernel only. */
int exm_set_gather_precision(int p);

/* Reads per workgroup of the band kernel: 256, 512, 1024 or 2048 (0 restores the default, 2048; the wave
 * and piece forms take tiles of 2048 in two parts of at most 1024).  Smaller tiles give more workgroups,
 * each sweeping fewer table rows of its reads' band.  A tuning knob (the autotuner exmarg_amd.tune
 * chooses it); no reference counterpart. */
int exm_set_gather_tile(int n);

/* Workspace of one batched transform of the matched filter (templates per transform = bytes / (16
 * n_time), / (8 n_time) for the float32 transform; the transform's input and output take this much
 * each): 0 restores the default 64 MiB, which keeps both in the last-level cache.  A tuning knob; no
 * reference counterpart. */
int exm_set_filter_workspace(int64_t bytes);

/* Arithmetic of the matched filter's transform: 0 float64 (Z2Z, default), 1 float32 (C2C; weighting,
 * modulus and phase stay float64).  The float32 transform is the reference GPU path's
 * (src/exmarg/search/bank.py:66-75 keeps complex64); part of the precision form. */
int exm_set_filter_precision(int p);

/* The matched filter's transform: 0 (default) the band transform, two passes that read only the band's
 * n_freq bins of each template and write ln|z| and arg z directly, in the arithmetic that
 * exm_set_filter_precision chooses, wherever n_time / 512 is a power of two in [16, 2048] (hipFFT
 * otherwise); 1 hipFFT (Z2Z, or C2C in float32) always.  The same transform as
 * src/exmarg/search/bank.py:70-75 (xp.fft.fft of the zero-padded band); no reference counterpart for
 * the choice. */
int exm_set_filter_transform(int t);

/* The current tuning: gather kernel choice, piece length, tile and precision (0 where the default
 * applies). */
int exm_get_gather_tuning(int* kernel, int* piece, int* tile, int* precision);

/* ln Lambda of one segment on the source-frame grid (m1, z) from its node values, and its prior
 * average: lamN and yN (nrow, nm) float64 (rows: the nodes and the boundary row, nrow <= 64), Mdet
 * (nm) ascending detector-frame masses, yq (nz) = D0 / D_L(z), m1 (nm1), z (nz); out (nm1, nz)
 * (may be NULL: not copied back); w (nm1, nz) normalized prior weights or NULL; ln_b = ln sum
 * w e^out over the finite cells (written when w is given).  float64 throughout.  Replaces
 * src/exmarg/search/pe.py:87-132 (_to_source_frame) and :135-142 (_log_mean) inside pe_redshift
 * (pe.py:236-257). */
int exm_source_frame(const double* lamN, const double* yN, int nrow, int nm, const double* Mdet, const double* yq,
                     const double* m1, int nm1, const double* z, int nz, const double* w, double* out, double* ln_b);

/* The same step for many segments of one pe_redshift call: exm_source_grid_create keeps the call's
 * constants on the device (yN, Mdet, yq, m1, z and w as for exm_source_frame; the arrays are copied, the
 * caller may free them) and exm_source_grid_run takes one segment's node values lamN (nrow, nm) and
 * writes out (nm1, nz, may be NULL) and ln_b (when the grid has w).  Release the grid with exm_free.
 * Replaces the per-segment loop of src/exmarg/search/pe.py:236-257 (_to_source_frame, pe.py:87-132, and
 * _log_mean, :135-142, once per segment with the call's grid and prior). */
int exm_source_grid_create(const double* yN, int nrow, int nm, const double* Mdet, const double* yq, const double* m1,
                           int nm1, const double* z, int nz, const double* w, exm_handle* out);
int exm_source_grid_run(exm_handle grid, const double* lamN, double* out, double* ln_b);
/* exm_source_grid_run over n segments in one call: lamN (n, nrow, nm), out (n, nm1, nz) or NULL, ln_b (n).
 * Each grid's copy to the host overlaps the next segment's source-frame step on the device. */
int exm_source_grid_run_many(exm_handle grid, int n, const double* lamN, double* out, double* ln_b);

/* ---------------------------------------------------------------- time-domain front end
 * The frequency-domain segment of each of n_rows time series x (n_rows, n) float64 sampled at fs:
 * d (n_rows, n_freq) complex128 = rfft(x)[nu0 .. nu0 + n_freq - 1] / fs, the segment convention of
 * src/exmarg/pp.py:38-72 (bins ceil(fmin T) .. floor(fmax T), E|d|^2 = T S / 2 for white noise of
 * one-sided density S).  Replaces the host transform a caller of read_segments
 * (src/exmarg/search/segments.py:11-54) would run on strain data.  nu0 + n_freq - 1 <= n / 2. */
int exm_strain_segments(const double* x, int n_rows, int64_t n, double fs, int nu0, int n_freq, double* d);

/* The noise variance per bin from n_blocks neighbouring blocks x (n_blocks, n) float64: sig2 (n_freq)
 * = scale / n_blocks * sum_b |rfft(x_b)[nu0 + k] / fs|^2, the blocks added in order.  scale N/(N-2)
 * gives the estimate of notebooks/07_estimated_psd.ipynb cell 5 (there formed from simulated
 * periodograms), which keeps the noise mean of the likelihood ratio. */
int exm_strain_psd(const double* x, int n_blocks, int64_t n, double fs, int nu0, int n_freq, double scale,
                   double* sig2);

/* ---------------------------------------------------------------- table build, stage 3
 * The sum over the orders of one node (scripts/build_tables/build_bd.py:52-79, build_node): col and
 * row are the last column (q2 = q_max) and last row (q1 = q_max) of every (q1, q2) block of the
 * stage-2 array lnK_n(tau, q1, q2, phi), each float32 (nord, ntau, nq, nphif) C order (build_bd.py:61-62);
 * lnq (nq) and beta (nb) float64 the node's ln q grid and the beta grid (build_bd.py:52-59).  For every
 * truncation order ladder[l] (ascending, < nord; BD_NMAX_LADDER of build_bd.py:27,38) out[l] receives
 * T (ntau, nphif, nb, 2 nq - 1) float32 = max_n E_n + ln sum_n exp(E_n - max) over n <= ladder[l],
 * E_n = n beta + g_n(Delta) (build_bd.py:63-76), computed in float64.  nord <= 128, at most 32 orders
 * in the ladder. */
int exm_tables_sum_orders(const float* col, const float* row, int nord, int ntau, int nq, int nphif,
                          const double* lnq, const double* beta, int nb, const int* ladder, int nlad, float* const* out);

/* ---------------------------------------------------------------- measurement
 * No reference counterpart: kernel timing with HIP events recorded on the library's stream around
 * every launch of a named stage ("filter", "select", "gather"), for the benchmark. */
int exm_profile_enable(int on);
int exm_profile_reset(void);
/* total milliseconds and launch count of stage `name` since the last reset (synchronizes). */
int exm_profile_get(const char* name, double* total_ms, int64_t* count);

/* Failure injection for tests of the error paths: the device allocation after `after` more
 * successful ones (pooled reuse included) fails once with EXM_ERR_NOMEM, as a refused hipMalloc
 * would; a negative value clears it.  No reference counterpart. */
int exm_debug_fail_alloc(int64_t after);

/* Copy `bytes` (a multiple of 16) between two fresh device buffers with a plain copy kernel and
 * return its time: the kernel of known traffic that calibrates the memory-traffic counters of a
 * profiler.  No reference counterpart. */
int exm_copy_calibration(int64_t bytes, double* ms);

#ifdef __cplusplus
}
#endif

#endif
