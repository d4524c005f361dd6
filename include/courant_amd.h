/*
 * courant_amd.h - MI355X engine extensions of libcourant.so.
 *
 * These entry points have no counterpart in the reference API; they serve
 * benchmarking and serving code that drives the GPU engine directly.
 */
#ifndef COURANT_AMD_H
#define COURANT_AMD_H

#include "courant.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Advance n steps on the GPU without refreshing host mirrors of fields or
 * spectra (timed loops).  Returns 0, or -1 with cem_last_error() set. */
CEM_API int cem_amd_step_device(cem_sim *s, long n);

/* Run the launch-shape autotuner now instead of on the first long enough
 * cem_step / cem_amd_step_device call; its steps advance the simulation like
 * ordinary steps.  Returns the number of steps it ran (0 when the shape was
 * already fixed), or -1 with cem_last_error() set. */
CEM_API long cem_amd_tune(cem_sim *s);

/* Both excitations of a 2-port as right-hand sides of shared launches.
 *
 * cem_amd_clone_excitation: a new simulation with a's plan for another excited port
 * of the same board.  Every set-up array is copied (no set-up step runs again; mode-port
 * cross-sections keep their solved state); its fields start from zero at step 0.  The
 * excitation moves to lumped port `port` with a's Gaussian (amplitude, t0, tau), or,
 * when a drives a mode port, to mode port `port` with a's modal waveform.  Destroy it
 * with cem_destroy.  Returns NULL with cem_last_error() set if a is not finalized or the
 * port does not exist.
 *
 * cem_amd_step_pair: advance a and its clone b (same step) n steps.  The H and E updates
 * of both run as one launch each per step (b's tiles next to a's, reading a's
 * coefficient classes, Debye tables and CPML profiles once for both); post-phase and
 * probe kernels run per simulation; chunks of steps replay as one graph on a's stream.
 * Each simulation's fields, spectra and states are bit-identical to stepping it alone.
 * Host mirrors are not refreshed (as cem_amd_step_device).  Returns 0, or -1 with
 * cem_last_error() set when b is not a clone of a, the two are at different steps, or
 * either runs the one-pass step. */
CEM_API cem_sim *cem_amd_clone_excitation(const cem_sim *a, int port);
CEM_API int cem_amd_step_pair(cem_sim *a, cem_sim *b, long n);

/* Queue the running DFT of every sample buffered since the last flush on the
 * simulation's stream (the engine otherwise flushes every 512 steps and when
 * spectra are read).  Returns 0, or -1 if the simulation has not run on the
 * GPU. */
CEM_API int cem_amd_flush_dft(cem_sim *s);

/* Wait until every queued kernel of this simulation has finished. */
CEM_API int cem_amd_synchronize(cem_sim *s);

/* Record HIP events around every kernel launch (on the engine's own stream)
 * and reset the statistics. */
CEM_API int cem_amd_set_profiling(cem_sim *s, int on);

/* Total GPU time (ms) and launch count of one kernel since profiling was
 * switched on: which = 0 H update, 1 E update, 2 probe gather, 3 DFT,
 * 4 post-phase edge physics (Debye, SIBC, sources), 5 one-pass step (H and E
 * updates of a step in one kernel; full steps use it in place of 0 and 1 when
 * cem_amd_step_shape reports rows > 0).  Profiling launches every kernel
 * directly (no graph replay). */
CEM_API int cem_amd_kernel_stats(cem_sim *s, int which, double *total_ms, long *count);

/* 2x2 S-matrix renormalisation of nf frequencies on the GPU in FP64, in place:
 * S holds 8 doubles per frequency (S11, S12, S21, S22 as re, im), z1[q], z2[q]
 * are the ports' reference impedances at frequency q.  Same formula as
 * cem_s2_renormalize (solver/include/courant.h, modeport.c:242-264), batched
 * over frequencies.  Returns 0, or -1 with cem_last_error() set. */
CEM_API int cem_amd_s2_renormalize(double *S, int nf, const double *z1, const double *z2, double zref);

/* The same renormalisation without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Its device staging
 * is a fixed array in the library (5 MiB per GPU, held until the process
 * ends), so nf is limited to 65536; larger nf returns -1 with an error.  The
 * array is shared by every call on one GPU: calls that overlap on one GPU
 * race.  Works on the device's null stream.  Returns 0, or -1 with
 * cem_last_error() set. */
CEM_API int cem_amd_s2_renormalize_noalloc(double *S, int nf, const double *z1, const double *z2, double zref);

/* Launch shape of the stencil kernels (x planes walked per thread by the H and
 * E updates).  Returns 1 once the autotuner (or COURANT_ISEG_H / _E) has fixed
 * it, 0 before, -1 if the simulation has not run on the GPU. */
CEM_API int cem_amd_launch_shape(cem_sim *s, int *iseg_h, int *iseg_e);

/* k points per thread of the H and E updates.  The H width is chosen by the
 * autotuner among 4, 2 and 1 (small grids step faster with more threads)
 * unless COURANT_LANES_H fixes it; the E width is COURANT_LANES_E (default 1).
 * Returns 1 once the autotuner has run (or the shape was fixed), 0 before, -1
 * if the simulation has not run on the GPU. */
CEM_API int cem_amd_lane_widths(cem_sim *s, int *lanes_h, int *lanes_e);

/* Tile order of the H and E updates: bit 0 set when each XCD runs one
 * contiguous band of tiles (tiles that share halo rows share its L2), clear in
 * dispatch order; bits 2 and up, in dispatch order, log2 of the tiles per
 * chunk (0: consecutive tiles on consecutive XCDs; s: runs of 2^s consecutive
 * tiles per XCD, the eight XCDs streaming one window); bit 1 set when the H
 * update runs its tiles from the last (each sweep then starts on the planes the
 * one before it touched last).  Chosen by the autotuner per grid shape unless
 * COURANT_TILE_ORDER=xcd|plain|plain<T> (T tiles per chunk) and
 * COURANT_H_REVERSE=0|1 fix them.  -1 if the simulation has not run on the
 * GPU. */
CEM_API int cem_amd_tile_order(cem_sim *s);

/* Bytes per grid point of the coefficient data the E update reads: 1 when the
 * grid holds at most 256 distinct (Ex, Ey, Ez) class triples (a class byte per
 * point and a 256-entry table; COURANT_CLASS_WORDS keeps the words), 4 for a
 * class word per point (three 8-bit classes), 24 for the ca and cb arrays
 * (more than 256 classes per component, or COURANT_COEF_ARRAYS).  -1 if the
 * simulation has not run on the GPU. */
CEM_API int cem_amd_class_bytes(cem_sim *s);

/* Shape of the one-pass step kernel (H and E updates of a step in one sweep,
 * field sets alternating), which full steps (cem_step, cem_run) use when
 * COURANT_STEP=onepass is in the environment at set-up: rows = grid rows (y)
 * owned by one workgroup, planes = x planes it walks.  rows = planes = 0 when
 * full steps run the separate H and E updates (the default).  The phase calls
 * cem_mt_H / cem_mt_E always run the separate updates.  Returns 0, or -1 if the
 * simulation cannot run on the GPU. */
CEM_API int cem_amd_step_shape(cem_sim *s, int *rows, int *planes);

/* Stream-copy bandwidth of the current GPU in GB/s (read + write bytes), best
 * of `iters` copies of `bytes` between two device buffers; -1 on error. */
CEM_API double cem_amd_copy_bandwidth(size_t bytes, int iters);

#ifdef __cplusplus
}
#endif
#endif
