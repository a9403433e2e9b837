/*
 * isce3_amd -- presumming / BLU resampling of non-uniformly sampled pulses
 * (dithered PRF with gaps) onto a uniform pulse grid, C ABI.  Plain C types
 * only.
 *
 * Replaces isce3::focus::getPresumWeights (cxx/isce3/focus/Presum.h:13-60,
 * Presum.icc:9-60), its Python accelerators (pybind focus/Presum.cpp:26-135:
 * apply_presum_weights, compute_ids_from_mask, get_unique_ids, fill_weights)
 * and the workflow's resampling loop nisar.workflows.focus.resample
 * (python/packages/nisar/workflows/focus.py:973-1062), whose per-sample
 * weighted, Doppler-deramped sums run here in one HIP kernel over all output
 * pulses.
 *
 * Return codes as in isce3_amd_backproject.h; bp_last_error_message()
 * returns the message of the calling thread's last failure.
 */
#ifndef ISCE3_AMD_PRESUM_H
#define ISCE3_AMD_PRESUM_H

#include <stdint.h>

#include "isce3_amd_backproject.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Autocorrelation models (core/Kernels.h): AzimuthKernel(scale) has width
 * 2 scale (the sinc^2 antenna pattern's piecewise-cubic autocorrelation,
 * Kernels.icc:95-107); BartlettKernel(width) is the triangle of that width
 * (Kernels.icc:16-23). */
enum { BP_ACORR_AZIMUTH = 0, BP_ACORR_BARTLETT = 1 };
typedef struct {
    int32_t kind;   /* BP_ACORR_* */
    int32_t pad_;
    double width;   /* kernel width (AzimuthKernel: 2 x scale) */
} bp_acorr;

/* Kernel value at x. */
double bp_acorr_eval(const bp_acorr* acorr, double x);

/* getPresumWeights (Presum.icc:9-60): BLU weights reconstructing a sample at
 * tout from samples at times t[0..n) (sorted): *offset = index of the first
 * sample within half the kernel width of tout, *nw = number of such samples
 * (0 when none), w[0..nw) the solution of the autocorrelation system by
 * pivoted LDL^T.  BP_ERR_LENGTH when nw would exceed max_w. */
int bp_presum_weights(const bp_acorr* acorr, const double* t, int64_t n, double tout,
                      int64_t* offset, double* w, int64_t max_w, int64_t* nw);

/* apply_presum_weights (pybind focus/Presum.cpp:26-66) on device buffers:
 * out[j] = sum_i float(weights[i][j]) * exp(-2 pi i pulse_times[i] doppler[j])
 *          * pulses[i][j], complex64 out [n], float64 pulse_times [m], doppler
 * [n], weights [m x n], complex64 pulses [m x n].  NULL stream = synchronous. */
int bp_presum_apply_weights(void* out, const double* pulse_times, const double* doppler,
                            const double* weights, const void* pulses, int64_t m, int64_t n,
                            void* stream);

/* compute_ids_from_mask (pybind focus/Presum.cpp:96-125): ids[i] has bit j
 * set where mask[i][j] (host, row-major nr x np bytes); np <= 63 else
 * BP_ERR_LENGTH. */
int bp_presum_compute_ids(const uint8_t* mask, int64_t nr, int64_t np, int64_t* ids);

/* nisar.workflows.focus.resample (focus.py:973-1062): output pulse i at
 * tout = t[0] + i / grid->prf (grid->length pulses, grid->width range bins
 * starting at grid->starting_range) is the BLU reconstruction from the input
 * pulses within L / |v(tout)| of tout, with per-range-bin weights for the
 * pattern of valid samples given by the sub-swath windows, Doppler-deramped
 * at the LUT's fd(tout, r).
 *   t       host float64 [nt] pulse times (sorted)
 *   swaths  host int32 [ns][nt][2], [start, stop) valid range bins
 *   raw     device complex64 [nt][grid->width]
 *   out     device complex64 [grid->length][grid->width]
 * Weights are solved on host threads per (output pulse, distinct mask
 * pattern) while the sums run on the GPU, in launches of up to 2048 output
 * pulses on `stream`; the call returns when they have finished.  The tables
 * go up through pinned staging buffers that a process-wide pool keeps for
 * later calls (a few MB per concurrent call at 16384 pulses).  BP_ERR_LENGTH
 * when a window holds more than 63 pulses (the mask hash's limit). */
int bp_presum_resample(int device, const double* t, int64_t nt, const int32_t* swaths,
                       int64_t ns, const bp_grid* grid, const bp_orbit* orbit,
                       const bp_lut2d* doppler, double antenna_length, const void* raw,
                       void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif
