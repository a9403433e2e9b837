/*
 * isce3_amd -- finishing focused RSLC blocks on the GPU, C ABI.  Plain C
 * types only.
 *
 * Replaces the per-block work of nisar.workflows.focus.BackgroundWriter.write
 * (python/packages/nisar/workflows/focus.py:883-914) up to the file write:
 * scale by the range correction, add the block already written (mixed-mode
 * accumulation), compute the block's statistics
 * (isce3::math::StatsRealImag<float>, cxx/isce3/math/Stats.h:15-77,
 * Stats.cpp:33-165) and encode for storage -- complex64 as is, complex32
 * (pairs of IEEE half, isce3.core.types.to_complex32, types.py:38-42) or
 * complex64 with truncated mantissas (isce3.core.types.truncate_mantissa,
 * types.py:116-171).  The HDF5 write itself stays with the caller.
 *
 * Return codes as in isce3_amd_backproject.h; bp_last_error_message()
 * returns the message of the calling thread's last failure.
 */
#ifndef ISCE3_AMD_RSLC_H
#define ISCE3_AMD_RSLC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BackgroundWriter data_type (focus.py:845-857) */
enum { BP_RSLC_COMPLEX64 = 0, BP_RSLC_COMPLEX32 = 1, BP_RSLC_COMPLEX64_ZERO_MANTISSA = 2 };

/* isce3::math::Stats<float> (Stats.h:15-49): min/max/mean of the valid
 * values, NaN min/max when none is valid; sample standard deviation =
 * sqrt(square_diff_sum / (n_valid - 1)) for n_valid > 1, else 0. */
typedef struct bp_stats_f32 {
    float min, max, mean;
    int32_t pad_;
    int64_t n_valid;
    double real_valued_mean;
    double square_diff_sum;
} bp_stats_f32;

/* isce3::math::StatsRealImag<float> (Stats.h:54-77): a sample counts when
 * neither part is NaN. */
typedef struct bp_stats_real_imag_f32 {
    bp_stats_f32 real, imag;
    int64_t n_valid;
} bp_stats_real_imag_f32;

typedef struct bp_rslc_finisher bp_rslc_finisher;

/* A finisher for blocks of up to max_pixels samples on `device` (it holds the
 * per-workgroup partial statistics and a pinned result buffer). */
int bp_rslc_finisher_create(int device, int64_t max_pixels, bp_rslc_finisher** f);
int bp_rslc_finisher_destroy(bp_rslc_finisher* f);

/* BackgroundWriter.write (focus.py:883-914) on device buffers, in order:
 *   z[i][j] *= range_cor[j]            (range_cor: complex64 [cols], NULL = no scaling;
 *                                        products rounded as (ac - bd, ad + bc) without FMA)
 *   z[i][j] += prior[i][j]             (prior: complex64 rows x cols, NULL = none)
 *   *stats = StatsRealImag of z        (NaN-skipping)
 *   encode: BP_RSLC_COMPLEX64: nothing more;
 *           BP_RSLC_COMPLEX32: encoded[i][j] = (half(re), half(im)), round to nearest
 *             even, 4 bytes per sample (z keeps the complex64 values);
 *           BP_RSLC_COMPLEX64_ZERO_MANTISSA: the low 23 - mantissa_nbits mantissa bits
 *             of both parts of z are cleared in place; 0 < mantissa_nbits <= 23, else
 *             BP_ERR_INVALID_ARGUMENT (truncate_mantissa's ValueError).
 * z is complex64 rows x cols, row-major, updated in place.  rows * cols above
 * the finisher's max_pixels gives BP_ERR_LENGTH.  The call returns after the
 * statistics are on the host (it synchronises `stream`; NULL = the null
 * stream).  Calls on one finisher share its partial-statistics buffer: the
 * caller serialises them (different finishers are independent). */
int bp_rslc_finish_block(bp_rslc_finisher* f, void* z, int64_t rows, int64_t cols,
                         const void* range_cor, const void* prior, int data_type,
                         int mantissa_nbits, void* encoded, bp_stats_real_imag_f32* stats,
                         void* stream);

#ifdef __cplusplus
}
#endif

#endif
