/*
 * isce3_amd -- range compression (frequency-domain matched filter) on the
 * GPU, C ABI.  Plain C types only.
 *
 * Replaces isce3::focus::RangeComp (cxx/isce3/focus/RangeComp.h:11-106,
 * RangeComp.cpp:39-131) and its binding isce3.focus.RangeComp
 * (python/extensions/pybind_isce3/focus/RangeComp.cpp:37-101): the output of
 * a pulse is the discrete convolution of the pulse with the time-reversed
 * complex conjugate of the chirp, cropped by the mode.  The reported
 * fft_size (bp_rangecomp_sizes) is the reference's FFT length, the smallest
 * 2^a 3^b 5^c >= chirp_size + input_size - 1 (RangeComp.cpp:60).  The
 * convolution is formed here with power-of-two transforms: L, the smallest
 * power of two >= max(1024, chirp_size + input_size - 1), gives the same
 * linear convolution.  The compressor runs each pulse as one fused transform
 * of L points (L <= 2^14) or as overlap-save blocks of 2^b points,
 * 10 <= b <= 14, for chirps of at most 2^(b-1) + 1 samples (any pulse
 * length), whichever its cost model of the kernel rates fastest; chirps longer
 * than 2^13 + 1 samples with L above 2^14 run as three four-step passes
 * through a device work buffer of at most 512 MiB that the compressor
 * allocates (L <= 2^19; a run of more pulses than fit goes through it in
 * sub-batches on the stream).  bp_rangecomp_plan reports the choice and its
 * transform length.  All plans give the same convolution to FP32 rounding.
 *
 * Return codes as in isce3_amd_backproject.h; bp_last_error_message()
 * returns the message of the calling thread's last failure.
 */
#ifndef ISCE3_AMD_RANGECOMP_H
#define ISCE3_AMD_RANGECOMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* RangeComp::Mode (RangeComp.h:13-40) */
enum { BP_RC_FULL = 0, BP_RC_VALID = 1, BP_RC_SAME = 2 };

typedef struct bp_rangecomp bp_rangecomp;

/* RangeComp::RangeComp (RangeComp.cpp:39-79): chirp is host complex64
 * (interleaved re, im) of chirp_size samples.  input_size < 1 and
 * max_batch < 1 give BP_ERR_DOMAIN, an unknown mode BP_ERR_RUNTIME. */
int bp_rangecomp_create(int device, const float* chirp, int chirp_size, int input_size,
                        int max_batch, int mode, bp_rangecomp** rc);
int bp_rangecomp_destroy(bp_rangecomp* rc);

/* chirpSize, inputSize, fftSize, maxBatch, mode, outputSize, firstValidSample
 * (RangeComp.h:57-91) into sizes[0..6]. */
int bp_rangecomp_sizes(const bp_rangecomp* rc, int32_t* sizes);

/* The compressor's plan (no counterpart in the reference; for tests and
 * tuning): plan[0] = 0 one fused transform per pulse, 1 overlap-save blocks,
 * 2 four steps; plan[1] = transform length; plan[2] = transforms per pulse
 * (1 for four steps). */
int bp_rangecomp_plan(const bp_rangecomp* rc, int32_t* plan);

/* RangeComp::rangecompress (RangeComp.cpp:93-131): in_dev device complex64
 * [batch x input_size], out_dev device complex64 [batch x output_size].
 * batch > max_batch gives BP_ERR_LENGTH.  Enqueued on `stream` (NULL: the
 * null stream, and the call returns when the work is done). */
int bp_rangecomp_run(bp_rangecomp* rc, void* out_dev, const void* in_dev, int batch,
                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ISCE3_AMD_RANGECOMP_H */
