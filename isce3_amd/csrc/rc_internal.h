// Device-side arguments of the range-compression kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rc {

constexpr int kMinLog2 = 10;      // FFT lengths 2^10 .. 2^14: one workgroup per pulse
constexpr int kMaxFusedLog2 = 14;
constexpr int kMaxLog2 = 19;      // 2^15 .. 2^19 with long chirps: four steps through a work buffer

struct ConvArgs {
    const float2* in;    // [batch x n_in] pulses
    float2* out;         // [batch x n_out] compressed pulses
    const float2* H;     // [L] matched-filter spectrum, scaled by 1/L (four steps: permuted,
                         // bin k1 + 32 k2 at k1 L2 + k2)
    const float2* tw;    // [L] W_L^k = exp(-2 pi i k / L)
    const float2* tw2;   // [L2] W_L2^k (four steps)
    float2* work;        // [batch x L] (four steps)
    int32_t n_in, n_out;
    int32_t offset;      // first output sample in the full convolution
    // Fused kernel blocks (overlap-save): block s of a pulse transforms input
    // samples [s valid - shift, s valid - shift + L) and yields full-convolution
    // samples s valid + i - shift for i in [shift, L).  One block per pulse:
    // shift 0, valid L, nblk 1, blk0 0.
    int32_t shift, valid, nblk, blk0;
};


hipError_t launch_conv(const ConvArgs& a, int log2l, int batch, hipStream_t stream);
hipError_t launch_four_step(const ConvArgs& a, int log2l, int batch, hipStream_t stream);

}  // namespace rc
