// Device-side arguments of the range-compression kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rc {

constexpr int kMinLog2 = 10;   // FFT lengths 2^10 .. 2^14 (one workgroup per pulse)
constexpr int kMaxLog2 = 14;

struct ConvArgs {
    const float2* in;    // [batch x n_in] pulses
    float2* out;         // [batch x n_out] compressed pulses
    const float2* H;     // [L] matched-filter spectrum, scaled by 1/L
    const float2* tw;    // [L] W_L^k = exp(-2 pi i k / L)
    int32_t n_in, n_out;
    int32_t offset;      // first output sample in the full convolution
    int32_t pad_;
};

hipError_t launch_conv(const ConvArgs& a, int log2l, int batch, hipStream_t stream);

}  // namespace rc
