// C-ABI of range compression (include/isce3_amd_rangecomp.h): argument rules
// of cxx/isce3/focus/RangeComp.cpp:39-79 and :93-99, the matched filter's
// spectrum formed on the host in FP64 (RangeComp.cpp:22-37), and the launch
// of the FFT-convolution kernels (rc_kernels.hip): one fused transform per
// pulse, or overlap-save blocks of 2^10 .. 2^14 points, whichever a cost
// model of the kernel on MI355X rates fastest, and three four-step passes
// through a work buffer for convolutions past 2^14 samples with chirps longer
// than 2^13 + 1 samples (transforms up to 2^19).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/isce3_amd_backproject.h"
#include "../../include/isce3_amd_rangecomp.h"
#include "rc_internal.h"

// Message of the last failing call of this thread (shared with bp_api.hip).
void bp_set_last_error(const std::string& msg);
int bp_host_exception() noexcept;

namespace {

int fail(int code, const std::string& msg)
{
    bp_set_last_error(msg);
    return code;
}

int output_size(int m, int n, int mode)
{
    switch (mode) {
    case BP_RC_FULL: return m + n - 1;
    case BP_RC_VALID: return std::max(m, n) - std::min(m, n) + 1;
    default: return n;
    }
}

int first_valid(int m, int mode)
{
    switch (mode) {
    case BP_RC_FULL: return m - 1;
    case BP_RC_VALID: return 0;
    default: return m / 2;
    }
}

// In-place radix-2 FFT (sign -1) of a power-of-two length, FP64.
void fft_host(std::vector<std::complex<double>>& x)
{
    const size_t n = x.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * M_PI / (double)len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < len / 2; ++k) {
                const std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
                const std::complex<double> u = x[i + k], v = x[i + k + len / 2] * w;
                x[i + k] = u + v;
                x[i + k + len / 2] = u - v;
            }
        }
    }
}

// How a compressor runs its pulses: one fused transform of 2^tlog2 points per
// pulse (method 0), overlap-save blocks of 2^tlog2 points (method 1, nblk per
// pulse), or the four-step passes (method 2).
struct Plan {
    int method, tlog2, nblk;
};

// Modelled cost of the fused kernel on MI355X, measured over 16384-pulse
// batches: picoseconds per transform point (forward and inverse transform and
// spectral product) for 2^10 .. 2^14-point transforms -- the smaller
// transforms run several workgroups per CU and hide their barriers and LDS
// latency -- and the rate at which the kernel streams pulses in and out when
// the transforms are cheap enough (bytes per picosecond = TB/s).  The two
// overlap imperfectly: a plan costs the larger of its streaming and transform
// times plus kOverlapLoss of the smaller.
constexpr double kPointPs[5] = {3.1, 3.0, 3.4, 4.1, 5.1};
constexpr double kStreamBytesPerPs = 4.1;
constexpr double kOverlapLoss = 0.15;

// Overlap-save blocks of 2^b points covering full-convolution samples
// [offset, offset + n_out) of a chirp of m samples.
int ols_blocks(int b, int m, int offset, int n_out)
{
    const int valid = (1 << b) - (m - 1);
    return (offset + n_out - 1) / valid - offset / valid + 1;
}

// The plan of least modelled time per pulse (streaming the pulse in and its
// output out against the transform work).  Overlap-save
// blocks of 2^b points need a chirp of at most 2^(b-1) + 1 samples; chirps
// longer than 2^13 + 1 samples with convolutions past 2^14 take four steps.
Plan plan_transform(int chirp_size, int input_size, int mode, int log2l)
{
    const int n_out = output_size(chirp_size, input_size, mode);
    const int offset = mode == BP_RC_FULL ? 0 : (mode == BP_RC_VALID ? chirp_size - 1 : chirp_size / 2);
    const double stream_ps = 8.0 * ((double)input_size + n_out) / kStreamBytesPerPs;
    Plan best{2, log2l, 1};
    double best_t = 0., best_pts = 0.;
    auto consider = [&](int method, int b, int nblk) {
        const double pts = (double)nblk * (double)(1 << b);
        const double work_ps = pts * kPointPs[b - rc::kMinLog2];
        const double t = std::max(stream_ps, work_ps) + kOverlapLoss * std::min(stream_ps, work_ps);
        if (best.method == 2 || t < best_t || (t == best_t && pts < best_pts)) {
            best = Plan{method, b, nblk};
            best_t = t;
            best_pts = pts;
        }
    };
    if (log2l <= rc::kMaxFusedLog2) consider(0, log2l, 1);
    for (int b = rc::kMinLog2; b <= std::min(rc::kMaxFusedLog2, log2l - 1); ++b)
        if (chirp_size - 1 <= (1 << (b - 1)))
            consider(1, b, ols_blocks(b, chirp_size, offset, n_out));
    return best;
}

}  // namespace

struct bp_rangecomp {
    int device = 0;
    int chirp_size = 0, input_size = 0, max_batch = 0, mode = 0;
    int fft_size = 0;       // the reference's FFT length (2^a 3^b 5^c), reported only
    int log2l = 0, out_size = 0, offset = 0;
    int method = 0;         // 0 one transform per pulse, 1 overlap-save blocks, 2 four steps
    int nblk = 1;           // transforms per pulse (methods 0 and 1)
    float2* H = nullptr;    // device [L] filter spectrum, L the transform length (four steps: permuted)
    float2* tw = nullptr;   // device [L]
    float2* tw2 = nullptr;  // device [L / 32] (four steps)
    float2* work = nullptr; // device [work_batch x L] (four steps)
    int work_batch = 0;     // pulses per four-step launch
};

namespace {

// Bytes of the four-step work buffer at most (it holds whole padded pulses).
constexpr size_t kWorkBytes = size_t(512) << 20;

// Smallest m = 2^a 3^b 5^c >= n (1 for n <= 1): the FFT length the reference's
// compressor uses (RangeComp.cpp:60, fft/FFTUtil.icc:37-73).  Only reported:
// the transforms here run at power-of-two lengths.
int64_t next_fast_power(int64_t n)
{
    if (n <= 1) return 1;
    int64_t best = INT64_MAX;
    for (int64_t p5 = 1; p5 < 2 * n; p5 *= 5)
        for (int64_t p35 = p5; p35 < 2 * n; p35 *= 3) {
            int64_t m = p35;
            while (m < n) m *= 2;
            best = std::min(best, m);
        }
    return best;
}

}  // namespace

extern "C" {

int bp_rangecomp_create(int device, const float* chirp, int chirp_size, int input_size,
                        int max_batch, int mode, bp_rangecomp** rc)
try {
    if (rc == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    *rc = nullptr;
    if (chirp_size < 1 || chirp == nullptr)
        return fail(BP_ERR_INVALID_ARGUMENT, "chirp must have at least one sample");
    if (input_size < 1) return fail(BP_ERR_DOMAIN, "number of samples must be > 0");
    if (max_batch < 1) return fail(BP_ERR_DOMAIN, "max batch size must be > 0");
    if (mode != BP_RC_FULL && mode != BP_RC_VALID && mode != BP_RC_SAME)
        return fail(BP_ERR_RUNTIME, "unexpected range compression mode");
    const int64_t full = (int64_t)chirp_size + input_size - 1;
    int log2l = rc::kMinLog2;
    while (log2l < 31 && ((int64_t)1 << log2l) < full) ++log2l;
    if (log2l >= 31) return fail(BP_ERR_INVALID_ARGUMENT, "pulse too long");
    const Plan plan = plan_transform(chirp_size, input_size, mode, log2l);
    const int method = plan.method;
    const int tlog2 = plan.tlog2;   // log2 of the transform length
    if (method == 2 && log2l > rc::kMaxLog2)
        return fail(BP_ERR_INVALID_ARGUMENT,
                    "range compression of " + std::to_string(full) + " convolution samples with a " +
                    std::to_string(chirp_size) + "-sample chirp exceeds the supported FFT length " +
                    std::to_string(1 << rc::kMaxLog2));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(BP_ERR_HIP, "no HIP device available (isce3_amd needs an MI355X GPU)");
    if (device < 0 || device >= ndev) return fail(BP_ERR_INVALID_ARGUMENT, "invalid device index");
    if (hipSetDevice(device) != hipSuccess) return fail(BP_ERR_HIP, "hipSetDevice failed");

    const int L = 1 << tlog2;
    // matched filter: time-reversed conjugate of the chirp, zero padded to L,
    // transformed, scaled by 1/L (RangeComp.cpp:22-37, :108-115)
    std::vector<std::complex<double>> h(L, 0.0);
    for (int i = 0; i < chirp_size; ++i) {
        const int j = chirp_size - 1 - i;
        h[i] = std::complex<double>(chirp[2 * j], -chirp[2 * j + 1]);
    }
    fft_host(h);
    const bool four_step = method == 2;
    const int L2 = L / 32;
    std::vector<float2> Hh(L), twh(L), tw2h(four_step ? L2 : 0);
    for (int i = 0; i < L; ++i) {
        // four steps read bin k1 + 32 k2 at k1 L2 + k2
        const int k = four_step ? (i / L2) + 32 * (i % L2) : i;
        Hh[i] = make_float2((float)(h[k].real() / L), (float)(h[k].imag() / L));
        const double ang = -2.0 * M_PI * (double)i / (double)L;
        twh[i] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    for (int i = 0; i < (int)tw2h.size(); ++i) {
        const double ang = -2.0 * M_PI * (double)i / (double)L2;
        tw2h[i] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    auto* r = new bp_rangecomp();
    r->device = device;
    r->chirp_size = chirp_size;
    r->input_size = input_size;
    r->max_batch = max_batch;
    r->mode = mode;
    r->fft_size = (int)next_fast_power(full);
    r->log2l = tlog2;
    r->method = method;
    r->nblk = plan.nblk;
    r->out_size = output_size(chirp_size, input_size, mode);
    r->offset = mode == BP_RC_FULL ? 0 : (mode == BP_RC_VALID ? chirp_size - 1 : chirp_size / 2);
    if (four_step)
        r->work_batch = (int)std::max<size_t>(
            1, std::min<size_t>((size_t)max_batch, kWorkBytes / (sizeof(float2) * (size_t)L)));
    if (hipMalloc(&r->H, sizeof(float2) * L) != hipSuccess ||
        hipMalloc(&r->tw, sizeof(float2) * L) != hipSuccess ||
        (four_step && (hipMalloc(&r->tw2, sizeof(float2) * L2) != hipSuccess ||
                       hipMalloc(&r->work, sizeof(float2) * (size_t)L * r->work_batch) !=
                           hipSuccess))) {
        bp_rangecomp_destroy(r);
        return fail(BP_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    if (hipMemcpy(r->H, Hh.data(), sizeof(float2) * L, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->tw, twh.data(), sizeof(float2) * L, hipMemcpyHostToDevice) != hipSuccess ||
        (four_step && hipMemcpy(r->tw2, tw2h.data(), sizeof(float2) * L2,
                                hipMemcpyHostToDevice) != hipSuccess)) {
        bp_rangecomp_destroy(r);
        return fail(BP_ERR_HIP, "hipMemcpy failed");
    }
    *rc = r;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_rangecomp_destroy(bp_rangecomp* rc)
try {
    if (rc == nullptr) return BP_SUCCESS;
    (void)hipSetDevice(rc->device);
    if (rc->H) (void)hipFree(rc->H);
    if (rc->tw) (void)hipFree(rc->tw);
    if (rc->tw2) (void)hipFree(rc->tw2);
    if (rc->work) (void)hipFree(rc->work);
    delete rc;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_rangecomp_sizes(const bp_rangecomp* rc, int32_t* sizes)
try {
    if (rc == nullptr || sizes == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    sizes[0] = rc->chirp_size;
    sizes[1] = rc->input_size;
    sizes[2] = rc->fft_size;
    sizes[3] = rc->max_batch;
    sizes[4] = rc->mode;
    sizes[5] = rc->out_size;
    sizes[6] = first_valid(rc->chirp_size, rc->mode);
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_rangecomp_plan(const bp_rangecomp* rc, int32_t* plan)
try {
    if (rc == nullptr || plan == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    plan[0] = rc->method;
    plan[1] = 1 << rc->log2l;
    plan[2] = rc->method == 2 ? 1 : rc->nblk;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_rangecomp_run(bp_rangecomp* rc, void* out_dev, const void* in_dev, int batch,
                     void* stream)
try {
    if (rc == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (batch > rc->max_batch) return fail(BP_ERR_LENGTH, "batch size exceeds max batch");
    if (batch < 0) return fail(BP_ERR_INVALID_ARGUMENT, "negative batch size");
    if (batch == 0) return BP_SUCCESS;
    if (out_dev == nullptr || in_dev == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null buffer");
    if (hipSetDevice(rc->device) != hipSuccess) return fail(BP_ERR_HIP, "hipSetDevice failed");
    rc::ConvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = static_cast<const float2*>(in_dev);
    a.out = static_cast<float2*>(out_dev);
    a.H = rc->H;
    a.tw = rc->tw;
    a.n_in = rc->input_size;
    a.n_out = rc->out_size;
    a.offset = rc->offset;
    a.tw2 = rc->tw2;
    a.work = rc->work;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    if (rc->method == 0) {
        a.shift = 0;
        a.valid = 1 << rc->log2l;
        a.nblk = 1;
        a.blk0 = 0;
        e = rc::launch_conv(a, rc->log2l, batch, s);
    } else if (rc->method == 1) {
        // blocks covering full-convolution samples [offset, offset + n_out)
        a.shift = rc->chirp_size - 1;
        a.valid = (1 << rc->log2l) - a.shift;
        a.blk0 = a.offset / a.valid;
        a.nblk = (a.offset + a.n_out - 1) / a.valid - a.blk0 + 1;
        e = rc::launch_conv(a, rc->log2l, batch, s);
    } else {
        // sub-batches of the work buffer, launched in order on one stream
        for (int b0 = 0; b0 < batch && e == hipSuccess; b0 += rc->work_batch) {
            rc::ConvArgs sub = a;
            sub.in = a.in + (int64_t)b0 * a.n_in;
            sub.out = a.out + (int64_t)b0 * a.n_out;
            e = rc::launch_four_step(sub, rc->log2l, std::min(rc->work_batch, batch - b0), s);
        }
    }
    if (e == hipSuccess && s == nullptr) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(BP_ERR_HIP, std::string("range compression launch: ") + hipGetErrorString(e));
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

}  // extern "C"
