// Finishing focused RSLC blocks on gfx950 (include/isce3_amd_rslc.h): the
// per-block work of BackgroundWriter.write (nisar/workflows/focus.py:883-914)
// in one pass over HBM -- range-correction product, accumulation of the block
// already written, statistics of isce3::math::StatsRealImag<float>
// (cxx/isce3/math/Stats.cpp:33-165) and the storage encoding -- plus a
// one-workgroup merge of the per-workgroup statistics.
//
// Each thread takes chunks of kPerThread samples, in pairs strided by the
// workgroup size (every load and store a coalesced 16-byte-per-lane wave
// access when the buffers allow, 8 bytes otherwise), keeps them
// in registers, and summarises its valid samples exactly in FP64 by two passes
// (mean, then the sum of squared deviations).  Summaries are combined with
// Chan's pairwise update -- over a thread's chunks, across the wave by lane
// shuffles, across the workgroup through LDS, across the (at most 8192)
// workgroups in a second kernel that merges the partials in a fixed order, so
// results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "../../include/isce3_amd_backproject.h"
#include "../../include/isce3_amd_rslc.h"

void bp_set_last_error(const std::string& msg);
int bp_host_exception() noexcept;

namespace rslc {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr int64_t kPerBlock = (int64_t)kThreads * kPerThread;
constexpr int kMaxGrid = 8192;   // workgroups (and partial summaries) at most

// Summary of the valid samples of one component: count, mean, sum of squared
// deviations, extremes.
struct Moments {
    double mean, m2;
    float lo, hi;
};

struct Partial {
    int64_t n;
    Moments re, im;
};

struct FinishArgs {
    float2* z;
    const float2* range_cor;   // [cols] or null
    const float2* prior;       // [n] or null
    uint32_t* encoded;         // [n] pairs of half (complex32) or null
    Partial* partials;         // [gridDim.x]
    int64_t n, cols, chunks;   // chunks of kPerBlock samples
    int32_t data_type;
    uint32_t mantissa_mask;
    int32_t store_z;
};

__device__ inline void merge(int64_t& na, Moments& a, int64_t nb, const Moments& b, int64_t n)
{
    // Chan et al.: combine (na, a) and (nb, b) into n = na + nb samples
    const double delta = b.mean - a.mean;
    const double fb = (double)nb / (double)n;
    a.mean += delta * fb;
    a.m2 += b.m2 + delta * delta * (double)na * fb;
    a.lo = fminf(a.lo, b.lo);
    a.hi = fmaxf(a.hi, b.hi);
}

__device__ inline void merge(Partial& a, const Partial& b)
{
    if (b.n == 0) return;
    if (a.n == 0) {
        a = b;
        return;
    }
    const int64_t n = a.n + b.n;
    merge(a.n, a.re, b.n, b.re, n);
    merge(a.n, a.im, b.n, b.im, n);
    a.n = n;
}

__device__ inline Partial shfl_xor(const Partial& p, int mask)
{
    Partial q;
    q.n = __shfl_xor(p.n, mask);
    q.re.mean = __shfl_xor(p.re.mean, mask);
    q.re.m2 = __shfl_xor(p.re.m2, mask);
    q.re.lo = __shfl_xor(p.re.lo, mask);
    q.re.hi = __shfl_xor(p.re.hi, mask);
    q.im.mean = __shfl_xor(p.im.mean, mask);
    q.im.m2 = __shfl_xor(p.im.m2, mask);
    q.im.lo = __shfl_xor(p.im.lo, mask);
    q.im.hi = __shfl_xor(p.im.hi, mask);
    return q;
}

__device__ inline Partial empty_partial()
{
    Partial p;
    p.n = 0;
    p.re = {0.0, 0.0, INFINITY, -INFINITY};
    p.im = p.re;
    return p;
}

// complex64 product rounded term by term, (ac - bd) + i (ad + bc), no FMA
__device__ inline float2 cmul_rounded(float2 a, float2 b)
{
#pragma clang fp contract(off)
    const float rr = a.x * b.x, ii = a.y * b.y, ri = a.x * b.y, ir = a.y * b.x;
    return make_float2(rr - ii, ri + ir);
}

__device__ inline uint32_t to_half_pair(float2 v)
{
    const _Float16 r = (_Float16)v.x, i = (_Float16)v.y;
    return (uint32_t)__builtin_bit_cast(uint16_t, r) |
           ((uint32_t)__builtin_bit_cast(uint16_t, i) << 16);
}

// Reduce the workgroup's summaries into one (wave shuffles, then LDS across
// the four waves, in a fixed order).
__device__ inline Partial block_reduce(Partial p)
{
    __shared__ Partial waves[kThreads / 64];
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) merge(p, shfl_xor(p, m));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) waves[wave] = p;
    __syncthreads();
    p = waves[0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) merge(p, waves[w]);
    return p;
}

// Native 4 x float vector for the streaming (non-temporal) accesses: the
// block is read once and written once, so its lines need not stay in L2.
typedef float v4f __attribute__((ext_vector_type(4)));

// One chunk of kPerBlock samples: scale, accumulate, encode and store; the
// thread's summary of its valid samples in *p (two passes over registers).
// kPair: each access moves two adjacent samples (16-byte loads and stores of
// z and prior, 8-byte stores of the encoding, all non-temporal); otherwise
// one sample.
template<bool kPair>
__device__ inline void finish_chunk(const FinishArgs& a, int64_t chunk, Partial* p)
{
    constexpr int P = kPair ? 2 : 1;
    constexpr int J = kPerThread / P;
    const int64_t e0 = chunk * kPerBlock + P * (int64_t)threadIdx.x;
    int64_t col = e0 % a.cols;
    const int64_t step = (int64_t)(P * kThreads) % a.cols;
    float2 v[kPerThread];
    bool ok[kPerThread];
    int64_t cnt = 0;
    double sre = 0.0, sim = 0.0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int64_t e = e0 + (int64_t)j * P * kThreads;
        float2 x[P], q[P];
        if (kPair && e + 1 < a.n) {
            const v4f z4 = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(a.z + e));
            x[0] = make_float2(z4.x, z4.y);
            x[P - 1] = make_float2(z4.z, z4.w);
            if (a.prior) {
                const v4f p4 =
                    __builtin_nontemporal_load(reinterpret_cast<const v4f*>(a.prior + e));
                q[0] = make_float2(p4.x, p4.y);
                q[P - 1] = make_float2(p4.z, p4.w);
            }
        } else {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                x[u] = e + u < a.n ? a.z[e + u] : make_float2(0.f, 0.f);
                if (a.prior) q[u] = e + u < a.n ? a.prior[e + u] : make_float2(0.f, 0.f);
            }
        }
        int64_t cu = col;
#pragma unroll
        for (int u = 0; u < P; ++u) {
            const int k = P * j + u;
            ok[k] = false;
            v[k] = make_float2(0.f, 0.f);
            if (e + u < a.n) {
                float2 y = x[u];
                if (a.range_cor) y = cmul_rounded(y, a.range_cor[cu]);
                if (a.prior) y = make_float2(y.x + q[u].x, y.y + q[u].y);
                v[k] = y;
                ok[k] = !(isnan(y.x) || isnan(y.y));
                if (ok[k]) {
                    ++cnt;
                    sre += (double)y.x;
                    sim += (double)y.y;
                }
            }
            if (++cu == a.cols) cu = 0;
        }
        col += step;
        if (col >= a.cols) col -= a.cols;
    }
    *p = empty_partial();
    p->n = cnt;
    if (cnt > 0) {
        p->re.mean = sre / (double)cnt;
        p->im.mean = sim / (double)cnt;
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            if (!ok[k]) continue;
            const double dr = (double)v[k].x - p->re.mean, di = (double)v[k].y - p->im.mean;
            p->re.m2 += dr * dr;
            p->im.m2 += di * di;
            p->re.lo = fminf(p->re.lo, v[k].x);
            p->re.hi = fmaxf(p->re.hi, v[k].x);
            p->im.lo = fminf(p->im.lo, v[k].y);
            p->im.hi = fmaxf(p->im.hi, v[k].y);
        }
    }
    // stores: statistics are of the values before truncation
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int64_t e = e0 + (int64_t)j * P * kThreads;
        float2 x[P];
        uint32_t h[P];
#pragma unroll
        for (int u = 0; u < P; ++u) {
            x[u] = v[P * j + u];
            if (a.data_type == BP_RSLC_COMPLEX32) {
                h[u] = to_half_pair(x[u]);
            } else if (a.data_type == BP_RSLC_COMPLEX64_ZERO_MANTISSA) {
                x[u].x = __uint_as_float(__float_as_uint(x[u].x) & a.mantissa_mask);
                x[u].y = __uint_as_float(__float_as_uint(x[u].y) & a.mantissa_mask);
            }
        }
        if (kPair && e + 1 < a.n) {
            if (a.data_type == BP_RSLC_COMPLEX32)
                __builtin_nontemporal_store(
                    (unsigned long long)h[0] | ((unsigned long long)h[P - 1] << 32),
                    reinterpret_cast<unsigned long long*>(a.encoded + e));
            if (a.store_z)
                __builtin_nontemporal_store(v4f{x[0].x, x[0].y, x[P - 1].x, x[P - 1].y},
                                            reinterpret_cast<v4f*>(a.z + e));
        } else {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                if (e + u >= a.n) continue;
                if (a.data_type == BP_RSLC_COMPLEX32) a.encoded[e + u] = h[u];
                if (a.store_z) a.z[e + u] = x[u];
            }
        }
    }
}

// Workgroup g takes chunks g, g + gridDim.x, ... (a fixed assignment, so the
// merge order and the result do not depend on scheduling) and leaves one
// partial summary.
template<bool kPair>
__global__ __launch_bounds__(kThreads) void finish_kernel(FinishArgs a)
{
    Partial acc = empty_partial();
    for (int64_t chunk = blockIdx.x; chunk < a.chunks; chunk += gridDim.x) {
        Partial p;
        finish_chunk<kPair>(a, chunk, &p);
        merge(acc, p);
    }
    acc = block_reduce(acc);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = acc;
}

__device__ inline void to_stats(const Moments& m, int64_t n, bp_stats_f32& s)
{
    const float qnan = __builtin_nanf("");
    s.min = n > 0 ? m.lo : qnan;
    s.max = n > 0 ? m.hi : qnan;
    s.mean = n > 0 ? (float)m.mean : 0.f;
    s.pad_ = 0;
    s.n_valid = n;
    s.real_valued_mean = n > 0 ? m.mean : 0.0;
    s.square_diff_sum = n > 0 ? m.m2 : 0.0;
}

__global__ __launch_bounds__(kThreads) void merge_kernel(const Partial* partials, int count,
                                                         bp_stats_real_imag_f32* out)
{
    Partial p = empty_partial();
    for (int i = threadIdx.x; i < count; i += kThreads) merge(p, partials[i]);
    p = block_reduce(p);
    if (threadIdx.x == 0) {
        to_stats(p.re, p.n, out->real);
        to_stats(p.im, p.n, out->imag);
        out->n_valid = p.n;
    }
}

int fail(int code, const std::string& msg)
{
    bp_set_last_error(msg);
    return code;
}

}  // namespace
}  // namespace rslc

struct bp_rslc_finisher {
    int device = 0;
    int64_t max_pixels = 0;
    int max_partials = 0;
    rslc::Partial* partials = nullptr;       // device
    bp_stats_real_imag_f32* d_stats = nullptr;
    bp_stats_real_imag_f32* h_stats = nullptr; // pinned
};

extern "C" {

int bp_rslc_finisher_create(int device, int64_t max_pixels, bp_rslc_finisher** f)
try {
    using rslc::fail;
    if (f == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    *f = nullptr;
    if (max_pixels < 1) return fail(BP_ERR_INVALID_ARGUMENT, "max_pixels must be > 0");
    const int64_t blocks = std::min<int64_t>((max_pixels + rslc::kPerBlock - 1) / rslc::kPerBlock,
                                             rslc::kMaxGrid);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(BP_ERR_HIP, "no HIP device available (isce3_amd needs an MI355X GPU)");
    if (device < 0 || device >= ndev) return fail(BP_ERR_INVALID_ARGUMENT, "invalid device index");
    if (hipSetDevice(device) != hipSuccess) return fail(BP_ERR_HIP, "hipSetDevice failed");
    auto* r = new bp_rslc_finisher();
    r->device = device;
    r->max_pixels = max_pixels;
    r->max_partials = (int)blocks;
    if (hipMalloc(&r->partials, sizeof(rslc::Partial) * blocks) != hipSuccess ||
        hipMalloc(&r->d_stats, sizeof(bp_stats_real_imag_f32)) != hipSuccess ||
        hipHostMalloc(&r->h_stats, sizeof(bp_stats_real_imag_f32), hipHostMallocDefault) !=
            hipSuccess) {
        bp_rslc_finisher_destroy(r);
        return fail(BP_ERR_OUT_OF_MEMORY, "device or pinned allocation failed");
    }
    *f = r;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_rslc_finisher_destroy(bp_rslc_finisher* f)
try {
    if (f == nullptr) return BP_SUCCESS;
    (void)hipSetDevice(f->device);
    if (f->partials) (void)hipFree(f->partials);
    if (f->d_stats) (void)hipFree(f->d_stats);
    if (f->h_stats) (void)hipHostFree(f->h_stats);
    delete f;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_rslc_finish_block(bp_rslc_finisher* f, void* z, int64_t rows, int64_t cols,
                         const void* range_cor, const void* prior, int data_type,
                         int mantissa_nbits, void* encoded, bp_stats_real_imag_f32* stats,
                         void* stream)
try {
    using rslc::fail;
    if (f == nullptr || stats == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (rows < 0 || cols < 0) return fail(BP_ERR_INVALID_ARGUMENT, "negative block shape");
    if (data_type != BP_RSLC_COMPLEX64 && data_type != BP_RSLC_COMPLEX32 &&
        data_type != BP_RSLC_COMPLEX64_ZERO_MANTISSA)
        return fail(BP_ERR_INVALID_ARGUMENT, "Requested invalid data_type");
    uint32_t mask = 0xffffffffu;
    if (data_type == BP_RSLC_COMPLEX64_ZERO_MANTISSA) {
        if (!(mantissa_nbits > 0 && mantissa_nbits <= 23))
            return fail(BP_ERR_INVALID_ARGUMENT, "Require 0 < significant_bits=" +
                                                     std::to_string(mantissa_nbits) + " <= 23");
        mask = 0xffffffffu << (23 - mantissa_nbits);
    }
    if (cols > 0 && rows > INT64_MAX / cols) return fail(BP_ERR_LENGTH, "block too large");
    const int64_t n = rows * cols;
    if (n > f->max_pixels)
        return fail(BP_ERR_LENGTH, "block of " + std::to_string(n) + " samples exceeds the "
                                   "finisher's maximum of " + std::to_string(f->max_pixels));
    if (n > 0 && z == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null block");
    if (n > 0 && data_type == BP_RSLC_COMPLEX32 && encoded == nullptr)
        return fail(BP_ERR_INVALID_ARGUMENT, "complex32 encoding needs an output buffer");
    if (n == 0) {
        bp_stats_real_imag_f32 s;
        std::memset(&s, 0, sizeof(s));
        s.real.min = s.real.max = s.imag.min = s.imag.max =
            std::numeric_limits<float>::quiet_NaN();
        *stats = s;
        return BP_SUCCESS;
    }
    if (hipSetDevice(f->device) != hipSuccess) return fail(BP_ERR_HIP, "hipSetDevice failed");
    rslc::FinishArgs a;
    std::memset(&a, 0, sizeof(a));
    a.z = static_cast<float2*>(z);
    a.range_cor = static_cast<const float2*>(range_cor);
    a.prior = static_cast<const float2*>(prior);
    a.encoded = static_cast<uint32_t*>(encoded);
    a.partials = f->partials;
    a.n = n;
    a.cols = cols;
    a.data_type = data_type;
    a.mantissa_mask = mask;
    a.store_z = range_cor != nullptr || prior != nullptr ||
                (data_type == BP_RSLC_COMPLEX64_ZERO_MANTISSA && mask != 0xffffffffu);
    a.chunks = (n + rslc::kPerBlock - 1) / rslc::kPerBlock;
    const int blocks = (int)std::min<int64_t>(a.chunks, rslc::kMaxGrid);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // two samples per access when every buffer is aligned for it
    const bool pair = reinterpret_cast<uintptr_t>(z) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(prior) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(encoded) % 8 == 0;
    if (pair)
        hipLaunchKernelGGL(rslc::finish_kernel<true>, dim3(blocks), dim3(rslc::kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(rslc::finish_kernel<false>, dim3(blocks), dim3(rslc::kThreads), 0, s, a);
    hipLaunchKernelGGL(rslc::merge_kernel, dim3(1), dim3(rslc::kThreads), 0, s,
                       (const rslc::Partial*)f->partials, blocks, f->d_stats);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(f->h_stats, f->d_stats, sizeof(bp_stats_real_imag_f32),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return fail(BP_ERR_HIP, std::string("RSLC block finishing: ") + hipGetErrorString(e));
    *stats = *f->h_stats;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

}  // extern "C"
