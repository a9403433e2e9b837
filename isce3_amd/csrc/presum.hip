// Presumming / BLU resampling of non-uniformly sampled pulses on gfx950
// (include/isce3_amd_presum.h).
//
// Host side: the BLU weights of isce3::focus::getPresumWeights
// (cxx/isce3/focus/Presum.icc:9-60) -- the samples within half the
// autocorrelation width of the output time, their autocorrelation matrix and
// right-hand side, solved by LDL^T with diagonal pivoting -- and, for the
// resampler, the distinct valid-sample patterns of each output pulse's window
// (taken from the sub-swath boundaries, which is where the pattern of a range
// bin can change) with one weight vector per pattern.
//
// Device side: one thread per range bin and 8 output pulses rebuilds the
// bin's pattern from the sub-swath table, picks its weights, evaluates the
// Doppler LUT at (tout, r) and accumulates
//   sum_k float(w_k) * exp(-2 pi i (t_k - tout) fd) * raw[k][r]
// in complex64 in pulse order, as nisar/workflows/focus.py:1011-1061 and
// pybind focus/Presum.cpp:26-66 do (the deramp from the phase in turns,
// reduced in FP64, by the hardware sine and cosine, where the reference
// rounds the phase in radians to float and calls sinf/cosf).  Consecutive
// output pulses share most of their input pulses: a thread loads its bin of
// every input pulse its output pulses reach (up to 12) at once, into its own
// column of an LDS array, and reads the windows from there (longer spans
// read them from global memory).
//
// The output pulses go in launches of 512, 1024, then 2048: host threads (a pool
// that lives across calls) plan the next launches' patterns and weights
// while the GPU resamples the earlier ones, and the launches' tables go up
// on a copy stream beside them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/isce3_amd_backproject.h"
#include "../../include/isce3_amd_presum.h"
#include "bp_math.h"

void bp_set_last_error(const std::string& msg);
int bp_host_exception() noexcept;
namespace bp {
int usable_cpus();  // bp_api.hip
}

namespace presum {
namespace {

int fail(int code, const std::string& msg)
{
    bp_set_last_error(msg);
    return code;
}

// core/Kernels.icc:16-23 (Bartlett), :95-107 (azimuth autocorrelation)
double acorr_eval(const bp_acorr& k, double t)
{
    const double hw = std::fabs(k.width / 2.0);
    if (k.kind == BP_ACORR_BARTLETT) {
        const double u = std::fabs(t / hw);
        return u > 1.0 ? 0.0 : 1.0 - u;
    }
    const double x = std::fabs(t * 2 / hw);
    if (x > 2.0) return 0.0;
    if (x < 1.0) return x * x * (0.75 * x - 1.5) + 1.0;
    return x * (x * (-0.25 * x + 1.5) - 3.0) + 2.0;
}

// Host buffers of the weight solves, reused from solve to solve.
struct SolveScratch {
    std::vector<double> g, r, d, tmp, y;
    std::vector<int> perm;
};

// Solve G w = r (G = s.g, n x n, lower triangle used; r = s.r) for symmetric
// positive semi-definite G by LDL^T with diagonal pivoting: at step k the
// largest remaining diagonal entry is swapped to k, column k of L is formed
// from the earlier columns (left-looking), and pivots not above the smallest
// normal number are treated as zero (their solution component is zero).
void ldlt_solve(SolveScratch& s, int n, std::vector<double>& w)
{
    std::vector<double>& g = s.g;
    auto G = [&](int i, int j) -> double& { return g[(size_t)i * n + j]; };
    s.perm.resize(n);
    for (int i = 0; i < n; ++i) s.perm[i] = i;
    // make the matrix full-symmetric for row/column swaps
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) G(i, j) = G(j, i);
    s.d.assign(n, 0.0);
    s.tmp.resize(n);
    std::vector<double>& d = s.d;
    std::vector<double>& tmp = s.tmp;
    const double cutoff = std::numeric_limits<double>::min();
    for (int k = 0; k < n; ++k) {
        // pivot: largest |diagonal| among k..n-1 (first on ties)
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(G(i, i)) > std::fabs(G(piv, piv))) piv = i;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(G(k, j), G(piv, j));
            for (int i = 0; i < n; ++i) std::swap(G(i, k), G(i, piv));
            std::swap(s.perm[k], s.perm[piv]);
        }
        // D_k = G_kk - sum_j L_kj^2 D_j ; L_ik = (G_ik - sum_j L_ij L_kj D_j) / D_k
        for (int j = 0; j < k; ++j) tmp[j] = G(k, j) * d[j];
        double dk = G(k, k);
        for (int j = 0; j < k; ++j) dk -= G(k, j) * tmp[j];
        d[k] = dk;
        for (int i = k + 1; i < n; ++i) {
            double v = G(i, k);
            for (int j = 0; j < k; ++j) v -= G(i, j) * tmp[j];
            G(i, k) = std::fabs(dk) > cutoff ? v / dk : v;
        }
    }
    // P b, L y = P b, D z = y, L^T x = z, w = P^T x
    std::vector<double>& y = s.y;
    y.resize(n);
    for (int i = 0; i < n; ++i) y[i] = s.r[s.perm[i]];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) y[i] -= G(i, j) * y[j];
    for (int i = 0; i < n; ++i) y[i] = std::fabs(d[i]) > cutoff ? y[i] / d[i] : 0.0;
    for (int i = n - 1; i >= 0; --i)
        for (int j = i + 1; j < n; ++j) y[i] -= G(j, i) * y[j];
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) w[s.perm[i]] = y[i];
}

// getPresumWeights' window: the samples within half the kernel width of tout,
// [*offset, *offset + return value).
int64_t presum_window(const bp_acorr& acorr, const double* t, int64_t n, double tout,
                      int64_t* offset)
{
    const double hw = 0.5 * acorr.width;
    const double* first = std::lower_bound(t, t + n, tout - hw);
    const double* last = std::upper_bound(first, t + n, tout + hw);
    *offset = first - t;
    return last - first;
}

// getPresumWeights: window [offset, offset + nw) and its weights
void presum_weights(const bp_acorr& acorr, const double* t, int64_t n, double tout,
                    int64_t* offset, std::vector<double>& w, SolveScratch& s)
{
    const int m = (int)presum_window(acorr, t, n, tout, offset);
    const double* first = t + *offset;
    s.g.assign((size_t)m * m, 0.0);
    s.r.resize(m);
    for (int i = 0; i < m; ++i) {
        const double xi = first[i];
        for (int j = i; j < m; ++j) s.g[(size_t)j * m + i] = acorr_eval(acorr, first[j] - xi);
        s.r[i] = acorr_eval(acorr, tout - xi);
    }
    ldlt_solve(s, m, w);
}

// Per output pulse: window, distinct patterns (bit k: range bin valid in
// window pulse k) and one weight vector per pattern.
// LUT2d<double>::eval (core/LUT2d.cpp:127-159, clamped bilinear) split into
// the range part, fixed per thread, and the azimuth part, fixed per output
// pulse (formed on the host).  floor and ceil of one coordinate differ by 0
// or exactly 1, so the divisions by (x2 - x1), (y2 - y1) and their product
// that the bilinear formula performs are divisions by 1.0 and are left out:
// the values are the same bits.
struct LutAxis {
    double v;       // clamped fractional index
    int64_t i1, i2; // floor, ceil
};

__host__ __device__ inline LutAxis lut_axis(double coord, double start, double step, int64_t n)
{
    double v = (coord - start) / step;
    v = fmin(fmax(v, 0.0), n - 1.0);
    return {v, (int64_t)floor(v), (int64_t)ceil(v)};
}

struct PulseDesc {
    double tout;
    int64_t offset;
    int32_t nw, pbase, npat, wbase;  // weights of pattern q: pat_w[wbase + q nw + k]
    LutAxis ly;                      // Doppler LUT azimuth axis at tout
    // device addresses, formed on the host when the launch's tables are placed
    const float2* raw_row;           // input pulse `offset`, bin 0
    float2* out_row;                 // this output pulse, bin 0
    const double* t_row;             // the window's pulse times
    const int64_t* ids;              // this pulse's npat pattern ids
    const float* w;                  // this pulse's weights, npat x nw
};

struct ResampleArgs {
    const float2* raw;
    float2* out;
    const double* t;           // [nt]
    const int32_t* swaths;     // [ns][nt][2]
    const PulseDesc* desc;     // [rows]: output pulses row0 .. row0 + rows - 1
    const int64_t* pat_id;     // the rows' patterns
    const float* pat_w;        // per pulse npat x nw
    bp::Lut2d doppler;
    double r0, dr;
    int64_t nt, ns, row0, rows, width;
    int32_t pulses_per_thread, pad_;
};

constexpr int kMaxWindow = 63;
constexpr int kThreads = 256;
constexpr int kPulsesPerThread = 8;    // output pulses per thread
constexpr int64_t kFirstChunk = 512;   // output pulses of the first launch,
constexpr int64_t kChunk = 2048;       // at most of the others (multiples of kPulsesPerThread)
constexpr int64_t kItem = 32;          // output pulses a host thread plans at a time
constexpr int kTapGroup = 4;           // window taps loaded together (longer windows)
constexpr int kFixedTaps = 8;          // windows up to this many taps: unrolled per count
constexpr int kStageSpan = 12;         // input pulses a thread stages in LDS
// readable bytes after each uploaded table: the sub-swath bounds of
// kStageSpan pulses from the last one
constexpr size_t kPartSlack = 2 * kStageSpan * sizeof(int32_t);

// complex64 product and sum rounded term by term (no fused multiply-add)
__device__ inline float2 cmul(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline float2 cadd(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return make_float2(a.x + b.x, a.y + b.y);
}

// Valid-sample bit of range bin j in input pulse p (any sub-swath).
// Native 2 x float vector for the non-temporal stores of the output samples
// (written once, not read again by the kernel).
typedef float v2f __attribute__((ext_vector_type(2)));

// A wave-uniform read-only pointer as a constant-address-space pointer: its
// loads go through the scalar data cache (one load per wavefront instead of
// one vector memory instruction of 64 addresses).  Only for data no kernel
// writes while it runs, read at wave-uniform indices.
// (The host compilation pass of the kernels keeps the generic pointer.)
template<class T>
__device__ inline auto uniform_ptr(const T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (const __attribute__((address_space(4))) T*)(p);
#else
    return p;
#endif
}

// Valid-sample bit of range bin j in input pulse p (any sub-swath); p is
// wave-uniform.
__device__ inline bool bin_valid(const ResampleArgs& a, int64_t j, int64_t p)
{
    bool valid = false;
    for (int64_t s = 0; s < a.ns; ++s) {
        const auto* sw = uniform_ptr(a.swaths) + (s * a.nt + p) * 2;
        valid |= (j >= sw[0]) & (j < sw[1]);
    }
    return valid;
}

// The four samples around (x, y) ...
struct LutCorners {
    double q11, q12, q21, q22;
};

__device__ inline LutCorners lut_corners(const bp::Lut2d& l, const LutAxis& x, const LutAxis& y)
{
    const double* z = l.data;
    return {z[y.i1 * l.width + x.i1], z[y.i2 * l.width + x.i1], z[y.i1 * l.width + x.i2],
            z[y.i2 * l.width + x.i2]};
}

// ... and the reference's bilinear combination of them
__device__ inline double lut_combine(const LutCorners& q, const LutAxis& x, const LutAxis& y)
{
#pragma clang fp contract(off)
    const double x1 = (double)x.i1, x2 = (double)x.i2, y1 = (double)y.i1, y2 = (double)y.i2;
    if (y.i1 == y.i2 && x.i1 == x.i2) return q.q11;
    if (y.i1 == y.i2) return ((x2 - x.v) * q.q11) + ((x.v - x1) * q.q21);
    if (x.i1 == x.i2) return ((y2 - y.v) * q.q11) + ((y.v - y1) * q.q12);
    return (q.q11 * ((x2 - x.v) * (y2 - y.v))) + (q.q21 * ((x.v - x1) * (y2 - y.v))) +
           (q.q12 * ((x2 - x.v) * (y.v - y1))) + (q.q22 * ((x.v - x1) * (y.v - y1)));
}

// One output sample from a window of N taps (N fixed at compile time): the
// taps' loads first, then the products, accumulated in tap order.
template<int N, class WU, class TU>
__device__ inline float2 window_sum(const float2* row, int64_t stride, uint32_t jj, const float* w,
                                    WU w0, bool uniform_pat, TU tp, double tout, double fd)
{
    float2 x[N];
    float wv[N];
    double tk[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        x[k] = row[k * stride + jj];
        wv[k] = uniform_pat ? w0[k] : w[k];
        tk[k] = tp[k];
    }
    float2 sum = make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        // deramp exp(-2 pi i (t_k - tout) fd): the phase in turns, reduced to
        // [0, 1) in FP64, then the hardware sine and cosine of 2 pi times it
        const double trel = tk[k] - tout;
        const float turns = (float)__builtin_amdgcn_fract(-trel * fd);
        const float sn = __builtin_amdgcn_sinf(turns), cs = __builtin_amdgcn_cosf(turns);
        sum = cadd(sum, cmul(make_float2(wv[k] * cs, wv[k] * sn), x[k]));
    }
    return sum;
}

__global__ __launch_bounds__(kThreads) void resample_kernel(ResampleArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= a.width) return;
    const uint32_t jj = (uint32_t)j;   // widths stay below 2^31
    const double r = a.r0 + a.dr * (double)j;
    // output pulses [i0, i1), i relative to row0
    const int64_t i0 = (int64_t)blockIdx.y * a.pulses_per_thread;
    const int64_t i1 = min(i0 + (int64_t)a.pulses_per_thread, a.rows);
    const bp::Lut2d& lut = a.doppler;
    LutAxis lx{};
    if (lut.have_data) lx = lut_axis(r, lut.xstart, lut.dx, lut.width);
    LutCorners q{};
    int64_t qy1 = -1, qy2 = -1;
    // validity of this bin in every input pulse the chunk's windows reach,
    // one bit per pulse from the earliest window start
    const auto* desc = uniform_ptr(a.desc);
    int64_t pfirst = desc[i0].offset, pend = pfirst;
    for (int64_t i = i0; i < i1; ++i) {
        pfirst = min(pfirst, desc[i].offset);
        pend = max(pend, desc[i].offset + (int64_t)desc[i].nw);
    }
    const int64_t span = pend - pfirst;
    // this thread's bin of every input pulse of the span, staged in LDS (its
    // own column: no other thread reads it, no barrier) with all loads in
    // flight together, so that the windows read it at LDS latency
    __shared__ float2 s_raw[kStageSpan * kThreads];
    const bool staged = span <= kStageSpan;
    if (staged) {
        const float2* col = a.raw + pfirst * a.width + j;
        float2 v[kStageSpan];
#pragma unroll
        for (int p = 0; p < kStageSpan; ++p)
            if (p < span) v[p] = col[(int64_t)p * a.width];
#pragma unroll
        for (int p = 0; p < kStageSpan; ++p)
            if (p < span) s_raw[p * kThreads + threadIdx.x] = v[p];
    }
    uint64_t bits = 0;
    if (staged) {
        // each sub-swath's bounds for the span in two wide scalar loads (the
        // table is padded so that kStageSpan pulses from any pulse are
        // readable; bounds past the span are ignored)
        const int32_t j32 = (int32_t)jj;
        uint32_t b = 0;
        for (int64_t sub = 0; sub < a.ns; ++sub) {
            const auto* sw = uniform_ptr(a.swaths) + (sub * a.nt + pfirst) * 2;
            int32_t lim[2 * kStageSpan];
#pragma unroll
            for (int k = 0; k < 2 * kStageSpan; ++k) lim[k] = sw[k];
#pragma unroll
            for (int p = 0; p < kStageSpan; ++p) {
                const bool in = (j32 >= lim[2 * p]) & (j32 < lim[2 * p + 1]);
                b |= in ? 1u << p : 0u;
            }
        }
        bits = b & ((1u << span) - 1u);
    } else if (span <= 64) {
        for (int64_t p = 0; p < span; ++p)
            if (bin_valid(a, j, pfirst + p)) bits |= (uint64_t)1 << p;
    }
    for (int64_t i = i0; i < i1; ++i) {
        const PulseDesc d = desc[i];
        // pattern of this range bin over the window
        int64_t id = 0;
        if (span <= 64) {
            const uint64_t m = d.nw >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << d.nw) - 1);
            id = (int64_t)((bits >> (d.offset - pfirst)) & m);
        } else {
            for (int k = 0; k < d.nw; ++k)
                if (bin_valid(a, j, d.offset + k)) id |= (int64_t)1 << k;
        }
        // the pulse's patterns are distinct: stop at the match (-1: none)
        int pat = -1;
        for (int q = 0; q < d.npat; ++q)
            if (uniform_ptr(d.ids)[q] == id) {
                pat = q;
                break;
            }
        // a wavefront whose bins share one pattern reads its weights through
        // the scalar cache
        const int pat0 = __builtin_amdgcn_readfirstlane(pat);
        const bool uniform_pat = __all(pat == pat0);
        double fd = lut.ref_value;
        if (lut.have_data) {
            // the corners are re-read only when the pulse moves to other LUT rows
            if (d.ly.i1 != qy1 || d.ly.i2 != qy2) {
                q = lut_corners(lut, lx, d.ly);
                qy1 = d.ly.i1;
                qy2 = d.ly.i2;
            }
            fd = lut_combine(q, lx, d.ly);
        }
        float2 sum = make_float2(0.f, 0.f);
        const float2* row = d.raw_row;
        const auto* tp = uniform_ptr(d.t_row);
        const float* w = d.w + (int64_t)max(pat, 0) * d.nw;
        const auto* w0 = uniform_ptr(d.w) + (int64_t)max(pat0, 0) * d.nw;
        if (pat >= 0 && d.nw <= kFixedTaps && staged) {
            const float2* srow = s_raw + (d.offset - pfirst) * kThreads;
            switch (d.nw) {
#define PRESUM_TAPS(N)                                                                        \
    case N:                                                                                   \
        sum = window_sum<N>(srow, kThreads, threadIdx.x, w, w0, uniform_pat, tp, d.tout, fd); \
        break;
                PRESUM_TAPS(1) PRESUM_TAPS(2) PRESUM_TAPS(3) PRESUM_TAPS(4)
                PRESUM_TAPS(5) PRESUM_TAPS(6) PRESUM_TAPS(7) PRESUM_TAPS(8)
#undef PRESUM_TAPS
            default: break;
            }
        } else if (pat >= 0 && d.nw <= kFixedTaps) {
            // windows of up to kFixedTaps pulses with the tap count fixed at
            // compile time
            switch (d.nw) {
#define PRESUM_TAPS(N)                                                                        \
    case N:                                                                                   \
        sum = window_sum<N>(row, a.width, jj, w, w0, uniform_pat, tp, d.tout, fd);            \
        break;
                PRESUM_TAPS(1) PRESUM_TAPS(2) PRESUM_TAPS(3) PRESUM_TAPS(4)
                PRESUM_TAPS(5) PRESUM_TAPS(6) PRESUM_TAPS(7) PRESUM_TAPS(8)
#undef PRESUM_TAPS
            default: break;
            }
        } else if (pat >= 0) {
            // taps in groups of kTapGroup: the group's loads first, then the
            // products, accumulated in tap order
            for (int k0 = 0; k0 < d.nw; k0 += kTapGroup) {
                float2 x[kTapGroup];
                float wv[kTapGroup];
                double tk[kTapGroup];
#pragma unroll
                for (int u = 0; u < kTapGroup; ++u)
                    if (k0 + u < d.nw) {
                        x[u] = row[jj];
                        row += a.width;
                        wv[u] = uniform_pat ? w0[k0 + u] : w[k0 + u];
                        tk[u] = tp[k0 + u];
                    }
#pragma unroll
                for (int u = 0; u < kTapGroup; ++u)
                    if (k0 + u < d.nw) {
                        // deramp exp(-2 pi i (t_k - tout) fd): the phase in
                        // turns, reduced to [0, 1) in FP64, then the hardware
                        // sine and cosine of 2 pi times it
                        const double trel = tk[u] - d.tout;
                        const float turns = (float)__builtin_amdgcn_fract(-trel * fd);
                        const float sn = __builtin_amdgcn_sinf(turns), cs = __builtin_amdgcn_cosf(turns);
                        sum = cadd(sum, cmul(make_float2(wv[u] * cs, wv[u] * sn), x[u]));
                    }
            }
        }
        __builtin_nontemporal_store(v2f{sum.x, sum.y}, reinterpret_cast<v2f*>(d.out_row + jj));
    }
}

struct ApplyArgs {
    float2* out;
    const double* t;
    const double* fd;
    const double* w;
    const float2* x;
    int64_t m, n;
};

__global__ __launch_bounds__(kThreads) void apply_kernel(ApplyArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= a.n) return;
    const double fd = a.fd[j];
    float2 sum = make_float2(0.f, 0.f);
    for (int64_t i = 0; i < a.m; ++i) {
        const float ph = (float)(-2 * M_PI * a.t[i] * fd);
        float s, c;
        sincosf(ph, &s, &c);
        const float wk = (float)a.w[i * a.n + j];
        sum = cadd(sum, cmul(make_float2(wk * c, wk * s), a.x[i * a.n + j]));
    }
    a.out[j] = sum;
}

// Pattern tables of a chunk of output pulses.
struct Patterns {
    std::vector<int64_t> id;
    std::vector<float> w;  // nw per pattern
    int status = BP_SUCCESS;
};

// Window, distinct patterns and their weights for output pulses [i0, i1);
// desc[i].pbase and .wbase are relative to out's tables.
int plan_pulses(const bp::Orbit& ho, const double* t, int64_t nt, const int32_t* swaths, int64_t ns,
                int64_t W, double prf, double antenna_length, const bp_lut2d& lut, int64_t i0,
                int64_t i1, PulseDesc* desc, Patterns& out, SolveScratch& scratch)
{
    std::vector<double> wsub, gfull, rfull;
    std::vector<int64_t> cuts;
    int idx[kMaxWindow + 1];
    for (int64_t i = i0; i < i1; ++i) {
        const double tout = t[0] + (double)i / prf;
        bp::V3 pos, vel;
        if (bp::orbit_interpolate(ho, tout, bp::kBorderError, &pos, &vel) != bp::kSuccess)
            return BP_ERR_OUT_OF_RANGE;
        const double v = bp::norm(vel);
        const bp_acorr acor{BP_ACORR_AZIMUTH, 0, 2.0 * (antenna_length / v)};
        // the window of the full weight vector (focus.py takes its length)
        int64_t offset = 0;
        const int64_t nw64 = presum_window(acor, t, nt, tout, &offset);
        if (nw64 > kMaxWindow) return BP_ERR_LENGTH;
        const int nw = (int)nw64;
        PulseDesc& d = desc[i];
        d.tout = tout;
        d.offset = offset;
        d.nw = nw;
        d.pbase = (int32_t)out.id.size();
        d.npat = 0;
        d.wbase = (int32_t)out.w.size();
        d.ly = lut.have_data ? lut_axis(tout, lut.ystart, lut.dy, lut.length) : LutAxis{0.0, 0, 0};
        // range bins where the pattern can change: every sub-swath boundary
        cuts.assign(1, 0);
        for (int k = 0; k < nw; ++k)
            for (int64_t s = 0; s < ns; ++s) {
                const int32_t* sw = swaths + (s * nt + offset + k) * 2;
                for (int e = 0; e < 2; ++e)
                    if (sw[e] > 0 && sw[e] < W) cuts.push_back(sw[e]);
            }
        std::sort(cuts.begin(), cuts.end());
        cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
        // autocorrelations among the window's pulses and with tout, computed
        // once: each pattern's system is the submatrix of its valid pulses
        // (the values getPresumWeights forms from those pulses alone)
        gfull.resize((size_t)nw * nw);
        rfull.resize(nw);
        for (int k = 0; k < nw; ++k) {
            const double xk = t[offset + k];
            for (int l = k; l < nw; ++l) gfull[(size_t)l * nw + k] = acorr_eval(acor, t[offset + l] - xk);
            rfull[k] = acorr_eval(acor, tout - xk);
        }
        for (int64_t c : cuts) {
            int64_t id = 0;
            for (int k = 0; k < nw; ++k)
                for (int64_t s = 0; s < ns; ++s) {
                    const int32_t* sw = swaths + (s * nt + offset + k) * 2;
                    if (c >= sw[0] && c < sw[1]) {
                        id |= (int64_t)1 << k;
                        break;
                    }
                }
            bool seen = false;
            for (int q = 0; q < d.npat; ++q) seen |= out.id[d.pbase + q] == id;
            if (seen) continue;
            // weights from the valid window pulses only, zeros elsewhere (all
            // of them lie within the half width of tout: the window's do)
            int m = 0;
            for (int k = 0; k < nw; ++k)
                if (id >> k & 1) idx[m++] = k;
            scratch.g.assign((size_t)m * m, 0.0);
            scratch.r.resize(m);
            for (int a = 0; a < m; ++a) {
                for (int b = a; b < m; ++b)
                    scratch.g[(size_t)b * m + a] = gfull[(size_t)idx[b] * nw + idx[a]];
                scratch.r[a] = rfull[idx[a]];
            }
            ldlt_solve(scratch, m, wsub);
            out.id.push_back(id);
            const size_t base = out.w.size();
            out.w.resize(base + nw, 0.f);
            for (int k = 0, u = 0; k < nw; ++k)
                if (id >> k & 1) out.w[base + k] = (float)wsub[u++];
            ++d.npat;
        }
    }
    return BP_SUCCESS;
}

struct Part {
    const void* src;
    size_t bytes;
};

// Pinned staging buffer and its device copy for a set of tables.  Buffers
// live in a process-wide pool for later calls on the same device -- pinning
// and freeing host memory costs more than the copies -- and the pool is never
// freed.
struct Staging {
    int device = 0;
    void* host = nullptr;
    void* dev = nullptr;
    size_t capacity = 0;
    size_t pending = 0;   // bytes placed by the last reserve()
    hipEvent_t ready = nullptr;   // the last upload's completion
    // Places the parts at 256-byte aligned offsets (into off), each followed
    // by at least kPartSlack bytes inside the buffer (so that a fixed-size
    // read may run past the end of a table), growing the buffers if needed
    // (the current device is `device`); then send() packs and enqueues the
    // upload on s.
    hipError_t reserve(const Part* parts, int n, size_t* off)
    {
        size_t total = 0;
        for (int k = 0; k < n; ++k) {
            off[k] = total;
            total += (parts[k].bytes + kPartSlack + 255) / 256 * 256;
        }
        pending = total;
        if (total == 0 || total <= capacity) return hipSuccess;
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
        dev = host = nullptr;
        capacity = 0;
        const size_t cap = std::max<size_t>(total + total / 2, size_t(1) << 20);
        hipError_t e = hipHostMalloc(&host, cap, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        e = hipMalloc(&dev, cap);
        if (e != hipSuccess) return e;
        capacity = cap;
        return hipSuccess;
    }
    // The copy goes on stream `copy`; `s` waits for it (an event kept with
    // the buffer), so copies for later launches overlap the kernels on s.
    hipError_t send(const Part* parts, int n, const size_t* off, hipStream_t copy, hipStream_t s)
    {
        if (pending == 0) return hipSuccess;
        for (int k = 0; k < n; ++k)
            if (parts[k].bytes) std::memcpy(static_cast<char*>(host) + off[k], parts[k].src, parts[k].bytes);
        hipError_t e = hipMemcpyAsync(dev, host, pending, hipMemcpyHostToDevice, copy);
        if (e != hipSuccess || copy == s) return e;
        if (ready == nullptr) {
            e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
        e = hipEventRecord(ready, copy);
        return e != hipSuccess ? e : hipStreamWaitEvent(s, ready, 0);
    }
    hipError_t upload(const Part* parts, int n, size_t* off, hipStream_t copy, hipStream_t s)
    {
        const hipError_t e = reserve(parts, n, off);
        return e != hipSuccess ? e : send(parts, n, off, copy, s);
    }
    const char* at(size_t off) const { return static_cast<const char*>(dev) + off; }
};

struct StagingPool {
    std::mutex mu;
    std::vector<Staging*> idle;
    Staging* take(int device)
    {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t k = 0; k < idle.size(); ++k)
            if (idle[k]->device == device) {
                Staging* b = idle[k];
                idle.erase(idle.begin() + k);
                return b;
            }
        auto* b = new Staging();
        b->device = device;
        return b;
    }
    void give(Staging* b)
    {
        std::lock_guard<std::mutex> lk(mu);
        idle.push_back(b);
    }
    // One non-blocking stream per device for the uploads (nullptr if it
    // cannot be created: the uploads then go on the caller's stream); the
    // current device is `device`.
    std::vector<std::pair<int, hipStream_t>> copy_streams;
    hipStream_t copy_stream(int device)
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const auto& cs : copy_streams)
            if (cs.first == device) return cs.second;
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
        copy_streams.emplace_back(device, st);
        return st;
    }
};

StagingPool& staging_pool()
{
    static auto* p = new StagingPool();   // lives until the process ends
    return *p;
}

// Host threads that plan the items of every call: created when a call first
// wants them (each inherits that caller's CPU affinity), never freed; between
// calls they wait on a condition variable.  A call posts its job; every idle
// thread takes items of the oldest job that has items left, by the job's
// atomic counter, alongside the calling thread.  Before it returns, the call
// retracts its job: no thread starts one of its items afterwards, and the
// call waits for the threads still inside one.  (A process forked after the
// pool started has no pool threads: its calls plan on the calling thread.)
struct PlanJob {
    std::function<void(int64_t, SolveScratch&)> run;
    std::atomic<int64_t>* next = nullptr;
    int64_t nitems = 0;
    int active = 0;        // pool threads inside the job (guarded by the pool's mutex)
    bool queued = false;   // in the pool's queue (guarded by the pool's mutex)
};

struct PlanPool {
    std::mutex mu;
    std::condition_variable work;   // a job was posted
    std::condition_variable left;   // a thread left a job
    std::deque<PlanJob*> jobs;
    int threads = 0;

    // At least n threads (fewer if no more can start).
    void ensure(int n)
    {
        std::lock_guard<std::mutex> lk(mu);
        while (threads < n) {
            try {
                std::thread(&PlanPool::loop, this).detach();
            } catch (...) {
                break;
            }
            ++threads;
        }
    }
    void post(PlanJob* j)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            j->queued = true;
            jobs.push_back(j);
        }
        work.notify_all();
    }
    void unqueue(PlanJob* j)  // mu held
    {
        if (!j->queued) return;
        jobs.erase(std::find(jobs.begin(), jobs.end(), j));
        j->queued = false;
    }
    void retract(PlanJob* j)
    {
        std::unique_lock<std::mutex> lk(mu);
        unqueue(j);
        left.wait(lk, [&] { return j->active == 0; });
    }
    void loop()
    {
        SolveScratch sc;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            work.wait(lk, [&] { return !jobs.empty(); });
            PlanJob* j = jobs.front();
            ++j->active;
            lk.unlock();
            for (int64_t k = (*j->next)++; k < j->nitems; k = (*j->next)++) j->run(k, sc);
            lk.lock();
            unqueue(j);  // no items left
            --j->active;
            left.notify_all();
        }
    }
};

PlanPool& plan_pool()
{
    static auto* p = new PlanPool();   // lives until the process ends
    return *p;
}

// The staging buffers one call holds; they go back to the pool once the
// call's stream and the copy stream have finished with them.
struct Leases {
    hipStream_t stream;
    hipStream_t copy;
    std::vector<Staging*> held;
    Staging* take(int device)
    {
        held.push_back(staging_pool().take(device));
        return held.back();
    }
    ~Leases()
    {
        if (!held.empty()) {
            (void)hipStreamSynchronize(stream);
            if (copy != stream) (void)hipStreamSynchronize(copy);
        }
        for (Staging* b : held) staging_pool().give(b);
    }
};

}  // namespace
}  // namespace presum

extern "C" {

double bp_acorr_eval(const bp_acorr* acorr, double x)
{
    return acorr ? presum::acorr_eval(*acorr, x) : 0.0;
}

int bp_presum_weights(const bp_acorr* acorr, const double* t, int64_t n, double tout,
                      int64_t* offset, double* w, int64_t max_w, int64_t* nw)
try {
    using presum::fail;
    if (acorr == nullptr || offset == nullptr || nw == nullptr || (n > 0 && t == nullptr))
        return fail(BP_ERR_INVALID_ARGUMENT, "no storage provided for index offset");
    if (acorr->kind != BP_ACORR_AZIMUTH && acorr->kind != BP_ACORR_BARTLETT)
        return fail(BP_ERR_INVALID_ARGUMENT, "unknown autocorrelation kernel");
    std::vector<double> ww;
    presum::SolveScratch scratch;
    presum::presum_weights(*acorr, t, n, tout, offset, ww, scratch);
    *nw = (int64_t)ww.size();
    if (*nw > max_w) return fail(BP_ERR_LENGTH, "weight buffer too short");
    if (*nw > 0) std::memcpy(w, ww.data(), sizeof(double) * ww.size());
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_presum_apply_weights(void* out, const double* pulse_times, const double* doppler,
                            const double* weights, const void* pulses, int64_t m, int64_t n,
                            void* stream)
try {
    using presum::fail;
    if (m < 0 || n < 0) return fail(BP_ERR_LENGTH, "negative dimensions");
    if (n == 0) return BP_SUCCESS;
    if (!out || !doppler || (m > 0 && (!pulse_times || !weights || !pulses)))
        return fail(BP_ERR_INVALID_ARGUMENT, "null buffer");
    presum::ApplyArgs a{static_cast<float2*>(out), pulse_times, doppler, weights,
                        static_cast<const float2*>(pulses), m, n};
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(presum::apply_kernel, dim3((unsigned)((n + presum::kThreads - 1) / presum::kThreads)),
                       dim3(presum::kThreads), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && s == nullptr) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(BP_ERR_HIP, std::string("apply_presum_weights: ") + hipGetErrorString(e));
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_presum_compute_ids(const uint8_t* mask, int64_t nr, int64_t np, int64_t* ids)
try {
    using presum::fail;
    if (np > 63)
        return fail(BP_ERR_LENGTH, "Hash algorithm requires number of pulses less than number "
                                   "of bits in identifier (64)");
    if (nr > 0 && (mask == nullptr || ids == nullptr)) return fail(BP_ERR_INVALID_ARGUMENT, "null buffer");
    for (int64_t i = 0; i < nr; ++i) {
        int64_t id = 0;
        for (int64_t j = 0; j < np; ++j)
            if (mask[i * np + j]) id |= (int64_t)1 << j;
        ids[i] = id;
    }
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_presum_resample(int device, const double* t, int64_t nt, const int32_t* swaths,
                       int64_t ns, const bp_grid* grid, const bp_orbit* orbit,
                       const bp_lut2d* doppler, double antenna_length, const void* raw,
                       void* out, void* stream)
try {
    using namespace presum;
    if (!grid || !orbit || !doppler || !t || (ns > 0 && !swaths) || nt < 1)
        return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (grid->length < 0 || grid->width < 0) return fail(BP_ERR_INVALID_ARGUMENT, "negative grid size");
    if (grid->length == 0 || grid->width == 0) return BP_SUCCESS;
    if (!raw || !out) return fail(BP_ERR_INVALID_ARGUMENT, "null buffer");
    if (!(grid->prf > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "PRF must be positive");
    if (!(antenna_length > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "antenna length must be positive");
    if (orbit->n < 2 || !orbit->pos || !orbit->vel || !(orbit->dt > 0))
        return fail(BP_ERR_INVALID_ARGUMENT, "orbit needs at least two state vectors");
    if (doppler->have_data && (doppler->width < 2 || doppler->length < 2 || !doppler->data))
        return fail(BP_ERR_INVALID_ARGUMENT, "Doppler LUT2d needs at least 2x2 samples");
    for (int64_t i = 1; i < nt; ++i)
        if (!(t[i] >= t[i - 1])) return fail(BP_ERR_INVALID_ARGUMENT, "pulse times must be sorted");

    bp::Orbit ho{orbit->t0, orbit->dt, orbit->n, orbit->method, orbit->pos, orbit->vel};
    const int64_t L = grid->length, W = grid->width;
    // every field of every descriptor is written by the planning or before
    // its launch's upload: no zero fill of the L descriptors
    std::unique_ptr<PulseDesc[]> desc_buf(new PulseDesc[L]);
    PulseDesc* const desc = desc_buf.get();
    // Output pulses in launches (chunks) of kFirstChunk pulses, then twice as
    // many each time up to kChunk (planning a chunk takes longer than the
    // GPU's work on the one before only while the chunks still grow), each
    // split into items of kItem pulses.  Host threads -- and this thread
    // while it waits for a chunk -- take the items in order, each item with
    // its own pattern tables, while the GPU resamples the chunks already
    // planned.
    std::vector<int64_t> chunk_start{0};
    for (int64_t c0 = 0, n = kFirstChunk;;) {
        c0 = std::min(L, c0 + n);
        if (c0 >= L) break;
        chunk_start.push_back(c0);
        n = std::min(2 * n, kChunk);
    }
    chunk_start.push_back(L);
    const int64_t nchunk = (int64_t)chunk_start.size() - 1;
    struct Item {
        int64_t i0, i1, chunk;
    };
    std::vector<Item> items;
    std::vector<int64_t> first_item(nchunk + 1);
    for (int64_t c = 0; c < nchunk; ++c) {
        first_item[c] = (int64_t)items.size();
        for (int64_t i0 = chunk_start[c]; i0 < chunk_start[c + 1]; i0 += kItem)
            items.push_back({i0, std::min(chunk_start[c + 1], i0 + kItem), c});
    }
    first_item[nchunk] = (int64_t)items.size();
    const int64_t nitems = (int64_t)items.size();
    // planners: the calling thread and pool threads, as many as the CPUs
    // the calling thread may run on, at most 16 (a caller pinned to one CPU
    // plans alone; pool threads started for one call serve later ones too)
    const int nworkers = (int)std::max<int64_t>(
        0, std::min<int64_t>({(int64_t)bp::usable_cpus(), 16, nitems}) - 1);
    std::vector<Patterns> item_pat(nitems);
    std::vector<int64_t> done(nchunk, 0);
    SolveScratch own_scratch;   // this thread's
    std::mutex mu;
    std::condition_variable cv;
    std::atomic<bool> abort{false};
    std::atomic<int64_t> next{0};
    auto run_item = [&](int64_t k, SolveScratch& sc) {
        const Item& it = items[k];
        Patterns& pt = item_pat[k];
        // an item always counts as done (an exception, on a pool thread too,
        // becomes its status), so no waiter is left waiting for it
        try {
            pt.status = abort ? BP_ERR_RUNTIME
                              : plan_pulses(ho, t, nt, swaths, ns, W, grid->prf, antenna_length,
                                            *doppler, it.i0, it.i1, desc, pt, sc);
        } catch (const std::bad_alloc&) {
            pt.status = BP_ERR_OUT_OF_MEMORY;
        } catch (...) {
            pt.status = BP_ERR_RUNTIME;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            ++done[it.chunk];
        }
        cv.notify_all();
    };
    // the job is retracted (no pool thread inside it) before anything it
    // refers to on this stack goes away, an exception included
    struct Retractor {
        PlanPool& pool;
        PlanJob& job;
        std::atomic<bool>& abort;
        ~Retractor()
        {
            abort = true;
            pool.retract(&job);
        }
    };
    PlanPool& planners = plan_pool();
    PlanJob job;
    job.run = run_item;
    job.next = &next;
    job.nitems = nitems;
    if (nworkers > 0) planners.ensure(nworkers);
    Retractor retractor{planners, job, abort};
    if (nworkers > 0) planners.post(&job);

    // every failure below stops the planners and waits for the launches
    // already made; the staging buffers go back to the pool after the wait
    // (destructors)
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool device_set = hipSetDevice(device) == hipSuccess;
    hipStream_t copy = device_set ? staging_pool().copy_stream(device) : nullptr;
    if (copy == nullptr) copy = s;
    Leases leases{s, copy, {}};
    auto finish = [&](int code, const std::string& msg) {
        abort = true;
        planners.retract(&job);
        const hipError_t e = hipStreamSynchronize(s);
        if (code == BP_SUCCESS && e != hipSuccess)
            return fail(BP_ERR_HIP, std::string("presum resample: ") + hipGetErrorString(e));
        return code == BP_SUCCESS ? BP_SUCCESS : fail(code, msg);
    };
    if (!device_set) return finish(BP_ERR_HIP, "hipSetDevice failed");
    // tables every chunk reads: pulse times, sub-swaths, Doppler samples
    const Part shared_parts[3] = {
        {t, sizeof(double) * nt},
        {swaths, sizeof(int32_t) * 2 * ns * nt},
        {doppler->have_data ? doppler->data : nullptr,
         doppler->have_data ? sizeof(double) * doppler->width * doppler->length : 0}};
    size_t shared_off[3];
    Staging* shared = leases.take(device);
    hipError_t e = shared->upload(shared_parts, 3, shared_off, copy, s);
    if (e != hipSuccess) return finish(BP_ERR_HIP, std::string("presum upload: ") + hipGetErrorString(e));

    ResampleArgs a;
    std::memset(&a, 0, sizeof(a));
    a.raw = static_cast<const float2*>(raw);
    a.out = static_cast<float2*>(out);
    a.t = reinterpret_cast<const double*>(shared->at(shared_off[0]));
    a.swaths = reinterpret_cast<const int32_t*>(shared->at(shared_off[1]));
    a.doppler = bp::Lut2d{doppler->have_data, doppler->bounds_error, doppler->ref_value,
                          doppler->xstart, doppler->ystart, doppler->dx, doppler->dy,
                          doppler->width, doppler->length,
                          reinterpret_cast<const double*>(shared->at(shared_off[2]))};
    a.r0 = grid->starting_range;
    a.dr = grid->range_pixel_spacing;
    a.nt = nt;
    a.ns = ns;
    a.width = W;
    a.pulses_per_thread = kPulsesPerThread;
    std::vector<int64_t> pat_id;
    std::vector<float> pat_w;
    for (int64_t c = 0; c < nchunk; ++c) {
        const int64_t need = first_item[c + 1] - first_item[c];
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (done[c] == need) break;
            }
            const int64_t k = next++;
            if (k < nitems) {
                run_item(k, own_scratch);
                continue;
            }
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return done[c] == need; });
            break;
        }
        pat_id.clear();
        pat_w.clear();
        for (int64_t k = first_item[c]; k < first_item[c + 1]; ++k) {
            const Patterns& pt = item_pat[k];
            if (pt.status == BP_ERR_OUT_OF_RANGE)
                return finish(BP_ERR_OUT_OF_RANGE, "output pulse time outside the orbit");
            if (pt.status == BP_ERR_OUT_OF_MEMORY)
                return finish(BP_ERR_OUT_OF_MEMORY, "host memory allocation failed in presum planning");
            if (pt.status == BP_ERR_LENGTH)
                return finish(BP_ERR_LENGTH, "Hash algorithm requires number of pulses less than "
                                             "number of bits in identifier (64)");
            if (pt.status != BP_SUCCESS)
                return finish(BP_ERR_RUNTIME, "presum planning of an output pulse failed");
            // the chunk's tables: its items' tables one after the other
            const int32_t base = (int32_t)pat_id.size(), wb = (int32_t)pat_w.size();
            for (int64_t i = items[k].i0; i < items[k].i1; ++i) {
                desc[i].pbase += base;
                desc[i].wbase += wb;
            }
            pat_id.insert(pat_id.end(), pt.id.begin(), pt.id.end());
            pat_w.insert(pat_w.end(), pt.w.begin(), pt.w.end());
        }
        const int64_t c0 = chunk_start[c], rows = chunk_start[c + 1] - c0;
        const Part parts[3] = {{desc + c0, sizeof(PulseDesc) * rows},
                               {pat_id.data(), sizeof(int64_t) * pat_id.size()},
                               {pat_w.data(), sizeof(float) * pat_w.size()}};
        size_t off[3];
        Staging* staged = leases.take(device);
        e = staged->reserve(parts, 3, off);
        if (e != hipSuccess)
            return finish(BP_ERR_HIP, std::string("presum upload: ") + hipGetErrorString(e));
        a.desc = reinterpret_cast<const PulseDesc*>(staged->at(off[0]));
        a.pat_id = reinterpret_cast<const int64_t*>(staged->at(off[1]));
        a.pat_w = reinterpret_cast<const float*>(staged->at(off[2]));
        // the device addresses each pulse's sum reads and writes
        for (int64_t i = c0; i < c0 + rows; ++i) {
            PulseDesc& d = desc[i];
            d.raw_row = a.raw + d.offset * W;
            d.out_row = a.out + i * W;
            d.t_row = a.t + d.offset;
            d.ids = a.pat_id + d.pbase;
            d.w = a.pat_w + d.wbase;
        }
        e = staged->send(parts, 3, off, copy, s);
        if (e != hipSuccess)
            return finish(BP_ERR_HIP, std::string("presum upload: ") + hipGetErrorString(e));
        a.row0 = c0;
        a.rows = rows;
        const dim3 blocks((unsigned)((W + kThreads - 1) / kThreads),
                          (unsigned)((rows + kPulsesPerThread - 1) / kPulsesPerThread));
        hipLaunchKernelGGL(resample_kernel, blocks, dim3(kThreads), 0, s, a);
        e = hipGetLastError();
        if (e != hipSuccess)
            return finish(BP_ERR_HIP, std::string("presum resample: ") + hipGetErrorString(e));
    }
    // the staging buffers go back to the pool on return, after the kernels
    return finish(BP_SUCCESS, "");
} catch (...) {
    return bp_host_exception();
}

}  // extern "C"
