// C-ABI of the backprojection path: argument checks that mirror the
// reference's (focus/Backproject.cpp:77-92, pybind focus/Backproject.cpp:97-135,
// DryTroposphereModel.cpp:20-33), device uploads into a per-context grow-only
// arena, and the three-pass launch (pulse orbit, per-pixel geometry,
// coherent sum) on the caller's stream.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/isce3_amd_backproject.h"
#include "bp_internal.h"

namespace {

thread_local std::string g_msg;

int fail(int code, const std::string& msg)
{
    g_msg = msg;
    return code;
}

}  // namespace

// Shared with the range-compression entry points (rc_api.hip).
void bp_set_last_error(const std::string& msg) { g_msg = msg; }

// Every int-valued entry point ends in a catch clause that returns this, so
// no C++ exception crosses the C ABI: a failed host allocation gives
// BP_ERR_OUT_OF_MEMORY, any other exception BP_ERR_RUNTIME, with the message
// set (or cleared, if even the message cannot be stored).
int bp_host_exception() noexcept
{
    int code = BP_ERR_RUNTIME;
    try {
        try {
            throw;
        } catch (const std::bad_alloc&) {
            code = BP_ERR_OUT_OF_MEMORY;
            g_msg = "host memory allocation failed";
        } catch (const std::exception& e) {
            g_msg = std::string("host exception: ") + e.what();
        } catch (...) {
            g_msg = "unknown host exception";
        }
    } catch (...) {
        g_msg.clear();
    }
    return code;
}

int bp::usable_cpus()
{
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
    return (int)std::max(1u, std::thread::hardware_concurrency());
}

namespace bp {

int check_out_block(const bp_out_block& b)
{
    const bp_grid& go = b.grid;
    if (go.length < 0 || go.width < 0) return fail(BP_ERR_INVALID_ARGUMENT, "negative grid size");
    if (go.length > INT32_MAX || go.width > INT32_MAX)
        return fail(BP_ERR_INVALID_ARGUMENT, "grid size exceeds int range");
    if (!(go.prf > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "PRF must be positive");
    if (go.length * go.width > 0 && b.out_dev == nullptr)
        return fail(BP_ERR_INVALID_ARGUMENT, "out is null");
    return BP_SUCCESS;
}

}  // namespace bp

namespace {

// A device buffer.  `fixed` buffers are views of a region the library does
// not own (the static scratch of bp_backproject_noalloc): they are never
// grown or freed, and a call that needs more than they hold fails.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool fixed = false;
};

}  // namespace

// Device copies of everything a call needs besides the pulses and the output
// blocks: orbits, Doppler tables, DEM, kernel tables and (plain precision)
// the platform state at every pulse.
struct StaticData {
    DevBuf table, table2, poly, poly2, in_pos, in_vel, out_pos, out_vel, in_dop, out_dop, dem,
        pulse_pv, sinc;
    std::vector<float> poly_host, poly2_host;
    std::vector<double> table2_host;
};

// Per-pixel records, batch header + output descriptors, and the flags of one
// call ([0] any failed, [1] out of range, [8..9] update count, [16 + g] grid g
// failed).
struct Scratch {
    DevBuf geom, split, flags, desc;
};

// Buffers of the host-input path (bp_backproject_host): device pulse windows,
// band descriptors and windows, pinned staging, a copy stream and its events.
// Created on first use and kept by the context.
struct HostPath {
    DevBuf win[2];            // device copies of band pulse windows
    DevBuf band_desc;         // batch header + descriptor of every band
    DevBuf band_win;          // [kmin, kmax) of every line band (int32 pairs)
    unsigned char* stage[2] = {nullptr, nullptr};  // pinned staging of pulse chunks
    size_t stage_bytes = 0;
    int32_t* h_win = nullptr; // pinned copy of band_win
    size_t h_win_words = 0;
    hipStream_t copy = nullptr;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};  // copy out of staging j finished
    hipEvent_t win_ev = nullptr;                  // band windows on the host
    hipEvent_t copied_ev = nullptr;               // a band's pulses on the device
    hipEvent_t sum_ev[2] = {nullptr, nullptr};    // sum reading window buffer w finished
    std::vector<hipEvent_t> timing;               // per band: copy start/end, sum start/end
};

struct bp_context {
    int device = 0;
    int n_cu = 0;
    unsigned char* scratch = nullptr;  // device address of this device's no-alloc scratch
    std::string scratch_err;           // why it is null
    StaticData st;
    Scratch sc;
    HostPath hp;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    std::vector<unsigned char> desc_host;
    std::vector<int> desc_block;
    std::vector<int32_t> flags_host;
};

namespace {

int ensure(DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return BP_SUCCESS;
    if (b.fixed)
        return fail(BP_ERR_LENGTH, "call needs " + std::to_string(bytes) +
                                       " bytes of a fixed scratch buffer of " +
                                       std::to_string(b.bytes));
    if (b.ptr) {
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(BP_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return BP_SUCCESS;
}

int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s)
{
    int rc = ensure(b, bytes);
    if (rc != BP_SUCCESS) return rc;
    if (bytes == 0 || src == nullptr) return BP_SUCCESS;
    hipError_t e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(BP_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
    return BP_SUCCESS;
}

#define BP_HIP_CHECK(expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(BP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

const char* orbit_err_string(int code)
{
    switch (code) {
    case BP_ORBIT_INTERP_SIZE_ERROR: return "insufficient orbit state vectors to form interpolant";
    case BP_ORBIT_INTERP_DOMAIN_ERROR: return "interpolation point outside orbit domain";
    case BP_ORBIT_INTERP_UNKNOWN_METHOD: return "unexpected orbit interpolation method";
    default: return "orbit interpolation error";
    }
}

int check_orbit(const bp_orbit& o, const char* what)
{
    if (o.n < 2 || o.pos == nullptr || o.vel == nullptr)
        return fail(BP_ERR_INVALID_ARGUMENT, std::string(what) + ": orbit needs at least two state vectors");
    if (!(o.dt > 0.0)) return fail(BP_ERR_INVALID_ARGUMENT, std::string(what) + ": orbit spacing must be positive");
    if (o.method != BP_INTERP_HERMITE && o.method != BP_INTERP_LEGENDRE)
        return fail(BP_ERR_OUT_OF_RANGE, orbit_err_string(BP_ORBIT_INTERP_UNKNOWN_METHOD));
    const int minsv = o.method == BP_INTERP_HERMITE ? 4 : 9;
    if (o.n < minsv) return fail(BP_ERR_OUT_OF_RANGE, orbit_err_string(BP_ORBIT_INTERP_SIZE_ERROR));
    return BP_SUCCESS;
}

int check_lut(const bp_lut2d& l, const char* what)
{
    if (!l.have_data) return BP_SUCCESS;
    if (l.width < 2 || l.length < 2 || l.data == nullptr || !(l.dx != 0.0) || !(l.dy != 0.0))
        return fail(BP_ERR_INVALID_ARGUMENT, std::string(what) + ": Doppler LUT2d needs at least 2x2 samples");
    return BP_SUCCESS;
}

bp::Orbit dev_orbit(const bp_orbit& o, const DevBuf& pos, const DevBuf& vel)
{
    bp::Orbit d;
    d.t0 = o.t0;
    d.dt = o.dt;
    d.n = o.n;
    d.method = o.method;
    d.pos = static_cast<const double*>(pos.ptr);
    d.vel = static_cast<const double*>(vel.ptr);
    return d;
}

bp::Lut2d dev_lut(const bp_lut2d& l, const DevBuf& data)
{
    bp::Lut2d d;
    d.have_data = l.have_data;
    d.bounds_error = l.bounds_error;
    d.ref_value = l.ref_value;
    d.xstart = l.xstart;
    d.ystart = l.ystart;
    d.dx = l.dx;
    d.dy = l.dy;
    d.width = l.width;
    d.length = l.length;
    d.data = static_cast<const double*>(data.ptr);
    return d;
}

bp::Orbit host_orbit(const bp_orbit& o)
{
    bp::Orbit d;
    d.t0 = o.t0;
    d.dt = o.dt;
    d.n = o.n;
    d.method = o.method;
    d.pos = o.pos;
    d.vel = o.vel;
    return d;
}

bp::Lut2d host_lut(const bp_lut2d& l)
{
    bp::Lut2d d;
    d.have_data = l.have_data;
    d.bounds_error = l.bounds_error;
    d.ref_value = l.ref_value;
    d.xstart = l.xstart;
    d.ystart = l.ystart;
    d.dx = l.dx;
    d.dy = l.dy;
    d.width = l.width;
    d.length = l.length;
    d.data = l.data;
    return d;
}

// The windowed-sinc kernel of the DEM factory's Sinc2dInterpolator(8, 8192)
// (core/Sinc2dInterpolator.cpp:14-41, coefficients :105-120): the 65536
// samples of a sinc under a raised-cosine window, regrouped into 8192
// fractional rows of 8 taps, each row divided by its sum; stored as float
// (the interpolator casts each coefficient to the raster's float type).
// Built once per process.
const std::vector<float>& sinc_kernel_table()
{
    static const std::vector<float> table = [] {
        constexpr int L = bp::kSincLen, D = bp::kSincSub, n = L * D;
        const double mid = (n - 1.0) / 2.0;
        std::vector<double> f(n);
        for (int i = 0; i < n; ++i) {
            const double window = 0.5 + 0.5 * std::cos(M_PI * (i - mid) / mid);
            const double s = std::floor(i - mid) / D;
            f[i] = window * (s != 0.0 ? std::sin(M_PI * s) / (M_PI * s) : 1.0);
        }
        std::vector<float> t((size_t)n);
        for (int r = 0; r < D; ++r) {
            double sum = 0.0;
            for (int j = 0; j < L; ++j) sum += f[r + D * j];
            for (int j = 0; j < L; ++j) t[(size_t)r * L + j] = (float)(f[r + D * j] / sum);
        }
        return t;
    }();
    return table;
}

bp::Dem host_dem(const bp_dem& m)
{
    bp::Dem d;
    d.have_raster = m.have_raster;
    d.epsg = m.epsg;
    d.method = m.method;
    d.pad_ = 0;
    d.ref_height = m.ref_height;
    d.xstart = m.xstart;
    d.ystart = m.ystart;
    d.dx = m.dx;
    d.dy = m.dy;
    d.width = m.width;
    d.length = m.length;
    d.data = m.data;
    d.sinc = m.method == BP_DEM_SINC ? sinc_kernel_table().data() : nullptr;
    if (!bp::proj_init(m.epsg, &d.proj)) bp::proj_init(4326, &d.proj);
    return d;
}

bool dem_method_ok(int m)
{
    return m == BP_DEM_BILINEAR || m == BP_DEM_BICUBIC || m == BP_DEM_NEAREST ||
           m == BP_DEM_BIQUINTIC || m == BP_DEM_SINC;
}

bp::TileMap make_map(int64_t length, int64_t width, int tile_rows, int order, int lane_rows)
{
    bp::TileMap m;
    m.tile_rows = tile_rows;
    m.tile_cols = 64 / tile_rows;
    m.lane_rows = lane_rows;
    const int64_t rows = (int64_t)tile_rows * lane_rows;
    m.tiles_x = (int32_t)((width + m.tile_cols - 1) / m.tile_cols);
    m.tiles_y = (int32_t)((length + rows - 1) / rows);
    m.length = length;
    m.width = width;
    m.n_tiles = (int64_t)m.tiles_x * m.tiles_y;
    m.order = order;
    return m;
}

// Polyphase weight rows for the split kernel: row q holds the W tap weights
// at fractional offset t0 = tmin + q/Q (evaluated from the float table with
// the reference's linear interpolation, in FP64) and the difference to row q+1.
void build_polyphase(const bp_kernel& k, int width, std::vector<float>& out)
{
    const int Q = bp::kPolyphaseRows;
    const int wp = (width + 3) / 4 * 4;
    const double tmin = (width % 2 == 0) ? 0.0 : -0.5;
    const double hw = k.halfwidth;
    const double dx = hw / (k.n - 1.0);
    const float rdx = (float)(1.0 / dx);
    auto weight = [&](double x) {
        double ax = std::fabs(x);
        if (ax > hw) return 0.0;
        double axn = ax * (double)rdx;
        int i = std::min((int)std::floor(axn), k.n - 2);
        return (double)k.table[i] + (axn - i) * (double)(k.table[i + 1] - k.table[i]);
    };
    std::vector<double> w((size_t)(Q + 1) * width);
    for (int q = 0; q <= Q; ++q) {
        const double t0 = tmin + (double)q / Q;
        for (int m = 0; m < width; ++m) w[(size_t)q * width + m] = weight((m - width / 2) + t0);
    }
    // Q + 1 rows so that offsets t0 = tmin + 1 (q = Q) index a valid row
    // (whose deltas are zero) without clamping in the kernel.
    const int row = bp::poly_row_floats(width);
    out.assign((size_t)(Q + 1) * row, 0.f);
    for (int q = 0; q <= Q; ++q) {
        for (int m = 0; m < width; ++m) {
            const float w0 = (float)w[(size_t)q * width + m];
            const float w1 = q < Q ? (float)w[(size_t)(q + 1) * width + m] : w0;
            out[(size_t)q * row + m] = w0;
            out[(size_t)q * row + wp + m] = w1 - w0;
        }
    }
}

// Two-line weight table of kernel width W: Q = 2^pair_q_log2(W) rows at
// fractional offsets t0 = tmin + q/Q (tmin = 0 for even W, -1/2 for odd W, as
// the one-line table), each row pair_row_floats(W) floats, in two copies
// stored one after the other (copy o of row q at floats (2 q + o) RF): copy 0
// holds the W weights in entries 0..W-1 and a zero in entry W, copy 1 a zero
// in entry 0 and the weights in entries 1..W.  Entries interpolated in
// the sum (pair_lerp_lo .. + pair_lerp_n) hold the weight at t0 and, after the
// W + 1 entries, its difference to the weight at t0 + 1/Q; the others hold the
// weight at t0 + 1/(2Q), the middle of the row's interval.
void build_pair_table(const bp_kernel& k, int W, std::vector<float>& out)
{
    const int Q = 1 << bp::pair_q_log2(W);
    const int RF = bp::pair_row_floats(W);
    const int lo = bp::pair_lerp_lo(W), nl = bp::pair_lerp_n(W);
    const double tmin = (W % 2 == 0) ? 0.0 : -0.5;
    const double hw = k.halfwidth;
    const double dx = hw / (k.n - 1.0);
    const float rdx = (float)(1.0 / dx);
    auto weight = [&](double x) {
        double ax = std::fabs(x);
        if (ax > hw) return 0.0;
        double axn = ax * (double)rdx;
        int i = std::min((int)std::floor(axn), k.n - 2);
        return (double)k.table[i] + (axn - i) * (double)(k.table[i + 1] - k.table[i]);
    };
    out.assign((size_t)bp::pair_table_floats(W), 0.f);
    for (int o = 0; o < 2; ++o) {
        for (int q = 0; q < Q; ++q) {
            float* r = &out[((size_t)q * 2 + o) * RF];
            const double t0 = tmin + (double)q / Q;
            for (int e = 0; e <= W; ++e) {
                const int m = e - o;  // tap of entry e in this copy
                if (m < 0 || m >= W) continue;
                const double x = (m - W / 2) + t0;
                if (e >= lo && e < lo + nl) {
                    const float w0 = (float)weight(x), w1 = (float)weight(x + 1.0 / Q);
                    r[e] = w0;
                    r[W + 1 + (e - lo)] = w1 - w0;
                } else {
                    r[e] = (float)weight(x + 0.5 / Q);
                }
            }
        }
    }
}

// Plain-sum table of TabulatedKernel<float> (core/Kernels.icc:137-154): entry
// i holds (t_i, t_{i+1} - t_i) as doubles with the difference formed in float,
// as the reference's evaluation forms it; entry n - 1 holds (t_{n-2} +
// (t_{n-1} - t_{n-2}), t_{n-1} - t_{n-2}), what the reference's clamp of the
// index to n - 2 evaluates for |x| at the half width.
void build_plain_table(const bp_kernel& k, std::vector<double>& out)
{
    const int n = k.n;
    out.assign((size_t)2 * n, 0.0);
    for (int i = 0; i + 1 < n; ++i) {
        out[2 * i] = (double)k.table[i];
        out[2 * i + 1] = (double)(float)(k.table[i + 1] - k.table[i]);
    }
    const double d = (double)(float)(k.table[n - 1] - k.table[n - 2]);
    out[2 * (n - 1)] = (double)k.table[n - 2] + d;
    out[2 * (n - 1) + 1] = d;
}

// The scratch region of bp_backproject_noalloc: one array per device, part of
// the library's code object, held until the process ends.  Host code never
// passes `scratch` itself (in host code the name is a stand-in, not a device
// address): bp_context_create looks up each device's address once, with that
// device current, and every use goes through that address.
constexpr size_t kNoAllocScratchBytes = size_t(BP_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device `device`.
int scratch_address(int device, unsigned char** addr)
{
    *addr = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return fail(BP_ERR_INVALID_ARGUMENT, "device index beyond the library's device table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            return fail(BP_ERR_HIP, std::string("hipGetSymbolAddress of the no-alloc scratch failed: ") +
                                        hipGetErrorString(e));
        }
        g_scratch_addr[device] = static_cast<unsigned char*>(p);
    }
    *addr = g_scratch_addr[device];
    return BP_SUCCESS;
}

}  // namespace

extern "C" {

const char* bp_last_error_message(void) { return g_msg.c_str(); }

int64_t bp_scratch_bytes_per_pixel(int precision)
{
    return (int64_t)sizeof(bp::PixGeom) +
           (precision == BP_PRECISION_SPLIT ? (int64_t)sizeof(bp::PixSplit) : 0);
}

int bp_context_create(int device, bp_context** ctx)
try {
    if (ctx == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "ctx is null");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return fail(BP_ERR_HIP, "no HIP device available (isce3_amd needs an MI355X GPU)");
    if (device < 0 || device >= ndev) return fail(BP_ERR_INVALID_ARGUMENT, "invalid device index");
    BP_HIP_CHECK(hipSetDevice(device));
    auto* c = new bp_context();
    c->device = device;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        c->n_cu < 1)
        c->n_cu = 256;
    for (auto& ev : c->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            for (auto& e2 : c->ev)
                if (e2) (void)hipEventDestroy(e2);
            delete c;
            return fail(BP_ERR_HIP, "hipEventCreate failed");
        }
    }
    // a failed lookup disables bp_backproject_noalloc only
    if (scratch_address(device, &c->scratch) != BP_SUCCESS) {
        c->scratch = nullptr;
        c->scratch_err = g_msg;
        g_msg.clear();
    }
    *ctx = c;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

}  // extern "C"

namespace {

void free_static(StaticData& s)
{
    DevBuf* bufs[] = {&s.table, &s.table2, &s.poly, &s.poly2, &s.in_pos, &s.in_vel, &s.out_pos,
                      &s.out_vel, &s.in_dop, &s.out_dop, &s.dem, &s.pulse_pv, &s.sinc};
    for (DevBuf* b : bufs)
        if (b->ptr && !b->fixed) (void)hipFree(b->ptr);
}

void free_scratch(Scratch& s)
{
    DevBuf* bufs[] = {&s.geom, &s.split, &s.flags, &s.desc};
    for (DevBuf* b : bufs)
        if (b->ptr && !b->fixed) (void)hipFree(b->ptr);
}

void free_host_path(HostPath& h)
{
    DevBuf* bufs[] = {&h.win[0], &h.win[1], &h.band_desc, &h.band_win};
    for (DevBuf* b : bufs)
        if (b->ptr && !b->fixed) (void)hipFree(b->ptr);
    for (auto* p : h.stage)
        if (p) (void)hipHostFree(p);
    if (h.h_win) (void)hipHostFree(h.h_win);
    hipEvent_t evs[] = {h.stage_ev[0], h.stage_ev[1], h.win_ev, h.copied_ev, h.sum_ev[0], h.sum_ev[1]};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : h.timing)
        if (e) (void)hipEventDestroy(e);
    if (h.copy) (void)hipStreamDestroy(h.copy);
}

}  // namespace

extern "C" int bp_context_destroy(bp_context* c)
try {
    if (c == nullptr) return BP_SUCCESS;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    free_static(c->st);
    free_scratch(c->sc);
    free_host_path(c->hp);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete c;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

namespace {

constexpr int kFlagWords = 16;  // flag words in front of the per-grid flags

// Arguments shared by every output block of a call (the raw set and its
// geometry, the output orbit/Doppler, DEM, kernel and parameters).
struct CallArgs {
    const bp_orbit* out_orbit;
    const bp_lut2d* out_doppler;
    const bp_radar_geometry* in_geometry;
    const bp_dem* dem;
    double fc, ds;
    const bp_kernel* kernel;
    int tropo;
    const bp_rdr2geo_params* r2g;
    const bp_geo2rdr_params* g2r;
    const void* in_dev;
    int precision;
};

// Validated launch options.
struct Opts {
    int tile_rows, chunk, order, blocks_per_cu, lane_rows;
    bool staged;
};

// Argument rules of focus/Backproject.cpp:77-92, the pybind binding
// (focus/Backproject.cpp:97-135) and DryTroposphereModel.cpp:20-33.
// Returns the kernel width on success through *width.
int check_call(const CallArgs& a, int* width)
{
    if (!a.out_orbit || !a.out_doppler || !a.in_geometry || !a.dem || !a.kernel || !a.r2g || !a.g2r)
        return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (a.tropo != BP_TROPO_NODELAY && a.tropo != BP_TROPO_TSX)
        return fail(BP_ERR_INVALID_ARGUMENT, "unexpected dry troposphere model");
    if (a.precision != BP_PRECISION_PLAIN && a.precision != BP_PRECISION_SPLIT)
        return fail(BP_ERR_INVALID_ARGUMENT, "unknown precision mode");
    const bp_grid& gi = a.in_geometry->grid;
    if (gi.length < 0 || gi.width < 0) return fail(BP_ERR_INVALID_ARGUMENT, "negative grid size");
    if (gi.length > INT32_MAX || gi.width > INT32_MAX)
        return fail(BP_ERR_INVALID_ARGUMENT, "grid size exceeds int range");
    if (!(gi.prf > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "PRF must be positive");
    if (!(a.fc > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "center frequency must be positive");
    if (!(a.ds > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "azimuth resolution must be positive");
    const bp_kernel* k = a.kernel;
    if (k->n < 2 || k->table == nullptr || !(k->halfwidth > 0))
        return fail(BP_ERR_INVALID_ARGUMENT, "kernel table needs at least two samples");
    *width = (int)std::ceil(2.0 * k->halfwidth);
    if (*width < bp::kMinKernelWidth || *width > bp::kMaxKernelWidth)
        return fail(BP_ERR_INVALID_ARGUMENT,
                    "kernel width must be in [" + std::to_string(bp::kMinKernelWidth) + ", " +
                        std::to_string(bp::kMaxKernelWidth) + "] taps (got " +
                        std::to_string(*width) + ")");
    if (k->n > (1 << 20)) return fail(BP_ERR_INVALID_ARGUMENT, "kernel table exceeds 2^20 samples");
    const bp_dem* dem = a.dem;
    if (dem->have_raster) {
        bp::Proj pj;
        if (!bp::proj_init(dem->epsg, &pj))
            return fail(BP_ERR_RUNTIME, "Unknown EPSG code (in factory): " + std::to_string(dem->epsg));
        if (!dem_method_ok(dem->method))
            return fail(BP_ERR_INVALID_ARGUMENT, "unsupported DEM interpolation method " + std::to_string(dem->method));
        if (dem->width < 4 || dem->length < 4 || dem->data == nullptr)
            return fail(BP_ERR_INVALID_ARGUMENT, "DEM raster must be at least 4x4");
    }
    int rc;
    if ((rc = check_orbit(*a.out_orbit, "output geometry")) != BP_SUCCESS) return rc;
    if ((rc = check_orbit(a.in_geometry->orbit, "input geometry")) != BP_SUCCESS) return rc;
    if ((rc = check_lut(*a.out_doppler, "output geometry")) != BP_SUCCESS) return rc;
    if ((rc = check_lut(a.in_geometry->doppler, "input geometry")) != BP_SUCCESS) return rc;
    // Pulse times must lie in the input orbit: Orbit::interpolate with the
    // default border mode throws OutOfRange (Backproject.cpp:101-106).
    const int64_t na = gi.length;
    if (na > 0) {
        const bp_orbit& io = a.in_geometry->orbit;
        const double ostart = io.t0, oend = io.t0 + (io.n - 1) * io.dt;
        const double in_dt = 1.0 / gi.prf;
        const double tf = gi.sensing_start, tl = gi.sensing_start + (na - 1) * in_dt;
        if (tf < ostart || tf > oend || tl < ostart || tl > oend)
            return fail(BP_ERR_OUT_OF_RANGE, orbit_err_string(BP_ORBIT_INTERP_DOMAIN_ERROR));
    }
    return BP_SUCCESS;
}

int check_options(const bp_options* options, Opts* o)
{
    bp_options opt{};
    if (options) opt = *options;
    o->tile_rows = opt.tile_rows == 0 ? 1 : opt.tile_rows;
    if (o->tile_rows != 1 && o->tile_rows != 2 && o->tile_rows != 4 && o->tile_rows != 8)
        return fail(BP_ERR_INVALID_ARGUMENT, "tile_rows must be 1, 2, 4 or 8");
    o->chunk = opt.chunk == 0 ? 32 : opt.chunk;
    // the split sum steps the phase by its first two differences between folds:
    // the third-difference term it leaves out grows as chunk^3
    if (o->chunk < 1 || o->chunk > BP_MAX_CHUNK)
        return fail(BP_ERR_INVALID_ARGUMENT, "chunk must be in [1, 64]");
    if (opt.fit_degree != 0 && opt.fit_degree != bp::kFitDegree)
        return fail(BP_ERR_INVALID_ARGUMENT, "fit_degree must be 4");
    if (opt.order < 0 || opt.order > 2) return fail(BP_ERR_INVALID_ARGUMENT, "order must be 0, 1 or 2");
    o->order = opt.order == 1 ? 0 : 1;
    if (opt.blocks_per_cu < 0 || opt.blocks_per_cu > 16)
        return fail(BP_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in [0, 16]");
    o->blocks_per_cu = opt.blocks_per_cu;
    if (opt.staged != 0 && opt.staged != 1) return fail(BP_ERR_INVALID_ARGUMENT, "staged must be 0 or 1");
    o->staged = opt.staged == 1;
    if (opt.lane_rows < 0 || opt.lane_rows > 2) return fail(BP_ERR_INVALID_ARGUMENT, "lane_rows must be 0, 1 or 2");
    o->lane_rows = opt.lane_rows;  // 0: resolved by check_lane_rows
    return BP_SUCCESS;
}

// Two lines per lane need the split kernel with gathered taps; lane_rows 0
// takes them where offered for kernel widths up to 9 (wider two-line kernels
// exceed the 128 registers of four waves per SIMD and spill, so one line per
// lane is their default).
int check_lane_rows(Opts& o, int precision, int width)
{
    const bool offered = precision == BP_PRECISION_SPLIT && !o.staged;
    if (o.staged && (precision != BP_PRECISION_SPLIT || (width != 8 && width != 9)))
        return fail(BP_ERR_INVALID_ARGUMENT, "staged 1 needs split precision and kernel width 8 or 9");
    if (o.lane_rows == 0) o.lane_rows = offered && width <= 9 ? 2 : 1;
    if (o.lane_rows == 2 && !offered)
        return fail(BP_ERR_INVALID_ARGUMENT,
                    "lane_rows 2 needs split precision and staged 0");
    return BP_SUCCESS;
}

int check_blocks(bp_out_block* blocks, int n_blocks, int64_t* npix)
{
    if (n_blocks < 0 || (n_blocks > 0 && blocks == nullptr))
        return fail(BP_ERR_INVALID_ARGUMENT, "invalid output block list");
    *npix = 0;
    for (int b = 0; b < n_blocks; ++b) {
        blocks[b].failed = 0;
        const int rc = bp::check_out_block(blocks[b]);
        if (rc != BP_SUCCESS) return rc;
        *npix += blocks[b].grid.length * blocks[b].grid.width;
    }
    return BP_SUCCESS;
}

// Header and descriptors of the non-empty blocks, written to `dst` (room for
// sizeof(BatchHeader) + max_grids descriptors); block_of_grid maps grids back.
int plan_batch(const bp_out_block* blocks, int n_blocks, const Opts& o, int32_t* d_failed,
               unsigned char* dst, int max_grids, bp::BatchHeader* hdr,
               std::vector<int>& block_of_grid)
{
    block_of_grid.clear();
    auto* descs = reinterpret_cast<bp::OutDesc*>(dst + sizeof(bp::BatchHeader));
    int64_t n_seq = 0, n_rec = 0;
    int g = 0;
    for (int b = 0; b < n_blocks; ++b) {
        const bp_grid& go = blocks[b].grid;
        if (go.length * go.width == 0) continue;
        if (g >= max_grids) return fail(BP_ERR_INVALID_ARGUMENT, "too many output blocks");
        bp::OutDesc d;
        std::memset(&d, 0, sizeof(d));
        d.map = make_map(go.length, go.width, o.tile_rows, o.order, o.lane_rows);
        d.t0 = go.sensing_start;
        d.dt = 1.0 / go.prf;
        d.r0 = go.starting_range;
        d.dr = go.range_pixel_spacing;
        d.side = go.look_side;
        d.seq_base = n_seq;
        d.rec_base = n_rec;
        d.out = static_cast<float2*>(blocks[b].out_dev);
        d.height = blocks[b].height_dev;
        d.failed = d_failed + g;
        n_seq += d.map.n_tiles;
        n_rec += go.length * go.width;
        std::memcpy(descs + g, &d, sizeof(d));
        block_of_grid.push_back(b);
        ++g;
    }
    std::memset(hdr, 0, sizeof(*hdr));
    hdr->n_seq = n_seq;
    hdr->n_groups = (n_seq + 3) / 4;
    hdr->n_geom_groups = (n_seq * o.lane_rows + 3) / 4;
    hdr->n_grids = g;
    hdr->lane_rows = o.lane_rows;
    if (hdr->n_geom_groups > INT32_MAX / 2) return fail(BP_ERR_INVALID_ARGUMENT, "too many output pixels");
    std::memcpy(dst, hdr, sizeof(*hdr));
    return BP_SUCCESS;
}

// Upload the orbits, Doppler tables, DEM and kernel tables of a call.
int upload_static(StaticData& st, const CallArgs& a, int width, hipStream_t s)
{
    int rc;
    const bp_orbit& oo = *a.out_orbit;
    const bp_orbit& io = a.in_geometry->orbit;
    const bp_lut2d& od = *a.out_doppler;
    const bp_lut2d& id = a.in_geometry->doppler;
    if ((rc = upload(st.out_pos, oo.pos, sizeof(double) * 3 * oo.n, s))) return rc;
    if ((rc = upload(st.out_vel, oo.vel, sizeof(double) * 3 * oo.n, s))) return rc;
    if ((rc = upload(st.in_pos, io.pos, sizeof(double) * 3 * io.n, s))) return rc;
    if ((rc = upload(st.in_vel, io.vel, sizeof(double) * 3 * io.n, s))) return rc;
    if ((rc = upload(st.out_dop, od.have_data ? od.data : nullptr,
                     od.have_data ? sizeof(double) * od.width * od.length : 0, s)))
        return rc;
    if ((rc = upload(st.in_dop, id.have_data ? id.data : nullptr,
                     id.have_data ? sizeof(double) * id.width * id.length : 0, s)))
        return rc;
    if ((rc = upload(st.dem, a.dem->have_raster ? a.dem->data : nullptr,
                     a.dem->have_raster ? sizeof(float) * a.dem->width * a.dem->length : 0, s)))
        return rc;
    if ((rc = upload(st.table, a.kernel->table, sizeof(float) * a.kernel->n, s))) return rc;
    const bool sinc = a.dem->have_raster && a.dem->method == BP_DEM_SINC;
    if ((rc = upload(st.sinc, sinc ? sinc_kernel_table().data() : nullptr,
                     sinc ? sizeof(float) * sinc_kernel_table().size() : 0, s)))
        return rc;
    if (a.precision == BP_PRECISION_SPLIT) {
        build_polyphase(*a.kernel, width, st.poly_host);
        if ((rc = upload(st.poly, st.poly_host.data(), sizeof(float) * st.poly_host.size(), s)))
            return rc;
        build_pair_table(*a.kernel, width, st.poly2_host);
        if ((rc = upload(st.poly2, st.poly2_host.data(), sizeof(float) * st.poly2_host.size(), s)))
            return rc;
    } else {
        build_plain_table(*a.kernel, st.table2_host);
        if ((rc = upload(st.table2, st.table2_host.data(),
                         sizeof(double) * st.table2_host.size(), s)))
            return rc;
    }
    return BP_SUCCESS;
}

// Platform position/velocity at every pulse (plain precision).
int pulse_states(StaticData& st, const CallArgs& a, int32_t* d_flags, hipStream_t s)
{
    if (a.precision != BP_PRECISION_PLAIN) return BP_SUCCESS;
    const bp_grid& gi = a.in_geometry->grid;
    int rc;
    if ((rc = ensure(st.pulse_pv, sizeof(double) * bp::kPulseRecDoubles * std::max<int64_t>(gi.length, 1))))
        return rc;
    if (gi.length > 0)
        BP_HIP_CHECK(bp::launch_pulse_orbit(dev_orbit(a.in_geometry->orbit, st.in_pos, st.in_vel),
                                            gi.sensing_start, 1.0 / gi.prf, (int)gi.length,
                                            static_cast<double*>(st.pulse_pv.ptr), d_flags, s));
    return BP_SUCCESS;
}

// Everything the launches of one step need, with device pointers.
struct Step {
    bp::GeomArgs ga;
    bp::SumArgs sa;
    int width, precision;
    bool staged, pair;
    int geom_grid, fit_grid, sum_grid;
    int32_t* flags;
    int flag_words;
    hipEvent_t ev[3];  // null: no timing
};

void build_step(Step& sp, const StaticData& st, const Scratch& sc, const CallArgs& a, int width,
                const Opts& o, int n_cu)
{
    const bp_grid& gi = a.in_geometry->grid;
    const double swst = 2. * gi.starting_range / bp::kC;
    const double dtau = 2. * gi.range_pixel_spacing / bp::kC;
    const auto* batch = static_cast<const bp::BatchHeader*>(sc.desc.ptr);
    int32_t* d_flags = static_cast<int32_t*>(sc.flags.ptr);

    bp::Dem ddem = host_dem(*a.dem);
    ddem.data = static_cast<const float*>(st.dem.ptr);
    ddem.sinc = static_cast<const float*>(st.sinc.ptr);

    bp::GeomArgs& ga = sp.ga;
    std::memset(&ga, 0, sizeof(ga));
    ga.batch = batch;
    ga.in_side = gi.look_side;
    ga.kwidth = width;
    ga.nr = (int32_t)gi.width;
    ga.in_t0 = gi.sensing_start;
    ga.in_dt = 1.0 / gi.prf;
    ga.na = (int32_t)gi.length;
    ga.tropo = a.tropo;
    ga.swst = swst;
    ga.dtau = dtau;
    ga.wvl = bp::kC / a.fc;
    ga.fc = a.fc;
    ga.ds = a.ds;
    ga.out_orbit = dev_orbit(*a.out_orbit, st.out_pos, st.out_vel);
    ga.in_orbit = dev_orbit(a.in_geometry->orbit, st.in_pos, st.in_vel);
    ga.out_dop = dev_lut(*a.out_doppler, st.out_dop);
    ga.in_dop = dev_lut(a.in_geometry->doppler, st.in_dop);
    ga.dem = ddem;
    ga.r2g = {a.r2g->tol_height, a.r2g->look_min, a.r2g->look_max};
    ga.g2r.tol_aztime = a.g2r->tol_aztime;
    ga.g2r.have_time_start = a.g2r->have_time_start;
    ga.g2r.have_time_end = a.g2r->have_time_end;
    ga.g2r.time_start = a.g2r->time_start;
    ga.g2r.time_end = a.g2r->time_end;
    ga.want_split = a.precision == BP_PRECISION_SPLIT;
    ga.geom = static_cast<bp::PixGeom*>(sc.geom.ptr);
    ga.split = static_cast<bp::PixSplit*>(sc.split.ptr);
    ga.flags = d_flags;
    ga.updates = reinterpret_cast<unsigned long long*>(d_flags + 8);

    bp::SumArgs& sa = sp.sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.batch = batch;
    sa.nr = gi.width;
    sa.swst = swst;
    sa.dtau = dtau;
    sa.fc = a.fc;
    sa.inv_k = 1.0 / (a.fc * dtau);
    sa.zscale = (float)(-(double)bp::kPolyphaseRows * sa.inv_k);
    sa.halfwidth = a.kernel->halfwidth;
    sa.rdx = (float)(1.0 / (a.kernel->halfwidth / (a.kernel->n - 1.0)));
    sa.ntab = a.kernel->n;
    sa.chunk = o.chunk;
    {
        const double n = (double)o.chunk;
        sa.fold_c[0] = n;
        sa.fold_c[1] = n * (n - 1.) / 2.;
        sa.fold_c[2] = n * (n - 1.) * (n - 2.) / 6.;
        sa.fold_c[3] = n * (n - 1.) * (n - 2.) * (n - 3.) / 24.;
    }
    sa.rc = static_cast<const float2*>(a.in_dev);
    sa.pulse_pv = static_cast<const double*>(st.pulse_pv.ptr);
    sa.geom = static_cast<const bp::PixGeom*>(sc.geom.ptr);
    sa.split = static_cast<const bp::PixSplit*>(sc.split.ptr);
    sa.table = static_cast<const float*>(st.table.ptr);
    sa.poly = static_cast<const float*>(st.poly.ptr);
    sa.poly2 = static_cast<const float*>(st.poly2.ptr);
    sa.zscale2 = (float)(-(double)(1 << bp::pair_q_log2(width)) * sa.inv_k);
    sa.inv_dtau = 1.0 / dtau;
    sa.exact_width = 2.0 * a.kernel->halfwidth == (double)width;
    sa.tab2 = static_cast<const double2*>(st.table2.ptr);

    sp.width = width;
    sp.precision = a.precision;
    sp.staged = o.staged;
    sp.pair = o.lane_rows == 2;
    const int per_cu = o.blocks_per_cu > 0 ? o.blocks_per_cu
                                           : bp::sum_blocks_per_cu(a.precision, width, o.staged, sp.pair);
    sp.sum_grid = (n_cu * per_cu + 7) / 8 * 8;  // persistent split grid
    sp.geom_grid = 0;
    sp.fit_grid = 0;
    sp.flags = d_flags;
    sp.flag_words = 0;
    sp.ev[0] = sp.ev[1] = sp.ev[2] = nullptr;
}

// The coherent sum of a step over the descriptors at `batch`, reading pulse k
// at rc + k * nr.
int enqueue_sum(const Step& sp, const bp::BatchHeader* batch, const float2* rc, int plain_grid,
                hipStream_t s)
{
    bp::SumArgs sa = sp.sa;
    sa.batch = batch;
    sa.rc = rc;
    if (sp.precision == BP_PRECISION_PLAIN)
        BP_HIP_CHECK(bp::launch_sum_plain(sa, sp.width, plain_grid, s));
    else
        BP_HIP_CHECK(bp::launch_sum_split(sa, sp.width, sp.sum_grid, sp.staged, sp.pair, s));
    return BP_SUCCESS;
}

// Enqueue the launches of one step: clear the flags, per-pixel geometry,
// coherent sum (with optional event records around the two passes).
int enqueue_step(const Step& sp, int plain_grid, hipStream_t s)
{
    BP_HIP_CHECK(bp::launch_clear(sp.flags, sp.flag_words, s));
    if (sp.ev[0]) BP_HIP_CHECK(hipEventRecord(sp.ev[0], s));
    BP_HIP_CHECK(bp::launch_geometry(sp.ga, sp.geom_grid, s));
    if (sp.precision == BP_PRECISION_SPLIT) BP_HIP_CHECK(bp::launch_fit(sp.ga, sp.fit_grid, s));
    if (sp.ev[1]) BP_HIP_CHECK(hipEventRecord(sp.ev[1], s));
    const int rc = enqueue_sum(sp, sp.sa.batch, sp.sa.rc, plain_grid, s);
    if (rc != BP_SUCCESS) return rc;
    if (sp.ev[2]) BP_HIP_CHECK(hipEventRecord(sp.ev[2], s));
    return BP_SUCCESS;
}

// Read the flags of a finished step into the caller's outputs.
int finish_step(const int32_t* hflags, int n_grids, const std::vector<int>& block_of_grid,
                bp_out_block* blocks, const hipEvent_t* ev, int* any_failed, bp_stats* stats)
{
    unsigned long long hupd = 0;
    std::memcpy(&hupd, hflags + 8, sizeof(hupd));
    for (int g = 0; g < n_grids; ++g) blocks[block_of_grid[g]].failed = hflags[kFlagWords + g] != 0;
    if (stats) {
        float ms = 0.f;
        if (ev[1]) {
            BP_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
            stats->geometry_ms = ms;
            BP_HIP_CHECK(hipEventElapsedTime(&ms, ev[1], ev[2]));
            stats->sum_ms = ms;
        }
        if (ev[0] && ev[2]) {  // calls without events report no times
            BP_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[2]));
            stats->step_ms = ms;
        }
        stats->updates = (int64_t)hupd;
        stats->any_failed = hflags[0] != 0;
    }
    if (hflags[1]) return fail(BP_ERR_OUT_OF_RANGE, orbit_err_string(BP_ORBIT_INTERP_DOMAIN_ERROR));
    if (any_failed) *any_failed = hflags[0] != 0;
    return hflags[0] ? BP_FAILED_TO_CONVERGE : BP_SUCCESS;
}

// Shared implementation of bp_backproject and bp_backproject_blocks: every
// output block shares the output orbit and Doppler, the input geometry, DEM,
// kernel and parameters; the pixels of all blocks are focused by one geometry
// launch and one coherent-sum launch.
// `carve`, when given, places the call's buffers in a fixed region instead of
// the context's grow-only arena (bp_backproject_noalloc).
struct Carve;
int carve_call(Carve& cv, const CallArgs& a, int width, int n_blocks, int64_t npix,
               StaticData& st, Scratch& sc);

int run_blocks(bp_context* c, const CallArgs& a, bp_out_block* blocks, int n_blocks,
               const bp_options* options, void* stream, int* any_failed, bp_stats* stats,
               Carve* carve = nullptr)
{
    if (any_failed) *any_failed = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!c) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    int width = 0, rc;
    if ((rc = check_call(a, &width)) != BP_SUCCESS) return rc;
    int64_t npix = 0;
    if ((rc = check_blocks(blocks, n_blocks, &npix)) != BP_SUCCESS) return rc;
    Opts o;
    if ((rc = check_options(options, &o)) != BP_SUCCESS) return rc;
    if ((rc = check_lane_rows(o, a.precision, width)) != BP_SUCCESS) return rc;
    const bp_grid& gi = a.in_geometry->grid;
    if (npix > 0 && gi.length * gi.width > 0 && a.in_dev == nullptr)
        return fail(BP_ERR_INVALID_ARGUMENT, "in is null");
    StaticData fixed_st;
    Scratch fixed_sc;
    if (carve && (rc = carve_call(*carve, a, width, n_blocks, npix, fixed_st, fixed_sc)) != BP_SUCCESS)
        return rc;
    if (npix == 0) return BP_SUCCESS;
    StaticData& st = carve ? fixed_st : c->st;
    Scratch& sc = carve ? fixed_sc : c->sc;

    BP_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);

    const int n_flag = kFlagWords + n_blocks;
    if ((rc = ensure(sc.flags, sizeof(int32_t) * n_flag))) return rc;
    c->desc_host.assign(sizeof(bp::BatchHeader) + sizeof(bp::OutDesc) * std::max(n_blocks, 1), 0);
    bp::BatchHeader hdr;
    if ((rc = plan_batch(blocks, n_blocks, o, static_cast<int32_t*>(sc.flags.ptr) + kFlagWords,
                         c->desc_host.data(), n_blocks, &hdr, c->desc_block)) != BP_SUCCESS)
        return rc;
    const int64_t n_rec = npix;

    if ((rc = upload_static(st, a, width, s))) return rc;
    if ((rc = upload(sc.desc, c->desc_host.data(), c->desc_host.size(), s))) return rc;
    if ((rc = ensure(sc.geom, sizeof(bp::PixGeom) * n_rec))) return rc;
    if (a.precision == BP_PRECISION_SPLIT && (rc = ensure(sc.split, sizeof(bp::PixSplit) * n_rec)))
        return rc;
    BP_HIP_CHECK(bp::launch_clear(static_cast<int32_t*>(sc.flags.ptr), n_flag, s));
    if ((rc = pulse_states(st, a, static_cast<int32_t*>(sc.flags.ptr), s))) return rc;

    Step sp;
    build_step(sp, st, sc, a, width, o, c->n_cu);
    sp.geom_grid = (int)hdr.n_geom_groups;
    sp.fit_grid = (int)hdr.n_groups;
    sp.flag_words = 0;  // cleared above, before the pulse states
    if (stats)
        for (int i = 0; i < 3; ++i) sp.ev[i] = c->ev[i];
    if ((rc = enqueue_step(sp, (int)hdr.n_groups, s))) return rc;

    c->flags_host.assign(n_flag, 0);
    BP_HIP_CHECK(hipMemcpyAsync(c->flags_host.data(), sc.flags.ptr, sizeof(int32_t) * n_flag,
                                hipMemcpyDeviceToHost, s));
    BP_HIP_CHECK(hipStreamSynchronize(s));
    return finish_step(c->flags_host.data(), hdr.n_grids, c->desc_block, blocks, c->ev,
                       any_failed, stats);
}

// A bump allocator over a fixed device region: every buffer of a call starts
// on a 256-byte boundary; a call that does not fit fails before any launch.
struct Carve {
    unsigned char* base;
    size_t cap;
    size_t used = 0;
    int take(DevBuf& b, size_t bytes)
    {
        if (bytes == 0) bytes = 16;  // as ensure()
        const size_t off = (used + 255) & ~size_t(255);
        if (off > cap || bytes > cap - off) {
            used = cap + 1;  // mark the call as too large
            return BP_ERR_LENGTH;
        }
        b.ptr = base + off;
        b.bytes = bytes;
        b.fixed = true;
        used = off + bytes;
        return BP_SUCCESS;
    }
};

// Buffers of one bp_backproject_noalloc call, of the sizes upload_static,
// pulse_states and run_blocks ask for (a buffer that came out smaller would
// fail in ensure(), never overrun).
int carve_call(Carve& cv, const CallArgs& a, int width, int n_blocks, int64_t npix,
               StaticData& st, Scratch& sc)
{
    const bp_orbit& oo = *a.out_orbit;
    const bp_orbit& io = a.in_geometry->orbit;
    const bp_lut2d& od = *a.out_doppler;
    const bp_lut2d& id = a.in_geometry->doppler;
    const bool split = a.precision == BP_PRECISION_SPLIT;
    const int64_t na = std::max<int64_t>(a.in_geometry->grid.length, 1);
    const struct {
        DevBuf* b;
        double bytes;  // double: sizes of absurd inputs must not wrap
    } want[] = {
        {&sc.flags, 4.0 * (kFlagWords + n_blocks)},
        {&sc.desc, (double)sizeof(bp::BatchHeader) + (double)sizeof(bp::OutDesc) * std::max(n_blocks, 1)},
        {&st.out_pos, 24.0 * oo.n},
        {&st.out_vel, 24.0 * oo.n},
        {&st.in_pos, 24.0 * io.n},
        {&st.in_vel, 24.0 * io.n},
        {&st.out_dop, od.have_data ? 8.0 * od.width * od.length : 0.0},
        {&st.in_dop, id.have_data ? 8.0 * id.width * id.length : 0.0},
        {&st.dem, a.dem->have_raster ? 4.0 * a.dem->width * a.dem->length : 0.0},
        {&st.sinc, a.dem->have_raster && a.dem->method == BP_DEM_SINC
                       ? 4.0 * bp::kSincLen * bp::kSincSub : 0.0},
        {&st.table, 4.0 * a.kernel->n},
        {&st.poly, split ? 4.0 * (bp::kPolyphaseRows + 1) * bp::poly_row_floats(width) : 0.0},
        {&st.poly2, split ? 4.0 * bp::pair_table_floats(width) : 0.0},
        {&st.table2, split ? 0.0 : 16.0 * a.kernel->n},
        {&st.pulse_pv, split ? 0.0 : 8.0 * bp::kPulseRecDoubles * na},
        {&sc.geom, (double)sizeof(bp::PixGeom) * npix},
        {&sc.split, split ? (double)sizeof(bp::PixSplit) * npix : 0.0},
    };
    double total = 0.0;
    for (const auto& w : want) total += std::ceil(std::max(w.bytes, 16.0) / 256.0) * 256.0;
    if (total > (double)cv.cap)
        return fail(BP_ERR_LENGTH,
                    "bp_backproject_noalloc: the call needs " + std::to_string((long long)total) +
                        " bytes of scratch (output records " +
                        std::to_string((long long)bp_scratch_bytes_per_pixel(a.precision)) +
                        " bytes per pixel, plus orbit, Doppler, DEM and kernel tables); the "
                        "library's fixed scratch holds " + std::to_string(cv.cap) + " bytes");
    for (const auto& w : want)
        if (cv.take(*w.b, (size_t)w.bytes) != BP_SUCCESS)
            return fail(BP_ERR_LENGTH, "bp_backproject_noalloc: scratch layout overflow");
    return BP_SUCCESS;
}

// ------------------------------------------------------------ host input
//
// Range-compressed pulses in host memory (an array or a memory map of a
// raster file): the geometry pass runs first, the union [kmin, kmax) of the
// pixels' pulse windows comes back per band of tile rows, and only those
// pulses are copied to the device, in chunks of at most `batch` pulses through
// two pinned staging buffers on a copy stream (the host reads chunk j + 1
// while chunk j is copied).  When the window of the whole block exceeds the
// device budget, the block is focused in bands of output lines with smaller
// windows; a band starts on a tile row, so its wavefront tiles, records and
// per-pixel arithmetic are those of the whole block and every result equals
// the resident call's.  Bands alternate between two window buffers when two
// fit the budget, so the copy of band b + 1 overlaps the sum of band b.

constexpr size_t kStageBytes = size_t(64) << 20;   // per pinned staging buffer

int host_path_init(HostPath& h)
{
    if (h.copy) return BP_SUCCESS;
    BP_HIP_CHECK(hipStreamCreateWithFlags(&h.copy, hipStreamNonBlocking));  // default priority
    hipEvent_t* evs[] = {&h.stage_ev[0], &h.stage_ev[1], &h.win_ev, &h.copied_ev, &h.sum_ev[0],
                         &h.sum_ev[1]};
    for (hipEvent_t* e : evs) BP_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    return BP_SUCCESS;
}

int ensure_pinned(unsigned char*& p, size_t& have, size_t bytes)
{
    if (have >= bytes && p) return BP_SUCCESS;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    have = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&p), bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        return fail(BP_ERR_OUT_OF_MEMORY, "hipHostMalloc of pulse staging failed");
    }
    have = bytes;
    return BP_SUCCESS;
}

// memcpy of `bytes` from src to dst, split over up to 8 host threads for
// large copies (reading a memory-mapped raster is the slow side).
void host_copy(unsigned char* dst, const unsigned char* src, size_t bytes)
{
    const size_t kPerThread = size_t(16) << 20;
    unsigned nt = (unsigned)std::min<size_t>(8, (bytes + kPerThread - 1) / kPerThread);
    nt = std::max(1u, std::min(nt, (unsigned)bp::usable_cpus()));
    if (nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) {
        const size_t o = std::min(bytes, t * part), n = std::min(bytes - o, part);
        try {
            th.emplace_back([=] { std::memcpy(dst + o, src + o, n); });
        } catch (...) {
            std::memcpy(dst + o, src + o, n);  // no thread: copy here
        }
    }
    std::memcpy(dst, src, std::min(bytes, part));
    for (auto& x : th) x.join();
}

struct Band {
    int64_t r0, r1;   // output lines
    int32_t k0, k1;   // pulses [k0, k1) (k0 >= k1: none)
};

int run_host(bp_context* c, const CallArgs& a, bp_out_block* blk, int64_t batch,
             int64_t max_window_bytes, const bp_options* options, void* stream, int* any_failed,
             bp_stats* stats, bp_host_stats* hstats)
{
    if (any_failed) *any_failed = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (hstats) std::memset(hstats, 0, sizeof(*hstats));
    if (!c) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    int width = 0, rc;
    if ((rc = check_call(a, &width)) != BP_SUCCESS) return rc;
    if (batch < 1) return fail(BP_ERR_DOMAIN, "batch size must be > 0");
    if (max_window_bytes < 0) return fail(BP_ERR_INVALID_ARGUMENT, "max_window_bytes must be >= 0");
    int64_t npix = 0;
    if ((rc = check_blocks(blk, 1, &npix)) != BP_SUCCESS) return rc;
    Opts o;
    if ((rc = check_options(options, &o)) != BP_SUCCESS) return rc;
    if ((rc = check_lane_rows(o, a.precision, width)) != BP_SUCCESS) return rc;
    const bp_grid& gi = a.in_geometry->grid;
    if (npix > 0 && gi.length * gi.width > 0 && a.in_dev == nullptr)
        return fail(BP_ERR_INVALID_ARGUMENT, "in is null");
    if (npix == 0) return BP_SUCCESS;

    BP_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    HostPath& hp = c->hp;
    if ((rc = host_path_init(hp))) return rc;
    // a call that fails after enqueueing work drains both streams before it
    // returns, so that the next call never refills buffers still being read
    struct Drain {
        hipStream_t a, b;
        bool armed = true;
        ~Drain()
        {
            if (armed) {
                (void)hipStreamSynchronize(a);
                (void)hipStreamSynchronize(b);
            }
        }
    } drain{s, hp.copy};

    const int n_flag = kFlagWords + 1;
    if ((rc = ensure(c->sc.flags, sizeof(int32_t) * n_flag))) return rc;
    c->desc_host.assign(sizeof(bp::BatchHeader) + sizeof(bp::OutDesc), 0);
    bp::BatchHeader hdr;
    if ((rc = plan_batch(blk, 1, o, static_cast<int32_t*>(c->sc.flags.ptr) + kFlagWords,
                         c->desc_host.data(), 1, &hdr, c->desc_block)) != BP_SUCCESS)
        return rc;
    bp::OutDesc whole;
    std::memcpy(&whole, c->desc_host.data() + sizeof(bp::BatchHeader), sizeof(whole));

    if ((rc = upload_static(c->st, a, width, s))) return rc;
    if ((rc = upload(c->sc.desc, c->desc_host.data(), c->desc_host.size(), s))) return rc;
    if ((rc = ensure(c->sc.geom, sizeof(bp::PixGeom) * npix))) return rc;
    if (a.precision == BP_PRECISION_SPLIT && (rc = ensure(c->sc.split, sizeof(bp::PixSplit) * npix)))
        return rc;
    BP_HIP_CHECK(bp::launch_clear(static_cast<int32_t*>(c->sc.flags.ptr), n_flag, s));
    if ((rc = pulse_states(c->st, a, static_cast<int32_t*>(c->sc.flags.ptr), s))) return rc;

    Step sp;
    build_step(sp, c->st, c->sc, a, width, o, c->n_cu);
    sp.geom_grid = (int)hdr.n_geom_groups;
    sp.fit_grid = (int)hdr.n_groups;

    // geometry, the band windows (read back while the fit runs), the fit
    const int band_rows = whole.map.tile_rows * whole.map.lane_rows;
    const int64_t length = whole.map.length, owidth = whole.map.width;
    const int64_t n_tr = (length + band_rows - 1) / band_rows;
    if (stats) BP_HIP_CHECK(hipEventRecord(c->ev[0], s));
    BP_HIP_CHECK(bp::launch_geometry(sp.ga, sp.geom_grid, s));
    if ((rc = ensure(hp.band_win, sizeof(int32_t) * 2 * n_tr))) return rc;
    BP_HIP_CHECK(bp::launch_band_window(static_cast<const bp::PixGeom*>(c->sc.geom.ptr), length,
                                        owidth, band_rows, n_tr,
                                        static_cast<int32_t*>(hp.band_win.ptr), s));
    if (hp.h_win_words < (size_t)(2 * n_tr)) {
        unsigned char* p = reinterpret_cast<unsigned char*>(hp.h_win);
        size_t have = hp.h_win_words * sizeof(int32_t);
        if ((rc = ensure_pinned(p, have, sizeof(int32_t) * 2 * n_tr))) {
            hp.h_win = nullptr;
            hp.h_win_words = 0;
            return rc;
        }
        hp.h_win = reinterpret_cast<int32_t*>(p);
        hp.h_win_words = have / sizeof(int32_t);
    }
    BP_HIP_CHECK(hipMemcpyAsync(hp.h_win, hp.band_win.ptr, sizeof(int32_t) * 2 * n_tr,
                                hipMemcpyDeviceToHost, s));
    BP_HIP_CHECK(hipEventRecord(hp.win_ev, s));
    if (a.precision == BP_PRECISION_SPLIT) BP_HIP_CHECK(bp::launch_fit(sp.ga, sp.fit_grid, s));
    if (stats) BP_HIP_CHECK(hipEventRecord(c->ev[1], s));
    BP_HIP_CHECK(hipEventSynchronize(hp.win_ev));

    // bands of tile rows whose union window fits the budget
    const size_t line_bytes = sizeof(float) * 2 * (size_t)gi.width;
    size_t budget = (size_t)max_window_bytes;
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        BP_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
        budget = free_b / 2;
    }
    std::vector<Band> bands;
    {
        Band cur{0, 0, INT32_MAX, INT32_MIN};
        for (int64_t t = 0; t < n_tr; ++t) {
            const int32_t lo = hp.h_win[2 * t], hi = hp.h_win[2 * t + 1];
            if (lo < hi && (size_t)(hi - lo) * line_bytes > budget)
                return fail(BP_ERR_LENGTH, "the pulse window of " + std::to_string(band_rows) +
                                               " output lines (" + std::to_string(hi - lo) +
                                               " pulses) exceeds max_window_bytes");
            const int32_t nlo = lo < hi ? std::min(cur.k0, lo) : cur.k0;
            const int32_t nhi = lo < hi ? std::max(cur.k1, hi) : cur.k1;
            if (cur.r1 > cur.r0 && nlo < nhi && (size_t)(nhi - nlo) * line_bytes > budget) {
                bands.push_back(cur);
                cur = Band{cur.r1, cur.r1, INT32_MAX, INT32_MIN};
                if (lo < hi) {
                    cur.k0 = lo;
                    cur.k1 = hi;
                }
            } else {
                cur.k0 = nlo;
                cur.k1 = nhi;
            }
            cur.r1 = std::min(length, (t + 1) * (int64_t)band_rows);
        }
        bands.push_back(cur);
    }
    size_t max_band = 16;
    for (const Band& b : bands)
        if (b.k0 < b.k1) max_band = std::max(max_band, (size_t)(b.k1 - b.k0) * line_bytes);
    const int nbuf = bands.size() > 1 && 2 * max_band <= budget ? 2 : 1;
    for (int w = 0; w < nbuf; ++w)
        if ((rc = ensure(hp.win[w], max_band))) return rc;
    const int64_t chunk_lines = std::max<int64_t>(
        1, std::min<int64_t>(batch, (int64_t)(kStageBytes / std::max<size_t>(line_bytes, 1))));
    {
        size_t have = hp.stage_bytes, have1 = hp.stage_bytes;
        const size_t want = (size_t)chunk_lines * line_bytes;
        if ((rc = ensure_pinned(hp.stage[0], have, want))) return rc;
        if ((rc = ensure_pinned(hp.stage[1], have1, want))) return rc;
        hp.stage_bytes = std::min(have, have1);
    }

    // descriptors of every band: the whole block's, cut to the band's lines
    const size_t stride = (sizeof(bp::BatchHeader) + sizeof(bp::OutDesc) + 255) & ~size_t(255);
    std::vector<unsigned char> bdesc(stride * bands.size(), 0);
    std::vector<int> band_groups(bands.size());
    for (size_t b = 0; b < bands.size(); ++b) {
        bp::OutDesc d = whole;
        const int64_t r0 = bands[b].r0;
        d.map = make_map(bands[b].r1 - r0, owidth, o.tile_rows, o.order, o.lane_rows);
        d.t0 = whole.t0 + r0 * whole.dt;
        d.seq_base = 0;
        d.rec_base = whole.rec_base + r0 * owidth;
        d.out = whole.out + r0 * owidth;
        d.height = whole.height ? whole.height + r0 * owidth : nullptr;
        bp::BatchHeader h;
        std::memset(&h, 0, sizeof(h));
        h.n_seq = d.map.n_tiles;
        h.n_groups = (h.n_seq + 3) / 4;
        h.n_geom_groups = (h.n_seq * o.lane_rows + 3) / 4;
        h.n_grids = 1;
        h.lane_rows = o.lane_rows;
        std::memcpy(bdesc.data() + b * stride, &h, sizeof(h));
        std::memcpy(bdesc.data() + b * stride + sizeof(h), &d, sizeof(d));
        band_groups[b] = (int)h.n_groups;
    }
    if ((rc = upload(hp.band_desc, bdesc.data(), bdesc.size(), s))) return rc;

    const bool timed = stats != nullptr || hstats != nullptr;
    while (timed && hp.timing.size() < 4 * bands.size()) {
        hipEvent_t e;
        BP_HIP_CHECK(hipEventCreate(&e));
        hp.timing.push_back(e);
    }
    // copies on the copy stream (they write only the context's window
    // buffers, which no earlier call still reads: every call ends with both
    // streams drained), sums on the caller's stream
    const unsigned char* src = static_cast<const unsigned char*>(a.in_dev);
    int64_t h2d = 0, pulses = 0, chunk_no = 0;
    bool used[2] = {false, false};
    for (size_t b = 0; b < bands.size(); ++b) {
        const Band& bd = bands[b];
        const int w = (int)(b % nbuf);
        unsigned char* dwin = static_cast<unsigned char*>(hp.win[w].ptr);
        if (used[w]) BP_HIP_CHECK(hipStreamWaitEvent(hp.copy, hp.sum_ev[w], 0));
        if (timed) BP_HIP_CHECK(hipEventRecord(hp.timing[4 * b], hp.copy));
        for (int64_t k = bd.k0; k < bd.k1; k += chunk_lines) {
            const int64_t n = std::min<int64_t>(chunk_lines, bd.k1 - k);
            const int j = (int)(chunk_no++ & 1);
            // staging j is free once its previous copy has finished
            BP_HIP_CHECK(hipEventSynchronize(hp.stage_ev[j]));
            const size_t bytes = (size_t)n * line_bytes;
            host_copy(hp.stage[j], src + (size_t)k * line_bytes, bytes);
            BP_HIP_CHECK(hipMemcpyAsync(dwin + (size_t)(k - bd.k0) * line_bytes, hp.stage[j],
                                        bytes, hipMemcpyHostToDevice, hp.copy));
            BP_HIP_CHECK(hipEventRecord(hp.stage_ev[j], hp.copy));
            h2d += (int64_t)bytes;
        }
        if (bd.k0 < bd.k1) pulses += bd.k1 - bd.k0;
        if (timed) BP_HIP_CHECK(hipEventRecord(hp.timing[4 * b + 1], hp.copy));
        BP_HIP_CHECK(hipEventRecord(hp.copied_ev, hp.copy));
        BP_HIP_CHECK(hipStreamWaitEvent(s, hp.copied_ev, 0));
        // pulse k of the band at dwin + (k - k0) * line_bytes; the sum reads
        // only pulses of its pixels' windows, all inside [k0, k1)
        const int64_t k0 = bd.k0 < bd.k1 ? bd.k0 : 0;
        const float2* rc_base = reinterpret_cast<const float2*>(
            reinterpret_cast<uintptr_t>(dwin) - (uintptr_t)((size_t)k0 * line_bytes));
        const auto* batch_b = reinterpret_cast<const bp::BatchHeader*>(
            static_cast<unsigned char*>(hp.band_desc.ptr) + b * stride);
        if (timed) BP_HIP_CHECK(hipEventRecord(hp.timing[4 * b + 2], s));
        if ((rc = enqueue_sum(sp, batch_b, rc_base, band_groups[b], s))) return rc;
        if (timed) BP_HIP_CHECK(hipEventRecord(hp.timing[4 * b + 3], s));
        BP_HIP_CHECK(hipEventRecord(hp.sum_ev[w], s));
        used[w] = true;
    }
    if (stats) BP_HIP_CHECK(hipEventRecord(c->ev[2], s));

    c->flags_host.assign(n_flag, 0);
    BP_HIP_CHECK(hipMemcpyAsync(c->flags_host.data(), c->sc.flags.ptr, sizeof(int32_t) * n_flag,
                                hipMemcpyDeviceToHost, s));
    BP_HIP_CHECK(hipStreamSynchronize(s));
    BP_HIP_CHECK(hipStreamSynchronize(hp.copy));
    drain.armed = false;
    double sum_ms = 0., copy_ms = 0.;
    if (timed) {
        for (size_t b = 0; b < bands.size(); ++b) {
            float ms = 0.f;
            BP_HIP_CHECK(hipEventElapsedTime(&ms, hp.timing[4 * b], hp.timing[4 * b + 1]));
            copy_ms += ms;
            BP_HIP_CHECK(hipEventElapsedTime(&ms, hp.timing[4 * b + 2], hp.timing[4 * b + 3]));
            sum_ms += ms;
        }
    }
    if (hstats) {
        hstats->h2d_bytes = h2d;
        hstats->window_pulses = pulses;
        hstats->bands = (int32_t)bands.size();
        hstats->copy_ms = copy_ms;
    }
    const hipEvent_t ev[3] = {stats ? c->ev[0] : nullptr, nullptr, stats ? c->ev[2] : nullptr};
    rc = finish_step(c->flags_host.data(), hdr.n_grids, c->desc_block, blk, ev, any_failed, stats);
    if (stats && c->ev[1]) {
        float ms = 0.f;
        BP_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        stats->geometry_ms = ms;
        stats->sum_ms = sum_ms;
    }
    return rc;
}

}  // namespace

extern "C" int bp_backproject_host(bp_context* c, const bp_radar_geometry* og,
                                   const bp_radar_geometry* ig, const bp_dem* dem, double fc,
                                   double ds, const bp_kernel* kernel, int dry_tropo_model,
                                   const bp_rdr2geo_params* r2g, const bp_geo2rdr_params* g2r,
                                   const void* in_host, int64_t batch, int64_t max_window_bytes,
                                   void* out_dev, float* height_dev, int precision,
                                   const bp_options* options, void* stream, int* any_failed,
                                   bp_stats* stats, bp_host_stats* host_stats)
try {
    if (any_failed) *any_failed = 0;
    if (!og) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    bp_out_block blk;
    std::memset(&blk, 0, sizeof(blk));
    blk.grid = og->grid;
    blk.out_dev = out_dev;
    blk.height_dev = height_dev;
    CallArgs a{&og->orbit, &og->doppler, ig, dem, fc, ds, kernel, dry_tropo_model, r2g, g2r,
               in_host, precision};
    return run_host(c, a, &blk, batch, max_window_bytes, options, stream, any_failed, stats,
                    host_stats);
} catch (...) {
    return bp_host_exception();
}

// ------------------------------------------------------------------ scenes

// One raw set kept resident with everything its tile requests share, scratch
// for up to max_pixels output pixels in up to max_blocks blocks, and one HIP
// graph per launch configuration.  The graph of a step copies the batch
// header and descriptors from pinned host memory, clears the flags with a
// kernel, runs the geometry and sum kernels (persistent grids that read the
// batch counts from the device header) between event records, and copies the
// flags back to pinned host memory; a request only rewrites the pinned
// descriptors and replays the graph.
struct bp_scene {
    bp_context* ctx = nullptr;
    CallArgs args{};
    int width = 0;
    int64_t max_pixels = 0;
    int max_blocks = 0;
    StaticData st;
    Scratch sc;
    // copies of the caller's structs that CallArgs points to
    bp_orbit out_orbit{};
    bp_lut2d out_doppler{};
    bp_radar_geometry in_geometry{};
    bp_dem dem{};
    bp_kernel kernel{};
    bp_rdr2geo_params r2g{};
    bp_geo2rdr_params g2r{};
    unsigned char* h_desc = nullptr;   // pinned: header + max_blocks descriptors
    int32_t* h_flags = nullptr;        // pinned: kFlagWords + max_blocks words
    hipStream_t cap = nullptr;         // capture stream (default priority)
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    std::vector<int> block_of_grid;
    struct Graph {
        int chunk, blocks_per_cu, lane_rows;
        bool staged;
        hipGraph_t graph;
        hipGraphExec_t exec;
        // kernel nodes whose grid follows the batch (geometry: geometry
        // groups; fit and plain sum: tile groups), with their captured
        // launch parameters
        hipGraphNode_t node[2];
        hipKernelNodeParams params[2];
        bool geom_groups[2];
        int n_nodes;
    };
    std::vector<Graph> graphs;
};

namespace {

void destroy_scene(bp_scene* sc)
{
    for (auto& g : sc->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    free_static(sc->st);
    free_scratch(sc->sc);
    if (sc->h_desc) (void)hipHostFree(sc->h_desc);
    if (sc->h_flags) (void)hipHostFree(sc->h_flags);
    for (auto& ev : sc->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (sc->cap) (void)hipStreamDestroy(sc->cap);
    delete sc;
}

// Capture the step of one launch configuration.
int capture_graph(bp_scene* sc, const Opts& o, bp_scene::Graph* out)
{
    bp_context* c = sc->ctx;
    Step sp;
    build_step(sp, sc->st, sc->sc, sc->args, sc->width, o, c->n_cu);
    // the geometry, fit and plain-sum grids are set per replay to the batch's
    // group counts (bp_scene_focus); one workgroup at capture
    const int cap_grid = 1;
    sp.geom_grid = cap_grid;
    sp.fit_grid = cap_grid;
    sp.flag_words = kFlagWords + sc->max_blocks;
    // no event records inside the graph: a replay is timed as a whole by
    // events recorded around the launch
    const size_t desc_bytes = sizeof(bp::BatchHeader) + sizeof(bp::OutDesc) * sc->max_blocks;

    hipStream_t s = sc->cap;
    BP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = BP_SUCCESS;
    hipError_t e = hipMemcpyAsync(sc->sc.desc.ptr, sc->h_desc, desc_bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = fail(BP_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
    if (rc == BP_SUCCESS) rc = enqueue_step(sp, cap_grid, s);
    if (rc == BP_SUCCESS) {
        e = hipMemcpyAsync(sc->h_flags, sc->sc.flags.ptr, sizeof(int32_t) * (kFlagWords + sc->max_blocks),
                           hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = fail(BP_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc != BP_SUCCESS) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(BP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    // find the kernel nodes whose grid follows the batch
    out->n_nodes = 0;
    size_t n_nodes = 0;
    e = hipGraphGetNodes(g, nullptr, &n_nodes);
    std::vector<hipGraphNode_t> nodes(n_nodes);
    if (e == hipSuccess && n_nodes) e = hipGraphGetNodes(g, nodes.data(), &n_nodes);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(BP_ERR_HIP, std::string("hipGraphGetNodes: ") + hipGetErrorString(e));
    }
    const bool pair = o.lane_rows == 2;
    const void* geo_fn = bp::kernel_fn(bp::dem_generic(host_dem(sc->dem)) ? 4 : 0, sc->width,
                                       o.staged, pair);
    const void* tile_fn = sc->args.precision == BP_PRECISION_PLAIN
                              ? bp::kernel_fn(1, sc->width, o.staged, pair)
                              : bp::kernel_fn(3, sc->width, o.staged, pair);
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType type;
        if (hipGraphNodeGetType(nd, &type) != hipSuccess || type != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams p{};
        if (hipGraphKernelNodeGetParams(nd, &p) != hipSuccess) continue;
        if ((p.func == geo_fn || p.func == tile_fn) && out->n_nodes < 2) {
            out->node[out->n_nodes] = nd;
            out->params[out->n_nodes] = p;
            out->geom_groups[out->n_nodes] = p.func == geo_fn;
            ++out->n_nodes;
        }
    }
    const int want = 2;
    if (out->n_nodes != want) {
        (void)hipGraphDestroy(g);
        return fail(BP_ERR_HIP, "captured step is missing its kernel nodes");
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(BP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    out->chunk = o.chunk;
    out->blocks_per_cu = o.blocks_per_cu;
    out->lane_rows = o.lane_rows;
    out->staged = o.staged;
    out->graph = g;
    out->exec = x;
    return BP_SUCCESS;
}

}  // namespace

extern "C" {

int bp_scene_create(bp_context* c, const bp_orbit* out_orbit, const bp_lut2d* out_doppler,
                    const bp_radar_geometry* in_geometry, const bp_dem* dem, double fc,
                    double ds, const bp_kernel* kernel, int dry_tropo_model,
                    const bp_rdr2geo_params* rdr2geo, const bp_geo2rdr_params* geo2rdr,
                    const void* in_dev, int precision, int64_t max_pixels, int max_blocks,
                    bp_scene** scene)
try {
    if (scene == nullptr || c == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    *scene = nullptr;
    CallArgs a{out_orbit, out_doppler, in_geometry, dem, fc, ds, kernel, dry_tropo_model,
               rdr2geo, geo2rdr, in_dev, precision};
    int width = 0, rc;
    if ((rc = check_call(a, &width)) != BP_SUCCESS) return rc;
    if (max_pixels < 1 || max_blocks < 1 || max_blocks > (1 << 20))
        return fail(BP_ERR_INVALID_ARGUMENT, "max_pixels and max_blocks must be positive");
    const bp_grid& gi = in_geometry->grid;
    if (gi.length * gi.width > 0 && in_dev == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "in is null");
    BP_HIP_CHECK(hipSetDevice(c->device));

    auto* sc = new bp_scene();
    sc->ctx = c;
    sc->width = width;
    sc->max_pixels = max_pixels;
    sc->max_blocks = max_blocks;
    sc->out_orbit = *out_orbit;
    sc->out_doppler = *out_doppler;
    sc->in_geometry = *in_geometry;
    sc->dem = *dem;
    sc->kernel = *kernel;
    sc->r2g = *rdr2geo;
    sc->g2r = *geo2rdr;
    sc->args = CallArgs{&sc->out_orbit, &sc->out_doppler, &sc->in_geometry, &sc->dem, fc, ds,
                        &sc->kernel, dry_tropo_model, &sc->r2g, &sc->g2r, in_dev, precision};
    auto bail = [&](int code) {
        destroy_scene(sc);
        return code;
    };
    if (hipStreamCreateWithFlags(&sc->cap, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(BP_ERR_HIP, "hipStreamCreate failed"));
    for (auto& ev : sc->ev)
        if (hipEventCreate(&ev) != hipSuccess) return bail(fail(BP_ERR_HIP, "hipEventCreate failed"));
    const size_t desc_bytes = sizeof(bp::BatchHeader) + sizeof(bp::OutDesc) * max_blocks;
    const size_t flag_bytes = sizeof(int32_t) * (kFlagWords + max_blocks);
    if (hipHostMalloc(reinterpret_cast<void**>(&sc->h_desc), desc_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sc->h_flags), flag_bytes, hipHostMallocDefault) != hipSuccess)
        return bail(fail(BP_ERR_OUT_OF_MEMORY, "hipHostMalloc failed"));
    std::memset(sc->h_desc, 0, desc_bytes);
    std::memset(sc->h_flags, 0, flag_bytes);
    if ((rc = ensure(sc->sc.desc, desc_bytes))) return bail(rc);
    if ((rc = ensure(sc->sc.flags, flag_bytes))) return bail(rc);
    if ((rc = ensure(sc->sc.geom, sizeof(bp::PixGeom) * max_pixels))) return bail(rc);
    if (precision == BP_PRECISION_SPLIT && (rc = ensure(sc->sc.split, sizeof(bp::PixSplit) * max_pixels)))
        return bail(rc);
    hipStream_t s = sc->cap;
    if ((rc = upload_static(sc->st, sc->args, width, s))) return bail(rc);
    if (hipMemsetAsync(sc->sc.flags.ptr, 0, flag_bytes, s) != hipSuccess)
        return bail(fail(BP_ERR_HIP, "hipMemsetAsync failed"));
    if ((rc = pulse_states(sc->st, sc->args, static_cast<int32_t*>(sc->sc.flags.ptr), s))) return bail(rc);
    int32_t oor = 0;
    if (hipMemcpyAsync(&oor, static_cast<int32_t*>(sc->sc.flags.ptr) + 1, sizeof(oor),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return bail(fail(BP_ERR_HIP, "scene upload failed"));
    if (oor) return bail(fail(BP_ERR_OUT_OF_RANGE, orbit_err_string(BP_ORBIT_INTERP_DOMAIN_ERROR)));
    *scene = sc;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_scene_destroy(bp_scene* sc)
try {
    if (sc == nullptr) return BP_SUCCESS;
    (void)hipSetDevice(sc->ctx->device);
    (void)hipStreamSynchronize(sc->cap);
    destroy_scene(sc);
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

}  // extern "C"

namespace bp {
int scene_device(const bp_scene* s) { return s->ctx->device; }
int64_t scene_max_pixels(const bp_scene* s) { return s->max_pixels; }
int scene_max_blocks(const bp_scene* s) { return s->max_blocks; }
}  // namespace bp

extern "C" {

int bp_scene_focus(bp_scene* sc, bp_out_block* blocks, int n_blocks, const bp_options* options,
                   void* stream, int* any_failed, bp_stats* stats)
try {
    if (any_failed) *any_failed = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (sc == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    int rc;
    int64_t npix = 0;
    if ((rc = check_blocks(blocks, n_blocks, &npix)) != BP_SUCCESS) return rc;
    if (npix > sc->max_pixels)
        return fail(BP_ERR_INVALID_ARGUMENT, "request of " + std::to_string(npix) +
                                                 " pixels exceeds the scene's max_pixels");
    Opts o;
    if ((rc = check_options(options, &o)) != BP_SUCCESS) return rc;
    if ((rc = check_lane_rows(o, sc->args.precision, sc->width)) != BP_SUCCESS) return rc;
    if (npix == 0) return BP_SUCCESS;
    BP_HIP_CHECK(hipSetDevice(sc->ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);

    bp_scene::Graph* g = nullptr;
    for (auto& x : sc->graphs)
        if (x.chunk == o.chunk && x.blocks_per_cu == o.blocks_per_cu && x.staged == o.staged &&
            x.lane_rows == o.lane_rows)
            g = &x;
    if (g == nullptr) {
        bp_scene::Graph ng{};
        if ((rc = capture_graph(sc, o, &ng)) != BP_SUCCESS) return rc;
        sc->graphs.push_back(ng);
        g = &sc->graphs.back();
    }
    // the previous replay has finished (every focus call waits for its step),
    // so the pinned descriptors are free to rewrite
    bp::BatchHeader hdr;
    if ((rc = plan_batch(blocks, n_blocks, o, static_cast<int32_t*>(sc->sc.flags.ptr) + kFlagWords,
                         sc->h_desc, sc->max_blocks, &hdr, sc->block_of_grid)) != BP_SUCCESS)
        return rc;
    for (int i = 0; i < g->n_nodes; ++i) {
        hipKernelNodeParams p = g->params[i];
        p.gridDim = dim3((unsigned)(g->geom_groups[i] ? hdr.n_geom_groups : hdr.n_groups));
        BP_HIP_CHECK(hipGraphExecKernelNodeSetParams(g->exec, g->node[i], &p));
    }
    if (stats) BP_HIP_CHECK(hipEventRecord(sc->ev[0], s));
    BP_HIP_CHECK(hipGraphLaunch(g->exec, s));
    if (stats) BP_HIP_CHECK(hipEventRecord(sc->ev[2], s));
    BP_HIP_CHECK(hipStreamSynchronize(s));
    const hipEvent_t ev[3] = {sc->ev[0], nullptr, sc->ev[2]};
    return finish_step(sc->h_flags, hdr.n_grids, sc->block_of_grid, blocks, ev, any_failed, stats);
} catch (...) {
    return bp_host_exception();
}

int bp_backproject(bp_context* c, const bp_radar_geometry* og, const bp_radar_geometry* ig,
                   const bp_dem* dem, double fc, double ds, const bp_kernel* kernel,
                   int dry_tropo_model, const bp_rdr2geo_params* r2g,
                   const bp_geo2rdr_params* g2r, const void* in_dev, void* out_dev,
                   float* height_dev, int precision, const bp_options* options, void* stream,
                   int* any_failed, bp_stats* stats)
try {
    if (any_failed) *any_failed = 0;
    if (!og) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    bp_out_block blk;
    std::memset(&blk, 0, sizeof(blk));
    blk.grid = og->grid;
    blk.out_dev = out_dev;
    blk.height_dev = height_dev;
    CallArgs a{&og->orbit, &og->doppler, ig, dem, fc, ds, kernel, dry_tropo_model, r2g, g2r,
               in_dev, precision};
    return run_blocks(c, a, &blk, 1, options, stream, any_failed, stats);
} catch (...) {
    return bp_host_exception();
}

int bp_backproject_noalloc(bp_context* c, const bp_radar_geometry* og,
                           const bp_radar_geometry* ig, const bp_dem* dem, double fc, double ds,
                           const bp_kernel* kernel, int dry_tropo_model,
                           const bp_rdr2geo_params* r2g, const bp_geo2rdr_params* g2r,
                           const void* in_dev, void* out_dev, float* height_dev, int precision,
                           const bp_options* options, void* stream, int* any_failed,
                           bp_stats* stats)
try {
    if (any_failed) *any_failed = 0;
    if (!og || !c) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (c->scratch == nullptr)
        return fail(BP_ERR_HIP, "bp_backproject_noalloc: no scratch address for this context's "
                                "device (" + c->scratch_err + ")");
    bp_out_block blk;
    std::memset(&blk, 0, sizeof(blk));
    blk.grid = og->grid;
    blk.out_dev = out_dev;
    blk.height_dev = height_dev;
    CallArgs a{&og->orbit, &og->doppler, ig, dem, fc, ds, kernel, dry_tropo_model, r2g, g2r,
               in_dev, precision};
    Carve cv{c->scratch, kNoAllocScratchBytes};
    return run_blocks(c, a, &blk, 1, options, stream, any_failed, stats, &cv);
} catch (...) {
    return bp_host_exception();
}

int bp_backproject_blocks(bp_context* c, const bp_orbit* out_orbit, const bp_lut2d* out_doppler,
                          bp_out_block* blocks, int n_blocks, const bp_radar_geometry* ig,
                          const bp_dem* dem, double fc, double ds, const bp_kernel* kernel,
                          int dry_tropo_model, const bp_rdr2geo_params* r2g,
                          const bp_geo2rdr_params* g2r, const void* in_dev, int precision,
                          const bp_options* options, void* stream, int* any_failed,
                          bp_stats* stats)
try {
    CallArgs a{out_orbit, out_doppler, ig, dem, fc, ds, kernel, dry_tropo_model, r2g, g2r,
               in_dev, precision};
    return run_blocks(c, a, blocks, n_blocks, options, stream, any_failed, stats);
} catch (...) {
    return bp_host_exception();
}

// ------------------------------------------------------------ host helpers

int bp_orbit_interpolate(const bp_orbit* orbit, double t, int border, double* pos, double* vel)
try {
    if (!orbit || !pos || !vel) return BP_NULL_DEREFERENCE;
    bp::V3 p{bp::qnan(), bp::qnan(), bp::qnan()}, v = p;
    int st = bp::orbit_interpolate(host_orbit(*orbit), t, border, &p, &v);
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
    return st;
} catch (...) {
    return bp_host_exception();
}

void bp_lonlat_to_xyz(double a, double e2, const double* llh, double* xyz)
{
    bp::V3 r = bp::lonlat_to_xyz(a, e2, bp::V3{llh[0], llh[1], llh[2]});
    xyz[0] = r.x; xyz[1] = r.y; xyz[2] = r.z;
}

void bp_xyz_to_lonlat(double a, double e2, const double* xyz, double* llh)
{
    bp::V3 r = bp::xyz_to_lonlat(a, e2, bp::V3{xyz[0], xyz[1], xyz[2]});
    llh[0] = r.x; llh[1] = r.y; llh[2] = r.z;
}

double bp_dem_interpolate_lonlat(const bp_dem* dem, double lon, double lat)
{
    return bp::dem_interpolate_lonlat(host_dem(*dem), lon, lat);
}

double bp_dem_interpolate_xy(const bp_dem* dem, double x, double y)
{
    return bp::dem_interpolate_xy(host_dem(*dem), x, y);
}

int bp_proj_forward(int epsg, const double* llh, double* xyz)
try {
    bp::Proj p;
    if (llh == nullptr || xyz == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (!bp::proj_init(epsg, &p))
        return fail(BP_ERR_INVALID_ARGUMENT, "Unknown EPSG code (in factory): " + std::to_string(epsg));
    bp::V3 out;
    if (!bp::proj_forward(p, bp::ld3(llh), &out)) return 1;
    xyz[0] = out.x; xyz[1] = out.y; xyz[2] = out.z;
    return 0;
} catch (...) {
    return bp_host_exception();
}

int bp_proj_inverse(int epsg, const double* xyz, double* llh)
try {
    bp::Proj p;
    if (llh == nullptr || xyz == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (!bp::proj_init(epsg, &p))
        return fail(BP_ERR_INVALID_ARGUMENT, "Unknown EPSG code (in factory): " + std::to_string(epsg));
    bp::V3 out;
    if (!bp::proj_inverse(p, bp::ld3(xyz), &out)) return 1;
    llh[0] = out.x; llh[1] = out.y; llh[2] = out.z;
    return 0;
} catch (...) {
    return bp_host_exception();
}

double bp_lut2d_eval(const bp_lut2d* lut, double y, double x)
{
    return bp::lut2d_eval(host_lut(*lut), y, x);
}

void bp_sincos_turns(double t, double* s, double* c)
{
    double ss = 0., cc = 0.;
    bp::sincos_turns(t, &ss, &cc);
    if (s) *s = ss;
    if (c) *c = cc;
}

double bp_knab_eval(double width, double bandwidth, double x)
{
    return bp::knab_eval(x, std::fabs(width / 2.0), bandwidth);
}

int bp_tabulate_knab(double width, double bandwidth, int n, float* table)
try {
    if (n < 2 || table == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "Require table size >= 2.");
    if (!(bandwidth > 0.0 && bandwidth < 1.0))
        return fail(BP_ERR_RUNTIME, "Require 0 < bandwidth < 1");
    const double hw = std::fabs(width / 2.0);
    const double dx = hw / (n - 1.0);
    for (int i = 0; i < n; ++i) table[i] = (float)bp::knab_eval(i * dx, hw, bandwidth);
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_rdr2geo_bracket(double aztime, double slant_range, double doppler, const bp_orbit* orbit,
                       const bp_dem* dem, double wavelength, int side,
                       const bp_rdr2geo_params* params, double* xyz)
try {
    bp::V3 x{bp::qnan(), bp::qnan(), bp::qnan()};
    bp::R2GParams p{params->tol_height, params->look_min, params->look_max};
    int st = bp::rdr2geo_bracket(aztime, slant_range, doppler, host_orbit(*orbit), host_dem(*dem),
                                 wavelength, side, p, &x);
    xyz[0] = x.x; xyz[1] = x.y; xyz[2] = x.z;
    return st;
} catch (...) {
    return bp_host_exception();
}

int bp_geo2rdr_bracket(const double* xyz, const bp_orbit* orbit, const bp_lut2d* doppler,
                       double wavelength, int side, const bp_geo2rdr_params* params,
                       double* aztime, double* range)
try {
    bp::G2RParams p;
    p.tol_aztime = params->tol_aztime;
    p.have_time_start = params->have_time_start;
    p.have_time_end = params->have_time_end;
    p.time_start = params->time_start;
    p.time_end = params->time_end;
    return bp::geo2rdr_bracket(bp::V3{xyz[0], xyz[1], xyz[2]}, host_orbit(*orbit),
                               host_lut(*doppler), wavelength, side, p, aztime, range);
} catch (...) {
    return bp_host_exception();
}

}  // extern "C"
