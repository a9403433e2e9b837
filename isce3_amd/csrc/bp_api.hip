// C-ABI of the backprojection path: argument checks that mirror the
// reference's (focus/Backproject.cpp:77-92, pybind focus/Backproject.cpp:97-135,
// DryTroposphereModel.cpp:20-33), device uploads into a per-context grow-only
// arena, and the three-pass launch (pulse orbit, per-pixel geometry,
// coherent sum) on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
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

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct bp_context {
    int device = 0;
    int n_cu = 0;
    DevBuf table, poly, in_pos, in_vel, out_pos, out_vel, in_dop, out_dop, dem;
    DevBuf pulse_pv, geom, split, flags, desc;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    std::vector<float> poly_host;
    std::vector<bp::OutDesc> desc_host;
    std::vector<int> desc_block;
    std::vector<int32_t> flags_host;
};

namespace {

int ensure(DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return BP_SUCCESS;
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
    return d;
}

bp::TileMap make_map(int64_t length, int64_t width, int tile_rows, int order)
{
    bp::TileMap m;
    m.tile_rows = tile_rows;
    m.tile_cols = 64 / tile_rows;
    m.tiles_x = (int32_t)((width + m.tile_cols - 1) / m.tile_cols);
    m.tiles_y = (int32_t)((length + m.tile_rows - 1) / m.tile_rows);
    m.length = length;
    m.width = width;
    m.n_tiles = (int64_t)m.tiles_x * m.tiles_y;
    m.order = order;
    m.pad_ = 0;
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
    out.assign((size_t)(Q + 1) * 2 * wp, 0.f);
    for (int q = 0; q <= Q; ++q) {
        for (int m = 0; m < width; ++m) {
            const float w0 = (float)w[(size_t)q * width + m];
            const float w1 = q < Q ? (float)w[(size_t)(q + 1) * width + m] : w0;
            out[(size_t)q * 2 * wp + m] = w0;
            out[(size_t)q * 2 * wp + wp + m] = w1 - w0;
        }
    }
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
{
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
            delete c;
            return fail(BP_ERR_HIP, "hipEventCreate failed");
        }
    }
    *ctx = c;
    return BP_SUCCESS;
}

int bp_context_destroy(bp_context* c)
{
    if (c == nullptr) return BP_SUCCESS;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    DevBuf* bufs[] = {&c->table, &c->poly, &c->in_pos, &c->in_vel, &c->out_pos, &c->out_vel,
                      &c->in_dop, &c->out_dop, &c->dem, &c->pulse_pv, &c->geom, &c->split,
                      &c->flags, &c->desc};
    for (DevBuf* b : bufs)
        if (b->ptr) (void)hipFree(b->ptr);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete c;
    return BP_SUCCESS;
}

}  // extern "C"

namespace {

// Shared implementation of bp_backproject and bp_backproject_blocks: every
// output block shares the output orbit and Doppler, the input geometry, DEM,
// kernel and parameters; the pixels of all blocks are focused by one geometry
// launch and one coherent-sum launch.
int run_blocks(bp_context* c, const bp_orbit* oo_p, const bp_lut2d* odop_p, bp_out_block* blocks,
               int n_blocks, const bp_radar_geometry* ig, const bp_dem* dem, double fc, double ds,
               const bp_kernel* kernel, int dry_tropo_model, const bp_rdr2geo_params* r2g,
               const bp_geo2rdr_params* g2r, const void* in_dev, int precision,
               const bp_options* options, void* stream, int* any_failed, bp_stats* stats)
{
    if (any_failed) *any_failed = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!c || !oo_p || !odop_p || !ig || !dem || !kernel || !r2g || !g2r)
        return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    if (n_blocks < 0 || (n_blocks > 0 && blocks == nullptr))
        return fail(BP_ERR_INVALID_ARGUMENT, "invalid output block list");
    if (dry_tropo_model != BP_TROPO_NODELAY && dry_tropo_model != BP_TROPO_TSX)
        return fail(BP_ERR_INVALID_ARGUMENT, "unexpected dry troposphere model");
    if (precision != BP_PRECISION_PLAIN && precision != BP_PRECISION_SPLIT)
        return fail(BP_ERR_INVALID_ARGUMENT, "unknown precision mode");
    const bp_grid& gi = ig->grid;
    if (gi.length < 0 || gi.width < 0) return fail(BP_ERR_INVALID_ARGUMENT, "negative grid size");
    if (gi.length > INT32_MAX || gi.width > INT32_MAX)
        return fail(BP_ERR_INVALID_ARGUMENT, "grid size exceeds int range");
    if (!(gi.prf > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "PRF must be positive");
    int64_t npix = 0;
    for (int b = 0; b < n_blocks; ++b) {
        const bp_grid& go = blocks[b].grid;
        blocks[b].failed = 0;
        if (go.length < 0 || go.width < 0) return fail(BP_ERR_INVALID_ARGUMENT, "negative grid size");
        if (go.length > INT32_MAX || go.width > INT32_MAX)
            return fail(BP_ERR_INVALID_ARGUMENT, "grid size exceeds int range");
        if (!(go.prf > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "PRF must be positive");
        if (go.length * go.width > 0 && blocks[b].out_dev == nullptr)
            return fail(BP_ERR_INVALID_ARGUMENT, "out is null");
        npix += go.length * go.width;
    }
    if (!(fc > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "center frequency must be positive");
    if (!(ds > 0)) return fail(BP_ERR_INVALID_ARGUMENT, "azimuth resolution must be positive");
    if (kernel->n < 2 || kernel->table == nullptr || !(kernel->halfwidth > 0))
        return fail(BP_ERR_INVALID_ARGUMENT, "kernel table needs at least two samples");
    const int width = (int)std::ceil(2.0 * kernel->halfwidth);
    if (width != 8 && width != 9)
        return fail(BP_ERR_INVALID_ARGUMENT, "kernel width must be 8 or 9 (got " + std::to_string(width) + ")");
    if (dem->have_raster) {
        if (dem->epsg != 4326) return fail(BP_ERR_INVALID_ARGUMENT, "only EPSG:4326 DEM rasters are supported");
        if (dem->method != BP_DEM_BILINEAR) return fail(BP_ERR_INVALID_ARGUMENT, "only bilinear DEM interpolation is supported");
        if (dem->width < 4 || dem->length < 4 || dem->data == nullptr)
            return fail(BP_ERR_INVALID_ARGUMENT, "DEM raster must be at least 4x4");
    }
    int rc;
    if ((rc = check_orbit(*oo_p, "output geometry")) != BP_SUCCESS) return rc;
    if ((rc = check_orbit(ig->orbit, "input geometry")) != BP_SUCCESS) return rc;
    if ((rc = check_lut(*odop_p, "output geometry")) != BP_SUCCESS) return rc;
    if ((rc = check_lut(ig->doppler, "input geometry")) != BP_SUCCESS) return rc;
    bp_options opt{};
    if (options) opt = *options;
    int tile_rows = opt.tile_rows == 0 ? 1 : opt.tile_rows;
    if (tile_rows != 1 && tile_rows != 2 && tile_rows != 4 && tile_rows != 8)
        return fail(BP_ERR_INVALID_ARGUMENT, "tile_rows must be 1, 2, 4 or 8");
    int chunk = opt.chunk == 0 ? 32 : opt.chunk;
    if (chunk < 1 || chunk > 4096) return fail(BP_ERR_INVALID_ARGUMENT, "chunk must be in [1, 4096]");
    if (opt.fit_degree != 0 && opt.fit_degree != bp::kFitDegree)
        return fail(BP_ERR_INVALID_ARGUMENT, "fit_degree must be 4");
    if (opt.order < 0 || opt.order > 2) return fail(BP_ERR_INVALID_ARGUMENT, "order must be 0, 1 or 2");
    if (opt.blocks_per_cu < 0 || opt.blocks_per_cu > 16)
        return fail(BP_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in [0, 16]");

    const int64_t na = gi.length, nr = gi.width;
    if (npix > 0 && na * nr > 0 && in_dev == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "in is null");

    // Pulse times must lie in the input orbit: Orbit::interpolate with the
    // default border mode throws OutOfRange (Backproject.cpp:101-106).
    const double in_dt = 1.0 / gi.prf;
    if (na > 0) {
        const double ostart = ig->orbit.t0, oend = ig->orbit.t0 + (ig->orbit.n - 1) * ig->orbit.dt;
        const double tf = gi.sensing_start, tl = gi.sensing_start + (na - 1) * in_dt;
        if (tf < ostart || tf > oend || tl < ostart || tl > oend)
            return fail(BP_ERR_OUT_OF_RANGE, orbit_err_string(BP_ORBIT_INTERP_DOMAIN_ERROR));
    }
    if (npix == 0) return BP_SUCCESS;

    BP_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);

    // Output descriptors: blocks with pixels, tiles numbered consecutively.
    const int order = opt.order == 1 ? 0 : 1;
    c->desc_host.clear();
    c->desc_block.clear();
    int64_t n_seq = 0, n_rec = 0;
    for (int b = 0; b < n_blocks; ++b) {
        const bp_grid& go = blocks[b].grid;
        if (go.length * go.width == 0) continue;
        bp::OutDesc d;
        std::memset(&d, 0, sizeof(d));
        d.map = make_map(go.length, go.width, tile_rows, order);
        d.t0 = go.sensing_start;
        d.dt = 1.0 / go.prf;
        d.r0 = go.starting_range;
        d.dr = go.range_pixel_spacing;
        d.side = go.look_side;
        d.seq_base = n_seq;
        d.rec_base = n_rec;
        d.out = static_cast<float2*>(blocks[b].out_dev);
        d.height = blocks[b].height_dev;
        n_seq += d.map.n_tiles;
        n_rec += go.length * go.width;
        c->desc_host.push_back(d);
        c->desc_block.push_back(b);
    }
    const int n_grids = (int)c->desc_host.size();
    if ((n_seq + 3) / 4 > INT32_MAX) return fail(BP_ERR_INVALID_ARGUMENT, "too many output pixels");
    if ((rc = ensure(c->flags, 64 + sizeof(int32_t) * n_grids))) return rc;
    int32_t* d_flags = static_cast<int32_t*>(c->flags.ptr);
    auto* d_updates = reinterpret_cast<unsigned long long*>(static_cast<char*>(c->flags.ptr) + 32);
    int32_t* d_failed = reinterpret_cast<int32_t*>(static_cast<char*>(c->flags.ptr) + 64);
    for (int g = 0; g < n_grids; ++g) c->desc_host[g].failed = d_failed + g;

    // Uploads.
    const bp_orbit& oo = *oo_p;
    const bp_orbit& io = ig->orbit;
    if ((rc = upload(c->out_pos, oo.pos, sizeof(double) * 3 * oo.n, s))) return rc;
    if ((rc = upload(c->out_vel, oo.vel, sizeof(double) * 3 * oo.n, s))) return rc;
    if ((rc = upload(c->in_pos, io.pos, sizeof(double) * 3 * io.n, s))) return rc;
    if ((rc = upload(c->in_vel, io.vel, sizeof(double) * 3 * io.n, s))) return rc;
    if ((rc = upload(c->out_dop, odop_p->have_data ? odop_p->data : nullptr,
                     odop_p->have_data ? sizeof(double) * odop_p->width * odop_p->length : 0, s)))
        return rc;
    if ((rc = upload(c->in_dop, ig->doppler.have_data ? ig->doppler.data : nullptr,
                     ig->doppler.have_data ? sizeof(double) * ig->doppler.width * ig->doppler.length : 0, s)))
        return rc;
    if ((rc = upload(c->dem, dem->have_raster ? dem->data : nullptr,
                     dem->have_raster ? sizeof(float) * dem->width * dem->length : 0, s)))
        return rc;
    if ((rc = upload(c->table, kernel->table, sizeof(float) * kernel->n, s))) return rc;
    if (precision == BP_PRECISION_SPLIT) {
        build_polyphase(*kernel, width, c->poly_host);
        if ((rc = upload(c->poly, c->poly_host.data(), sizeof(float) * c->poly_host.size(), s))) return rc;
    }
    if ((rc = upload(c->desc, c->desc_host.data(), sizeof(bp::OutDesc) * n_grids, s))) return rc;
    if ((rc = ensure(c->geom, sizeof(bp::PixGeom) * n_rec))) return rc;
    if (precision == BP_PRECISION_SPLIT && (rc = ensure(c->split, sizeof(bp::PixSplit) * n_rec))) return rc;
    if (precision == BP_PRECISION_PLAIN && (rc = ensure(c->pulse_pv, sizeof(double) * 6 * std::max<int64_t>(na, 1))))
        return rc;
    BP_HIP_CHECK(hipMemsetAsync(c->flags.ptr, 0, 64 + sizeof(int32_t) * n_grids, s));

    const double swst = 2. * gi.starting_range / bp::kC;
    const double dtau = 2. * gi.range_pixel_spacing / bp::kC;
    const int per_cu = opt.blocks_per_cu > 0 ? opt.blocks_per_cu
                                             : bp::sum_blocks_per_cu(precision, width);
    int grid = c->n_cu * per_cu;
    grid = (grid + 7) / 8 * 8;
    const auto* d_desc = static_cast<const bp::OutDesc*>(c->desc.ptr);
    const int32_t wblocks = (int32_t)((n_seq + 3) / 4);

    bp::Dem ddem = host_dem(*dem);
    ddem.data = static_cast<const float*>(c->dem.ptr);

    if (precision == BP_PRECISION_PLAIN && na > 0) {
        BP_HIP_CHECK(bp::launch_pulse_orbit(dev_orbit(io, c->in_pos, c->in_vel), gi.sensing_start,
                                            in_dt, (int)na, static_cast<double*>(c->pulse_pv.ptr),
                                            d_flags, s));
    }

    bp::GeomArgs ga;
    std::memset(&ga, 0, sizeof(ga));
    ga.grids = d_desc;
    ga.n_grids = n_grids;
    ga.n_seq = n_seq;
    ga.n_blocks = wblocks;
    ga.in_side = gi.look_side;
    ga.kwidth = width;
    ga.nr = (int32_t)nr;
    ga.in_t0 = gi.sensing_start;
    ga.in_dt = in_dt;
    ga.na = (int32_t)na;
    ga.tropo = dry_tropo_model;
    ga.swst = swst;
    ga.dtau = dtau;
    ga.wvl = bp::kC / fc;
    ga.fc = fc;
    ga.ds = ds;
    ga.out_orbit = dev_orbit(oo, c->out_pos, c->out_vel);
    ga.in_orbit = dev_orbit(io, c->in_pos, c->in_vel);
    ga.out_dop = dev_lut(*odop_p, c->out_dop);
    ga.in_dop = dev_lut(ig->doppler, c->in_dop);
    ga.dem = ddem;
    ga.r2g = {r2g->tol_height, r2g->look_min, r2g->look_max};
    ga.g2r.tol_aztime = g2r->tol_aztime;
    ga.g2r.have_time_start = g2r->have_time_start;
    ga.g2r.have_time_end = g2r->have_time_end;
    ga.g2r.time_start = g2r->time_start;
    ga.g2r.time_end = g2r->time_end;
    ga.want_split = precision == BP_PRECISION_SPLIT;
    ga.geom = static_cast<bp::PixGeom*>(c->geom.ptr);
    ga.split = static_cast<bp::PixSplit*>(c->split.ptr);
    ga.flags = d_flags;
    ga.updates = d_updates;

    if (stats) BP_HIP_CHECK(hipEventRecord(c->ev[0], s));
    BP_HIP_CHECK(bp::launch_geometry(ga, s));
    if (stats) BP_HIP_CHECK(hipEventRecord(c->ev[1], s));

    bp::SumArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.grids = d_desc;
    sa.n_grids = n_grids;
    sa.n_seq = n_seq;
    sa.n_blocks = wblocks;
    sa.nr = nr;
    sa.swst = swst;
    sa.dtau = dtau;
    sa.fc = fc;
    sa.inv_k = 1.0 / (fc * dtau);
    sa.zscale = (float)(-(double)bp::kPolyphaseRows * sa.inv_k);
    sa.halfwidth = kernel->halfwidth;
    sa.rdx = (float)(1.0 / (kernel->halfwidth / (kernel->n - 1.0)));
    sa.ntab = kernel->n;
    sa.chunk = chunk;
    sa.rc = static_cast<const float2*>(in_dev);
    sa.pulse_pv = static_cast<const double*>(c->pulse_pv.ptr);
    sa.geom = static_cast<const bp::PixGeom*>(c->geom.ptr);
    sa.split = static_cast<const bp::PixSplit*>(c->split.ptr);
    sa.table = static_cast<const float*>(c->table.ptr);
    sa.poly = static_cast<const float*>(c->poly.ptr);
    if (precision == BP_PRECISION_PLAIN)
        BP_HIP_CHECK(bp::launch_sum_plain(sa, width, s));
    else
        BP_HIP_CHECK(bp::launch_sum_split(sa, width, grid, s));
    if (stats) BP_HIP_CHECK(hipEventRecord(c->ev[2], s));

    c->flags_host.assign(16 + n_grids, 0);
    BP_HIP_CHECK(hipMemcpyAsync(c->flags_host.data(), c->flags.ptr, 64 + sizeof(int32_t) * n_grids,
                                hipMemcpyDeviceToHost, s));
    BP_HIP_CHECK(hipStreamSynchronize(s));
    const int32_t* hflags = c->flags_host.data();
    unsigned long long hupd = 0;
    std::memcpy(&hupd, hflags + 8, sizeof(hupd));
    for (int g = 0; g < n_grids; ++g) blocks[c->desc_block[g]].failed = hflags[16 + g] != 0;
    if (stats) {
        float ms = 0.f;
        BP_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        stats->geometry_ms = ms;
        BP_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
        stats->sum_ms = ms;
        stats->updates = (int64_t)hupd;
        stats->any_failed = hflags[0] != 0;
    }
    if (hflags[1]) return fail(BP_ERR_OUT_OF_RANGE, orbit_err_string(BP_ORBIT_INTERP_DOMAIN_ERROR));
    if (any_failed) *any_failed = hflags[0] != 0;
    return hflags[0] ? BP_FAILED_TO_CONVERGE : BP_SUCCESS;
}

}  // namespace

extern "C" {

int bp_backproject(bp_context* c, const bp_radar_geometry* og, const bp_radar_geometry* ig,
                   const bp_dem* dem, double fc, double ds, const bp_kernel* kernel,
                   int dry_tropo_model, const bp_rdr2geo_params* r2g,
                   const bp_geo2rdr_params* g2r, const void* in_dev, void* out_dev,
                   float* height_dev, int precision, const bp_options* options, void* stream,
                   int* any_failed, bp_stats* stats)
{
    if (any_failed) *any_failed = 0;
    if (!og) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    bp_out_block blk;
    std::memset(&blk, 0, sizeof(blk));
    blk.grid = og->grid;
    blk.out_dev = out_dev;
    blk.height_dev = height_dev;
    return run_blocks(c, &og->orbit, &og->doppler, &blk, 1, ig, dem, fc, ds, kernel,
                      dry_tropo_model, r2g, g2r, in_dev, precision, options, stream, any_failed,
                      stats);
}

int bp_backproject_blocks(bp_context* c, const bp_orbit* out_orbit, const bp_lut2d* out_doppler,
                          bp_out_block* blocks, int n_blocks, const bp_radar_geometry* ig,
                          const bp_dem* dem, double fc, double ds, const bp_kernel* kernel,
                          int dry_tropo_model, const bp_rdr2geo_params* r2g,
                          const bp_geo2rdr_params* g2r, const void* in_dev, int precision,
                          const bp_options* options, void* stream, int* any_failed,
                          bp_stats* stats)
{
    return run_blocks(c, out_orbit, out_doppler, blocks, n_blocks, ig, dem, fc, ds, kernel,
                      dry_tropo_model, r2g, g2r, in_dev, precision, options, stream, any_failed,
                      stats);
}

// ------------------------------------------------------------ host helpers

int bp_orbit_interpolate(const bp_orbit* orbit, double t, int border, double* pos, double* vel)
{
    if (!orbit || !pos || !vel) return BP_NULL_DEREFERENCE;
    bp::V3 p{bp::qnan(), bp::qnan(), bp::qnan()}, v = p;
    int st = bp::orbit_interpolate(host_orbit(*orbit), t, border, &p, &v);
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
    return st;
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

double bp_lut2d_eval(const bp_lut2d* lut, double y, double x)
{
    return bp::lut2d_eval(host_lut(*lut), y, x);
}

double bp_knab_eval(double width, double bandwidth, double x)
{
    return bp::knab_eval(x, std::fabs(width / 2.0), bandwidth);
}

int bp_tabulate_knab(double width, double bandwidth, int n, float* table)
{
    if (n < 2 || table == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "Require table size >= 2.");
    if (!(bandwidth > 0.0 && bandwidth < 1.0))
        return fail(BP_ERR_RUNTIME, "Require 0 < bandwidth < 1");
    const double hw = std::fabs(width / 2.0);
    const double dx = hw / (n - 1.0);
    for (int i = 0; i < n; ++i) table[i] = (float)bp::knab_eval(i * dx, hw, bandwidth);
    return BP_SUCCESS;
}

int bp_rdr2geo_bracket(double aztime, double slant_range, double doppler, const bp_orbit* orbit,
                       const bp_dem* dem, double wavelength, int side,
                       const bp_rdr2geo_params* params, double* xyz)
{
    bp::V3 x{bp::qnan(), bp::qnan(), bp::qnan()};
    bp::R2GParams p{params->tol_height, params->look_min, params->look_max};
    int st = bp::rdr2geo_bracket(aztime, slant_range, doppler, host_orbit(*orbit), host_dem(*dem),
                                 wavelength, side, p, &x);
    xyz[0] = x.x; xyz[1] = x.y; xyz[2] = x.z;
    return st;
}

int bp_geo2rdr_bracket(const double* xyz, const bp_orbit* orbit, const bp_lut2d* doppler,
                       double wavelength, int side, const bp_geo2rdr_params* params,
                       double* aztime, double* range)
{
    bp::G2RParams p;
    p.tol_aztime = params->tol_aztime;
    p.have_time_start = params->have_time_start;
    p.have_time_end = params->have_time_end;
    p.time_start = params->time_start;
    p.time_end = params->time_end;
    return bp::geo2rdr_bracket(bp::V3{xyz[0], xyz[1], xyz[2]}, host_orbit(*orbit),
                               host_lut(*doppler), wavelength, side, p, aztime, range);
}

}  // extern "C"
