// HIP kernels of the backprojection path for gfx950 (wave64):
//   pulse_orbit_kernel  FP64 platform position/velocity at every pulse
//   geometry_kernel     FP64 per-pixel rdr2geo/geo2rdr, CPI window, TSX delay,
//                       and (split mode) the per-pixel delay polynomial fit
//   sum_plain_kernel    reference-semantics coherent sum, FP64 delay & phase
//   sum_split_kernel    split-precision coherent sum: FP64 polynomial phase,
//                       FP32 range-bin delta, polyphase weights from LDS,
//                       FP32 complex MACs with Kahan-compensated chunk sums
#include <hip/hip_runtime.h>
#include <limits.h>

#include <utility>

#include "bp_internal.h"

namespace bp {
namespace {

__device__ inline int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline long long wave_sum_ll(long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap group `it` of
// `ng` (it = blockIdx.x, or blockIdx.x plus multiples of a grid that is a
// multiple of 8, so that it % 8 is the XCD) so that each XCD (own L2) walks a
// contiguous run of tiles, which share pulse data.  Returns the global
// wavefront-tile index of this wavefront (4 per group).
__device__ inline int64_t group_tile(int64_t it, int64_t ng, int* lane)
{
    const int64_t per = ng / 8, rem = ng % 8;
    const int64_t xcd = it % 8, idx = it / 8;
    const int64_t lb = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    *lane = threadIdx.x & 63;
    return lb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Descriptor words read through vector loads are not known to be uniform;
// readfirstlane moves them into scalar registers (every lane of the
// wavefront reads the same descriptor).
__device__ inline int64_t uniform64(int64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// A read-only buffer read at wave-uniform indices as a constant-address-space
// pointer: its loads go through the scalar data cache into scalar registers
// (one load per wavefront, no vector memory instruction).  Only for data no
// kernel writes while the reading kernel runs.
template<class T>
__device__ inline const __attribute__((address_space(4))) T* uniform_ptr(const T* p)
{
    return (const __attribute__((address_space(4))) T*)(p);
}

// Output grid owning global wavefront tile s (s uniform across the wavefront).
__device__ inline int find_grid(const OutDesc* g, int n, int64_t s)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (uniform64(g[mid].seq_base) <= s)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Copy of descriptor i held in scalar registers.
__device__ inline OutDesc load_desc(const OutDesc* g, int i)
{
    static_assert(sizeof(OutDesc) % 4 == 0, "descriptor must be whole dwords");
    OutDesc d;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(g + i);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&d);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(OutDesc) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    return d;
}

// Pixel j (0 <= j < lane_rows) of `lane` in wavefront tile `tile`.
__device__ inline bool tile_pixel(const TileMap& m, int64_t tile, int lane, int j, int64_t* row,
                                  int64_t* col)
{
    if (tile >= m.n_tiles) return false;
    const int64_t ty = tile / m.tiles_x, tx = tile % m.tiles_x;
    *row = (ty * m.tile_rows + lane / m.tile_cols) * m.lane_rows + j;
    *col = tx * m.tile_cols + lane % m.tile_cols;
    return *row < m.length && *col < m.width;
}

// TabulatedKernel<float>::operator() with the reference's FP64 index math
// (core/Kernels.icc:137-154): axn = |x| * float(1/dx) in FP64, i = floor(axn),
// K = t_i + (axn - i) (t_{i+1} - t_i) rounded to float.  The caller passes
// a = x * float(1/dx) (signed); the table holds each entry as the pair
// (t_i, t_{i+1} - t_i) in FP64 (the difference formed in float, as the
// reference does), so one 16-byte read gives both; entry n - 1 holds the
// value the reference's clamp to i = n - 2 gives at |x| = hw.
// kCheck: x may lie past the half width (|a| > an_max = hw * float(1/dx):
// then K = 0, and the entry index is clamped into the table).  Without it
// |x| <= hw, so axn <= (n - 1)(1 + 2^-23) and floor(axn) <= n - 1 for tables
// of up to 2^20 entries.
template<bool kCheck>
__device__ inline float tab_weight(const double2* tab2, int ntab, double an_max, double a)
{
    const double axn = fabs(a);
    int i = (int)axn;  // axn >= 0: truncation is floor
    if (kCheck) i = min(i, ntab - 1);
    const double2 e = tab2[i];
    const float w = (float)fma(__builtin_amdgcn_fract(axn), e.y, e.x);
    return kCheck && axn > an_max ? 0.f : w;
}

}  // namespace

__global__ void pulse_orbit_kernel(Orbit orbit, double t0, double dt, int na, double* pv,
                                   int32_t* flags)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= na) return;
    V3 p, v;
    const int st = orbit_interpolate(orbit, t0 + k * dt, kBorderError, &p, &v);
    if (st != kSuccess) atomicOr(&flags[1], 1);
    double* o = pv + kPulseRecDoubles * (int64_t)k;
    o[0] = p.x;
    o[1] = p.y;
    o[2] = p.z;
    o[3] = v.x;
    o[4] = v.y;
    o[5] = v.z;
    // bistatic delay tau = (r.v - c |r|) * 2 / (|v|^2 - c^2) (BistaticDelay.icc:9-17):
    // the factor once per pulse instead of a division per pixel and pulse
    o[6] = 2. / (dot(v, v) - kC * kC);
    o[7] = 0.;
}

__global__ void clear_kernel(int32_t* words, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        words[i] = 0;
}

// Per-pixel geometry of geometry wavefront `gw`: pixel gw % lane_rows of
// every lane of sum tile gw / lane_rows, following Backproject.cpp:127-199.
template<bool kGenericDem>
__device__ inline void geometry_tile(const GeomArgs& a, const OutDesc* grids, int n_grids,
                                     int64_t n_geom, int lane_rows, int64_t gw, int lane)
{
    const bool in_seq = gw < n_geom;
    const int64_t gseq = gw / lane_rows;
    const int j = (int)(gw - gseq * lane_rows);
    const OutDesc d = load_desc(grids, in_seq ? find_grid(grids, n_grids, gseq) : 0);
    int64_t row = 0, col = 0;
    const bool valid = in_seq && tile_pixel(d.map, gseq - d.seq_base, lane, j, &row, &col);
    const int64_t pix = row * d.map.width + col;
    const int64_t rec = d.rec_base + pix;

    PixGeom g;
    g.x[0] = g.x[1] = g.x[2] = qnan();
    g.tau_atm = 0.;
    g.kstart = 0;
    g.kstop = 0;
    g.status = kPixOutside;
    g.pad_ = 0;
    V3 x{qnan(), qnan(), qnan()};
    V3 p{0., 0., 0.};
    double t = 0., r = 0.;

    if (valid) {
        const double ta = d.t0 + row * d.dt;
        const double ra = d.r0 + col * d.dr;
        const double fd = lut2d_eval(a.out_dop, ta, ra);
        int st = rdr2geo_bracket<kGenericDem>(ta, ra, fd, a.out_orbit, a.dem, a.wvl, d.side,
                                              a.r2g, &x);
        if (st == kOrbitInterpDomainError || st == kOrbitInterpSizeError ||
            st == kOrbitInterpUnknownMethod) {
            atomicOr(&a.flags[1], 1);
            g.status = kPixOutOfRange;
        } else {
            const V3 llh = xyz_to_lonlat(kWgsA, kWgsE2, x);
            if (d.height) d.height[pix] = (float)llh.z;
            if (st != kSuccess) {
                g.status = kPixRdr2GeoFailed;
                if (d.height) d.height[pix] = __builtin_nanf("");
            } else {
                int st2 = geo2rdr_bracket(x, a.in_orbit, a.in_dop, a.wvl, a.in_side, a.g2r, &t, &r);
                if (st2 != kSuccess) {
                    g.status = kPixGeo2RdrFailed;
                } else {
                    V3 v;
                    if (orbit_interpolate(a.in_orbit, t, kBorderError, &p, &v) != kSuccess) {
                        atomicOr(&a.flags[1], 1);
                        g.status = kPixOutOfRange;
                    } else {
                        const double l = a.wvl * r * (norm(p) / norm(x)) / (2. * a.ds);
                        const double cpi = l / norm(v);
                        const double tstart = t - 0.5 * cpi;
                        const double tstop = t + 0.5 * cpi;
                        int kstart = (int)floor((tstart - a.in_t0) / a.in_dt);
                        int kstop = (int)ceil((tstop - a.in_t0) / a.in_dt);
                        kstart = max(kstart, 0);
                        kstop = min(kstop, a.na);
                        g.kstart = kstart;
                        g.kstop = kstop;
                        if (a.tropo == 1) g.tau_atm = dry_tropo_tsx(p, llh);
                        g.status = kPixOk;
                    }
                }
            }
            if (g.status == kPixRdr2GeoFailed || g.status == kPixGeo2RdrFailed) {
                atomicOr(&a.flags[0], 1);
                atomicOr(d.failed, 1);
            }
        }
        g.x[0] = x.x;
        g.x[1] = x.y;
        g.x[2] = x.z;
        a.geom[rec] = g;
    }

    const bool has_win = valid && g.status == kPixOk && g.kstop > g.kstart;
    const long long nupd = has_win ? (long long)(g.kstop - g.kstart) : 0;
    const long long wsum = wave_sum_ll(nupd);
    if (lane == 0 && wsum > 0) atomicAdd(a.updates, (unsigned long long)wsum);

}

// One workgroup per group of four tiles (workgroups past the batch's groups
// return at once).  No persistent loop: a loop around the body makes the
// compiler keep the arguments live across iterations and spill more.
// kGenericDem: DEM rasters in other projections or interpolated other than
// bilinearly; the EPSG:4326 bilinear variant keeps the lookup compact and is
// held to 168 VGPRs, three wavefronts per SIMD, with 18 registers spilled to
// scratch (80 bytes per lane): 5 % faster than its 191 VGPRs at two per SIMD.
// The generic variant keeps the compiler's allocation.
template<bool kGenericDem>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kGenericDem ? 1 : 3)))
void geometry_kernel(GeomArgs a)
{
    const int64_t ng = uniform64(a.batch->n_geom_groups);
    if ((int64_t)blockIdx.x >= ng) return;
    const OutDesc* grids = batch_grids(a.batch);
    const int n_grids = __builtin_amdgcn_readfirstlane(a.batch->n_grids);
    const int lane_rows = __builtin_amdgcn_readfirstlane(a.batch->lane_rows);
    const int64_t n_geom = uniform64(a.batch->n_seq) * lane_rows;
    int lane;
    const int64_t gw = group_tile(blockIdx.x, ng, &lane);
    geometry_tile<kGenericDem>(a, grids, n_grids, n_geom, lane_rows, gw, lane);
}

constexpr int kFitNodes = kFitDegree + 1;

// Chebyshev nodes cos(pi (m + 1/2) / M), M = 5
__device__ inline double fit_node(int m)
{
    return m == 0 ? 0.95105651629515353
                  : m == 1 ? 0.58778525229247314
                           : m == 2 ? 0.0 : m == 3 ? -0.58778525229247314 : -0.95105651629515353;
}

__device__ inline double readlane_f64(double x, int lane)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ inline V3 readlane_v3(V3 x, int lane)
{
    return {readlane_f64(x.x, lane), readlane_f64(x.y, lane), readlane_f64(x.z, lane)};
}

// Split-precision record of one pixel: the pulse phase in cycles fitted
// over the sum wavefront's union window [kw0, kw1) with a degree-4 Chebyshev
// interpolant evaluated in FP64, re-expanded as forward differences at kw0.
// tau_c: the delay at the reference pulse kc; gm[m]: the phase in cycles at
// Chebyshev node m relative to the reference pulse's.
__device__ inline PixSplit fit_pixel(const GeomArgs& a, const PixGeom& g, bool has_win, int kw0,
                                     int kw1, double q_rows, double tau_c,
                                     const double (&gm)[kFitNodes])
{
    PixSplit sp;
    sp.p0 = sp.d1 = sp.d2 = sp.d3 = sp.d4 = 0.;
    sp.zoff = 0.f;
    sp.cbase = 0;
    sp.kstart = has_win ? g.kstart : 0;
    sp.kstop = has_win ? g.kstop : 0;
    sp.status = g.status;
    sp.edge = 0;
    if (has_win) {
        const double kc = 0.5 * ((double)kw0 + (double)(kw1 - 1));
        const double h = fmax(0.5 * (double)(kw1 - 1 - kw0), 0.5);
        constexpr int M = kFitNodes;
        // Chebyshev nodes cos(pi (m + 1/2) / M) and T_j at the nodes
        constexpr double kNode[M] = {0.95105651629515353, 0.58778525229247314, 0.0,
                                     -0.58778525229247314, -0.95105651629515353};
        // Chebyshev coefficients c_j = 2/M sum_m g_m T_j(s_m), c_0 halved,
        // with T_j(s_m) = cos(j pi (m + 1/2) / M).
        double cc[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double acc = 0.;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                // T_j(s_m) by the recurrence (exact to rounding at these nodes)
                double t0 = 1., t1 = kNode[m], tj = j == 0 ? 1. : kNode[m];
#pragma unroll
                for (int i = 2; i <= j; ++i) {
                    tj = 2. * kNode[m] * t1 - t0;
                    t0 = t1;
                    t1 = tj;
                }
                acc += gm[m] * tj;
            }
            cc[j] = acc * (2.0 / M);
        }
        cc[0] *= 0.5;
        // Chebyshev -> monomial in s = (k - kc) / h: T_{j+1} = 2 s T_j - T_{j-1}
        double mono[M], tjm1[M], tj[M], tjp1[M];
#pragma unroll
        for (int n = 0; n < M; ++n) {
            mono[n] = 0.;
            tjm1[n] = 0.;
            tj[n] = 0.;
        }
        tjm1[0] = 1.;  // T_0
        tj[1] = 1.;    // T_1
        mono[0] += cc[0];
        mono[1] += cc[1];
#pragma unroll
        for (int j = 1; j < M - 1; ++j) {
#pragma unroll
            for (int n = 0; n < M; ++n) tjp1[n] = -tjm1[n] + (n > 0 ? 2.0 * tj[n - 1] : 0.);
#pragma unroll
            for (int n = 0; n < M; ++n) {
                mono[n] += cc[j + 1] * tjp1[n];
                tjm1[n] = tj[n];
                tj[n] = tjp1[n];
            }
        }
        // Reference phase (cycles, reduced) and range-bin position at s = 0.
        const double cyc_c = a.fc * tau_c;
        const double c_ref = (cyc_c - floor(cyc_c)) + mono[0];
        const double u_c = (tau_c - a.swst) / a.dtau + mono[0] / (a.fc * a.dtau);
        const double n_ref = floor(u_c);
        const double f_ref = u_c - n_ref;
        // Re-expand dc(s) = sum_{n>=1} mono[n] s^n about the first pulse kw0:
        // s = s0 + j / h, j = k - kw0; coefficients a_i of j^i.
        const double s0 = ((double)kw0 - kc) / h;
        const double ih = 1.0 / h;
        double av[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double acc = 0., binom = 1.;  // C(n, i) s0^(n - i), n = i..4
            double sp0 = 1.;
#pragma unroll
            for (int n = i; n < M; ++n) {
                if (n >= 1) acc += mono[n] * binom * sp0;
                // next n: C(n+1, i) = C(n, i) (n + 1) / (n + 1 - i)
                binom = binom * (double)(n + 1) / (double)(n + 1 - i);
                sp0 *= s0;
            }
            av[i] = acc;
        }
        double ihp = 1.;
#pragma unroll
        for (int i = 1; i < M; ++i) {
            ihp *= ih;
            av[i] *= ihp;
        }
        // forward differences of P(j) = c_ref + sum a_i j^i at j = 0
        sp.p0 = c_ref + av[0];
        sp.d1 = av[1] + av[2] + av[3] + av[4];
        sp.d2 = 2. * av[2] + 6. * av[3] + 14. * av[4];
        sp.d3 = 6. * av[3] + 36. * av[4];
        sp.d4 = 24. * av[4];
        // polyphase coordinate z = -Q (u + shift), u = f_ref + (P - c_ref) / (fc dtau);
        // shift 0 for even W (first tap ceil(u) - W/2), 1/2 for odd W (round(u) - W/2)
        const double inv_k = 1.0 / (a.fc * a.dtau);
        const bool odd = (a.kwidth & 1) != 0;
        const double shift = odd ? 0.5 : 0.0;
        // Conservative range of the first tap over the window (|s| <= 1 on the
        // wavefront's window, so |dc| <= sum |mono[n]|), one bin of margin.
        const double bnd = inv_k * (fabs(mono[1]) + fabs(mono[2]) + fabs(mono[3]) + fabs(mono[4]));
        // koff whole bins of offset keep z >= 0 over the window, so that the
        // sum kernel takes floor(z) and fract(z) by truncation and v_fract
        const double koff = ceil(bnd) + 2.0;
        sp.zoff = (float)(-q_rows * (f_ref - c_ref * inv_k + shift - koff));
        sp.cbase = (int32_t)n_ref - a.kwidth / 2 - (odd ? 1 : 0) + (int32_t)koff;
        const double c_lo = n_ref + ceil(f_ref - bnd + shift) - (a.kwidth / 2) - (odd ? 1 : 0) - 1;
        const double c_hi = n_ref + ceil(f_ref + bnd + shift) - (a.kwidth / 2) - (odd ? 1 : 0) + 1;
        sp.edge = (c_lo < 0. || c_hi + a.kwidth > (double)a.nr) ? 1 : 0;
    }
    return sp;
}

// One sum wavefront tile: read the geometry records of its pixels (lane_rows
// per lane), take the union window of the tile and fit every pixel.
__device__ inline void fit_tile(const GeomArgs& a, const OutDesc* grids, int n_grids,
                                int64_t n_seq, int64_t gseq, int lane)
{
    const bool in_seq = gseq < n_seq;
    const OutDesc d = load_desc(grids, in_seq ? find_grid(grids, n_grids, gseq) : 0);
    const int lr = d.map.lane_rows;  // 1 or 2
    bool valid[2] = {false, false}, has_win[2] = {false, false};
    int64_t rec[2] = {0, 0};
    PixGeom g[2];
    int ks = INT_MAX, ke = INT_MIN;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        int64_t row = 0, col = 0;
        valid[j] = j < lr && in_seq && tile_pixel(d.map, gseq - d.seq_base, lane, j, &row, &col);
        rec[j] = d.rec_base + row * d.map.width + col;
        if (valid[j]) {
            g[j] = a.geom[rec[j]];
            has_win[j] = g[j].status == kPixOk && g[j].kstop > g[j].kstart;
            if (has_win[j]) {
                ks = min(ks, g[j].kstart);
                ke = max(ke, g[j].kstop);
            }
        } else {
            g[j].x[0] = g[j].x[1] = g[j].x[2] = 0.;
            g[j].tau_atm = 0.;
        }
    }
    const int kw0 = __builtin_amdgcn_readfirstlane(wave_min_i(ks));
    const int kw1 = __builtin_amdgcn_readfirstlane(wave_max_i(ke));
    double tau_c[2] = {0., 0.};
    double gm[2][kFitNodes] = {};
    if (kw0 < kw1) {
        // The platform states at the reference pulse kc and the Chebyshev
        // nodes kc + h s_m are the same for every pixel of the tile: one orbit
        // evaluation per lane on lanes 0-5 (lane m < 5: node m, lane 5: kc),
        // each state read back from its lane in turn.
        const double kc = 0.5 * ((double)kw0 + (double)(kw1 - 1));
        const double h = fmax(0.5 * (double)(kw1 - 1 - kw0), 0.5);
        const int m = min(lane, kFitNodes);
        const double km = m == kFitNodes ? kc : kc + h * fit_node(m);
        V3 p, v;
        orbit_interpolate<true>(a.in_orbit, a.in_t0 + km * a.in_dt, kBorderExtrapolate, &p, &v);
        {
            const V3 pc = readlane_v3(p, kFitNodes), vc = readlane_v3(v, kFitNodes);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                if (j < lr)
                    tau_c[j] =
                        g[j].tau_atm + bistatic_delay(pc, vc, V3{g[j].x[0], g[j].x[1], g[j].x[2]});
        }
#pragma unroll
        for (int n = 0; n < kFitNodes; ++n) {
            const V3 pm = readlane_v3(p, n), vm = readlane_v3(v, n);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (j >= lr) continue;
                const double tau_m =
                    g[j].tau_atm + bistatic_delay(pm, vm, V3{g[j].x[0], g[j].x[1], g[j].x[2]});
                gm[j][n] = a.fc * (tau_m - tau_c[j]);  // cycles relative to the reference pulse
            }
        }
    }
    // polyphase rows of the weight table the sum reads for this tiling
    const double q_rows = lr == 2 ? (double)(1 << pair_q_log2(a.kwidth)) : (double)kPolyphaseRows;
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (valid[j])
            a.split[rec[j]] = fit_pixel(a, g[j], has_win[j], kw0, kw1, q_rows, tau_c[j], gm[j]);
}

// Split-precision records of a batch, one wavefront per sum tile.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void fit_kernel(GeomArgs a)
{
    const int64_t ng = uniform64(a.batch->n_groups);
    if ((int64_t)blockIdx.x >= ng) return;
    const OutDesc* grids = batch_grids(a.batch);
    const int n_grids = __builtin_amdgcn_readfirstlane(a.batch->n_grids);
    const int64_t n_seq = uniform64(a.batch->n_seq);
    int lane;
    const int64_t gseq = group_tile(blockIdx.x, ng, &lane);
    fit_tile(a, grids, n_grids, n_seq, gseq, lane);
}

// Persistent tile schedule of the sum kernels.  A workgroup holds 8
// wavefronts; the grid is a multiple of 8 workgroups, spread round-robin over
// the 8 XCDs.  Each XCD walks a contiguous chunk of the tile sequence, and with
// column-major order (map.order == 1) the wavefronts running together cover a
// narrow band of range columns across many output lines: their pulse windows
// overlap, so the pulse lines they gather stay resident in that XCD's L2.
constexpr int kSumWaves = 8;

__device__ inline int64_t seq_to_tile(const TileMap& m, int64_t s)
{
    if (m.order == 0) return s;
    const int64_t tx = s / m.tiles_y, ty = s % m.tiles_y;
    return ty * m.tiles_x + tx;
}

// Calls body(grid descriptor, tile of that grid, lane) for every global
// wavefront tile this wavefront owns; tiles past the last grid are passed as
// tile == map.n_tiles (no valid pixel).
template<typename Body>
__device__ inline void for_each_tile(const BatchHeader* batch, Body&& body)
{
    const OutDesc* grids = batch_grids(batch);
    const int n_grids = __builtin_amdgcn_readfirstlane(batch->n_grids);
    const int64_t n_seq = uniform64(batch->n_seq);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x & 7;
    const int64_t j = blockIdx.x >> 3;
    const int64_t n_local = gridDim.x >> 3;
    const int64_t n_wgseq = (n_seq + kSumWaves - 1) / kSumWaves;
    const int64_t chunk = (n_wgseq + 7) / 8;
    const int64_t end = min(n_wgseq, (xcd + 1) * chunk);
    for (int64_t w = xcd * chunk + j; w < end; w += n_local) {
        const int64_t s = w * kSumWaves + wave;
        if (s < n_seq) {
            const OutDesc d = load_desc(grids, find_grid(grids, n_grids, s));
            body(d, seq_to_tile(d.map, s - d.seq_base), lane);
        } else {
            const OutDesc d = load_desc(grids, n_grids - 1);
            body(d, d.map.n_tiles, lane);
        }
    }
}

// Reference-semantics coherent sum of one wavefront tile (Backproject.cpp:30-63):
// FP64 bistatic delay, range-bin position and phase per pulse; float tabulated
// taps with FP64 index math and a complex<float> inner product (interp1d,
// core/detail/Interp1d.h:23-106: first tap ceil(u) - W/2 for even W, round(u)
// - W/2 for odd W, taps outside the line zero); complex<double> sum.
//
// Per pulse, everything the W taps share is formed once: the delay from the
// pulse record (its 2/(|v|^2 - c^2) factor precomputed), u by a product with
// 1/dtau, and the first tap's offset x0 = low - u (exact, as the reference's
// (m + low) - u: both terms are within a few bins of each other for u >= W);
// tap m's table coordinate is then x0 * float(1/dx) + m * float(1/dx) (the
// reference rounds (x0 + m) * float(1/dx) once instead of twice: the index
// and fraction agree to an ulp of the coordinate).  The phase 2 pi fc tau
// comes from sincos_turns of fc tau in FP64.
// kExact: 2 * halfwidth == W, so no tap lies past the half width.
// `tab2` is the table in LDS or in global memory: each caller passes one of
// the two, so that the reads compile to LDS or global loads.
template<int W, bool kExact>
__device__ inline void plain_tile(const SumArgs& a, const double2* tab2, const OutDesc& od,
                                  int64_t tile, int lane)
{
    int64_t row = 0, col = 0;
    const bool valid = tile_pixel(od.map, tile, lane, 0, &row, &col);
    const int64_t pix = row * od.map.width + col;

    PixGeom g;
    if (valid) g = a.geom[od.rec_base + pix];
    const bool ok = valid && g.status == kPixOk;
    const int ks = ok ? g.kstart : INT_MAX;
    const int ke = ok ? g.kstop : INT_MIN;
    const int kw0 = __builtin_amdgcn_readfirstlane(wave_min_i(ks));
    const int kw1 = __builtin_amdgcn_readfirstlane(wave_max_i(ke));
    const V3 x{g.x[0], g.x[1], g.x[2]};
    const double tau_atm = ok ? g.tau_atm : 0.;
    const double nr_d = (double)a.nr;
    const double rdx = (double)a.rdx;
    const double an_max = a.halfwidth * rdx;
    const int ntab = a.ntab;

    double acc_re = 0., acc_im = 0.;
    for (int k = kw0; k < kw1; ++k) {
        if (k < ks || k >= ke) continue;
        // the pulse record (k is wave-uniform) through the scalar cache
        const auto* pv = uniform_ptr(a.pulse_pv) + kPulseRecDoubles * (int64_t)k;
        const V3 r{x.x - pv[0], x.y - pv[1], x.z - pv[2]};
        const double rv = r.x * pv[3] + r.y * pv[4] + r.z * pv[5];
        const double rn = sqrt_normal(r.x * r.x + r.y * r.y + r.z * r.z);
        const double tau = tau_atm + (rv - kC * rn) * pv[6];
        const double u = (tau - a.swst) * a.inv_dtau;
        const double low = ((W % 2 == 0) ? ceil(u) : round(u)) - (double)(W / 2);
        const double a0 = (low - u) * rdx;
        float w[W];
#pragma unroll
        for (int m = 0; m < W; ++m) {
            const double am = a0 + (double)m * rdx;  // m * rdx exact
            // only the end taps can reach past a half width below W / 2
            w[m] = (!kExact && (m == 0 || m == W - 1))
                       ? tab_weight<true>(tab2, ntab, an_max, am)
                       : tab_weight<false>(tab2, ntab, an_max, am);
        }
        float sre = 0.f, sim = 0.f;
        if (low >= 0. && low + (double)W <= nr_d) {
            // 0 <= low < nr, an int: one conversion (a double-to-int64 one
            // takes six instructions)
            const float2* p = a.rc + (int64_t)k * a.nr + (int)low;
#pragma unroll
            for (int m = 0; m < W; ++m) {
                const float2 d = p[m];
                sre = fmaf(w[m], d.x, sre);
                sim = fmaf(w[m], d.y, sim);
            }
        } else {
            const float2* line = a.rc + (int64_t)k * a.nr;
#pragma unroll
            for (int m = 0; m < W; ++m) {
                const double j = low + (double)m;
                const float2 d = (j >= 0. && j < nr_d) ? line[(int)j] : make_float2(0.f, 0.f);
                sre = fmaf(w[m], d.x, sre);
                sim = fmaf(w[m], d.y, sim);
            }
        }
        double sn, cs;
        sincos_turns(a.fc * tau, &sn, &cs);
        const double dre = (double)sre, dim = (double)sim;
        acc_re = fma(dre, cs, acc_re);
        acc_re = fma(-dim, sn, acc_re);
        acc_im = fma(dre, sn, acc_im);
        acc_im = fma(dim, cs, acc_im);
    }
    if (valid) {
        const float nanf_ = __builtin_nanf("");
        od.out[pix] = ok ? make_float2((float)acc_re, (float)acc_im) : make_float2(nanf_, nanf_);
    }
}

// One 256-thread workgroup per group of four tiles (measured faster than the
// persistent walk of the split kernel for this FP64-bound kernel); workgroups
// past the batch's groups return at once.  The (t_i, t_{i+1} - t_i) table is
// copied to LDS when it fits kMaxTab2Lds entries.
constexpr int kMaxTab2Lds = 4096;

template<int W>
__global__ __launch_bounds__(256) void sum_plain_kernel(SumArgs a)
{
    const int64_t ng = uniform64(a.batch->n_groups);
    if ((int64_t)blockIdx.x >= ng) return;
    extern __shared__ double2 s_tab2[];
    const bool tab_in_lds = a.ntab <= kMaxTab2Lds;
    if (tab_in_lds)
        for (int i = threadIdx.x; i < a.ntab; i += blockDim.x) s_tab2[i] = a.tab2[i];
    __syncthreads();
    const OutDesc* grids = batch_grids(a.batch);
    const int n_grids = __builtin_amdgcn_readfirstlane(a.batch->n_grids);
    const int64_t n_seq = uniform64(a.batch->n_seq);
    int lane;
    const int64_t gseq = group_tile(blockIdx.x, ng, &lane);
    if (gseq < n_seq) {
        const OutDesc d = load_desc(grids, find_grid(grids, n_grids, gseq));
        const int64_t t = gseq - d.seq_base;
        if (tab_in_lds) {
            if (a.exact_width)
                plain_tile<W, true>(a, s_tab2, d, t, lane);
            else
                plain_tile<W, false>(a, s_tab2, d, t, lane);
        } else {
            if (a.exact_width)
                plain_tile<W, true>(a, a.tab2, d, t, lane);
            else
                plain_tile<W, false>(a, a.tab2, d, t, lane);
        }
    }
}

// Split-precision coherent sum of one wavefront tile.
//
// Per pixel (FP64, from the geometry pass): the pulse phase in cycles P(k) as
// a degree-4 polynomial stepped by forward differences from the wavefront's
// first pulse, and the polyphase coordinate offset.  Per pulse: P advanced by
// two FP64 adds (full rate on CDNA; the higher differences enter at each
// fold), its fraction for the phase, and one FP32
// polyphase coordinate z that gives the first tap bin, the weight row and the
// fraction between rows; weights linearly interpolated between polyphase rows
// in LDS, FP32 complex MACs, hardware sin/cos of the phase in revolutions, and
// a fused rotate-accumulate into an FP32 partial sum that is folded every
// `chunk` pulses into a Kahan-compensated accumulator.
//
// The pulses every lane of the wavefront integrates, [max kstart, min kstop),
// form the core loop; the ragged ends of the lanes' windows run in a masked
// loop that gathers its taps from global memory.  In the core loop the taps
// either come from the wavefront's LDS ring of staged pulse windows (one
// coalesced 512-byte load per pulse; see split_core_staged) or are gathered
// from global memory, one pulse ahead (split_core_gather).
//
// FP64 phase of one lane in cycles, P(k), a degree-4 polynomial in k held as
// forward differences.  Per pulse only P and its first difference advance
// (step(): two adds), the second difference being held; after every `chunk`
// steps restore() adds what those n steps left out, so that the state is the
// polynomial's again: after n two-term steps from the exact state
// (p, d1, d2, d3, d4),
//   p  += C(n,3) d3 + C(n,4) d4,   d1 += C(n,2) d3 + C(n,3) d4,
//   d2 += n d3 + C(n,2) d4,        d3 += n d4.
// Inside a fold interval P differs from the polynomial by C(j,3) d3 +
// C(j,4) d4 after j steps: at 850 km range, 7.5 km/s and a 1520 Hz PRF,
// 2050 pulses off the closest approach, d3 = 5e-11 cycles per pulse^3 and the
// difference stays under 2.1e-6 cycles (1.3e-5 rad) for j <= 64.
struct PhaseState {
    double p, d1, d2, d3, d4;
    __device__ inline void step()
    {
        p += d1;
        d1 += d2;
    }
    // c = {n, C(n,2), C(n,3), C(n,4)}
    __device__ inline void restore(const double (&c)[4])
    {
        p = fma(c[2], d3, fma(c[3], d4, p));
        d1 = fma(c[1], d3, fma(c[2], d4, d1));
        d2 = fma(c[0], d3, fma(c[1], d4, d2));
        d3 = fma(c[0], d4, d3);
    }
};

struct SplitPulse {
    float ph;   // phase, revolutions in [0, 1)
    float qf;   // fraction between weight rows
    int row;    // byte offset of the weight row in LDS
    int c0;     // first tap's range bin
};

template<int W>
__device__ inline SplitPulse split_pulse(const PhaseState& st, float zscale, float zoff, int cbase)
{
    constexpr int ROWB = 4 * poly_row_floats(W);  // bytes of one weight row
    SplitPulse r;
    r.ph = (float)__builtin_amdgcn_fract(st.p);
    // z >= 0 over the wavefront's window (the geometry pass offsets it), so
    // truncation is floor
    const float z = fmaf((float)st.p, zscale, zoff);
    const uint32_t m = (uint32_t)z;
    r.qf = __builtin_amdgcn_fractf(z);
    r.row = (int)(m & (kPolyphaseRows - 1)) * ROWB;
    r.c0 = cbase - (int)(m >> kPolyphaseLog2);
    return r;
}

// Tap loads from global memory.  Unguarded: the window is known to lie
// inside the line (the clamp only keeps every load inside the line).
// Guarded: taps outside [0, nr) are zero (core/detail/Interp1d.h:62-83).
template<int W, bool kGuard>
__device__ inline void split_taps(const float2* line, int c0, int nr, float2 (&d)[W])
{
    if (kGuard) {
#pragma unroll
        for (int m = 0; m < W; ++m) {
            const int j = c0 + m;
            d[m] = (j >= 0 && j < nr) ? line[j] : make_float2(0.f, 0.f);
        }
    } else {
        // 32-bit byte offset from the wave-uniform line base
        const uint32_t off = (uint32_t)min(max(c0, 0), nr - W) * (uint32_t)sizeof(float2);
        const float2* p = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(line) + off);
#pragma unroll
        for (int m = 0; m < W; ++m) d[m] = p[m];
    }
}

template<int W, int N>
__device__ inline void split_update(const float* s_poly, const SplitPulse& pu,
                                    const float2 (&d)[N], float& cre, float& cim)
{
    static_assert(N >= W, "tap array");
    constexpr int WP = (W + 3) / 4 * 4;
    // one polyphase row: WP weights then WP deltas, read as float4
    const float4* row4 =
        reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_poly) + pu.row);
    float w[WP];
#pragma unroll
    for (int i = 0; i < WP / 4; ++i) {
        const float4 a = row4[i];
        const float4 b = row4[WP / 4 + i];
        w[4 * i + 0] = fmaf(pu.qf, b.x, a.x);
        w[4 * i + 1] = fmaf(pu.qf, b.y, a.y);
        w[4 * i + 2] = fmaf(pu.qf, b.z, a.z);
        w[4 * i + 3] = fmaf(pu.qf, b.w, a.w);
    }
    float sre = 0.f, sim = 0.f;
#pragma unroll
    for (int m = 0; m < W; ++m) {
        sre = fmaf(w[m], d[m].x, sre);
        sim = fmaf(w[m], d[m].y, sim);
    }
    const float sn = __builtin_amdgcn_sinf(pu.ph);
    const float cs = __builtin_amdgcn_cosf(pu.ph);
    cre = fmaf(sre, cs, cre);
    cre = fmaf(-sim, sn, cre);
    cim = fmaf(sre, sn, cim);
    cim = fmaf(sim, cs, cim);
}

__device__ inline void kahan_fold(float& acc, float& cmp, float& part)
{
    const float y = part - cmp;
    const float t = acc + y;
    cmp = (t - acc) - y;
    acc = t;
    part = 0.f;
}

// State of one wavefront's split sum shared by its pulse loops.
struct SplitAcc {
    PhaseState st;
    float zoff;
    int cbase;
    float cre, cim;          // partial sums since the last fold
    float acc_re, acc_im;    // compensated sums
    float cmp_re, cmp_im;    // compensation terms
    int count;               // pulses accumulated since the last fold
    int steps;               // phase steps since the last restore
    __device__ inline void fold()
    {
        kahan_fold(acc_re, cmp_re, cre);
        kahan_fold(acc_im, cmp_im, cim);
        count = 0;
    }
    __device__ inline void tick(const SumArgs& a)
    {
        if (++count >= a.chunk) fold();
    }
    // The phase to the next pulse.  The one-line loops take a pulse's phase
    // one or two pulses before they accumulate it, so the restores keep their
    // own count; the two-line loops, which do not, restore at each fold.
    __device__ inline void advance(const SumArgs& a)
    {
        st.step();
        if (++steps >= a.chunk) {
            st.restore(a.fold_c);
            steps = 0;
        }
    }
};

// Pulses [k_lo, k_hi) of which each lane integrates [ks, ke): masked taps
// gathered from global memory.  Every pulse advances every lane's phase.
template<int W, bool kGuard>
__device__ inline void split_masked(const SumArgs& a, const float* s_poly, SplitAcc& s, int ks,
                                    int ke, int k_lo, int k_hi)
{
    const int nr = (int)a.nr;
    for (int k = k_lo; k < k_hi; ++k) {
        const SplitPulse pu = split_pulse<W>(s.st, a.zscale, s.zoff, s.cbase);
        s.advance(a);
        if (k >= ks && k < ke) {
            float2 d[W];
            split_taps<W, kGuard>(a.rc + (int64_t)k * nr, pu.c0, nr, d);
            split_update<W>(s_poly, pu, d, s.cre, s.cim);
        }
        s.tick(a);
    }
}

// Core pulses [k0, k1), taps gathered from global memory with the loads of
// pulse k+1 issued before pulse k is accumulated.
template<int W, bool kGuard>
__device__ inline void split_core_gather(const SumArgs& a, const float* s_poly, SplitAcc& s,
                                         int k0, int k1)
{
    const int nr = (int)a.nr;
    const float2* rc = a.rc;
    SplitPulse pa = split_pulse<W>(s.st, a.zscale, s.zoff, s.cbase), pb;
    s.advance(a);
    float2 da[W], db[W];
    split_taps<W, kGuard>(rc + (int64_t)k0 * nr, pa.c0, nr, da);
    int k = k0;
    for (; k + 2 < k1; k += 2) {
        pb = split_pulse<W>(s.st, a.zscale, s.zoff, s.cbase);
        s.advance(a);
        split_taps<W, kGuard>(rc + (int64_t)(k + 1) * nr, pb.c0, nr, db);
        split_update<W>(s_poly, pa, da, s.cre, s.cim);
        s.tick(a);
        pa = split_pulse<W>(s.st, a.zscale, s.zoff, s.cbase);
        s.advance(a);
        split_taps<W, kGuard>(rc + (int64_t)(k + 2) * nr, pa.c0, nr, da);
        split_update<W>(s_poly, pb, db, s.cre, s.cim);
        s.tick(a);
    }
    // the last one or two pulses (pulse k's taps are loaded)
    if (k + 1 < k1) {
        pb = split_pulse<W>(s.st, a.zscale, s.zoff, s.cbase);
        s.advance(a);
        split_taps<W, kGuard>(rc + (int64_t)(k + 1) * nr, pb.c0, nr, db);
        split_update<W>(s_poly, pa, da, s.cre, s.cim);
        s.tick(a);
        split_update<W>(s_poly, pb, db, s.cre, s.cim);
        s.tick(a);
    } else {
        split_update<W>(s_poly, pa, da, s.cre, s.cim);
        s.tick(a);
    }
}

// LDS-staged pulse windows.  Each wavefront owns a ring of kWinSlots slots;
// a slot holds the kWinBins range bins [b, b + kWinBins) of one pulse twice:
// copy A with bin b + i at byte 8 i, and copy B with bin b + i at byte
// kWinCopyB + 8 i, kWinCopyB = 8 (mod 16).  The N = 2 ceil(W / 2) bins from
// bin b + r on then start on a 16-byte boundary in copy A for even r and in
// copy B for odd r, and are read with N / 2 ds_read_b128.  The window base b
// is the first tap of a reference lane less kWinMargin bins, clamped into the
// line; lanes whose taps leave the window (never for tiles narrower than
// about 40 range bins) gather them from global memory instead.
constexpr int kWinBins = 64;
constexpr int kWinSlots = 3;
constexpr int kWinCopyB = 520;
constexpr int kWinSlotBytes = 1040;
constexpr int kWinWaveBytes = kWinSlots * kWinSlotBytes;
constexpr int kWinMargin = 4;
static_assert(kWinCopyB % 16 == 8 && kWinCopyB >= 8 * kWinBins, "copy B placement");
static_assert(kWinWaveBytes % 16 == 0, "rings must start on 16-byte boundaries");
static_assert(kWinCopyB + 8 * kWinBins <= kWinSlotBytes && kWinSlotBytes % 16 == 0, "slot size");

template<int W>
struct StagedPulse {
    SplitPulse pu;
    int b;      // window base bin (wave-uniform)
    float2 g;   // this lane's bin b + lane of the window, in flight
};

// Core pulses [k0, k1) with taps from the wavefront's LDS ring.  Pulse k's
// window is loaded at iteration k - 2, written to its slot at iteration k - 1
// and read at iteration k.  `ring` is the byte offset of the wavefront's ring
// in the shared array `s_win`; `active` marks lanes whose output is used.
template<int W>
__device__ inline void split_core_staged(const SumArgs& a, const float* s_poly, char* s_win,
                                         int ring, SplitAcc& s, int k0, int k1, int ref,
                                         bool active)
{
    constexpr int N = 2 * ((W + 1) / 2);          // taps read per pulse
    constexpr int kMaxRel = kWinBins - N;         // largest window-relative first tap
    const int nr = (int)a.nr;
    const int lane = threadIdx.x & 63;
    const uint32_t lane8 = 8u * (uint32_t)lane;
    const float2* rc = a.rc;
    const uint64_t act = __builtin_amdgcn_ballot_w64(active);

    auto start = [&](StagedPulse<W>& x, int k) {
        x.pu = split_pulse<W>(s.st, a.zscale, s.zoff, s.cbase);
        s.advance(a);
        const int c = __builtin_amdgcn_readlane(x.pu.c0, ref) - kWinMargin;
        x.b = min(max(c, 0), nr - kWinBins);
        // wave-uniform window start plus a 32-bit lane offset (the empty asm
        // keeps the offset from being folded into a 64-bit per-lane address)
        const char* w = reinterpret_cast<const char*>(rc + (int64_t)k * nr + x.b);
        uint32_t lo = lane8;
        asm volatile("" : "+v"(lo));
        x.g = *reinterpret_cast<const float2*>(w + lo);
    };
    auto stage = [&](const StagedPulse<W>& x, int slot) {
        char* base = s_win + ring + slot * kWinSlotBytes + lane8;
        *reinterpret_cast<float2*>(base) = x.g;
        *reinterpret_cast<float2*>(base + kWinCopyB) = x.g;
    };
    auto finish = [&](const StagedPulse<W>& x, int slot, int k) {
        const uint32_t rel = (uint32_t)(x.pu.c0 - x.b);
        const uint32_t r = min(rel, (uint32_t)kMaxRel);
        // byte offset 8 r in copy A (even r) or kWinCopyB + 8 r in copy B
        // (odd r): y + (y & 8) * (kWinCopyB / 8) with y = 8 r + ring, ring a
        // multiple of 16
        const uint32_t y = 8u * r + (uint32_t)ring;
        uint32_t off;
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(off) : "v"(y & 8u), "s"(kWinCopyB / 8), "v"(y));
        const float4* p = reinterpret_cast<const float4*>(s_win + slot * kWinSlotBytes + off);
        float2 d[N];
#pragma unroll
        for (int i = 0; i < N / 2; ++i) {
            const float4 v = p[i];
            d[2 * i] = make_float2(v.x, v.y);
            d[2 * i + 1] = make_float2(v.z, v.w);
        }
        // some active lane's taps leave the window: all lanes take this
        // pulse's taps from global memory (same values as the window holds)
        if ((__builtin_amdgcn_ballot_w64(rel > (uint32_t)kMaxRel) & act) != 0) {
            float2 e[W];
            split_taps<W, false>(rc + (int64_t)k * nr, x.pu.c0, nr, e);
#pragma unroll
            for (int m = 0; m < W; ++m) d[m] = e[m];
        }
        split_update<W>(s_poly, x.pu, d, s.cre, s.cim);
        s.tick(a);
    };

    StagedPulse<W> A, B, C;
    start(A, k0);
    if (k0 + 1 < k1) start(B, k0 + 1);
    stage(A, 0);
    // three pulses per trip: the ring slot of pulse k is (k - k0) % 3
    for (int k = k0;;) {
        if (k + 2 < k1) start(C, k + 2);
        if (k + 1 < k1) stage(B, 1);
        finish(A, 0, k);
        if (++k >= k1) break;
        if (k + 2 < k1) start(A, k + 2);
        if (k + 1 < k1) stage(C, 2);
        finish(B, 1, k);
        if (++k >= k1) break;
        if (k + 2 < k1) start(B, k + 2);
        if (k + 1 < k1) stage(A, 0);
        finish(C, 2, k);
        if (++k >= k1) break;
    }
}

template<int W, bool kStaged>
__device__ inline void split_tile(const SumArgs& a, const float* s_poly, char* s_win,
                                  const OutDesc& od, int64_t tile, int lane)
{
    int64_t row = 0, col = 0;
    const bool valid = tile_pixel(od.map, tile, lane, 0, &row, &col);
    const int64_t pix = row * od.map.width + col;

    PixSplit sp;
    if (valid) sp = a.split[od.rec_base + pix];
    const bool ok = valid && sp.status == kPixOk;
    const bool win = ok && sp.kstop > sp.kstart;
    const int ks = win ? sp.kstart : INT_MAX;
    const int ke = win ? sp.kstop : INT_MIN;
    // union window [kw0, kw1) of the wavefront; empty (kw0 >= kw1) when no
    // lane has pulses, including wavefronts whose tile lies past the grid
    const int kw0 = __builtin_amdgcn_readfirstlane(wave_min_i(ks));
    const int kw1 = __builtin_amdgcn_readfirstlane(wave_max_i(ke));
    // pulses integrated by every lane of the wavefront, [kf0, kf1) inside
    // [kw0, kw1): run unmasked only when every valid lane has a window
    int kf0 = __builtin_amdgcn_readfirstlane(wave_max_i(win ? ks : INT_MIN));
    int kf1 = __builtin_amdgcn_readfirstlane(wave_min_i(win ? ke : INT_MAX));
    const bool all_win = __builtin_amdgcn_readfirstlane(wave_min_i(valid ? (int)win : 1)) != 0;
    if (kw0 >= kw1 || !all_win || kf0 < kw0 || kf1 > kw1 || kf0 >= kf1) {
        kf0 = kw1;
        kf1 = kw1;
    }
    SplitAcc s;
    s.st.p = win ? sp.p0 : 0.;
    s.st.d1 = win ? sp.d1 : 0.;
    s.st.d2 = win ? sp.d2 : 0.;
    s.st.d3 = win ? sp.d3 : 0.;
    s.st.d4 = win ? sp.d4 : 0.;
    s.zoff = win ? sp.zoff : 0.f;
    s.cbase = win ? sp.cbase : 0;
    s.cre = s.cim = s.acc_re = s.acc_im = s.cmp_re = s.cmp_im = 0.f;
    s.count = 0;
    s.steps = 0;
    // guarded taps when some lane's window may reach past a line end, or the
    // lines are narrower than the kernel (wavefront-uniform choice)
    const bool guard = a.nr < W ||
                       __builtin_amdgcn_readfirstlane(wave_max_i(win && sp.edge ? 1 : 0)) != 0;

    if (kw0 < kw1) {
        if (guard) {
            split_masked<W, true>(a, s_poly, s, ks, ke, kw0, kf0);
            if (kf0 < kf1) split_core_gather<W, true>(a, s_poly, s, kf0, kf1);
            split_masked<W, true>(a, s_poly, s, ks, ke, max(kf1, kw0), kw1);
        } else {
            split_masked<W, false>(a, s_poly, s, ks, ke, kw0, kf0);
            if (kf0 < kf1) {
                if (kStaged && a.nr >= kWinBins) {
                    // reference lane: the first lane with pulses
                    const uint64_t m = __builtin_amdgcn_ballot_w64(win);
                    const int ref = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(m));
                    const int ring = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * kWinWaveBytes;
                    split_core_staged<W>(a, s_poly, s_win, ring, s, kf0, kf1, ref, win);
                } else {
                    split_core_gather<W, false>(a, s_poly, s, kf0, kf1);
                }
            }
            split_masked<W, false>(a, s_poly, s, ks, ke, max(kf1, kw0), kw1);
        }
        s.fold();
    }
    if (valid) {
        const float nanf_ = __builtin_nanf("");
        od.out[pix] = ok ? make_float2(s.acc_re - s.cmp_re, s.acc_im - s.cmp_im)
                         : make_float2(nanf_, nanf_);
    }
}

// Two output lines per lane (lane_rows = 2).  The lane's pixels A and B are
// vertically adjacent: at any pulse their first taps differ by at most one
// range bin (adjacent lines are one azimuth sample apart, so their ranges to
// a pulse differ by a few centimetres), and one (W + 1)-bin gather from the
// smaller first tap serves both.  The weights come from a table of
// (W + 1)-entry rows in two copies, the W weights followed (copy 0) or
// preceded (copy 1) by a zero, so that each pixel picks its alignment inside
// the shared bins by its row offset alone.  This halves the gather traffic
// per update.  Where the first taps of some lane's pixels are further apart,
// that pulse is gathered per pixel.  Each pixel reads table row floor(Q t)
// of its fractional offset t, so the window of taps is the reference's, and
// interpolates the weights linearly to t (reading the nearest row instead
// leaves an offset error of up to 1/(2Q) bin: about 8e-4 relative error per
// interpolated sample at Q = 512, against 4e-6 interpolated at Q = 256).
struct PairPulse {
    float ph;   // phase, revolutions in [0, 1)
    float qf;   // fraction between weight rows
    int row;    // byte offset of the weight row (copy 0)
    int c0;     // first tap's range bin
};

template<int W>
struct PairTable {
    static constexpr int kNE = W + 1;                  // entries per row
    static constexpr int kLo = pair_lerp_lo(W);        // first interpolated entry
    static constexpr int kNL = pair_lerp_n(W);         // interpolated entries
    static constexpr int kRF = pair_row_floats(W);     // floats per row
    static constexpr int kQL = pair_q_log2(W);
    static constexpr int kCopyBytes = 4 * kRF;         // copy 1 of a row follows copy 0
    static constexpr int kRowBytes = 2 * kCopyBytes;   // stride of the rows
    static constexpr int kFloats = pair_table_floats(W);
};

template<int W>
__device__ inline PairPulse pair_pulse(const PhaseState& st, float zscale, float zoff, int cbase)
{
    using T = PairTable<W>;
    PairPulse r;
    r.ph = (float)__builtin_amdgcn_fract(st.p);
    const float z = fmaf((float)st.p, zscale, zoff);  // >= 0 over the window
    const uint32_t m = (uint32_t)z;
    r.qf = __builtin_amdgcn_fractf(z);
    r.row = (int)(m & ((1u << T::kQL) - 1)) * T::kRowBytes;
    r.c0 = cbase - (int)(m >> T::kQL);
    return r;
}

// Interpolate the W + 1 bins `d` with the weight row at byte offset `row`
// (its interpolated entries moved the fraction qf towards the next row),
// rotate by the phase and accumulate.  kGate: a pixel with `on` false (outside
// its own pulse window) adds zero, its phasor being zero.
template<int W, bool kGate = false>
__device__ inline void pair_update(const float* s_pair, int row, float qf, float ph,
                                   const float2 (&d)[W + 1], float& cre, float& cim,
                                   bool on = true)
{
    using T = PairTable<W>;
    const float4* r4 = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_pair) + row);
    float v[T::kRF];
#pragma unroll
    for (int i = 0; i < T::kRF / 4; ++i) {
        const float4 x = r4[i];
        v[4 * i + 0] = x.x;
        v[4 * i + 1] = x.y;
        v[4 * i + 2] = x.z;
        v[4 * i + 3] = x.w;
    }
    float w[T::kNE];
#pragma unroll
    for (int e = 0; e < T::kNE; ++e)
        w[e] = (e >= T::kLo && e < T::kLo + T::kNL) ? fmaf(qf, v[T::kNE + e - T::kLo], v[e]) : v[e];
    float sre = 0.f, sim = 0.f;
#pragma unroll
    for (int m = 0; m < T::kNE; ++m) {
        sre = fmaf(w[m], d[m].x, sre);
        sim = fmaf(w[m], d[m].y, sim);
    }
    float sn = __builtin_amdgcn_sinf(ph);
    float cs = __builtin_amdgcn_cosf(ph);
    if (kGate) {
        sn = on ? sn : 0.f;
        cs = on ? cs : 0.f;
    }
    cre = fmaf(sre, cs, cre);
    cre = fmaf(-sim, sn, cre);
    cim = fmaf(sre, sn, cim);
    cim = fmaf(sim, cs, cim);
}

// Byte offset of copy `o` (0 or 1) of the weight row at byte offset `row`:
// one shift-add for power-of-two copies, else one 24-bit multiply-add.
template<int W>
__device__ inline int pair_row(int row, uint32_t o)
{
    constexpr int kCB = PairTable<W>::kCopyBytes;
    if constexpr ((kCB & (kCB - 1)) == 0)
        return (int)(o << __builtin_ctz((unsigned)kCB)) + row;
    else
        return (int)__umul24(o, (uint32_t)kCB) + row;
}

// The W taps from bin c0 (guarded: zero outside [0, nr)), and a zero.
template<int W, bool kGuard>
__device__ inline void pair_taps_one(const float2* line, int c0, int nr, float2 (&d)[W + 1])
{
    float2 e[W];
    split_taps<W, kGuard>(line, c0, nr, e);
#pragma unroll
    for (int m = 0; m < W; ++m) d[m] = e[m];
    d[W] = make_float2(0.f, 0.f);
}

// The W + 1 bins from `base` (base + W + 1 <= nr).
template<int W>
__device__ inline void pair_taps_shared(const float2* line, int base, float2 (&d)[W + 1])
{
    const float2* p = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(line) +
                                                      (uint32_t)base * (uint32_t)sizeof(float2));
#pragma unroll
    for (int m = 0; m <= W; ++m) d[m] = p[m];
}

struct PairLine {
    SplitAcc s;
    int ks, ke;     // this pixel's pulses (empty: ks = INT_MAX)
};

// The fold every chunk pulses of both lines.  The two-line loops step the
// phase of a pulse when they accumulate it, so the phase has taken exactly
// `chunk` steps since the last fold and is restored here (A.s.count counts
// for both lines; the `steps` counters are unused).
__device__ inline void pair_fold(const SumArgs& a, PairLine& A, PairLine& B)
{
    A.s.fold();
    B.s.fold();
    A.s.st.restore(a.fold_c);
    B.s.st.restore(a.fold_c);
}

// Pulses [k_lo, k_hi): each pixel gathers its own taps where it integrates.
template<int W, bool kGuard>
__device__ inline void pair_masked(const SumArgs& a, const float* s_pair, PairLine& A,
                                   PairLine& B, int k_lo, int k_hi)
{
    const int nr = (int)a.nr;
    for (int k = k_lo; k < k_hi; ++k) {
        const float2* line = a.rc + (int64_t)k * nr;
        const PairPulse pa = pair_pulse<W>(A.s.st, a.zscale2, A.s.zoff, A.s.cbase);
        const PairPulse pb = pair_pulse<W>(B.s.st, a.zscale2, B.s.zoff, B.s.cbase);
        A.s.st.step();
        B.s.st.step();
        if (k >= A.ks && k < A.ke) {
            float2 d[W + 1];
            pair_taps_one<W, kGuard>(line, pa.c0, nr, d);
            pair_update<W>(s_pair, pa.row, pa.qf, pa.ph, d, A.s.cre, A.s.cim);
        }
        if (k >= B.ks && k < B.ke) {
            float2 d[W + 1];
            pair_taps_one<W, kGuard>(line, pb.c0, nr, d);
            pair_update<W>(s_pair, pb.row, pb.qf, pb.ph, d, B.s.cre, B.s.cim);
        }
        if (++A.s.count >= a.chunk) {
            pair_fold(a, A, B);
        }
    }
}

// Core pulses [k0, k1), every pixel of every lane integrating: one shared
// (W + 1)-bin gather per lane and pulse (no software prefetch: the two lines'
// state leaves no registers for a second pulse in flight).
//
// A lane whose pixel has no pulses follows its partner (same first tap) or
// holds zero state (first tap 0, the smallest base), so every lane's offset
// from the shared base is 0 or 1 except where an integrating pixel's first
// taps are further apart -- the ballot catches those, and then the pulse is
// gathered per pixel, outside the loop of shared gathers, which runs up to
// the next fold or such a pulse and so carries no other branch.
//
// kRagged: pulses at the ends of the wavefront's window that some pixels do
// not integrate.  Every pixel's phase fit spans the whole window, so its first
// tap stays on its range curve there and the shared gather serves it too; a
// pixel outside its own window adds zero (gated phasor) instead of taking a
// per-pixel branch.
template<int W, bool kRagged>
__device__ inline void pair_core(const SumArgs& a, const float* s_pair, PairLine& A, PairLine& B,
                                 int k0, int k1, uint64_t act_a, uint64_t act_b)
{
    const int nr = (int)a.nr;
    const int base_hi = nr - (W + 1);
    int k = k0;
    while (k < k1) {
        const int kr = min(k1, k + (a.chunk - A.s.count));
        const int k_run = k;
        PairPulse pa, pb;
        bool apart = false;  // pulse k needs per-pixel gathers
        for (; k < kr; ++k) {
            pa = pair_pulse<W>(A.s.st, a.zscale2, A.s.zoff, A.s.cbase);
            pb = pair_pulse<W>(B.s.st, a.zscale2, B.s.zoff, B.s.cbase);
            const int base = min(min(pa.c0, pb.c0), base_hi);
            const uint32_t oa = (uint32_t)(pa.c0 - base), ob = (uint32_t)(pb.c0 - base);
            const uint64_t bad = (__builtin_amdgcn_ballot_w64(oa > 1u) & act_a) |
                                 (__builtin_amdgcn_ballot_w64(ob > 1u) & act_b);
            if (bad != 0) {
                apart = true;
                break;
            }
            A.s.st.step();
            B.s.st.step();
            float2 d[W + 1];
            pair_taps_shared<W>(a.rc + (int64_t)k * nr, base, d);
            // copy oa of the row: the weights aligned to this pixel's first tap
            pair_update<W, kRagged>(s_pair, pair_row<W>(pa.row, oa), pa.qf, pa.ph, d, A.s.cre,
                                    A.s.cim, k >= A.ks && k < A.ke);
            pair_update<W, kRagged>(s_pair, pair_row<W>(pb.row, ob), pb.qf, pb.ph, d, B.s.cre,
                                    B.s.cim, k >= B.ks && k < B.ke);
        }
        if (apart) {
            A.s.st.step();
            B.s.st.step();
            const float2* line = a.rc + (int64_t)k * nr;
            float2 e[W + 1];
            pair_taps_one<W, false>(line, pa.c0, nr, e);
            pair_update<W, kRagged>(s_pair, pa.row, pa.qf, pa.ph, e, A.s.cre, A.s.cim,
                                    k >= A.ks && k < A.ke);
            pair_taps_one<W, false>(line, pb.c0, nr, e);
            pair_update<W, kRagged>(s_pair, pb.row, pb.qf, pb.ph, e, B.s.cre, B.s.cim,
                                    k >= B.ks && k < B.ke);
            ++k;
        }
        A.s.count += k - k_run;
        if (A.s.count >= a.chunk) {
            pair_fold(a, A, B);
        }
    }
}

template<int W>
__device__ inline void pair_tile(const SumArgs& a, const float* s_pair, const OutDesc& od,
                                 int64_t tile, int lane)
{
    int64_t row[2] = {0, 0}, col[2] = {0, 0};
    bool valid[2], ok[2], win[2];
    PixSplit sp[2];
    PairLine L[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        valid[j] = tile_pixel(od.map, tile, lane, j, &row[j], &col[j]);
        if (valid[j]) sp[j] = a.split[od.rec_base + row[j] * od.map.width + col[j]];
        ok[j] = valid[j] && sp[j].status == kPixOk;
        win[j] = ok[j] && sp[j].kstop > sp[j].kstart;
        L[j].ks = win[j] ? sp[j].kstart : INT_MAX;
        L[j].ke = win[j] ? sp[j].kstop : INT_MIN;
    }
    // a pixel without pulses follows its partner's phase and taps (its sum
    // is discarded), so that the shared gather always serves the lane
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        // selects, not an index into sp[] (a run-time index would place the
        // records in scratch memory)
        const bool own = win[j], other = !win[j] && win[1 - j];
        const PixSplit& r0 = sp[j];
        const PixSplit& r1 = sp[1 - j];
        SplitAcc& s = L[j].s;
        s.st.p = own ? r0.p0 : (other ? r1.p0 : 0.);
        s.st.d1 = own ? r0.d1 : (other ? r1.d1 : 0.);
        s.st.d2 = own ? r0.d2 : (other ? r1.d2 : 0.);
        s.st.d3 = own ? r0.d3 : (other ? r1.d3 : 0.);
        s.st.d4 = own ? r0.d4 : (other ? r1.d4 : 0.);
        s.zoff = own ? r0.zoff : (other ? r1.zoff : 0.f);
        s.cbase = own ? r0.cbase : (other ? r1.cbase : 0);
        s.cre = s.cim = s.acc_re = s.acc_im = s.cmp_re = s.cmp_im = 0.f;
        s.count = 0;
        s.steps = 0;
    }
    const int kw0 = __builtin_amdgcn_readfirstlane(wave_min_i(min(L[0].ks, L[1].ks)));
    const int kw1 = __builtin_amdgcn_readfirstlane(wave_max_i(max(L[0].ke, L[1].ke)));
    int kf0 = __builtin_amdgcn_readfirstlane(
        wave_max_i(max(win[0] ? L[0].ks : INT_MIN, win[1] ? L[1].ks : INT_MIN)));
    int kf1 = __builtin_amdgcn_readfirstlane(
        wave_min_i(min(win[0] ? L[0].ke : INT_MAX, win[1] ? L[1].ke : INT_MAX)));
    const bool all_win = __builtin_amdgcn_readfirstlane(
        wave_min_i((!valid[0] || win[0]) && (!valid[1] || win[1]) ? 1 : 0)) != 0;
    if (kw0 >= kw1 || !all_win || kf0 < kw0 || kf1 > kw1 || kf0 >= kf1) {
        kf0 = kw1;
        kf1 = kw1;
    }
    const bool edge = (win[0] && sp[0].edge) || (win[1] && sp[1].edge);
    const bool guard = a.nr < W + 1 || __builtin_amdgcn_readfirstlane(wave_max_i(edge ? 1 : 0)) != 0;
    if (kw0 < kw1) {
        if (guard) {
            pair_masked<W, true>(a, s_pair, L[0], L[1], kw0, kw1);
        } else {
            const uint64_t act_a = __builtin_amdgcn_ballot_w64(win[0]);
            const uint64_t act_b = __builtin_amdgcn_ballot_w64(win[1]);
            pair_core<W, true>(a, s_pair, L[0], L[1], kw0, kf0, act_a, act_b);
            if (kf0 < kf1) pair_core<W, false>(a, s_pair, L[0], L[1], kf0, kf1, act_a, act_b);
            pair_core<W, true>(a, s_pair, L[0], L[1], max(kf1, kw0), kw1, act_a, act_b);
        }
        L[0].s.fold();
        L[1].s.fold();
    }
    const float nanf_ = __builtin_nanf("");
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (valid[j])
            od.out[row[j] * od.map.width + col[j]] =
                ok[j] ? make_float2(L[j].s.acc_re - L[j].s.cmp_re, L[j].s.acc_im - L[j].s.cmp_im)
                      : make_float2(nanf_, nanf_);
}

// Two workgroups per CU need at most 128 VGPRs (4 waves per SIMD).
template<int W, bool kStaged, bool kPair>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
void sum_split_kernel(SumArgs a)
{
    static_assert(!(kPair && kStaged), "two-line lanes gather their taps");
    constexpr int ROW = poly_row_floats(W);
    constexpr int kTable = kPair ? PairTable<W>::kFloats : (kPolyphaseRows + 1) * ROW;
    // static LDS arrays: row and slot offsets fold into the ds_read offsets
    __shared__ float4 s_poly4[kTable / 4];
    __shared__ float4 s_win4[kStaged ? kSumWaves * kWinWaveBytes / 16 : 1];
    float* s_poly = reinterpret_cast<float*>(s_poly4);
    const float* src = kPair ? a.poly2 : a.poly;
    for (int i = threadIdx.x; i < kTable; i += blockDim.x) s_poly[i] = src[i];
    __syncthreads();
    char* s_win = reinterpret_cast<char*>(s_win4);
    for_each_tile(a.batch, [&](const OutDesc& d, int64_t tile, int lane) {
        if constexpr (kPair)
            pair_tile<W>(a, s_poly, d, tile, lane);
        else
            split_tile<W, kStaged>(a, s_poly, s_win, d, tile, lane);
    });
}

// Pulse window [kmin, kmax) of every band of `band_rows` output lines of one
// grid whose records start at `geom`: the union of the windows of the band's
// pixels that integrate pulses (kmin > kmax when none does).  One workgroup
// of 256 threads per band.
__global__ __launch_bounds__(256) void band_window_kernel(const PixGeom* geom, int64_t length,
                                                          int64_t width, int band_rows,
                                                          int32_t* win)
{
    const int64_t b = blockIdx.x;
    const int64_t r0 = b * band_rows;
    const int64_t r1 = min(length, r0 + band_rows);
    int lo = INT_MAX, hi = INT_MIN;
    for (int64_t i = r0 * width + threadIdx.x; i < r1 * width; i += blockDim.x) {
        const PixGeom& g = geom[i];
        if (g.status == kPixOk && g.kstop > g.kstart) {
            lo = min(lo, g.kstart);
            hi = max(hi, g.kstop);
        }
    }
    lo = wave_min_i(lo);
    hi = wave_max_i(hi);
    __shared__ int s_lo[4], s_hi[4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_lo[wave] = lo;
        s_hi[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        win[2 * b] = lo;
        win[2 * b + 1] = hi;
    }
}

// ------------------------------------------------------------------ launchers

hipError_t launch_band_window(const PixGeom* geom, int64_t length, int64_t width, int band_rows,
                              int64_t n_bands, int32_t* win, hipStream_t stream)
{
    if (n_bands <= 0) return hipSuccess;
    hipLaunchKernelGGL(band_window_kernel, dim3((unsigned)n_bands), dim3(256), 0, stream, geom,
                       length, width, band_rows, win);
    return hipGetLastError();
}

hipError_t launch_pulse_orbit(const Orbit& orbit, double t0, double dt, int na, double* pv,
                              int32_t* flags, hipStream_t stream)
{
    if (na <= 0) return hipSuccess;
    const int threads = 256;
    const int blocks = (na + threads - 1) / threads;
    hipLaunchKernelGGL(pulse_orbit_kernel, dim3(blocks), dim3(threads), 0, stream, orbit, t0, dt,
                       na, pv, flags);
    return hipGetLastError();
}

hipError_t launch_clear(int32_t* words, int n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(clear_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, words, n);
    return hipGetLastError();
}

hipError_t launch_geometry(const GeomArgs& args, int grid, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    if (dem_generic(args.dem))
        hipLaunchKernelGGL(geometry_kernel<true>, dim3(grid), dim3(256), 0, stream, args);
    else
        hipLaunchKernelGGL(geometry_kernel<false>, dim3(grid), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launch_fit(const GeomArgs& args, int grid, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(fit_kernel, dim3(grid), dim3(256), 0, stream, args);
    return hipGetLastError();
}

namespace {

// Kernels of every width kMinKernelWidth..kMaxKernelWidth, indexed by width.
using Widths = std::make_integer_sequence<int, kMaxKernelWidth - kMinKernelWidth + 1>;

template<int... I>
const void* plain_fn(int w, std::integer_sequence<int, I...>)
{
    static const void* const fns[] = {
        reinterpret_cast<const void*>(&sum_plain_kernel<I + kMinKernelWidth>)...};
    return fns[w - kMinKernelWidth];
}

template<int... I>
const void* split_fn(int w, bool pair, std::integer_sequence<int, I...>)
{
    static const void* const one[] = {
        reinterpret_cast<const void*>(&sum_split_kernel<I + kMinKernelWidth, false, false>)...};
    static const void* const two[] = {
        reinterpret_cast<const void*>(&sum_split_kernel<I + kMinKernelWidth, false, true>)...};
    return (pair ? two : one)[w - kMinKernelWidth];
}

bool width_ok(int w) { return w >= kMinKernelWidth && w <= kMaxKernelWidth; }

// Dynamic LDS of the plain sum: the (t_i, t_{i+1} - t_i) table when it fits.
size_t plain_lds(int ntab) { return ntab <= kMaxTab2Lds ? sizeof(double2) * ntab : 0; }

}  // namespace

// Host-side handle of a kernel (0 geometry, 4 geometry for generic DEMs,
// 1 plain sum, 2 split sum, 3 fit), as graph kernel nodes report it; null for
// a width or variant that is not built.
const void* kernel_fn(int which, int width, bool staged, bool pair)
{
    switch (which) {
    case 0: return reinterpret_cast<const void*>(&geometry_kernel<false>);
    case 4: return reinterpret_cast<const void*>(&geometry_kernel<true>);
    case 3: return reinterpret_cast<const void*>(&fit_kernel);
    case 1: return width_ok(width) ? plain_fn(width, Widths{}) : nullptr;
    default:
        if (staged) {
            // LDS-staged pulse windows: widths 8 and 9, one line per lane
            if (pair) return nullptr;
            if (width == 8) return reinterpret_cast<const void*>(&sum_split_kernel<8, true, false>);
            if (width == 9) return reinterpret_cast<const void*>(&sum_split_kernel<9, true, false>);
            return nullptr;
        }
        return width_ok(width) ? split_fn(width, pair, Widths{}) : nullptr;
    }
}

hipError_t launch_sum_plain(const SumArgs& args, int width, int grid, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    const void* fn = kernel_fn(1, width, false, false);
    if (fn == nullptr) return hipErrorInvalidValue;
    void* kargs[] = {const_cast<SumArgs*>(&args)};
    return hipLaunchKernel(fn, dim3(grid), dim3(256), kargs, plain_lds(args.ntab), stream);
}

// Workgroups one CU can hold for the chosen sum kernel (256 threads for the
// plain kernel, 512 for the split kernel).
int sum_blocks_per_cu(int precision, int width, bool staged, bool pair)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (precision == 0) {
        const void* fn = kernel_fn(1, width, false, false);
        if (fn) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 256, plain_lds(2048));
    } else {
        const void* fn = kernel_fn(2, width, staged, pair);
        if (fn) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 512, 0);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n < 1 ? 1 : n;
}

// Persistent grid: `grid` workgroups of 512 threads (a multiple of 8, one
// share per XCD), sized by the caller from the device's CU count and
// occupancy.  The weight tables and window rings are static LDS arrays.
hipError_t launch_sum_split(const SumArgs& args, int width, int grid, bool staged, bool pair,
                            hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    const void* fn = kernel_fn(2, width, staged, pair);
    if (fn == nullptr) return hipErrorInvalidValue;
    void* kargs[] = {const_cast<SumArgs*>(&args)};
    return hipLaunchKernel(fn, dim3(grid), dim3(512), kargs, 0, stream);
}

}  // namespace bp
