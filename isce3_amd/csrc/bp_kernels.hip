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

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD (own L2) walks a contiguous run of tiles, which share pulse data.
// Returns the global wavefront-tile index of this wavefront (4 per block).
__device__ inline int64_t wave_tile(int nb, int* lane)
{
    const int b = blockIdx.x;
    const int per = nb / 8, rem = nb % 8;
    const int xcd = b % 8, idx = b / 8;
    const int lb = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    *lane = threadIdx.x & 63;
    return (int64_t)lb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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

__device__ inline bool tile_pixel(const TileMap& m, int64_t tile, int lane, int64_t* row,
                                  int64_t* col)
{
    if (tile >= m.n_tiles) return false;
    const int64_t ty = tile / m.tiles_x, tx = tile % m.tiles_x;
    *row = ty * m.tile_rows + lane / m.tile_cols;
    *col = tx * m.tile_cols + lane % m.tile_cols;
    return *row < m.length && *col < m.width;
}

// TabulatedKernel<float>::operator() with the reference's FP64 index math
// (core/Kernels.icc:137-154).
__device__ inline float tab_eval(const float* tab, int ntab, double hw, float rdx, double x)
{
    const double ax = fabs(x);
    if (ax > hw) return 0.f;
    const double axn = ax * (double)rdx;
    int i = (int)floor(axn);
    i = min(i, ntab - 2);
    const float t0 = tab[i];
    const float dt = tab[i + 1] - t0;
    return (float)((double)t0 + (axn - i) * (double)dt);
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
    double* o = pv + 6 * (int64_t)k;
    o[0] = p.x;
    o[1] = p.y;
    o[2] = p.z;
    o[3] = v.x;
    o[4] = v.y;
    o[5] = v.z;
}

// Per-pixel geometry, following Backproject.cpp:127-199.
__global__ __launch_bounds__(256) void geometry_kernel(GeomArgs a)
{
    int lane;
    const int64_t gseq = wave_tile(a.n_blocks, &lane);
    const bool in_seq = gseq < a.n_seq;
    const OutDesc d = load_desc(a.grids, in_seq ? find_grid(a.grids, a.n_grids, gseq) : 0);
    int64_t row = 0, col = 0;
    const bool valid = in_seq && tile_pixel(d.map, gseq - d.seq_base, lane, &row, &col);
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
        int st = rdr2geo_bracket(ta, ra, fd, a.out_orbit, a.dem, a.wvl, d.side, a.r2g, &x);
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

    if (!a.want_split) return;

    // Split mode: fit the pulse phase (cycles) over the wavefront's union
    // window with a degree-4 Chebyshev interpolant evaluated in FP64.
    const int kw0 = wave_min_i(has_win ? g.kstart : INT_MAX);
    const int kw1 = wave_max_i(has_win ? g.kstop : INT_MIN);
    if (!valid) return;
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
        // delay at the reference pulse kc
        V3 pc, vc;
        orbit_interpolate(a.in_orbit, a.in_t0 + kc * a.in_dt, kBorderExtrapolate, &pc, &vc);
        const double tau_c = g.tau_atm + bistatic_delay(pc, vc, x);
        constexpr int M = kFitDegree + 1;
        // Chebyshev nodes cos(pi (m + 1/2) / M) and T_j at the nodes
        constexpr double kNode[M] = {0.95105651629515353, 0.58778525229247314, 0.0,
                                     -0.58778525229247314, -0.95105651629515353};
        double gm[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const double km = kc + h * kNode[m];
            V3 pm, vm;
            orbit_interpolate(a.in_orbit, a.in_t0 + km * a.in_dt, kBorderExtrapolate, &pm, &vm);
            const double tau_m = g.tau_atm + bistatic_delay(pm, vm, x);
            gm[m] = a.fc * (tau_m - tau_c);  // cycles relative to the reference pulse
        }
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
        sp.zoff = (float)(-(double)kPolyphaseRows * (f_ref - c_ref * inv_k + shift));
        sp.cbase = (int32_t)n_ref - a.kwidth / 2 - (odd ? 1 : 0);
        // Conservative range of the first tap over the window (|s| <= 1 on the
        // wavefront's window, so |dc| <= sum |mono[n]|), one bin of margin.
        const double bnd = inv_k * (fabs(mono[1]) + fabs(mono[2]) + fabs(mono[3]) + fabs(mono[4]));
        const double c_lo = n_ref + ceil(f_ref - bnd + shift) - (a.kwidth / 2) - (odd ? 1 : 0) - 1;
        const double c_hi = n_ref + ceil(f_ref + bnd + shift) - (a.kwidth / 2) - (odd ? 1 : 0) + 1;
        sp.edge = (c_lo < 0. || c_hi + a.kwidth > (double)a.nr) ? 1 : 0;
    }
    a.split[rec] = sp;
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
__device__ inline void for_each_tile(const OutDesc* grids, int n_grids, int64_t n_seq, Body&& body)
{
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
// taps with FP64 index math and a complex<float> inner product; complex<double> sum.
template<int W>
__device__ inline void plain_tile(const SumArgs& a, const float* tab, const OutDesc& od,
                                  int64_t tile, int lane)
{
    int64_t row = 0, col = 0;
    const bool valid = tile_pixel(od.map, tile, lane, &row, &col);
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

    double acc_re = 0., acc_im = 0.;
    for (int k = kw0; k < kw1; ++k) {
        if (k < ks || k >= ke) continue;
        const double* pv = a.pulse_pv + 6 * (int64_t)k;
        const V3 p{pv[0], pv[1], pv[2]};
        const V3 v{pv[3], pv[4], pv[5]};
        const double tau = tau_atm + bistatic_delay(p, v, x);
        const double u = (tau - a.swst) / a.dtau;
        const long i0 = (W % 2 == 0) ? (long)ceil(u) : (long)round(u);
        const long low = i0 - W / 2;
        const float2* line = a.rc + (int64_t)k * a.nr;
        float sre = 0.f, sim = 0.f;
        if (low >= 0 && low + W <= a.nr) {
#pragma unroll
            for (int m = 0; m < W; ++m) {
                const float w = tab_eval(tab, a.ntab, a.halfwidth, a.rdx, (double)(m + low) - u);
                const float2 d = line[low + m];
                sre += w * d.x;
                sim += w * d.y;
            }
        } else {
#pragma unroll
            for (int m = 0; m < W; ++m) {
                const long j = low + m;
                const float w = tab_eval(tab, a.ntab, a.halfwidth, a.rdx, (double)j - u);
                const float2 d = (j >= 0 && j < a.nr) ? line[j] : make_float2(0.f, 0.f);
                sre += w * d.x;
                sim += w * d.y;
            }
        }
        double sn, cs;
        sincospi(2.0 * a.fc * tau, &sn, &cs);
        acc_re += (double)sre * cs - (double)sim * sn;
        acc_im += (double)sre * sn + (double)sim * cs;
    }
    if (valid) {
        const float nanf_ = __builtin_nanf("");
        od.out[pix] = ok ? make_float2((float)acc_re, (float)acc_im) : make_float2(nanf_, nanf_);
    }
}

// One 256-thread workgroup per four tiles (measured faster than the
// persistent walk for this FP64-bound kernel, which needs ~100 VGPRs).
template<int W>
__global__ __launch_bounds__(256) void sum_plain_kernel(SumArgs a)
{
    extern __shared__ float s_tab[];
    const bool tab_in_lds = a.ntab <= kMaxTableLds;
    if (tab_in_lds)
        for (int i = threadIdx.x; i < a.ntab; i += blockDim.x) s_tab[i] = a.table[i];
    __syncthreads();
    const float* tab = tab_in_lds ? s_tab : a.table;
    int lane;
    const int64_t gseq = wave_tile(a.n_blocks, &lane);
    if (gseq < a.n_seq) {
        const OutDesc d = load_desc(a.grids, find_grid(a.grids, a.n_grids, gseq));
        plain_tile<W>(a, tab, d, gseq - d.seq_base, lane);
    }
}

// Split-precision coherent sum of one wavefront tile.
//
// Per pixel (FP64, from the geometry pass): phase polynomial coefficients
// b0..b3 of s = (k - kc)/h, the reduced reference phase c_ref and the
// reference range bin n_ref + f_ref.  Per pulse: s by one FP64 FMA, the phase
// offset dc by FP64 Horner (full-rate FP64 FMA on CDNA), the phase fraction by
// v_fract_f64, the fractional range-bin position f_ref + dc/(fc dtau) in FP32,
// tap weights linearly interpolated between polyphase rows in LDS, FP32 complex
// MACs, hardware sin/cos of the phase in revolutions, and a fused
// rotate-accumulate into an FP32 partial sum that is folded every `chunk`
// pulses into a Kahan-compensated accumulator.
//
// The pulses every lane of the wavefront integrates, [max kstart, min kstop),
// run in an unmasked loop whose tap loads are issued one pulse ahead; the
// ragged ends of the lanes' windows run in a masked loop.
// FP64 phase of one lane in cycles, P(k), stepped by forward differences.
struct PhaseState {
    double p, d1, d2, d3, d4;
    __device__ inline void advance()
    {
        p += d1;
        d1 += d2;
        d2 += d3;
        d3 += d4;
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
    constexpr int ROWB = 4 * 2 * ((W + 3) / 4 * 4);  // bytes of one weight row
    SplitPulse r;
    r.ph = (float)__builtin_amdgcn_fract(st.p);
    const float z = fmaf((float)st.p, zscale, zoff);
    const float zf = floorf(z);
    const int m = (int)zf;
    r.qf = z - zf;
    r.row = (m & (kPolyphaseRows - 1)) * ROWB;
    r.c0 = cbase - (m >> kPolyphaseLog2);
    return r;
}

// Tap loads.  Unguarded: the window is known to lie inside the line (the
// clamp only keeps every load inside the line).  Guarded: taps outside
// [0, nr) are zero (core/detail/Interp1d.h:62-83).
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

template<int W>
__device__ inline void split_update(const float* s_poly, const SplitPulse& pu,
                                    const float2 (&d)[W], float& cre, float& cim)
{
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

// The pulse loops of one wavefront over its union window [kw0, kw1): the
// pulses every lane integrates, [kf0, kf1), run unmasked with the tap loads
// of pulse k+1 issued before pulse k is accumulated; the ragged ends run
// masked.  Every pulse advances every lane's phase state exactly once.
template<int W, bool kGuard>
__device__ inline void split_loops(const SumArgs& a, const float* s_poly, PhaseState st,
                                   float zoff, int cbase, int ks, int ke, int kw0, int kw1,
                                   int kf0, int kf1, float& acc_re, float& acc_im)
{
    const int nr = (int)a.nr;
    const int chunk = a.chunk;
    const float zscale = a.zscale;
    const float2* rc = a.rc;
    float cmp_re = 0.f, cmp_im = 0.f, cre = 0.f, cim = 0.f;

    auto masked = [&](int k_lo, int k_hi) {
        for (int k0 = k_lo; k0 < k_hi; k0 += chunk) {
            const int k1 = min(k0 + chunk, k_hi);
            for (int k = k0; k < k1; ++k) {
                const SplitPulse pu = split_pulse<W>(st, zscale, zoff, cbase);
                st.advance();
                if (k >= ks && k < ke) {
                    float2 d[W];
                    split_taps<W, kGuard>(rc + (int64_t)k * nr, pu.c0, nr, d);
                    split_update<W>(s_poly, pu, d, cre, cim);
                }
            }
            kahan_fold(acc_re, cmp_re, cre);
            kahan_fold(acc_im, cmp_im, cim);
        }
    };

    if (kw0 < kw1) masked(kw0, kf0);
    if (kf0 < kf1) {
        // software pipeline over the whole range: the taps of the next pulse
        // are in flight while the current one is accumulated; partial sums
        // are folded every `chunk` pulses (rounded up to a pair)
        const int pairs = (chunk + 1) / 2;
        SplitPulse pa = split_pulse<W>(st, zscale, zoff, cbase), pb;
        st.advance();
        float2 da[W], db[W];
        split_taps<W, kGuard>(rc + (int64_t)kf0 * nr, pa.c0, nr, da);
        int k = kf0, it = 0;
        for (; k + 2 < kf1; k += 2) {
            pb = split_pulse<W>(st, zscale, zoff, cbase);
            st.advance();
            split_taps<W, kGuard>(rc + (int64_t)(k + 1) * nr, pb.c0, nr, db);
            split_update<W>(s_poly, pa, da, cre, cim);
            pa = split_pulse<W>(st, zscale, zoff, cbase);
            st.advance();
            split_taps<W, kGuard>(rc + (int64_t)(k + 2) * nr, pa.c0, nr, da);
            split_update<W>(s_poly, pb, db, cre, cim);
            if (++it == pairs) {
                it = 0;
                kahan_fold(acc_re, cmp_re, cre);
                kahan_fold(acc_im, cmp_im, cim);
            }
        }
        // the last one or two pulses (pulse k's taps are loaded)
        if (k + 1 < kf1) {
            pb = split_pulse<W>(st, zscale, zoff, cbase);
            st.advance();
            split_taps<W, kGuard>(rc + (int64_t)(k + 1) * nr, pb.c0, nr, db);
            split_update<W>(s_poly, pa, da, cre, cim);
            split_update<W>(s_poly, pb, db, cre, cim);
        } else {
            split_update<W>(s_poly, pa, da, cre, cim);
        }
        kahan_fold(acc_re, cmp_re, cre);
        kahan_fold(acc_im, cmp_im, cim);
    }
    if (kw0 < kw1) masked(max(kf1, kw0), kw1);
    acc_re -= cmp_re;
    acc_im -= cmp_im;
}

template<int W>
__device__ inline void split_tile(const SumArgs& a, const float* s_poly, const OutDesc& od,
                                  int64_t tile, int lane)
{
    int64_t row = 0, col = 0;
    const bool valid = tile_pixel(od.map, tile, lane, &row, &col);
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
    PhaseState st;
    st.p = win ? sp.p0 : 0.;
    st.d1 = win ? sp.d1 : 0.;
    st.d2 = win ? sp.d2 : 0.;
    st.d3 = win ? sp.d3 : 0.;
    st.d4 = win ? sp.d4 : 0.;
    const float zoff = win ? sp.zoff : 0.f;
    const int cbase = win ? sp.cbase : 0;
    // guarded taps when some lane's window may reach past a line end, or the
    // lines are narrower than the kernel (wavefront-uniform choice)
    const bool guard = a.nr < W ||
                       __builtin_amdgcn_readfirstlane(wave_max_i(win && sp.edge ? 1 : 0)) != 0;

    float acc_re = 0.f, acc_im = 0.f;
    if (guard)
        split_loops<W, true>(a, s_poly, st, zoff, cbase, ks, ke, kw0, kw1, kf0, kf1, acc_re,
                             acc_im);
    else
        split_loops<W, false>(a, s_poly, st, zoff, cbase, ks, ke, kw0, kw1, kf0, kf1, acc_re,
                              acc_im);
    if (valid) {
        const float nanf_ = __builtin_nanf("");
        od.out[pix] = ok ? make_float2(acc_re, acc_im) : make_float2(nanf_, nanf_);
    }
}

template<int W>
__global__ __launch_bounds__(512) void sum_split_kernel(SumArgs a)
{
    constexpr int ROW = 2 * ((W + 3) / 4 * 4);
    // static LDS array: row offsets fold into the ds_read offsets
    __shared__ float4 s_poly4[(kPolyphaseRows + 1) * ROW / 4];
    float* s_poly = reinterpret_cast<float*>(s_poly4);
    for (int i = threadIdx.x; i < (kPolyphaseRows + 1) * ROW; i += blockDim.x) s_poly[i] = a.poly[i];
    __syncthreads();
    for_each_tile(a.grids, a.n_grids, a.n_seq, [&](const OutDesc& d, int64_t tile, int lane) {
        split_tile<W>(a, s_poly, d, tile, lane);
    });
}

// ------------------------------------------------------------------ launchers

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

hipError_t launch_geometry(const GeomArgs& args, hipStream_t stream)
{
    if (args.n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(geometry_kernel, dim3(args.n_blocks), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launch_sum_plain(const SumArgs& args, int width, hipStream_t stream)
{
    if (args.n_seq <= 0) return hipSuccess;
    const size_t lds = args.ntab <= kMaxTableLds ? sizeof(float) * args.ntab : 0;
    switch (width) {
    case 8:
        hipLaunchKernelGGL(sum_plain_kernel<8>, dim3(args.n_blocks), dim3(256), lds, stream, args);
        break;
    case 9:
        hipLaunchKernelGGL(sum_plain_kernel<9>, dim3(args.n_blocks), dim3(256), lds, stream, args);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Workgroups of 512 threads one CU can hold for the chosen sum kernel.
int sum_blocks_per_cu(int precision, int width)
{
    int n = 0;
    hipError_t e = hipSuccess;
    if (precision == 0) {
        const size_t lds = sizeof(float) * kMaxTableLds;
        e = width == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sum_plain_kernel<8>, 256, lds)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sum_plain_kernel<9>, 256, lds);
    } else {
        e = width == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sum_split_kernel<8>, 512, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sum_split_kernel<9>, 512, 0);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n < 1 ? 1 : n;
}

// Persistent grid: `grid` workgroups of 512 threads (a multiple of 8, one
// share per XCD), sized by the caller from the device's CU count and occupancy.
hipError_t launch_sum_split(const SumArgs& args, int width, int grid, hipStream_t stream)
{
    if (args.n_seq <= 0) return hipSuccess;
    const size_t lds = 0;  // the weight table is a static LDS array
    switch (width) {
    case 8:
        hipLaunchKernelGGL(sum_split_kernel<8>, dim3(grid), dim3(512), lds, stream, args);
        break;
    case 9:
        hipLaunchKernelGGL(sum_split_kernel<9>, dim3(grid), dim3(512), lds, stream, args);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace bp
