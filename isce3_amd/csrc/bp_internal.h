// Internal device-side data layout shared by the geometry pass, the
// coherent-sum kernels and the host launcher.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/isce3_amd_backproject.h"
#include "bp_math.h"

namespace bp {

// Per-pixel status values written by the geometry pass.
enum PixStatus : int32_t {
    kPixOk = 0,
    kPixRdr2GeoFailed = 1,   // output NaN, height NaN (Backproject.cpp:146-152)
    kPixGeo2RdrFailed = 2,   // output NaN, height valid (Backproject.cpp:167-170)
    kPixOutOfRange = 3,      // orbit domain error; the call raises OutOfRange
    kPixOutside = 4,         // padding lane outside the output grid
};

// Result of the per-pixel geometry pass consumed by the plain FP64 sum:
// target ECEF position, dry-troposphere delay and pulse window [kstart, kstop).
struct alignas(16) PixGeom {
    double x[3];
    double tau_atm;
    int32_t kstart, kstop;
    int32_t status;
    int32_t pad_;
};

// Per-pixel record of the split-precision sum.  The pulse phase in cycles,
// P(k), is a degree-4 polynomial in the pulse index fitted in FP64; the sum
// kernel steps it pulse by pulse from the wavefront's first pulse kw0 by FP64
// forward differences, P += d1; d1 += d2, and every `chunk` pulses adds the
// terms of d3 and d4 those steps left out (binomial coefficients of chunk).
// The polyphase coordinate z = fma((float)P, zscale, zoff) >= 0, with the
// global zscale = -Q/(fc dtau), gives the first tap c0 = cbase - (floor(z) >>
// log2 Q) and the weight row floor(z) & (Q - 1) + fract(z).
struct alignas(16) PixSplit {
    double p0;               // P(kw0)
    double d1, d2, d3, d4;   // forward differences of P at kw0
    float zoff;
    int32_t cbase;
    int32_t kstart, kstop;
    int32_t status;
    int32_t edge;            // nonzero: some tap of the window may leave [0, nr)
};

static_assert(sizeof(PixGeom) == 48, "PixGeom layout");
static_assert(sizeof(PixSplit) == 64, "PixSplit layout");

constexpr int kFitDegree = 4;        // delay polynomial degree of the split path
constexpr int kPolyphaseLog2 = 9;
constexpr int kPolyphaseRows = 1 << kPolyphaseLog2;  // fractional-offset rows of the split weight table
constexpr int kMaxTableLds = 8192;   // plain path keeps tables up to this size in LDS

// One polyphase weight row of the split sum: the W tap weights (padded to a
// multiple of 4), the same number of deltas to the next row, and for W <= 8
// four floats of padding.  Rows are read as float4 at data-dependent row
// offsets; an 80-byte stride spreads their starts over all LDS banks.
__host__ __device__ constexpr int poly_row_floats(int w)
{
    return 2 * ((w + 3) / 4 * 4) + (((w + 3) / 4 * 4) % 8 == 0 ? 4 : 0);
}

// Pixel tiling of one output grid: each 64-lane wavefront of the sum owns a
// tile of (tile_rows * lane_rows) x (64 / tile_rows) output pixels; lane l
// holds the lane_rows vertically adjacent pixels of lane row l / tile_cols and
// column l % tile_cols.
struct TileMap {
    int32_t tile_rows, tile_cols;
    int32_t tiles_x, tiles_y;
    int64_t length, width;      // output grid
    int64_t n_tiles;            // wavefront tiles of this grid
    int32_t order;              // sum-kernel tile sequence: 0 row-major, 1 column-major
    int32_t lane_rows;          // output lines per lane: 1, or 2 (split sum, W = 8)
};

// Interpolation kernel widths (taps, ceil of the kernel width) with built
// coherent-sum kernels.
constexpr int kMinKernelWidth = 2;
constexpr int kMaxKernelWidth = 16;

// Two-line weight table of kernel width W (split sum with lane_rows = 2):
// Q = 2^pair_q_log2(W) rows at fractional offsets tmin + q/Q, each holding
// the W + 1 entries of a (W + 1)-bin window followed by the differences to the
// next row of entries [pair_lerp_lo(W), pair_lerp_lo(W) + pair_lerp_n(W)),
// padded to whole float4s, in two copies -- the W weights followed by a zero
// (copy 0) or preceded by one (copy 1), copy 1 of a row right after its copy
// 0.  A pixel reads row floor(Q t) and
// interpolates the entries that have differences linearly between the rows;
// entries without differences hold the weight at the middle of the row's
// interval.  Q is the largest power of two up to 512 that keeps the table
// within 64 KB of LDS, so that two 512-thread workgroups fit one CU.
// Entries interpolated between rows: the BP_PAIR_LERP_N central ones (all
// W + 1 when W + 1 <= BP_PAIR_LERP_N).  The outer taps of the kernel change
// slowly with the offset, so their half-row offset error costs little: on the
// config-1 block (width 8, 512 rows) 5 interpolated entries give 5.4e-5
// relative and 4.5e-5 rad RMS phase error against the FP64 oracle, 7 give
// 2.0e-5, all 9 (256 rows) 2.7e-6, at 91.1 / 93.4 / 96.3 ms for the sum.
#ifndef BP_PAIR_LERP_N
#define BP_PAIR_LERP_N 5
#endif
__host__ __device__ constexpr int pair_lerp_n(int w) { return BP_PAIR_LERP_N < w + 1 ? BP_PAIR_LERP_N : w + 1; }
__host__ __device__ constexpr int pair_lerp_lo(int w) { return (w + 1 - pair_lerp_n(w)) / 2; }
__host__ __device__ constexpr int pair_row_floats(int w)
{
    return (w + 1 + pair_lerp_n(w) + 3) / 4 * 4;
}
__host__ __device__ constexpr int pair_q_log2(int w)
{
    int q = 9;
    while (q > 4 && 2 * (1 << q) * pair_row_floats(w) * 4 > 65536) --q;
    return q;
}
__host__ __device__ constexpr int pair_table_floats(int w)
{
    return 2 * (1 << pair_q_log2(w)) * pair_row_floats(w);
}

// Doubles per pulse of the plain sum's per-pulse record: platform position
// (3), velocity (3), 2 / (|v|^2 - c^2) (the bistatic delay's factor) and one
// of padding.
constexpr int kPulseRecDoubles = 8;

// One output grid of a (batched) call.  Wavefront tiles of all grids are
// numbered consecutively: grid g owns [seq_base, seq_base + map.n_tiles).
struct OutDesc {
    TileMap map;
    double t0, dt, r0, dr;      // sensing start, 1/PRF, starting range, spacing
    int32_t side;               // look side
    int32_t pad_;
    int64_t seq_base;           // first global wavefront tile of this grid
    int64_t rec_base;           // first per-pixel record of this grid
    float2* out;                // complex64 [length x width]
    float* height;              // optional float32 [length x width]
    int32_t* failed;            // per-grid flag: some pixel failed to converge
    void* pad2_;
};

// Counts of one (batched) call, stored on the device directly in front of its
// output descriptors, so that the kernels read them at run time and a
// captured step replays with new counts.
struct BatchHeader {
    int64_t n_seq;              // wavefront tiles of all grids
    int64_t n_groups;           // groups of four tiles (plain-sum and fit workgroups)
    int64_t n_geom_groups;      // groups of four geometry wavefronts (lane_rows per tile)
    int32_t n_grids;
    int32_t lane_rows;
};
static_assert(sizeof(BatchHeader) == 32, "BatchHeader layout");

__host__ __device__ inline const OutDesc* batch_grids(const BatchHeader* b)
{
    return reinterpret_cast<const OutDesc*>(b + 1);
}

struct GeomArgs {
    const BatchHeader* batch;   // device: header then OutDesc[n_grids]
    int32_t want_split;
    int32_t in_side;
    int32_t kwidth;             // interpolation kernel width W (taps)
    int32_t nr;                 // range samples per pulse
    double in_t0, in_dt;
    int32_t na, tropo;
    double swst, dtau;
    double wvl, fc, ds;
    Orbit out_orbit, in_orbit;  // device pointers
    Lut2d out_dop, in_dop;      // device pointers
    Dem dem;                    // device pointers
    R2GParams r2g;
    G2RParams g2r;
    PixGeom* geom;              // plain records (always written)
    PixSplit* split;            // split records (when want_split)
    int32_t* flags;             // [0] any failed, [1] out of range
    unsigned long long* updates;
};

struct SumArgs {
    const BatchHeader* batch;   // device: header then OutDesc[n_grids]
    int32_t chunk;
    int32_t ntab;
    int64_t nr;                 // range samples per pulse
    double swst, dtau, fc;
    double inv_k;               // 1 / (fc * dtau): cycles -> range bins
    double halfwidth;           // kernel half width
    float rdx;                  // float(1/dx) of the tabulated kernel
    float zscale;               // -Q * inv_k (split polyphase coordinate per cycle)
    const float2* rc;           // range-compressed pulses [na x nr]
    const double* pulse_pv;     // per-pulse records [na x kPulseRecDoubles]
    const PixGeom* geom;
    const PixSplit* split;
    const float* table;         // tabulated kernel (plain)
    const float* poly;          // polyphase weights (split): (rows+1) x poly_row_floats(W)
    const float* poly2;         // two-line weights (split, lane_rows 2): pair_table_floats(W)
    float zscale2;              // -2^pair_q_log2(W) * inv_k
    int32_t exact_width;        // plain: 2 * halfwidth == W (no tap lies past the half width)
    double inv_dtau;            // 1 / dtau
    const double2* tab2;        // plain: (t_i, t_{i+1} - t_i) per table entry, as doubles
    double fold_c[4];           // split: n, C(n,2), C(n,3), C(n,4) for n = chunk
};

// Host launchers (defined in bp_kernels.hip).  `grid` is the number of
// workgroups: the geometry and plain-sum kernels take one group of four tiles
// per workgroup (grid >= n_groups; extra workgroups return at once), the
// split kernel walks the wavefront tiles persistently.
hipError_t launch_pulse_orbit(const Orbit& orbit, double t0, double dt, int na, double* pv,
                              int32_t* flags, hipStream_t stream);
hipError_t launch_clear(int32_t* words, int n, hipStream_t stream);
hipError_t launch_geometry(const GeomArgs& args, int grid, hipStream_t stream);
hipError_t launch_fit(const GeomArgs& args, int grid, hipStream_t stream);
hipError_t launch_sum_plain(const SumArgs& args, int width, int grid, hipStream_t stream);
hipError_t launch_sum_split(const SumArgs& args, int width, int grid, bool staged, bool pair,
                            hipStream_t stream);
int sum_blocks_per_cu(int precision, int width, bool staged, bool pair);
// [kmin, kmax) of each band of band_rows output lines (n_bands workgroups).
hipError_t launch_band_window(const PixGeom* geom, int64_t length, int64_t width, int band_rows,
                              int64_t n_bands, int32_t* win, hipStream_t stream);
// Host-side handle of a kernel (0 geometry, 4 geometry for generic DEMs,
// 1 plain sum, 2 split sum, 3 fit),
// as graph kernel nodes report it.
const void* kernel_fn(int which, int width, bool staged, bool pair);

// CPUs the calling thread may run on (its affinity mask; threads it starts
// inherit the mask), at least 1.
int usable_cpus();

// Scene facts the serving queue needs (bp_api.hip).
int scene_device(const struct bp_scene* scene);
// Argument rules of one output block (shape, PRF, output buffer), shared by
// the batched calls and the serving queue's admission; returns BP_SUCCESS or
// a negative code with the thread's message set.
int check_out_block(const bp_out_block& block);
int64_t scene_max_pixels(const struct bp_scene* scene);
int scene_max_blocks(const struct bp_scene* scene);

}  // namespace bp
