/*
 * isce3_amd -- MI355X-native time-domain backprojection, C ABI.
 *
 * This is the drop-in boundary for isce3's backprojection focuser.  Plain C
 * types only: pointers, sizes and POD structs; no C++ or torch types.  All
 * times are seconds relative to the orbit's reference epoch (the caller
 * aligns radar-grid epochs to the orbit epoch, as RadarGeometry does in
 * cxx/isce3/container/RadarGeometry.icc:15-26).
 *
 * Return codes: values 0..11 mirror isce3::error::ErrorCode
 * (cxx/isce3/error/ErrorCode.h:8-21).  Negative values are argument/runtime
 * errors that a host binding raises as exceptions; bp_last_error_message()
 * returns the message for the calling thread.  No C++ exception leaves an
 * entry point: a failed host allocation returns BP_ERR_OUT_OF_MEMORY.
 *
 * Threading: every call is synchronous with respect to the host unless a
 * non-null HIP stream is passed, in which case device work is enqueued on
 * that stream and the call returns after enqueueing (outputs that are read
 * back to the host, such as any_failed and the update count, are then
 * synchronised on that stream before return).  A context may be used from
 * one host thread at a time.
 */
#ifndef ISCE3_AMD_BACKPROJECT_H
#define ISCE3_AMD_BACKPROJECT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
enum {
    BP_SUCCESS = 0,
    BP_ORBIT_INTERP_SIZE_ERROR = 1,
    BP_ORBIT_INTERP_DOMAIN_ERROR = 2,
    BP_ORBIT_INTERP_UNKNOWN_METHOD = 3,
    BP_OUT_OF_BOUNDS_DEM = 4,
    BP_INVALID_DEM = 5,
    BP_FAILED_TO_CONVERGE = 6,
    BP_WRONG_LOOK_SIDE = 7,
    BP_OUT_OF_BOUNDS_LOOKUP = 8,
    BP_NULL_DEREFERENCE = 9,
    BP_INVALID_TOLERANCE = 10,
    BP_INVALID_INTERVAL = 11,
    BP_ERR_INVALID_ARGUMENT = -1, /* isce3::except::InvalidArgument */
    BP_ERR_OUT_OF_RANGE = -2,     /* isce3::except::OutOfRange */
    BP_ERR_DOMAIN = -3,           /* isce3::except::DomainError */
    BP_ERR_HIP = -4,              /* HIP runtime failure */
    BP_ERR_RUNTIME = -5,          /* isce3::except::RuntimeError */
    BP_ERR_OUT_OF_MEMORY = -6,
    BP_ERR_LENGTH = -7            /* isce3::except::LengthError */
};

enum { BP_INTERP_HERMITE = 0, BP_INTERP_LEGENDRE = 1 };   /* OrbitInterpMethod */
enum { BP_LOOK_LEFT = 1, BP_LOOK_RIGHT = -1 };            /* isce3 LookSide */
enum { BP_TROPO_NODELAY = 0, BP_TROPO_TSX = 1 };          /* DryTroposphereModel */
/* DEM interpolation: isce3::core::dataInterpMethod values (core/Constants.h:23-29);
 * sinc is the factory's 8-tap, 8192-row Sinc2dInterpolator. */
enum { BP_DEM_SINC = 0, BP_DEM_BILINEAR = 1, BP_DEM_BICUBIC = 2, BP_DEM_NEAREST = 3,
       BP_DEM_BIQUINTIC = 4 };
enum {
    BP_PRECISION_PLAIN = 0, /* FP64 delay/phase/taps index, complex<double> sum */
    BP_PRECISION_SPLIT = 1  /* FP64 per-pixel reference + FP32 deltas, Kahan sum */
};

/* Uniformly spaced orbit state vectors (isce3::core::Orbit, core/Orbit.h). */
typedef struct {
    double t0;          /* time of first state vector (s) */
    double dt;          /* state vector spacing (s) */
    int32_t n;          /* number of state vectors */
    int32_t method;     /* BP_INTERP_* */
    const double* pos;  /* host, n x 3 ECEF (m) */
    const double* vel;  /* host, n x 3 ECEF (m/s) */
} bp_orbit;

/* Bilinear LUT2d<double> (core/LUT2d.h); have_data == 0 -> constant ref_value. */
typedef struct {
    int32_t have_data;
    int32_t bounds_error;
    double ref_value;
    double xstart, ystart, dx, dy;  /* x = slant range (m), y = azimuth time (s) */
    int64_t width, length;
    const double* data;             /* host, length x width */
} bp_lut2d;

/* RadarGridParameters subset used by the path (product/RadarGridParameters.h). */
typedef struct {
    double sensing_start;        /* s, relative to the orbit epoch */
    double prf;                  /* Hz; azimuth spacing is 1/prf */
    double starting_range;       /* m */
    double range_pixel_spacing;  /* m */
    int64_t length;              /* lines (azimuth) */
    int64_t width;               /* samples (range) */
    int32_t look_side;           /* BP_LOOK_* */
    int32_t pad_;
} bp_grid;

/* isce3::container::RadarGeometry: grid + orbit + Doppler. */
typedef struct {
    bp_grid grid;
    bp_orbit orbit;
    bp_lut2d doppler;
} bp_radar_geometry;

/* DEMInterpolator (geometry/DEMInterpolator.h): constant height or a float
 * raster in the coordinates of `epsg` -- 4326 (lon/lat degrees), 4978
 * (geocentric), 32601-32660 / 32701-32760 (UTM north/south), 3031 / 3413
 * (polar stereographic), 6933 (EASE-Grid 2.0) -- interpolated by `method`. */
typedef struct {
    int32_t have_raster;
    int32_t epsg;
    int32_t method;      /* BP_DEM_* */
    int32_t pad_;
    double ref_height;
    double xstart, ystart, dx, dy;  /* raster coordinates of the projection */
    int64_t width, length;
    const float* data;   /* host, length x width */
} bp_dem;

/* TabulatedKernel<float> (core/Kernels.h:133-151). */
typedef struct {
    double halfwidth;    /* kernel.width() / 2 */
    int32_t n;           /* table size (>= 2) */
    int32_t pad_;
    const float* table;  /* host, n samples of K(x), x in [0, halfwidth] */
} bp_kernel;

/* Rdr2GeoBracketParams (geometry/detail/Rdr2Geo.h:87-97). */
typedef struct {
    double tol_height, look_min, look_max;
} bp_rdr2geo_params;

/* Geo2RdrBracketParams (geometry/detail/Geo2Rdr.h:57-68); the optional
 * times are present when have_* is nonzero. */
typedef struct {
    double tol_aztime;
    int32_t have_time_start, have_time_end;
    double time_start, time_end;
} bp_geo2rdr_params;

/* Largest `chunk`: between folds the split sum steps each pixel's phase by
 * its first two forward differences and adds the third and fourth at the
 * fold, so the phase it uses inside a fold interval of n pulses is short by
 * up to C(n,3) d3 (2e-6 cycles at n = 64 for a spaceborne aperture; the term
 * grows as n^3). */
#define BP_MAX_CHUNK 64

/* Launch/tiling options; zero-initialise for defaults. */
typedef struct {
    int32_t tile_rows;     /* output rows per 64-pixel wavefront tile: 1,2,4,8 (0 = 1) */
    int32_t chunk;         /* pulses per compensated FP32 partial sum and per
                              phase-step interval (split; 0 = 32; at most
                              BP_MAX_CHUNK) */
    int32_t fit_degree;    /* split: delay polynomial degree (0 = 4; only 4 is built) */
    int32_t order;         /* tile walk of the sum kernels: 0 auto, 1 row-major, 2 column-major */
    int32_t blocks_per_cu; /* persistent 512-thread workgroups per CU (0 = occupancy limit) */
    int32_t staged;        /* split: 0 gathers the taps from global memory (default),
                              1 stages each pulse's range window in LDS first */
    int32_t lane_rows;     /* output lines per lane of the split sum: 2 (adjacent lines
                              share their taps; kernel width 8, staged 0 only), 1, or 0 =
                              2 where offered, else 1 (default) */
    int32_t reserved[1];
} bp_options;

/* Optional per-call measurements (filled when non-null). */
typedef struct {
    int64_t updates;       /* sum over pixels of (kstop - kstart) */
    double geometry_ms;    /* device time of the geometry pass (HIP events; 0 in a
                              graph-replayed scene step, which is timed as a whole) */
    double sum_ms;         /* device time of the coherent-sum kernel (likewise) */
    int32_t any_failed;    /* some pixel failed to converge */
    int32_t pad_;
    double step_ms;        /* device time of the whole step */
} bp_stats;

typedef struct bp_context bp_context;

/* Create a context bound to a HIP device (holds device scratch buffers). */
int bp_context_create(int device, bp_context** ctx);
int bp_context_destroy(bp_context* ctx);

/* Thread-local message of the last failing call. */
const char* bp_last_error_message(void);

/* Device scratch bytes per output pixel that a context holds for a call of
 * the given precision (per-pixel geometry records, plus the split-precision
 * records); a context grows its scratch to pixels x this value and keeps it.
 * No reference counterpart: used by callers that budget device memory. */
int64_t bp_scratch_bytes_per_pixel(int precision);

/*
 * Focus in azimuth by time-domain backprojection.
 *
 * Replaces isce3::focus::backproject (cxx/isce3/focus/Backproject.h:37-47,
 * Backproject.cpp:66-212) and isce3::cuda::focus::backproject
 * (cxx/isce3/cuda/focus/Backproject.h:45-56).
 *
 *   in_dev      device, complex64 [in.grid.length x in.grid.width], row-major
 *   out_dev     device, complex64 [out.grid.length x out.grid.width], written
 *   height_dev  device, float32 [out shape] or NULL
 *   any_failed  host int, set nonzero when some pixel failed to converge
 *               (those pixels are NaN), as the reference's return value
 *   stream      hipStream_t or NULL for the null stream (synchronous)
 *
 * Returns BP_SUCCESS / BP_FAILED_TO_CONVERGE, or a negative error code.
 */
int bp_backproject(bp_context* ctx, const bp_radar_geometry* out_geometry,
                   const bp_radar_geometry* in_geometry, const bp_dem* dem, double fc,
                   double ds, const bp_kernel* kernel, int dry_tropo_model,
                   const bp_rdr2geo_params* rdr2geo, const bp_geo2rdr_params* geo2rdr,
                   const void* in_dev, void* out_dev, float* height_dev, int precision,
                   const bp_options* options, void* stream, int* any_failed,
                   bp_stats* stats);

/*
 * bp_backproject without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during a
 * call -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.
 *
 * Same arguments, checks, kernels and results as bp_backproject.  Every
 * device buffer of the call (orbit, Doppler, DEM and kernel tables, the
 * per-pulse states, the per-pixel records, the flags and the descriptor)
 * is placed in one fixed scratch array of BP_NOALLOC_SCRATCH_MIB MiB per
 * device, part of the library and held until the process ends; its device
 * address is looked up by bp_context_create.  A call whose buffers do not fit
 * (the records take bp_scratch_bytes_per_pixel(precision) bytes per output
 * pixel: 48 plain, 112 split, so about 2.7 M plain or 1.1 M split pixels with
 * small tables) fails with BP_ERR_LENGTH before any launch.  The call makes
 * no hipMalloc, hipFree, pinned or library allocation on the device; its
 * table uploads are pageable host-to-device copies on `stream`.
 *
 * All calls on one GPU share that GPU's array, and each call's launches read
 * what the call itself wrote there: calls that overlap in time on one GPU
 * (other streams, other contexts, other threads) race and must not be issued.
 */
#define BP_NOALLOC_SCRATCH_MIB 128
int bp_backproject_noalloc(bp_context* ctx, const bp_radar_geometry* out_geometry,
                           const bp_radar_geometry* in_geometry, const bp_dem* dem, double fc,
                           double ds, const bp_kernel* kernel, int dry_tropo_model,
                           const bp_rdr2geo_params* rdr2geo, const bp_geo2rdr_params* geo2rdr,
                           const void* in_dev, void* out_dev, float* height_dev, int precision,
                           const bp_options* options, void* stream, int* any_failed,
                           bp_stats* stats);

/* What bp_backproject_host moved (filled when non-null). */
typedef struct {
    int64_t h2d_bytes;      /* pulse bytes copied host to device */
    int64_t window_pulses;  /* pulses copied, summed over bands */
    int32_t bands;          /* bands of output lines focused (1 unless the window
                               exceeded max_window_bytes) */
    int32_t pad_;
    double copy_ms;         /* device time of the pulse copies (copy stream) */
} bp_host_stats;

/*
 * Focus from range-compressed pulses in host memory -- an array or a memory
 * map of a raster file -- moving only the pulses the output needs.
 *
 * Replaces isce3::cuda::focus::backproject with host input
 * (cxx/isce3/cuda/focus/Backproject.cu:641-688: the geometry first, then the
 * pulses [kstart_min, kstop_max) in batches of `batch`), as bound by
 * python/extensions/pybind_isce3/cuda/focus/Backproject.cpp:87-116 and called
 * per block with the whole raster by python/packages/nisar/workflows/
 * focus.py:2001.  The union [kmin, kmax) of the output pixels' pulse windows
 * is copied to the device in chunks of at most `batch` pulses (and at most
 * 64 MiB) through pinned staging on a copy stream; no pulse outside it is
 * read.  When that window exceeds max_window_bytes of device memory, the
 * block is focused in bands of output lines, each with its own window.  The
 * results equal bp_backproject's with the whole raster on the device, bit for
 * bit, in every case.
 *
 *   in_host            host, complex64 [in.grid.length x in.grid.width]
 *   batch              pulses per host-to-device copy (>= 1, else BP_ERR_DOMAIN)
 *   max_window_bytes   device bytes for pulse windows (0: half the free memory);
 *                      BP_ERR_LENGTH when the window of one tile row of output
 *                      lines exceeds it
 *   out_dev, height_dev  device buffers as in bp_backproject
 *   host_stats         bytes and pulses moved, or NULL
 */
int bp_backproject_host(bp_context* ctx, const bp_radar_geometry* out_geometry,
                        const bp_radar_geometry* in_geometry, const bp_dem* dem, double fc,
                        double ds, const bp_kernel* kernel, int dry_tropo_model,
                        const bp_rdr2geo_params* rdr2geo, const bp_geo2rdr_params* geo2rdr,
                        const void* in_host, int64_t batch, int64_t max_window_bytes,
                        void* out_dev, float* height_dev, int precision,
                        const bp_options* options, void* stream, int* any_failed,
                        bp_stats* stats, bp_host_stats* host_stats);

/* One output block of bp_backproject_blocks. */
typedef struct {
    bp_grid grid;        /* output radar grid of this block */
    void* out_dev;       /* device, complex64 [grid.length x grid.width], written */
    float* height_dev;   /* device, float32 [same shape] or NULL */
    int32_t failed;      /* set on return: some pixel of this block failed to converge */
    int32_t pad_;
} bp_out_block;

/*
 * Focus many output blocks of one range-compressed raw set in one pass: one
 * geometry launch and one coherent-sum launch over the pixels of all blocks
 * (blocks may differ in shape, start time, range and spacing).
 *
 * Replaces the per-block loop of isce3.focus.backproject calls in the RSLC
 * workflow (python/packages/nisar/workflows/focus.py:1986-2004), whose blocks
 * share the raw data, input geometry, output orbit and Doppler, DEM, kernel
 * and parameters.  Each block's result equals bp_backproject on that block
 * alone (the per-pixel arithmetic does not depend on the batch).
 *
 *   out_orbit, out_doppler  orbit and Doppler of every output block
 *   blocks[n_blocks]        output grids and device buffers; failed is set
 *
 * Returns BP_SUCCESS / BP_FAILED_TO_CONVERGE (some block failed), or a
 * negative error code.
 */
int bp_backproject_blocks(bp_context* ctx, const bp_orbit* out_orbit,
                          const bp_lut2d* out_doppler, bp_out_block* blocks, int n_blocks,
                          const bp_radar_geometry* in_geometry, const bp_dem* dem, double fc,
                          double ds, const bp_kernel* kernel, int dry_tropo_model,
                          const bp_rdr2geo_params* rdr2geo, const bp_geo2rdr_params* geo2rdr,
                          const void* in_dev, int precision, const bp_options* options,
                          void* stream, int* any_failed, bp_stats* stats);

/* ------------------------------------------------------------ scenes
 * A scene keeps one range-compressed raw set resident together with
 * everything its tile requests share: the input geometry, the output orbit
 * and Doppler, DEM, kernel tables and parameters (uploaded once), per-pixel
 * scratch for up to max_pixels output pixels in up to max_blocks blocks, and
 * pinned host staging.  bp_scene_focus focuses a batch of output blocks by
 * replaying a HIP graph of the whole step (descriptor upload, flag clear,
 * geometry pass, coherent sum, flag read-back), captured on first use of each
 * launch configuration.  The pulses at in_dev must stay valid and unchanged
 * for the scene's lifetime.  A scene is used from one host thread at a time.
 *
 * No single reference counterpart: a scene replaces the per-block calls of
 * isce3.focus.backproject that share one raw set in the RSLC workflow
 * (python/packages/nisar/workflows/focus.py:1986-2004), with the same
 * arguments and the same per-block results as bp_backproject_blocks.
 */
typedef struct bp_scene bp_scene;

int bp_scene_create(bp_context* ctx, const bp_orbit* out_orbit, const bp_lut2d* out_doppler,
                    const bp_radar_geometry* in_geometry, const bp_dem* dem, double fc,
                    double ds, const bp_kernel* kernel, int dry_tropo_model,
                    const bp_rdr2geo_params* rdr2geo, const bp_geo2rdr_params* geo2rdr,
                    const void* in_dev, int precision, int64_t max_pixels, int max_blocks,
                    bp_scene** scene);
int bp_scene_destroy(bp_scene* scene);

/* Focus blocks[n_blocks] (at most max_pixels pixels in at most max_blocks
 * non-empty blocks) on `stream`; returns after the step has finished, like
 * bp_backproject_blocks, whose return values and per-block flags it shares. */
int bp_scene_focus(bp_scene* scene, bp_out_block* blocks, int n_blocks,
                   const bp_options* options, void* stream, int* any_failed, bp_stats* stats);

/* ------------------------------------------------------------ serving
 * A continuous-batching server over one scene: requests (one output block
 * each) are queued by bp_serve_submit from any host thread; a worker thread
 * takes every pending request that fits the scene (max_pixels, max_blocks),
 * focuses them with one bp_scene_focus call on its own stream, completes
 * their tickets and starts on the requests that arrived meanwhile.  The
 * server owns the scene while it exists (no other calls on the scene).
 * bp_server_destroy finishes every accepted request first.
 *
 * No reference counterpart: the tile-serving entry points of the build's
 * serving workload (many tile requests against resident raw sets). */
typedef struct bp_server bp_server;

int bp_server_create(bp_scene* scene, const bp_options* options, int max_pending,
                     bp_server** server);
int bp_server_destroy(bp_server* server);
/* Queue one block (its out_dev / height_dev device buffers must stay valid
 * until the ticket completes).  The block is checked as bp_backproject_blocks
 * checks its blocks (shape, PRF > 0, output buffer): a bad request is refused
 * here with BP_ERR_INVALID_ARGUMENT and never joins a batch.  BP_ERR_LENGTH
 * when max_pending requests are waiting or the block alone exceeds the
 * scene's max_pixels. */
int bp_serve_submit(bp_server* server, const bp_out_block* block, uint64_t* ticket);
/* Non-blocking: *done = 1 once the request has been focused (its result code
 * returned, *failed = some pixel failed to converge), else 0.  A completed
 * ticket is forgotten after it has been reported. */
int bp_serve_poll(bp_server* server, uint64_t ticket, int* done, int* failed);
/* Block until the request has been focused; returns its result code.  A
 * ticket has one consumer: when another thread's wait or poll reports it
 * first, this wait returns BP_ERR_INVALID_ARGUMENT ("unknown ticket").
 * bp_server_destroy lets every thread blocked here return before it frees the
 * server; calls that start after destroy has begun are not allowed. */
int bp_serve_wait(bp_server* server, uint64_t ticket, int* failed);

/* ---------------------------------------------------- host-side helpers
 * Host implementations of the same FP64 primitives the kernels use; they back
 * the Python mirror's Orbit.interpolate, Ellipsoid and DEM/LUT lookups. */

/* isce3::core::Orbit::interpolate (core/Orbit.cpp:71-86); border: 0 error,
 * 1 extrapolate, 2 fill NaN.  Returns an ErrorCode value. */
int bp_orbit_interpolate(const bp_orbit* orbit, double t, int border, double* pos,
                         double* vel);
/* Ellipsoid::lonLatToXyz / xyzToLonLat (core/Ellipsoid.h:176-224). */
void bp_lonlat_to_xyz(double a, double e2, const double* llh, double* xyz);
void bp_xyz_to_lonlat(double a, double e2, const double* xyz, double* llh);
/* DEMInterpolator::interpolateLonLat / interpolateXY (DEMInterpolator.cpp:593-662). */
double bp_dem_interpolate_lonlat(const bp_dem* dem, double lon, double lat);
double bp_dem_interpolate_xy(const bp_dem* dem, double x, double y);
/* ProjectionBase::forward / inverse of the EPSG code (core/Projections.cpp):
 * (lon, lat, h) in radians <-> projected (x, y, h).  Return 0, 1 where the
 * projection reports failure, BP_ERR_INVALID_ARGUMENT for an unknown code. */
int bp_proj_forward(int epsg, const double* llh, double* xyz);
int bp_proj_inverse(int epsg, const double* xyz, double* llh);
/* LUT2d<double>::eval (core/LUT2d.cpp:127-159). */
double bp_lut2d_eval(const bp_lut2d* lut, double y, double x);
/* sin and cos of 2 pi t in FP64 for a phase t in turns: the phase rotation
 * exp(j 2 pi fc tau) of the plain sum (Backproject.cpp:53-55 forms cos/sin of
 * 2 pi fc tau in double). */
void bp_sincos_turns(double t, double* s, double* c);
/* KnabKernel<double>::operator() (core/Kernels.icc:29-52). */
double bp_knab_eval(double width, double bandwidth, double x);
/* TabulatedKernel<float> table fill from a Knab kernel (core/Kernels.icc:113-135). */
int bp_tabulate_knab(double width, double bandwidth, int n, float* table);
/* rdr2geo_bracket / geo2rdr_bracket on the host (geometry/rdr2geo_roots.h:34-40,
 * geometry/geo2rdr_roots.h:28-33); return ErrorCode values. */
int bp_rdr2geo_bracket(double aztime, double slant_range, double doppler,
                       const bp_orbit* orbit, const bp_dem* dem, double wavelength,
                       int side, const bp_rdr2geo_params* params, double* xyz);
int bp_geo2rdr_bracket(const double* xyz, const bp_orbit* orbit, const bp_lut2d* doppler,
                       double wavelength, int side, const bp_geo2rdr_params* params,
                       double* aztime, double* range);

#ifdef __cplusplus
}
#endif

#endif /* ISCE3_AMD_BACKPROJECT_H */
