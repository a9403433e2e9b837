// Host+device FP64 geometry primitives for the backprojection path.
//
// Each routine restates the reference behaviour it cites (paths relative to
// the isce3 source tree).  They are written once as __host__ __device__ so
// the same code serves the GPU kernels and the host-side helpers exposed
// through the C-ABI (orbit interpolation, ellipsoid, DEM and LUT lookups).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BP_HD __host__ __device__ inline

namespace bp {

constexpr double kC = 299792458.0;               // core/Constants.h:50
constexpr double kWgsA = 6378137.0;              // core/Constants.h:41
constexpr double kWgsE2 = 0.006694379990141317;  // core/Constants.h:44
constexpr double kPi = 3.14159265358979323846;

// isce3::error::ErrorCode values (error/ErrorCode.h:8-21)
enum Err : int {
    kSuccess = 0,
    kOrbitInterpSizeError = 1,
    kOrbitInterpDomainError = 2,
    kOrbitInterpUnknownMethod = 3,
    kOutOfBoundsDem = 4,
    kInvalidDem = 5,
    kFailedToConverge = 6,
    kWrongLookSide = 7,
    kOutOfBoundsLookup = 8,
    kNullDereference = 9,
    kInvalidTolerance = 10,
    kInvalidInterval = 11,
};

enum Border : int { kBorderError = 0, kBorderExtrapolate = 1, kBorderFillNaN = 2 };
enum InterpMethod : int { kHermite = 0, kLegendre = 1 };
enum LookSideV : int { kLookLeft = 1, kLookRight = -1 };

struct V3 {
    double x, y, z;
};
BP_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
BP_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
BP_HD V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
BP_HD V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
BP_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BP_HD double norm(V3 a) { return sqrt(dot(a, a)); }
BP_HD V3 normalized(V3 a)
{
    double n = norm(a);
    return {a.x / n, a.y / n, a.z / n};
}
BP_HD V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
BP_HD V3 ld3(const double* p) { return {p[0], p[1], p[2]}; }
BP_HD double qnan() { return __builtin_nan(""); }

// Uniformly sampled orbit (state vectors every dt seconds from t0).
struct Orbit {
    double t0, dt;
    int32_t n, method;
    const double* pos;  // n x 3
    const double* vel;  // n x 3
};

// Bilinear 2-D lookup table (Doppler), or a constant when have_data == 0.
struct Lut2d {
    int32_t have_data, bounds_error;
    double ref_value;
    double xstart, ystart, dx, dy;
    int64_t width, length;
    const double* data;
};

// DEM: constant ref_height, or an EPSG:4326 float raster (bilinear).
struct Dem {
    int32_t have_raster, epsg, method, pad_;
    double ref_height;
    double xstart, ystart, dx, dy;
    int64_t width, length;
    const float* data;
};

// core/Linspace.icc:77-89
BP_HD int linspace_search(double first, double spacing, int size, double val)
{
    double last = first + (size - 1) * spacing;
    if (spacing >= 0) {
        if (val < first) return 0;
        if (val > last) return size;
    } else {
        if (val > first) return 0;
        if (val < last) return size;
    }
    return static_cast<int>((val - first) / spacing + 1);
}

BP_HD double orbit_time(const Orbit& o, int i) { return o.t0 + i * o.dt; }
BP_HD double orbit_start(const Orbit& o) { return o.t0; }
BP_HD double orbit_end(const Orbit& o) { return orbit_time(o, o.n - 1); }

// 4-point Hermite interpolation with position and velocity samples
// (core/detail/InterpolateOrbit.icc:15-109).
BP_HD void orbit_hermite(const Orbit& o, double t, V3* pos, V3* vel)
{
    int idx = linspace_search(o.t0, o.dt, o.n, t) - 2;
    idx = idx < 0 ? 0 : idx;
    idx = idx > o.n - 4 ? o.n - 4 : idx;
    double ti[4], f0[4], f1[4], h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        ti[i] = orbit_time(o, idx + i);
        f1[i] = t - ti[i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double sum = 0.;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j != i) sum += 1. / (ti[i] - ti[j]);
        f0[i] = 1. - 2. * sum * (t - ti[i]);
        double hi = 1.;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j != i) hi *= (t - ti[j]) / (ti[i] - ti[j]);
        h[i] = hi;
    }
    if (pos) {
        V3 p{0., 0., 0.};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            V3 pi = ld3(o.pos + 3 * (idx + i));
            V3 vi = ld3(o.vel + 3 * (idx + i));
            double hh = h[i] * h[i];
            V3 term{pi.x * f0[i] + vi.x * f1[i], pi.y * f0[i] + vi.y * f1[i],
                    pi.z * f0[i] + vi.z * f1[i]};
            p = p + hh * term;
        }
        *pos = p;
    }
    if (!vel) return;
    double hdot[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double acc = 0.;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            double prod = 1. / (ti[i] - ti[j]);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k != i && k != j) prod *= (t - ti[k]) / (ti[i] - ti[k]);
            acc += prod;
        }
        hdot[i] = acc;
    }
    V3 v{0., 0., 0.};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double g1 = h[i] + 2. * hdot[i] * (t - ti[i]);
        double sum = 0.;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j != i) sum += 1. / (ti[i] - ti[j]);
        double g0 = 2. * (f0[i] * hdot[i] - sum * h[i]);
        V3 pi = ld3(o.pos + 3 * (idx + i));
        V3 vi = ld3(o.vel + 3 * (idx + i));
        V3 term{pi.x * g0 + vi.x * g1, pi.y * g0 + vi.y * g1, pi.z * g0 + vi.z * g1};
        v = v + h[i] * term;
    }
    *vel = v;
}

// 9-point Legendre interpolation (core/detail/InterpolateOrbit.icc:111-155).
BP_HD void orbit_legendre(const Orbit& o, double t, V3* pos, V3* vel)
{
    int idx = linspace_search(o.t0, o.dt, o.n, t) - 5;
    idx = idx < 0 ? 0 : idx;
    idx = idx > o.n - 9 ? o.n - 9 : idx;
    double trel = 8. * (t - orbit_time(o, idx)) / (orbit_time(o, idx + 8) - orbit_time(o, idx));
    double teller = 1.;
    for (int i = 0; i < 9; ++i) teller *= trel - i;
    if (teller == 0.) {
        int i = static_cast<int>(trel);
        if (pos) *pos = ld3(o.pos + 3 * (idx + i));
        if (vel) *vel = ld3(o.vel + 3 * (idx + i));
        return;
    }
    const double noemer[9] = {40320.0, -5040.0, 1440.0, -720.0, 576.0,
                              -720.0, 1440.0, -5040.0, 40320.0};
    V3 p{0., 0., 0.}, v{0., 0., 0.};
    for (int i = 0; i < 9; ++i) {
        double coeff = (teller / noemer[i]) / (trel - i);
        p = p + coeff * ld3(o.pos + 3 * (idx + i));
        v = v + coeff * ld3(o.vel + 3 * (idx + i));
    }
    if (pos) *pos = p;
    if (vel) *vel = v;
}

// Border handling and method dispatch (core/detail/InterpolateOrbit.icc:157-193).
BP_HD int orbit_interpolate(const Orbit& o, double t, int border, V3* pos, V3* vel)
{
    int minsv = o.method == kHermite ? 4 : (o.method == kLegendre ? 9 : -1);
    if (o.n < minsv) return kOrbitInterpSizeError;
    if (t < orbit_start(o) || t > orbit_end(o)) {
        if (border == kBorderFillNaN) {
            V3 nan3{qnan(), qnan(), qnan()};
            if (pos) *pos = nan3;
            if (vel) *vel = nan3;
        }
        if (border != kBorderExtrapolate) return kOrbitInterpDomainError;
    }
    if (o.method == kHermite) {
        orbit_hermite(o, t, pos, vel);
        return kSuccess;
    }
    if (o.method == kLegendre) {
        orbit_legendre(o, t, pos, vel);
        return kSuccess;
    }
    return kOrbitInterpUnknownMethod;
}

// core/Ellipsoid.h:176-190 (with rEast, :117-120)
BP_HD V3 lonlat_to_xyz(double a, double e2, V3 llh)
{
    double sl = sin(llh.y);
    double cl = cos(llh.y);
    double re = a / sqrt(1.0 - (e2 * (sl * sl)));
    return {(re + llh.z) * cl * cos(llh.x), (re + llh.z) * cl * sin(llh.x),
            ((re * (1.0 - e2)) + llh.z) * sl};
}

// Vermeille (2002) xyz -> lon/lat/h (core/Ellipsoid.h:196-224)
BP_HD V3 xyz_to_lonlat(double a, double e2, V3 xyz)
{
    const double e4 = e2 * e2;
    const double a2 = a * a;
    const double xy2 = xyz.x * xyz.x + xyz.y * xyz.y;
    double p = xy2 / a2;
    double q = ((1. - e2) * (xyz.z * xyz.z)) / a2;
    double r = (p + q - e4) / 6.;
    // pow(r, 3) and pow(., 1/3) of the reference as a product and a cube
    // root (agree to a few ulp; the FP64 pow costs ~150 instructions on the GPU)
    double s = (e4 * p * q) / (4. * (r * r * r));
    double t = cbrt(1. + s + sqrt(s * (2. + s)));
    double u = r * (1. + t + (1. / t));
    double rv = sqrt(u * u + (e4 * q));
    double w = (e2 * (u + rv - q)) / (2. * rv);
    double k = sqrt(u + rv + w * w) - w;
    double d = (k * sqrt(xy2)) / (k + e2);
    V3 llh;
    llh.y = atan2(xyz.z, d);
    llh.x = atan2(xyz.y, xyz.x);
    llh.z = ((k + e2 - 1.) * sqrt(d * d + xyz.z * xyz.z)) / k;
    return llh;
}

// core/Ellipsoid.h:151-157
BP_HD V3 nvector(double lon, double lat)
{
    double clat = cos(lat);
    return {clat * cos(lon), clat * sin(lon), sin(lat)};
}

// Bilinear on a float raster with float weights
// (core/BilinearInterpolator.cpp:13-47, U = float).
BP_HD float bilinear_f32(double x, double y, const float* z, int64_t width)
{
    int x1 = static_cast<int>(floor(x));
    int x2 = static_cast<int>(ceil(x));
    int y1 = static_cast<int>(floor(y));
    int y2 = static_cast<int>(ceil(y));
    float q11 = z[(int64_t)y1 * width + x1];
    float q12 = z[(int64_t)y2 * width + x1];
    float q21 = z[(int64_t)y1 * width + x2];
    float q22 = z[(int64_t)y2 * width + x2];
    if ((y1 == y2) && (x1 == x2)) return q11;
    if (y1 == y2)
        return (static_cast<float>((x2 - x) / (x2 - x1)) * q11) +
               (static_cast<float>((x - x1) / (x2 - x1)) * q21);
    if (x1 == x2)
        return (static_cast<float>((y2 - y) / (y2 - y1)) * q11) +
               (static_cast<float>((y - y1) / (y2 - y1)) * q12);
    const float den = static_cast<float>((x2 - x1) * (y2 - y1));
    return ((q11 * static_cast<float>((x2 - x) * (y2 - y))) / den) +
           ((q21 * static_cast<float>((x - x1) * (y2 - y))) / den) +
           ((q12 * static_cast<float>((x2 - x) * (y - y1))) / den) +
           ((q22 * static_cast<float>((x - x1) * (y - y1))) / den);
}

// Same formula on a double table (LUT2d<double>).
BP_HD double bilinear_f64(double x, double y, const double* z, int64_t width)
{
    int x1 = static_cast<int>(floor(x));
    int x2 = static_cast<int>(ceil(x));
    int y1 = static_cast<int>(floor(y));
    int y2 = static_cast<int>(ceil(y));
    double q11 = z[(int64_t)y1 * width + x1];
    double q12 = z[(int64_t)y2 * width + x1];
    double q21 = z[(int64_t)y1 * width + x2];
    double q22 = z[(int64_t)y2 * width + x2];
    if ((y1 == y2) && (x1 == x2)) return q11;
    if (y1 == y2) return (((x2 - x) / (x2 - x1)) * q11) + (((x - x1) / (x2 - x1)) * q21);
    if (x1 == x2) return (((y2 - y) / (y2 - y1)) * q11) + (((y - y1) / (y2 - y1)) * q12);
    const double den = static_cast<double>((x2 - x1) * (y2 - y1));
    return ((q11 * ((x2 - x) * (y2 - y))) / den) + ((q21 * ((x - x1) * (y2 - y))) / den) +
           ((q12 * ((x2 - x) * (y - y1))) / den) + ((q22 * ((x - x1) * (y - y1))) / den);
}

// DEM lookup in native (lon/lat degree) coordinates
// (geometry/DEMInterpolator.cpp:615-662).
BP_HD double dem_interpolate_xy(const Dem& dem, double x, double y)
{
    if (!dem.have_raster) return dem.ref_height;
    if (dem.epsg == 4326 && (x > 360 || x < -360)) x = fmod(x, 360.0);
    if (dem.epsg == 4326 && x < -180) x += 360;
    if (dem.epsg == 4326 && x - 360 >= dem.xstart) {
        x -= 360;
    } else if (dem.epsg == 4326 && x < dem.xstart && x + 360 >= dem.xstart) {
        x += 360;
    } else if (x < dem.xstart) {
        return dem.ref_height;
    }
    const double row = (y - dem.ystart) / dem.dy;
    const double col = (x - dem.xstart) / dem.dx;
    const int irow = static_cast<int>(floor(row));
    const int icol = static_cast<int>(floor(col));
    if (irow < 2 || irow >= int(dem.length - 1)) return dem.ref_height;
    if (icol < 2 || icol >= int(dem.width - 1)) return dem.ref_height;
    return bilinear_f32(col, row, dem.data, dem.width);
}

// geometry/DEMInterpolator.cpp:593-610 through the LonLat projection
// (core/Projections.h:127-133: radians -> degrees).
BP_HD double dem_interpolate_lonlat(const Dem& dem, double lon, double lat)
{
    if (!dem.have_raster) return dem.ref_height;
    return dem_interpolate_xy(dem, lon * 180.0 / kPi, lat * 180.0 / kPi);
}

// core/LUT2d.cpp:127-159 (bilinear, clamped to the table).
BP_HD double lut2d_eval(const Lut2d& l, double y, double x)
{
    if (!l.have_data) return l.ref_value;
    double xi = (x - l.xstart) / l.dx;
    double yi = (y - l.ystart) / l.dy;
    xi = fmin(fmax(xi, 0.0), l.width - 1.0);
    yi = fmin(fmax(yi, 0.0), l.length - 1.0);
    return bilinear_f64(xi, yi, l.data, l.width);
}

// math/RootFind1dBracket.icc:16-29
BP_HD bool opposite_sign(double a, double b) { return __builtin_signbit(a) != __builtin_signbit(b); }

// Brent's bracketing root finder, following the step-selection rules of
// math/RootFind1dBracket.icc:56-216 so that iterates agree with the CPU.
template<typename F>
BP_HD int find_zero_brent(double a, double b, F& f, const double tol, double* root)
{
    const double eps = 2.220446049250313e-16;
    if (tol < 0.) return kInvalidTolerance;
    double c, d, e, fa, fb, fc, p, q, r, s, tol1;
    fa = f(a);
    if (fa == 0.) {
        *root = a;
        return kSuccess;
    }
    fb = f(b);
    if (fb == 0.) {
        *root = b;
        return kSuccess;
    }
    if (!opposite_sign(fa, fb)) return kInvalidInterval;
    c = a;
    fc = fa;
    e = d = b - a;
    tol1 = tol > 0. ? tol : eps;
    const int maxiter = 3 * static_cast<int>(ceil(log2(fabs((a - b) / tol1))));
    for (int i = 0; i < maxiter; ++i) {
        if (fabs(fc) < fabs(fb)) {
            a = b;
            b = c;
            c = a;
            fa = fb;
            fb = fc;
            fc = fa;
        }
        tol1 = 2 * eps * fabs(b) + 0.5 * tol;
        const double xm = 0.5 * (c - b);
        if ((fabs(xm) <= tol1) || (fb == 0.)) {
            *root = b;
            return kSuccess;
        }
        if ((fabs(e) < tol1) || (fabs(fa) <= fabs(fb))) {
            e = d = xm;
        } else {
            s = fb / fa;
            if (a == c) {
                p = 2 * xm * s;
                q = 1. - s;
            } else {
                q = fa / fc;
                r = fb / fc;
                p = s * (2 * xm * q * (q - r) - (b - a) * (r - 1.));
                q = (q - 1.) * (r - 1.) * (s - 1.);
            }
            if (p > 0.)
                q = -q;
            else
                p = -p;
            s = e;
            e = d;
            if (((2 * p) >= (3 * xm * q - fabs(tol1 * q))) || (p >= fabs(0.5 * s * q))) {
                e = d = xm;
            } else {
                d = p / q;
            }
        }
        a = b;
        fa = fb;
        if (fabs(d) <= tol1) {
            b = (xm <= 0.) ? b - tol1 : b + tol1;
        } else {
            b = b + d;
        }
        fb = f(b);
        if (!opposite_sign(fb, fc)) {
            c = a;
            fc = fa;
            e = d = b - a;
        }
    }
    *root = b;
    return kFailedToConverge;
}

// Two-way bistatic delay (focus/BistaticDelay.icc:9-17).
BP_HD double bistatic_delay(V3 p, V3 v, V3 x)
{
    V3 r = x - p;
    return 2. * (dot(r, v) - kC * norm(r)) / (dot(v, v) - (kC * kC));
}

// TSX dry troposphere delay (focus/DryTroposphereModel.icc:9-29).
BP_HD double dry_tropo_tsx(V3 p, V3 llh)
{
    const double zpd = 2.3, H = 6000.;
    V3 x = lonlat_to_xyz(kWgsA, kWgsE2, llh);
    V3 r_hat = normalized(p - x);
    V3 n_hat = normalized(nvector(llh.x, llh.y));
    double cos_theta = dot(r_hat, n_hat);
    return 2. * zpd * exp(-llh.z / H) / (kC * cos_theta);
}

struct R2GParams {
    double tol_height, look_min, look_max;
};
struct G2RParams {
    double tol_aztime;
    int32_t have_time_start, have_time_end;
    double time_start, time_end;
};

// rdr2geo by Brent on the pseudo-look angle against the DEM
// (geometry/detail/Rdr2Geo.icc:172-242).  Returns kOrbitInterpDomainError
// when aztime is outside the orbit (the CPU raises OutOfRange there).
BP_HD int rdr2geo_bracket(double aztime, double slant_range, double doppler, const Orbit& orbit,
                          const Dem& dem, double wavelength, int side, const R2GParams& params,
                          V3* xyz)
{
    V3 radar, velocity;
    int st = orbit_interpolate(orbit, aztime, kBorderError, &radar, &velocity);
    if (st != kSuccess) return st;
    const double speed = norm(velocity);
    const V3 along{velocity.x / speed, velocity.y / speed, velocity.z / speed};
    const V3 right = normalized(cross(along, radar));
    const V3 down = cross(along, right);
    const V3 horizontal = (side == kLookRight) ? right : neg(right);
    const double sin_squint = doppler * wavelength / (2 * speed);
    const double cos_squint = sqrt(1.0 - sin_squint * sin_squint);
    const V3 center = radar + (sin_squint * slant_range) * along;
    const double radius = cos_squint * slant_range;
    auto get_xyz = [&](double look) {
        double sl, cl;
        sincos(look, &sl, &cl);
        const double hs = radius * sl;
        const double dc = radius * cl;
        return V3{center.x + hs * horizontal.x + dc * down.x,
                  center.y + hs * horizontal.y + dc * down.y,
                  center.z + hs * horizontal.z + dc * down.z};
    };
    auto dh = [&](double look) {
        V3 llh = xyz_to_lonlat(kWgsA, kWgsE2, get_xyz(look));
        return llh.z - dem_interpolate_lonlat(dem, llh.x, llh.y);
    };
    const double tol_look = params.tol_height / radius;
    double look = 0.0;
    const int err = find_zero_brent(params.look_min, params.look_max, dh, tol_look, &look);
    if (err != kSuccess) return err;
    *xyz = get_xyz(look);
    return kSuccess;
}

// geo2rdr by Brent on azimuth time against the Doppler model
// (geometry/detail/Geo2Rdr.icc:182-238).
BP_HD int geo2rdr_bracket(V3 x, const Orbit& orbit, const Lut2d& doppler, double wavelength,
                          int side, const G2RParams& params, double* aztime, double* range)
{
    double t0, t1;
    if (params.have_time_start)
        t0 = params.time_start;
    else
        t0 = doppler.have_data ? fmax(orbit_start(orbit), doppler.ystart) : orbit_start(orbit);
    if (params.have_time_end) {
        t1 = params.time_end;
    } else if (doppler.have_data) {
        const double dop_end = doppler.ystart + doppler.dy * (doppler.length - 1);
        t1 = fmin(orbit_end(orbit), dop_end);
    } else {
        t1 = orbit_end(orbit);
    }
    V3 xp, v, r;
    auto doppler_error = [&](double t) {
        orbit_interpolate(orbit, t, kBorderFillNaN, &xp, &v);
        r = x - xp;
        const double rnorm = norm(r);
        double fd = lut2d_eval(doppler, t, rnorm);
        return 2.0 / wavelength * dot(v, r) / rnorm - fd;
    };
    const int err = find_zero_brent(t0, t1, doppler_error, params.tol_aztime, aztime);
    if (err != kSuccess) return err;
    orbit_interpolate(orbit, *aztime, kBorderFillNaN, &xp, &v);
    r = x - xp;
    *range = norm(r);
    if ((side == kLookRight) != (dot(cross(r, v), xp) > 0)) return kWrongLookSide;
    return kSuccess;
}

// Knab window x sinc (core/Kernels.icc:29-52, math/Sinc.icc:68-91).  Only
// real arguments |t| <= halfwidth are tabulated, where the complex square
// root of the reference is real.
BP_HD double knab_eval(double t, double halfwidth, double bandwidth)
{
    const double c = kPi * halfwidth * (1.0 - bandwidth);
    const double tf = t / halfwidth;
    const double arg = 1.0 - tf * tf;
    double window;
    if (arg >= 0.0) {
        window = cosh(c * sqrt(arg)) / cosh(c);
    } else {
        // cosh(j*y) = cos(y) for the imaginary root outside the support
        window = cos(c * sqrt(-arg)) / cosh(c);
    }
    const double eps1 = 0.1490116119384765625e-7, eps2 = 0.0001220703125;
    double xs = kPi * fabs(t);
    double sc;
    if (xs < eps2) {
        sc = 1.;
        if (xs > eps1) sc -= xs * xs / 6.;
    } else {
        sc = sin(xs) / xs;
    }
    return window * sc;
}

}  // namespace bp
