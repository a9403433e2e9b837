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

// Map projection of a DEM raster (core/Projections.h, Projections.cpp), WGS84;
// every parameter follows from the EPSG code (proj_init).
enum ProjKind : int { kProjLonLat = 0, kProjGeocent = 1, kProjUtm = 2, kProjPolar = 3, kProjCea = 4 };
struct Proj {
    int32_t kind, north;
    double lon0;
    // UTM: transverse Mercator by Krueger series (Poder/Engsager)
    double qn, zb;
    double cgb[6], cbg[6], utg[6], gtu[6];
    // polar stereographic
    double akm1, e;
    // cylindrical equal area (EASE-Grid 2.0)
    double k0, one_es, qp;
    double apa[3];
};

// DEM interpolation methods: isce3::core::dataInterpMethod (core/Constants.h:23-29)
enum DemMethod : int {
    kDemSinc = 0, kDemBilinear = 1, kDemBicubic = 2, kDemNearest = 3, kDemBiquintic = 4
};

// DEM: constant ref_height, or a float raster in the projection of `epsg`
// interpolated by `method`.
struct Dem {
    int32_t have_raster, epsg, method, pad_;
    double ref_height;
    double xstart, ystart, dx, dy;
    int64_t width, length;
    const float* data;
    const float* sinc;   // kDemSinc: kernel [kSincSub][kSincLen] (sinc_kernel_table)
    Proj proj;
};

// The DEM factory's windowed-sinc interpolator (core/Interpolator.h:203-221,
// core/Constants.h:32-35): 8 taps, 8192 fractional rows.
constexpr int kSincLen = 8, kSincSub = 8192;

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

// The same interpolation with its divisions by the node differences t_i - t_j
// taken as products with the six distinct reciprocals (1 / (t_j - t_i) =
// -1 / (t_i - t_j) exactly): 30 % fewer instructions, but each quotient
// (t - t_j) / (t_i - t_j) may differ from the reference's by one rounding, and
// the velocity's cancelling terms amplify that to ~1e-14 relative.  Root
// finders fed by it can take another path to a root within their tolerance
// (micrometres for rdr2geo), so only the split-precision phase fit, which finds
// no roots and whose 1e-3 rad budget such errors do not touch, uses it.
BP_HD void orbit_hermite_reciprocal(const Orbit& o, double t, V3* pos, V3* vel)
{
    int idx = linspace_search(o.t0, o.dt, o.n, t) - 2;
    idx = idx < 0 ? 0 : idx;
    idx = idx > o.n - 4 ? o.n - 4 : idx;
    double ti[4], f1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        ti[i] = orbit_time(o, idx + i);
        f1[i] = t - ti[i];
    }
    double r[4][4];  // r[i][j] = 1 / (t_i - t_j), i != j
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r[i][i] = 0.;
#pragma unroll
        for (int j = i + 1; j < 4; ++j) {
            r[i][j] = 1. / (ti[i] - ti[j]);
            r[j][i] = -r[i][j];
        }
    }
    double S[4], f0[4], h[4], q[4][4];  // q[i][j] = (t - t_j) / (t_i - t_j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double sum = 0., hi = 1.;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            sum += r[i][j];
            q[i][j] = f1[j] * r[i][j];
            hi *= q[i][j];
        }
        S[i] = sum;
        f0[i] = 1. - 2. * sum * f1[i];
        h[i] = hi;
    }
    V3 pn[4], vn[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pn[i] = ld3(o.pos + 3 * (idx + i));
        vn[i] = ld3(o.vel + 3 * (idx + i));
    }
    if (pos) {
        V3 p{0., 0., 0.};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double hh = h[i] * h[i];
            V3 term{pn[i].x * f0[i] + vn[i].x * f1[i], pn[i].y * f0[i] + vn[i].y * f1[i],
                    pn[i].z * f0[i] + vn[i].z * f1[i]};
            p = p + hh * term;
        }
        *pos = p;
    }
    if (!vel) return;
    V3 v{0., 0., 0.};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        // hdot_i = sum_j 1 / (t_i - t_j) prod_{k != i, j} (t - t_k) / (t_i - t_k)
        double hdot = 0.;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            double prod = r[i][j];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k != i && k != j) prod *= q[i][k];
            hdot += prod;
        }
        const double g1 = h[i] + 2. * hdot * f1[i];
        const double g0 = 2. * (f0[i] * hdot - S[i] * h[i]);
        V3 term{pn[i].x * g0 + vn[i].x * g1, pn[i].y * g0 + vn[i].y * g1,
                pn[i].z * g0 + vn[i].z * g1};
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
// kReciprocal: Hermite by orbit_hermite_reciprocal (the split-precision fit).
template<bool kReciprocal = false>
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
        if (kReciprocal)
            orbit_hermite_reciprocal(o, t, pos, vel);
        else
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
// (core/BilinearInterpolator.cpp:13-47, U = float).  floor and ceil of a
// coordinate differ by 0 or exactly 1, so the formula's divisions by x2 - x1,
// y2 - y1 and their product, where it reaches them, divide by 1 and are left
// out (the same bits; products are not fused into the sums, as the
// divisions between them kept them apart).
BP_HD float bilinear_f32(double x, double y, const float* z, int64_t width)
{
#pragma clang fp contract(off)
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
        return (static_cast<float>(x2 - x) * q11) + (static_cast<float>(x - x1) * q21);
    if (x1 == x2)
        return (static_cast<float>(y2 - y) * q11) + (static_cast<float>(y - y1) * q12);
    return (q11 * static_cast<float>((x2 - x) * (y2 - y))) +
           (q21 * static_cast<float>((x - x1) * (y2 - y))) +
           (q12 * static_cast<float>((x2 - x) * (y - y1))) +
           (q22 * static_cast<float>((x - x1) * (y - y1)));
}

// Same formula on a double table (LUT2d<double>), the divisions by 1 left
// out in the same way.
BP_HD double bilinear_f64(double x, double y, const double* z, int64_t width)
{
#pragma clang fp contract(off)
    int x1 = static_cast<int>(floor(x));
    int x2 = static_cast<int>(ceil(x));
    int y1 = static_cast<int>(floor(y));
    int y2 = static_cast<int>(ceil(y));
    double q11 = z[(int64_t)y1 * width + x1];
    double q12 = z[(int64_t)y2 * width + x1];
    double q21 = z[(int64_t)y1 * width + x2];
    double q22 = z[(int64_t)y2 * width + x2];
    if ((y1 == y2) && (x1 == x2)) return q11;
    if (y1 == y2) return ((x2 - x) * q11) + ((x - x1) * q21);
    if (x1 == x2) return ((y2 - y) * q11) + ((y - y1) * q12);
    return (q11 * ((x2 - x) * (y2 - y))) + (q21 * ((x - x1) * (y2 - y))) +
           (q12 * ((x2 - x) * (y - y1))) + (q22 * ((x - x1) * (y - y1)));
}

// ------------------------------------------------------------ projections
// Clenshaw summations of core/Projections.cpp:23-77: real (sin series at
// `arg`) and complex (at arg_r + i arg_i, real part returned, imaginary in *im).
BP_HD double clenshaw_sin(const double* a, int n, double arg)
{
    const double c2 = 2. * cos(arg);
    double h = 0., h1 = a[n - 1], h2 = 0.;
    for (int k = n - 2; k >= 0; --k) {
        h = -h2 + c2 * h1 + a[k];
        h2 = h1;
        h1 = h;
    }
    if (n == 1) h = h1;
    return sin(arg) * h;
}

BP_HD double clenshaw_sin_complex(const double* a, int n, double ar, double ai, double* im)
{
    const double sr = sin(ar), cr = cos(ar), sh = sinh(ai), ch = cosh(ai);
    const double r = 2. * cr * ch, i = -2. * sr * sh;
    double hr = 0., hi = 0., hr1 = a[n - 1], hi1 = 0., hr2 = 0., hi2 = 0.;
    for (int k = n - 2; k >= 0; --k) {
        hr = -hr2 + (r * hr1 - i * hi1) + a[k];
        hi = -hi2 + (i * hr1 + r * hi1);
        hr2 = hr1;
        hi2 = hi1;
        hr1 = hr;
        hi1 = hi;
    }
    if (n == 1) {
        hr = hr1;
        hi = hi1;
    }
    // (sin(ar) cosh(ai) + i cos(ar) sinh(ai)) * (hr + i hi)
    const double zr = sr * ch, zi = cr * sh;
    *im = zr * hi + zi * hr;
    return zr * hr - zi * hi;
}

// Projection parameters of an EPSG code (Projections.cpp:79-168, 182-214,
// 262-281, 368-393); false for a code the reference's factory rejects.
inline bool proj_init(int epsg, Proj* p)
{
    *p = Proj{};
    const double e2 = kWgsE2;
    if (epsg == 4326) {
        p->kind = kProjLonLat;
        return true;
    }
    if (epsg == 4978) {
        p->kind = kProjGeocent;
        return true;
    }
    if ((epsg > 32600 && epsg <= 32660) || (epsg > 32700 && epsg <= 32760)) {
        p->kind = kProjUtm;
        p->north = epsg < 32700;
        const int zone = p->north ? epsg - 32600 : epsg - 32700;
        p->lon0 = ((zone - 0.5) * (kPi / 30.)) - kPi;
        const double f = e2 / (1. + sqrt(1. - e2));  // flattening
        const double n = f / (2. - f);                // third flattening
        const double n2 = n * n, n3 = n2 * n, n4 = n3 * n, n5 = n4 * n, n6 = n5 * n;
        // Gauss <-> geodetic latitude (Krueger, to n^6)
        p->cgb[0] = n * (2 + n * (-2. / 3. + n * (-2 + n * (116. / 45. + n * (26. / 45. + n * (-2854. / 675.))))));
        p->cbg[0] = n * (-2 + n * (2. / 3. + n * (4. / 3. + n * (-82. / 45. + n * (32. / 45. + n * (4642. / 4725.))))));
        p->cgb[1] = n2 * (7. / 3. + n * (-8. / 5. + n * (-227. / 45. + n * (2704. / 315. + n * (2323. / 945.)))));
        p->cbg[1] = n2 * (5. / 3. + n * (-16. / 15. + n * (-13. / 9. + n * (904. / 315. + n * (-1522. / 945.)))));
        p->cgb[2] = n3 * (56. / 15. + n * (-136. / 35. + n * (-1262. / 105. + n * (73814. / 2835.))));
        p->cbg[2] = n3 * (-26. / 15. + n * (34. / 21. + n * (8. / 5. + n * (-12686. / 2835.))));
        p->cgb[3] = n4 * (4279. / 630. + n * (-332. / 35. + n * (-399572. / 14175.)));
        p->cbg[3] = n4 * (1237. / 630. + n * (-12. / 5. + n * (-24832. / 14175.)));
        p->cgb[4] = n5 * (4174. / 315. + n * (-144838. / 6237.));
        p->cbg[4] = n5 * (-734. / 315. + n * (109598. / 31185.));
        p->cgb[5] = n6 * (601676. / 22275.);
        p->cbg[5] = n6 * (444337. / 155925.);
        // meridian quadrant scaled by k0 = 0.9996
        p->qn = (0.9996 / (1. + n)) * (1. + n2 * (1. / 4. + n2 * (1. / 64. + n2 / 256.)));
        // ellipsoidal <-> spherical northing/easting
        p->utg[0] = n * (-.5 + n * (2. / 3. + n * (-37. / 96. + n * (1. / 360. + n * (81. / 512. + n * (-96199. / 604800.))))));
        p->gtu[0] = n * (.5 + n * (-2. / 3. + n * (5. / 16. + n * (41. / 180. + n * (-127. / 288. + n * (7891. / 37800.))))));
        p->utg[1] = n2 * (-1. / 48. + n * (-1. / 15. + n * (437. / 1440. + n * (-46. / 105. + n * (1118711. / 3870720.)))));
        p->gtu[1] = n2 * (13. / 48. + n * (-3. / 5. + n * (557. / 1440. + n * (281. / 630. + n * (-1983433. / 1935360.)))));
        p->utg[2] = n3 * (-17. / 480. + n * (37. / 840. + n * (209. / 4480. + n * (-5569. / 90720.))));
        p->gtu[2] = n3 * (61. / 240. + n * (-103. / 140. + n * (15061. / 26880. + n * (167603. / 181440.))));
        p->utg[3] = n4 * (-4397. / 161280. + n * (11. / 504. + n * (830251. / 7257600.)));
        p->gtu[3] = n4 * (49561. / 161280. + n * (-179. / 168. + n * (6601661. / 7257600.)));
        p->utg[4] = n5 * (-4583. / 161280. + n * (108847. / 3991680.));
        p->gtu[4] = n5 * (34729. / 80640. + n * (-3418889. / 1995840.));
        p->utg[5] = n6 * (-20648693. / 638668800.);
        p->gtu[5] = n6 * (212378941. / 319334400.);
        // northing of the equator: Gauss latitude of 0 is 0
        const double z = clenshaw_sin(p->cbg, 6, 0.);
        p->zb = -p->qn * (z + clenshaw_sin(p->gtu, 6, 2 * z));
        return true;
    }
    if (epsg == 3031 || epsg == 3413) {
        p->kind = kProjPolar;
        p->north = epsg == 3413;
        const double lat_ts = p->north ? 70. * (kPi / 180.) : (71. * kPi) / 180.;
        p->lon0 = p->north ? -45. * (kPi / 180.) : 0.;
        p->e = sqrt(e2);
        const double es = p->e * sin(lat_ts);
        const double ts = tan(.5 * ((.5 * kPi) - lat_ts)) / pow((1. - es) / (1. + es), .5 * p->e);
        p->akm1 = cos(lat_ts) / ts;
        p->akm1 *= kWgsA / sqrt(1. - ((p->e * p->e) * (sin(lat_ts) * sin(lat_ts))));
        return true;
    }
    if (epsg == 6933) {
        p->kind = kProjCea;
        const double lat_ts = kPi / 6.;
        p->k0 = cos(lat_ts) / sqrt(1. - (e2 * (sin(lat_ts) * sin(lat_ts))));
        p->e = sqrt(e2);
        p->one_es = 1. - e2;
        p->apa[0] = e2 * ((1. / 3.) + (e2 * ((31. / 180.) + (e2 * (517. / 5040.)))));
        p->apa[1] = e2 * e2 * ((23. / 360.) + (e2 * (251. / 3780.)));
        p->apa[2] = e2 * e2 * e2 * (761. / 45360.);
        const double con = p->e;  // authalic q at the pole (sin(phi) = 1)
        p->qp = p->one_es * ((1. / (1. - con * con)) - ((.5 / p->e) * log((1. - con) / (1. + con))));
        return true;
    }
    return false;
}

// PROJ's t(phi) of the polar stereographic (Projections.cpp:252-257)
BP_HD double polar_tsfn(double phi, double sinphi, double e)
{
    sinphi *= e;
    return tan(.5 * ((.5 * kPi) - phi)) / pow((1. - sinphi) / (1. + sinphi), .5 * e);
}

// authalic q(phi) of the equal-area projection (Projections.cpp:353-358)
BP_HD double cea_qsfn(double sinphi, double e, double one_es)
{
    const double con = e * sinphi;
    return one_es * ((sinphi / (1. - con * con)) - ((.5 / e) * log((1. - con) / (1. + con))));
}

// (lon, lat, h) in radians -> projected (x, y, h); false where the reference's
// forward reports failure (UTM beyond its easting limit).
BP_HD bool proj_forward(const Proj& p, V3 llh, V3* out)
{
    switch (p.kind) {
    case kProjLonLat:  // Projections.h:127-133
        *out = {llh.x * 180.0 / kPi, llh.y * 180.0 / kPi, llh.z};
        return true;
    case kProjGeocent:  // Projections.cpp:10-16
        *out = lonlat_to_xyz(kWgsA, kWgsE2, llh);
        return true;
    case kProjUtm: {  // Projections.cpp:170-197
        const double gauss = clenshaw_sin(p.cbg, 6, 2. * llh.y) + llh.y;
        const double lam = llh.x - p.lon0;
        double cn = atan2(sin(gauss), cos(lam) * cos(gauss));
        double ce = atan2(sin(lam) * cos(gauss), hypot(sin(gauss), cos(gauss) * cos(lam)));
        ce = asinh(tan(ce));
        double dce;
        cn += clenshaw_sin_complex(p.gtu, 6, 2 * cn, 2 * ce, &dce);
        ce += dce;
        if (!(fabs(ce) <= 2.623395162778)) return false;
        *out = {(p.qn * ce * kWgsA) + 500000.,
                (((p.qn * cn) + p.zb) * kWgsA) + (p.north ? 0. : 10000000.), llh.z};
        return true;
    }
    case kProjPolar: {  // Projections.cpp:283-297
        const double lam = llh.x - p.lon0;
        const double phi = llh.y * (p.north ? 1. : -1.);
        const double t = p.akm1 * polar_tsfn(phi, sin(phi), p.e);
        *out = {t * sin(lam), -t * cos(lam) * (p.north ? 1. : -1.), llh.z};
        return true;
    }
    default:  // kProjCea, Projections.cpp:395-401
        *out = {p.k0 * llh.x * kWgsA, (.5 * kWgsA * cea_qsfn(sin(llh.y), p.e, p.one_es)) / p.k0,
                llh.z};
        return true;
    }
}

// Projected (x, y, h) -> (lon, lat, h) in radians; false where the reference's
// inverse reports failure.
inline bool proj_inverse(const Proj& p, V3 xyz, V3* llh);

// ------------------------------------------------------------ DEM interpolators
BP_HD int64_t clamp64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Catmull-Rom value between p1 and p2 (core/BicubicInterpolator.cpp:9-38, float)
BP_HD float catmull_rom(float p0, float p1, float p2, float p3, double t)
{
    const double u = 1. - t;
    return (float(t) * (p2 - p0 * float(u * u) + (p2 * float(u * 3. + 1.) - p3 * float(u)) * float(t)) +
            p1 * float(t * t * (t * 3. - 5.) + 2.)) / 2.f;
}

BP_HD float bicubic_f32(double x, double y, const float* z, int64_t width, int64_t length)
{
    // core/BicubicInterpolator.cpp:43-61; rows/columns past the raster (the
    // reference reads beyond it there) are clamped to its last row/column
    const int x0 = static_cast<int>(floor(x)), y0 = static_cast<int>(floor(y));
    float col[4];
#pragma unroll
    for (int i = -1; i < 3; ++i) {
        const float* row = z + clamp64(y0 + i, 0, length - 1) * width;
        col[i + 1] = catmull_rom(row[x0 - 1], row[x0], row[clamp64(x0 + 1, 0, width - 1)],
                                 row[clamp64(x0 + 2, 0, width - 1)], x - x0);
    }
    return catmull_rom(col[0], col[1], col[2], col[3], y - y0);
}

// Natural cubic spline through n samples (second derivatives R by the
// tridiagonal sweep of core/Spline2dInterpolator.cpp:109-126), evaluated at
// 1-based position t inside (1, n) (:80-106, middle branch).
// The sweep's pivots p_i = 1 / (Q_{i-1} / 2 + 2) and Q_i = -p_i / 2 do not
// depend on the samples: they are evaluated once, in float, at compile time.
template<int N>
struct SplinePivots {
    float p[N], q[N];
    constexpr SplinePivots() : p{}, q{}
    {
        for (int i = 1; i < N - 1; ++i) {
            p[i] = 1.f / (0.5f * q[i - 1] + 2.f);
            q[i] = -0.5f * p[i];
        }
    }
};

template<int N>
BP_HD void spline_init(const float (&Y)[N], float (&R)[N])
{
    constexpr SplinePivots<N> piv;
    float Q[N];
    Q[0] = 0.f;
    R[0] = 0.f;
#pragma unroll
    for (int i = 1; i < N - 1; ++i) {
        const float p = piv.p[i];
        Q[i] = piv.q[i];
        R[i] = (3.f * (Y[i + 1] - 2.f * Y[i] + Y[i - 1]) - 0.5f * R[i - 1]) * p;
    }
    R[N - 1] = 0.f;
#pragma unroll
    for (int i = N - 2; i > 0; --i) R[i] = Q[i] * R[i + 1] + R[i];
}

template<int N>
BP_HD float spline_eval(double t, const float (&Y)[N], const float (&R)[N])
{
    if (t < 1.0) return Y[0] + static_cast<float>(t - 1.0) * (Y[1] - Y[0] - (R[1] / 6.f));
    if (t > N) return Y[N - 1] + static_cast<float>(t - N) * (Y[N - 1] - Y[N - 2] + (R[N - 2] / 6.f));
    const int j = static_cast<int>(floor(t));
    const float u = static_cast<float>(t - j);
    // Y[j - 1], Y[j], R[j - 1], R[j] by unrolled selection (no indexed private arrays)
    float y0 = Y[0], y1 = Y[1], r0 = R[0], r1 = R[1];
#pragma unroll
    for (int k = 2; k < N; ++k) {
        if (j == k) {
            y0 = Y[k - 1];
            y1 = Y[k];
            r0 = R[k - 1];
            r1 = R[k];
        }
    }
    const float t0 = y1 - y0 - (r0 / 3.f) - (r1 / 6.f);
    const float t1 = u * ((r0 / 2.f) + (u * ((r1 - r0) / 6.f)));
    return y0 + (u * (t0 + t1));
}

// Order-6 spline of core/Spline2dInterpolator.cpp:31-64 (the biquintic
// method): rows and columns clamp(start + k, 0, n - 2) + 1, k < 6, with
// start = trunc(coordinate) - 2.
BP_HD float biquintic_f32(double x, double y, const float* z, int64_t width, int64_t length)
{
    constexpr int N = 6;
    const int i0 = static_cast<int>(y) - 2, j0 = static_cast<int>(x) - 2;
    float hc[N], a[N], r[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const float* row = z + (clamp64(i0 + i, 0, length - 2) + 1) * width;
#pragma unroll
        for (int j = 0; j < N; ++j) a[j] = row[clamp64(j0 + j, 0, width - 2) + 1];
        spline_init(a, r);
        hc[i] = spline_eval(x - j0, a, r);
    }
    spline_init(hc, r);
    return spline_eval(y - i0, hc, r);
}

// 8 x 8 windowed-sinc interpolation of a float raster at (x, y) with the
// kernel table k (core/Sinc2dInterpolator.cpp:45-97): zero where the window
// does not fit the raster; the window's rows run from y's floor + 4 down to
// floor - 3, its columns likewise; kernel rows picked by the fractions times
// 8192; products and sum in float.
BP_HD float sinc_f32(double x, double y, const float* z, int64_t width, int64_t length,
                     const float* k)
{
    constexpr int half = kSincLen / 2;
    const int ix = static_cast<int>(floor(x));
    const int iy = static_cast<int>(floor(y));
    const double fx = x - ix, fy = y - iy;
    if (ix < half - 1 || ix > width - half - 1) return 0.f;
    if (iy < half - 1 || iy > length - half - 1) return 0.f;
    const int fxi = min(max(0, (int)(fx * kSincSub)), kSincSub - 1);
    const int fyi = min(max(0, (int)(fy * kSincSub)), kSincSub - 1);
    const float* ky = k + (int64_t)fyi * kSincLen;
    const float* kx = k + (int64_t)fxi * kSincLen;
    float ret = 0.f;
    for (int i = 0; i < kSincLen; ++i) {
        const float* row = z + (int64_t)(iy + half - i) * width + (ix + half);
        for (int j = 0; j < kSincLen; ++j) ret += row[-j] * ky[i] * kx[j];
    }
    return ret;
}

// DEM lookup in native raster coordinates
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
    switch (dem.method) {
    case kDemBicubic: return bicubic_f32(col, row, dem.data, dem.width, dem.length);
    case kDemNearest:  // core/NearestNeighborInterpolator.cpp:11-22
        return dem.data[(int64_t)llround(row) * dem.width + llround(col)];
    case kDemBiquintic: return biquintic_f32(col, row, dem.data, dem.width, dem.length);
    case kDemSinc: return sinc_f32(col, row, dem.data, dem.width, dem.length, dem.sinc);
    default: return bilinear_f32(col, row, dem.data, dem.width);
    }
}

// geometry/DEMInterpolator.cpp:593-610: (lon, lat) through the raster's
// projection, then the native-coordinate lookup.
BP_HD double dem_interpolate_lonlat(const Dem& dem, double lon, double lat)
{
    if (!dem.have_raster) return dem.ref_height;
    V3 xy;
    if (!proj_forward(dem.proj, V3{lon, lat, 0.0}, &xy)) return dem.ref_height;
    return dem_interpolate_xy(dem, xy.x, xy.y);
}

// True when a DEM needs more than the lon/lat-degree bilinear lookup (the
// geometry kernel has a compact variant for that common case).
BP_HD bool dem_generic(const Dem& dem)
{
    return dem.have_raster && !(dem.proj.kind == kProjLonLat && dem.method == kDemBilinear);
}

// The lookup of dem_interpolate_lonlat, specialised: kGeneric false is the
// EPSG:4326 bilinear case only (callers check dem_generic first).
template<bool kGeneric>
BP_HD double dem_lookup(const Dem& dem, double lon, double lat)
{
    if (kGeneric) return dem_interpolate_lonlat(dem, lon, lat);
    if (!dem.have_raster) return dem.ref_height;
    double x = lon * 180.0 / kPi, y = lat * 180.0 / kPi;
    if (x > 360 || x < -360) x = fmod(x, 360.0);
    if (x < -180) x += 360;
    if (x - 360 >= dem.xstart) {
        x -= 360;
    } else if (x < dem.xstart && x + 360 >= dem.xstart) {
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

inline bool proj_inverse(const Proj& p, V3 xyz, V3* llh)
{
    switch (p.kind) {
    case kProjLonLat:  // Projections.h:135-141
        *llh = {xyz.x * kPi / 180.0, xyz.y * kPi / 180.0, xyz.z};
        return true;
    case kProjGeocent:  // Projections.cpp:18-24
        *llh = xyz_to_lonlat(kWgsA, kWgsE2, xyz);
        return true;
    case kProjUtm: {  // Projections.cpp:199-232
        double cn = (xyz.y - (p.north ? 0. : 10000000.)) / kWgsA;
        double ce = (xyz.x - 500000.) / kWgsA;
        cn = (cn - p.zb) / p.qn;
        ce /= p.qn;
        if (!(fabs(ce) <= 2.623395162778)) return false;
        double dce;
        cn += clenshaw_sin_complex(p.utg, 6, 2 * cn, 2 * ce, &dce);
        ce = atan(sinh(ce + dce));
        const double sce = sin(ce), cce = cos(ce);
        ce = atan2(sce, cce * cos(cn));
        cn = atan2(sin(cn) * cce, hypot(sce, cce * cos(cn)));
        *llh = {ce + p.lon0, clenshaw_sin(p.cgb, 6, 2 * cn) + cn, xyz.z};
        return true;
    }
    case kProjPolar: {  // Projections.cpp:299-322
        const double tp = -hypot(xyz.x, xyz.y) / p.akm1;
        const double fact = p.north ? 1. : -1.;
        double phi_l = (.5 * kPi) - (2. * atan(tp));
        for (int i = 0; i < 8; ++i) {
            const double sinphi = p.e * sin(phi_l);
            const double phi = 0.5 * kPi +
                               2. * atan(tp * pow((1. + sinphi) / (1. - sinphi), -0.5 * p.e));
            if (fabs(phi_l - phi) < 1.e-10) {
                const double lon = (xyz.x == 0. && xyz.y == 0.)
                                       ? 0.
                                       : atan2(xyz.x, -fact * xyz.y) + p.lon0;
                *llh = {lon, phi * fact, xyz.z};
                return true;
            }
            phi_l = phi;
        }
        return false;
    }
    default: {  // kProjCea, Projections.cpp:403-411
        const double beta = asin((2. * xyz.y * p.k0) / (kWgsA * p.qp));
        *llh = {xyz.x / (p.k0 * kWgsA),
                beta + (p.apa[0] * sin(2. * beta)) + (p.apa[1] * sin(4. * beta)) +
                    (p.apa[2] * sin(6. * beta)),
                xyz.z};
        return true;
    }
    }
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

// fma(a, b, c) as one v_fma_f64 on the device.  The compiler's two-address
// form of an FMA (v_fmac_f64, the result written over the addend) costs a copy
// of the addend when the addend is a constant kept in registers across loop
// trips, as the coefficients of a polynomial in a loop are: one 64-bit move
// per term.
BP_HD double fma_untied(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return fma(a, b, c);
#endif
}

// sqrt(x) for x well inside the normal range (x > 2^-900, finite): the
// reciprocal square root seed and two Newton steps of the correctly rounded
// FP64 square root, without its rescaling of tiny and special inputs.
BP_HD double sqrt_normal(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = 0.5 * y;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
#else
    return sqrt(x);
#endif
}

// sin(2 pi t) and cos(2 pi t) in FP64 for a phase t in turns, |t| < 2^49
// (where a double still resolves 1/16 turn; a phase fc tau is below 1e12).
// t is reduced to r = t - q/4 with q the nearest integer to 4t (ties to
// even), so that |2 pi r| <= pi/4; sin and cos of x = 2 pi r come from their
// Taylor series through x^17 and x^18 (truncation below 1e-16 on that
// interval), and the quadrant q mod 4 rotates the pair.  q comes from one
// FMA: 4t + 1.5 * 2^52 lies in [2^52, 2^53), where doubles are the integers,
// so the sum rounds 4t to the nearest integer (ties to even, as 1.5 * 2^52 is
// even), subtracting 1.5 * 2^52 again is exact, and the low bits of the sum's
// mantissa are those of q + 2^51, i.e. q mod 4.  Error about 2 ulp; no table,
// no division.
BP_HD void sincos_turns(double t, double* s, double* c)
{
    constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52
    const double tq = fma(4.0, t, kRound);
    const double q = tq - kRound;
    const double r = fma(-0.25, q, t);
    const double x = r * 6.283185307179586476925286766559;
    const double x2 = x * x;
    // sin x = x (1 + x^2 (-1/3! + x^2 (1/5! - ...)))
    double ps = 2.8114572543455207632e-15;                  //  1/17!
    ps = fma_untied(ps, x2, -7.6471637318198164759e-13);   // -1/15!
    ps = fma_untied(ps, x2, 1.6059043836821614599e-10);    //  1/13!
    ps = fma_untied(ps, x2, -2.5052108385441718775e-08);   // -1/11!
    ps = fma_untied(ps, x2, 2.7557319223985890653e-06);    //  1/9!
    ps = fma_untied(ps, x2, -1.9841269841269841270e-04);   // -1/7!
    ps = fma_untied(ps, x2, 8.3333333333333333333e-03);    //  1/5!
    ps = fma_untied(ps, x2, -1.6666666666666666667e-01);   // -1/3!
    const double sx = fma(x * x2, ps, x);
    // cos x = 1 + x^2 (-1/2! + x^2 (1/4! - ...))
    double pc = 1.5619206968586225686e-16;                  //  1/18!
    pc = fma_untied(pc, x2, -4.7794773323873852974e-14);   // -1/16!
    pc = fma_untied(pc, x2, 1.1470745597729724714e-11);    //  1/14!
    pc = fma_untied(pc, x2, -2.0876756987868098979e-09);   // -1/12!
    pc = fma_untied(pc, x2, 2.7557319223985890653e-07);    //  1/10!
    pc = fma_untied(pc, x2, -2.4801587301587301587e-05);   // -1/8!
    pc = fma_untied(pc, x2, 1.3888888888888888889e-03);    //  1/6!
    pc = fma_untied(pc, x2, -4.1666666666666666667e-02);   //  1/4!
    pc = fma(pc, x2, 0.5);                                  // (sign folded below)
    const double cx = fma(-x2, pc, 1.0);
    // quadrant: (sin, cos) of x + q pi/2, q mod 4 from the sum's low bits
    const int iq = (int)(__builtin_bit_cast(unsigned long long, tq) & 3u);
    const double sa = (iq & 1) ? cx : sx;
    const double ca = (iq & 1) ? sx : cx;
    *s = (iq & 2) ? -sa : sa;
    *c = ((iq + 1) & 2) ? -ca : ca;
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
template<bool kGenericDem = true>
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
        return llh.z - dem_lookup<kGenericDem>(dem, llh.x, llh.y);
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
