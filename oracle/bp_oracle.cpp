/*
 * CPU oracle for time-domain backprojection -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++/OpenMP restatement of the reference CPU path.  Every function
 * cites the reference file:line whose behaviour it restates (paths relative
 * to the isce3 source tree).  The product never links this file; tests, the
 * smoke check and the benchmark's CPU-baseline leg use it as the checker.
 */
#include "bp_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kC = 299792458.0;                 // core/Constants.h:50
constexpr double kWgsA = 6378137.0;                // core/Constants.h:41
constexpr double kWgsE2 = 0.006694379990141317;    // core/Constants.h:44
constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

thread_local std::string g_err;

struct V3 {
    double x, y, z;
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double sqnorm(V3 a) { return dot(a, a); }
inline double norm(V3 a) { return std::sqrt(sqnorm(a)); }
inline V3 normalized(V3 a) { double n = norm(a); return {a.x / n, a.y / n, a.z / n}; }
inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 load3(const double* p) { return {p[0], p[1], p[2]}; }
inline void store3(V3 v, double* p) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// ---------------------------------------------------------------- Linspace
// core/Linspace.icc:77-89
int linspace_search(double first, double spacing, int size, double val)
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

// ---------------------------------------------------------------- Orbit
inline double otime(const orc_orbit& o, int i) { return o.t0 + i * o.dt; }
inline double ostart(const orc_orbit& o) { return o.t0; }
inline double oend(const orc_orbit& o) { return otime(o, o.n - 1); }
inline double omid(const orc_orbit& o) { return o.t0 + 0.5 * (o.n - 1) * o.dt; }

// core/detail/InterpolateOrbit.icc:15-109
void hermite(const orc_orbit& o, double t, V3* pos, V3* vel)
{
    int idx = linspace_search(o.t0, o.dt, o.n, t) - 2;
    idx = std::min(std::max(idx, 0), o.n - 4);
    double ti[4];
    for (int i = 0; i < 4; ++i) ti[i] = otime(o, idx + i);
    double f1[4], f0[4], h[4];
    for (int i = 0; i < 4; ++i) f1[i] = t - ti[i];
    for (int i = 0; i < 4; ++i) {
        double sum = 0.;
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            sum += 1. / (ti[i] - ti[j]);
        }
        f0[i] = 1. - 2. * sum * (t - ti[i]);
    }
    for (int i = 0; i < 4; ++i) {
        h[i] = 1.;
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            h[i] *= (t - ti[j]) / (ti[i] - ti[j]);
        }
    }
    if (pos) {
        V3 p{0., 0., 0.};
        for (int i = 0; i < 4; ++i) {
            V3 pi = load3(o.pos + 3 * (idx + i));
            V3 vi = load3(o.vel + 3 * (idx + i));
            double hh = h[i] * h[i];
            V3 term{pi.x * f0[i] + vi.x * f1[i], pi.y * f0[i] + vi.y * f1[i],
                    pi.z * f0[i] + vi.z * f1[i]};
            p = p + hh * term;
        }
        *pos = p;
    }
    if (!vel) return;
    double hdot[4], g1[4], g0[4];
    for (int i = 0; i < 4; ++i) {
        hdot[i] = 0.;
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            double prod = 1. / (ti[i] - ti[j]);
            for (int k = 0; k < 4; ++k) {
                if (k == i || k == j) continue;
                prod *= (t - ti[k]) / (ti[i] - ti[k]);
            }
            hdot[i] += prod;
        }
    }
    for (int i = 0; i < 4; ++i) g1[i] = h[i] + 2. * hdot[i] * (t - ti[i]);
    for (int i = 0; i < 4; ++i) {
        double sum = 0.;
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            sum += 1. / (ti[i] - ti[j]);
        }
        g0[i] = 2. * (f0[i] * hdot[i] - sum * h[i]);
    }
    V3 v{0., 0., 0.};
    for (int i = 0; i < 4; ++i) {
        V3 pi = load3(o.pos + 3 * (idx + i));
        V3 vi = load3(o.vel + 3 * (idx + i));
        V3 term{pi.x * g0[i] + vi.x * g1[i], pi.y * g0[i] + vi.y * g1[i],
                pi.z * g0[i] + vi.z * g1[i]};
        v = v + h[i] * term;
    }
    *vel = v;
}

// core/detail/InterpolateOrbit.icc:111-155
void legendre(const orc_orbit& o, double t, V3* pos, V3* vel)
{
    int idx = linspace_search(o.t0, o.dt, o.n, t) - 5;
    idx = std::min(std::max(idx, 0), o.n - 9);
    double trel = 8. * (t - otime(o, idx)) / (otime(o, idx + 8) - otime(o, idx));
    double teller = 1.;
    for (int i = 0; i < 9; ++i) teller *= trel - i;
    if (teller == 0.) {
        int i = static_cast<int>(trel);
        if (pos) *pos = load3(o.pos + 3 * (idx + i));
        if (vel) *vel = load3(o.vel + 3 * (idx + i));
        return;
    }
    static const double noemer[9] = {40320.0, -5040.0, 1440.0, -720.0, 576.0,
                                     -720.0, 1440.0, -5040.0, 40320.0};
    V3 p{0., 0., 0.}, v{0., 0., 0.};
    for (int i = 0; i < 9; ++i) {
        double coeff = (teller / noemer[i]) / (trel - i);
        p = p + coeff * load3(o.pos + 3 * (idx + i));
        v = v + coeff * load3(o.vel + 3 * (idx + i));
    }
    if (pos) *pos = p;
    if (vel) *vel = v;
}

// core/detail/InterpolateOrbit.icc:157-193 and core/Orbit.cpp:71-86
int orbit_interp(const orc_orbit& o, double t, int border, V3* pos, V3* vel)
{
    int minsv = o.method == ORC_INTERP_HERMITE ? 4 : (o.method == ORC_INTERP_LEGENDRE ? 9 : -1);
    if (o.n < minsv) return ORC_ORBIT_INTERP_SIZE_ERROR;
    if (t < ostart(o) || t > oend(o)) {
        if (border == ORC_BORDER_FILLNAN) {
            if (pos) *pos = {kNaN, kNaN, kNaN};
            if (vel) *vel = {kNaN, kNaN, kNaN};
        }
        if (border != ORC_BORDER_EXTRAPOLATE) return ORC_ORBIT_INTERP_DOMAIN_ERROR;
    }
    if (o.method == ORC_INTERP_HERMITE) { hermite(o, t, pos, vel); return ORC_SUCCESS; }
    if (o.method == ORC_INTERP_LEGENDRE) { legendre(o, t, pos, vel); return ORC_SUCCESS; }
    return ORC_ORBIT_INTERP_UNKNOWN_METHOD;
}

// ---------------------------------------------------------------- Ellipsoid
// core/Ellipsoid.h:117-120 (rEast), 176-190 (lonLatToXyz)
V3 lonlat_to_xyz(double a, double e2, V3 llh)
{
    double sl = std::sin(llh[1]);
    double re = a / std::sqrt(1.0 - (e2 * (sl * sl)));
    V3 xyz;
    xyz.x = (re + llh[2]) * std::cos(llh[1]) * std::cos(llh[0]);
    xyz.y = (re + llh[2]) * std::cos(llh[1]) * std::sin(llh[0]);
    xyz.z = ((re * (1.0 - e2)) + llh[2]) * std::sin(llh[1]);
    return xyz;
}

// core/Ellipsoid.h:196-224 (Vermeille 2002)
V3 xyz_to_lonlat(double a, double e2, V3 xyz)
{
    const double e4 = e2 * e2;
    const double a2 = a * a;
    double p = (xyz.x * xyz.x + xyz.y * xyz.y) / a2;
    double q = ((1. - e2) * (xyz.z * xyz.z)) / a2;
    double r = (p + q - e4) / 6.;
    double s = (e4 * p * q) / (4. * std::pow(r, 3));
    double t = std::pow(1. + s + std::sqrt(s * (2. + s)), (1. / 3.));
    double u = r * (1. + t + (1. / t));
    double rv = std::sqrt(u * u + (e4 * q));
    double w = (e2 * (u + rv - q)) / (2. * rv);
    double k = std::sqrt(u + rv + w * w) - w;
    double d = (k * std::sqrt(xyz.x * xyz.x + xyz.y * xyz.y)) / (k + e2);
    V3 llh;
    llh.y = std::atan2(xyz.z, d);
    llh.x = std::atan2(xyz.y, xyz.x);
    llh.z = ((k + e2 - 1.) * std::sqrt(d * d + xyz.z * xyz.z)) / k;
    return llh;
}

// core/Ellipsoid.h:151-157
V3 nvector(double lon, double lat)
{
    double clat = std::cos(lat);
    return {clat * std::cos(lon), clat * std::sin(lon), std::sin(lat)};
}

// ---------------------------------------------------------------- DEM
// core/BilinearInterpolator.cpp:13-47 with U = float (DEM rasters are float)
float bilinear_f32(double x, double y, const float* z, int64_t width)
{
    int x1 = static_cast<int>(std::floor(x));
    int x2 = static_cast<int>(std::ceil(x));
    int y1 = static_cast<int>(std::floor(y));
    int y2 = static_cast<int>(std::ceil(y));
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

// Same formula with U = double (LUT2d<double>).
double bilinear_f64(double x, double y, const double* z, int64_t width)
{
    int x1 = static_cast<int>(std::floor(x));
    int x2 = static_cast<int>(std::ceil(x));
    int y1 = static_cast<int>(std::floor(y));
    int y2 = static_cast<int>(std::ceil(y));
    double q11 = z[(int64_t)y1 * width + x1];
    double q12 = z[(int64_t)y2 * width + x1];
    double q21 = z[(int64_t)y1 * width + x2];
    double q22 = z[(int64_t)y2 * width + x2];
    if ((y1 == y2) && (x1 == x2)) return q11;
    if (y1 == y2)
        return (((x2 - x) / (x2 - x1)) * q11) + (((x - x1) / (x2 - x1)) * q21);
    if (x1 == x2)
        return (((y2 - y) / (y2 - y1)) * q11) + (((y - y1) / (y2 - y1)) * q12);
    const double den = static_cast<double>((x2 - x1) * (y2 - y1));
    return ((q11 * ((x2 - x) * (y2 - y))) / den) +
           ((q21 * ((x - x1) * (y2 - y))) / den) +
           ((q12 * ((x2 - x) * (y - y1))) / den) +
           ((q22 * ((x - x1) * (y - y1))) / den);
}

// ---------------------------------------------------------------- 2-D interpolation
// Value type T of the raster; weights and intermediate sums in T, as the
// reference's Interpolator<U> templates compute them.
template<class T>
T raster_at(const T* z, int64_t width, int64_t r, int64_t c) { return z[r * width + c]; }

// core/NearestNeighborInterpolator.cpp:11-22
template<class T>
T nearest2d(double x, double y, const T* z, int64_t width)
{
    return raster_at(z, width, (int64_t)std::round(y), (int64_t)std::round(x));
}

// Catmull-Rom segment between p1 and p2 at fraction t
// (core/BicubicInterpolator.cpp:9-38).
template<class T>
T catmull_rom(T p0, T p1, T p2, T p3, double t)
{
    const double s = 1. - t;
    const T a = p2 - p0 * T(s * s);
    const T b = (p2 * T(s * 3. + 1.) - p3 * T(s)) * T(t);
    const T c = p1 * T(t * t * (t * 3. - 5.) + 2.);
    return (T(t) * (a + b) + c) / T(2.);
}

// core/BicubicInterpolator.cpp:43-61 over rows/columns floor-1 .. floor+2;
// indices past the raster (which the reference reads beyond it) are clamped
// to its last row/column.
template<class T>
T bicubic2d(double x, double y, const T* z, int64_t width, int64_t length)
{
    const int64_t xf = (int64_t)std::floor(x), yf = (int64_t)std::floor(y);
    auto cl = [](int64_t v, int64_t hi) { return v > hi ? hi : v; };
    T v[4];
    for (int k = 0; k < 4; ++k) {
        const int64_t r = cl(yf - 1 + k, length - 1);
        v[k] = catmull_rom<T>(raster_at(z, width, r, xf - 1), raster_at(z, width, r, xf),
                              raster_at(z, width, r, cl(xf + 1, width - 1)),
                              raster_at(z, width, r, cl(xf + 2, width - 1)), x - (double)xf);
    }
    return catmull_rom<T>(v[0], v[1], v[2], v[3], y - (double)yf);
}

// Natural cubic spline coefficients (second derivatives) of samples y by a
// forward elimination / back substitution on the tridiagonal system
// (core/Spline2dInterpolator.cpp:109-126).
template<class T>
std::vector<T> spline_second_derivs(const std::vector<T>& y)
{
    const size_t n = y.size();
    std::vector<T> r(n, T(0)), q(n, T(0));
    for (size_t i = 1; i + 1 < n; ++i) {
        const T piv = T(1.0) / (T(0.5) * q[i - 1] + T(2.0));
        q[i] = T(-0.5) * piv;
        r[i] = (T(3.0) * (y[i + 1] - T(2.0) * y[i] + y[i - 1]) - T(0.5) * r[i - 1]) * piv;
    }
    r[n - 1] = T(0);
    for (size_t i = n - 2; i > 0; --i) r[i] = q[i] * r[i + 1] + r[i];
    return r;
}

// Spline value at 1-based position t (core/Spline2dInterpolator.cpp:80-106).
template<class T>
T spline_value(double t, const std::vector<T>& y, const std::vector<T>& r)
{
    const int n = (int)y.size();
    if (t < 1.0) return y[0] + T(t - 1.0) * (y[1] - y[0] - (r[1] / T(6.0)));
    if (t > n) return y[n - 1] + T(t - n) * (y[n - 1] - y[n - 2] + (r[n - 2] / T(6.0)));
    const int j = (int)std::floor(t);
    const T u = T(t - j);
    const T lin = y[j] - y[j - 1] - (r[j - 1] / T(3.0)) - (r[j] / T(6.0));
    const T cur = u * ((r[j - 1] / T(2.0)) + (u * ((r[j] - r[j - 1]) / T(6.0))));
    return y[j - 1] + (u * (lin + cur));
}

// Separable spline of `order` points (core/Spline2dInterpolator.cpp:31-64;
// the biquintic method is order 6): window start trunc(coordinate)
// (minus 0.5 for odd orders) - order/2 + 1, indices clamped to [0, n-2], +1.
template<class T>
T spline2d(double x, double y, const T* z, int64_t width, int64_t length, int order = 6)
{
    int i0 = (order % 2) ? (int)(y - 0.5) : (int)y;
    int j0 = (order % 2) ? (int)(x - 0.5) : (int)x;
    i0 = i0 - order / 2 + 1;
    j0 = j0 - order / 2 + 1;
    std::vector<T> cols(order), line(order);
    for (int i = 0; i < order; ++i) {
        const int64_t r = std::min<int64_t>(std::max<int64_t>(i0 + i, 0), length - 2) + 1;
        for (int j = 0; j < order; ++j) {
            const int64_t c = std::min<int64_t>(std::max<int64_t>(j0 + j, 0), width - 2) + 1;
            line[j] = raster_at(z, width, r, c);
        }
        cols[i] = spline_value(x - j0, line, spline_second_derivs(line));
    }
    return spline_value(y - i0, cols, spline_second_derivs(cols));
}

// Windowed-sinc 2-D interpolation as the DEM factory creates it
// (core/Interpolator.h:203-221: Sinc2dInterpolator(SINC_LEN = 8,
// SINC_SUB = 8192), core/Constants.h:32-35).
constexpr int kSincLen = 8, kSincSub = 8192;

// The interpolator's kernel (core/Sinc2dInterpolator.cpp:14-41, coefficients
// :105-120): a raised-cosine-windowed sinc of kSincLen x kSincSub samples,
// each of the kSincSub fractional rows normalised to unit sum, [row][tap].
const std::vector<double>& sinc2d_kernel()
{
    static const std::vector<double> kernel = [] {
        const int n = kSincLen * kSincSub;
        const double wgthgt = 0.5;             // (1 - pedestal) / 2, pedestal 0
        const double soff = (n - 1.0) / 2.0;
        std::vector<double> f(n);
        for (int i = 0; i < n; ++i) {
            const double wgt = (1.0 - wgthgt) + wgthgt * std::cos((M_PI * (i - soff)) / soff);
            const double s = std::floor(i - soff) / (1.0 * kSincSub);   // beta = 1
            const double fct = (s != 0.0) ? std::sin(M_PI * s) / (M_PI * s) : 1.0;
            f[i] = fct * wgt;
        }
        std::vector<double> k((size_t)n);
        for (int i = 0; i < kSincSub; ++i) {
            double ssum = 0.0;
            for (int j = 0; j < kSincLen; ++j) ssum += f[i + kSincSub * j];
            for (int j = 0; j < kSincLen; ++j)
                k[(size_t)i * kSincLen + j] = f[i + kSincSub * j] / ssum;
        }
        return k;
    }();
    return kernel;
}

// core/Sinc2dInterpolator.cpp:45-97: zero where the 8 x 8 window leaves the
// raster; the sum in the data type, products in the order of the reference.
template<class T>
T sinc2d(double x, double y, const T* z, int64_t width, int64_t length)
{
    const int half = kSincLen / 2;
    const int ix = static_cast<int>(std::floor(x));
    const int iy = static_cast<int>(std::floor(y));
    const double fx = x - ix;
    const double fy = y - iy;
    if ((ix < half - 1) || (ix > width - half - 1)) return T(0);
    if ((iy < half - 1) || (iy > length - half - 1)) return T(0);
    const int xx = ix + half, yy = iy + half;
    const int ifx = std::min(std::max(0, int(fx * kSincSub)), kSincSub - 1);
    const int ify = std::min(std::max(0, int(fy * kSincSub)), kSincSub - 1);
    const std::vector<double>& k = sinc2d_kernel();
    T ret(0);
    for (int i = 0; i < kSincLen; ++i)
        for (int j = 0; j < kSincLen; ++j)
            ret += z[(int64_t)(yy - i) * width + (xx - j)] *
                   static_cast<T>(k[(size_t)ify * kSincLen + i]) *
                   static_cast<T>(k[(size_t)ifx * kSincLen + j]);
    return ret;
}

template<class T>
T interp2d(int method, double x, double y, const T* z, int64_t width, int64_t length);

template<>
float interp2d<float>(int method, double x, double y, const float* z, int64_t width,
                      int64_t length)
{
    switch (method) {
    case ORC_DEM_BICUBIC: return bicubic2d(x, y, z, width, length);
    case ORC_DEM_NEAREST: return nearest2d(x, y, z, width);
    case ORC_DEM_BIQUINTIC: return spline2d(x, y, z, width, length);
    case ORC_DEM_SINC: return sinc2d(x, y, z, width, length);
    default: return bilinear_f32(x, y, z, width);
    }
}

template<>
double interp2d<double>(int method, double x, double y, const double* z, int64_t width,
                        int64_t length)
{
    switch (method) {
    case ORC_DEM_BICUBIC: return bicubic2d(x, y, z, width, length);
    case ORC_DEM_NEAREST: return nearest2d(x, y, z, width);
    case ORC_DEM_BIQUINTIC: return spline2d(x, y, z, width, length);
    case ORC_DEM_SINC: return sinc2d(x, y, z, width, length);
    default: return bilinear_f64(x, y, z, width);
    }
}

// ---------------------------------------------------------------- projections
// core/Projections.cpp, WGS84.  Series and constants as there (the
// Poder/Engsager transverse Mercator of PROJ, PROJ's polar stereographic and
// equal-area cylindrical formulas).
struct OrcProj {
    int code = 4326;
    bool north = true;
    double lon0 = 0.;
    double Qn = 0., Zb = 0., cgb[6] = {}, cbg[6] = {}, utg[6] = {}, gtu[6] = {};
    double akm1 = 0., e = 0.;
    double k0 = 0., one_es = 0., qp = 0., apa[3] = {};
};

// sin(arg) * sum_k a_k U_k(cos arg) by Clenshaw's recurrence (real argument)
double clenshaw_real(const double* a, int n, double arg)
{
    double b1 = 0., b2 = 0., b = 0.;
    const double c = std::cos(arg);
    b1 = a[n - 1];
    for (int k = n - 2; k >= 0; --k) {
        b = -b2 + (2. * b1 * c) + a[k];
        b2 = b1;
        b1 = b;
    }
    return std::sin(arg) * b;
}

// Complex Clenshaw at (re, im): returns the real part, im part in *out_im
double clenshaw_cplx(const double* a, int n, double re, double im, double* out_im)
{
    const double s = std::sin(re), c = std::cos(re), sh = std::sinh(im), ch = std::cosh(im);
    double br = 0., bi = 0., br1 = a[n - 1], bi1 = 0., br2 = 0., bi2 = 0.;
    for (int k = n - 2; k >= 0; --k) {
        br = -br2 + (2. * br1 * c * ch) - (-2. * bi1 * s * sh) + a[k];
        bi = -bi2 + (-2. * br1 * s * sh) + (2. * bi1 * c * ch);
        br2 = br1;
        bi2 = bi1;
        br1 = br;
        bi1 = bi;
    }
    *out_im = (s * ch * bi) + (c * sh * br);
    return (s * ch * br) - (c * sh * bi);
}

double tsfn(double phi, double sinphi, double e)
{
    const double es = sinphi * e;
    return std::tan(.5 * ((.5 * M_PI) - phi)) / std::pow((1. - es) / (1. + es), .5 * e);
}

double qsfn(double sinphi, double e, double one_es)
{
    const double con = e * sinphi;
    return one_es * ((sinphi / (1. - std::pow(con, 2))) - ((.5 / e) * std::log((1. - con) / (1. + con))));
}

bool make_proj(int code, OrcProj& p)
{
    p = OrcProj{};
    p.code = code;
    const double e2 = kWgsE2;
    if (code == 4326 || code == 4978) return true;
    if ((code > 32600 && code <= 32660) || (code > 32700 && code <= 32760)) {
        p.north = code < 32700;
        const int zone = p.north ? code - 32600 : code - 32700;
        p.lon0 = ((zone - 0.5) * (M_PI / 30.)) - M_PI;
        const double f = e2 / (1. + std::sqrt(1 - e2));
        const double n = f / (2. - f);
        const double np[7] = {1., n, std::pow(n, 2), std::pow(n, 3), std::pow(n, 4),
                              std::pow(n, 5), std::pow(n, 6)};
        // polynomials in n, highest power last (Horner from the top)
        auto horner = [n](std::initializer_list<double> c) {
            double acc = 0.;
            const double* b = c.begin();
            for (int i = (int)c.size() - 1; i >= 0; --i) acc = (i == (int)c.size() - 1) ? b[i] : b[i] + n * acc;
            return acc;
        };
        p.cgb[0] = n * horner({2., -2. / 3., -2., 116. / 45., 26. / 45., -2854. / 675.});
        p.cbg[0] = n * horner({-2., 2. / 3., 4. / 3., -82. / 45., 32. / 45., 4642. / 4725.});
        p.cgb[1] = np[2] * horner({7. / 3., -8. / 5., -227. / 45., 2704. / 315., 2323. / 945.});
        p.cbg[1] = np[2] * horner({5. / 3., -16. / 15., -13. / 9., 904. / 315., -1522. / 945.});
        p.cgb[2] = np[3] * horner({56. / 15., -136. / 35., -1262. / 105., 73814. / 2835.});
        p.cbg[2] = np[3] * horner({-26. / 15., 34. / 21., 8. / 5., -12686. / 2835.});
        p.cgb[3] = np[4] * horner({4279. / 630., -332. / 35., -399572. / 14175.});
        p.cbg[3] = np[4] * horner({1237. / 630., -12. / 5., -24832. / 14175.});
        p.cgb[4] = np[5] * horner({4174. / 315., -144838. / 6237.});
        p.cbg[4] = np[5] * horner({-734. / 315., 109598. / 31185.});
        p.cgb[5] = np[6] * (601676. / 22275.);
        p.cbg[5] = np[6] * (444337. / 155925.);
        p.Qn = (0.9996 / (1. + n)) * (1. + n * n * ((1. / 4.) + n * n * ((1. / 64.) + ((n * n) / 256.))));
        p.utg[0] = n * horner({-.5, 2. / 3., -37. / 96., 1. / 360., 81. / 512., -96199. / 604800.});
        p.gtu[0] = n * horner({.5, -2. / 3., 5. / 16., 41. / 180., -127. / 288., 7891. / 37800.});
        p.utg[1] = np[2] * horner({-1. / 48., -1. / 15., 437. / 1440., -46. / 105., 1118711. / 3870720.});
        p.gtu[1] = np[2] * horner({13. / 48., -3. / 5., 557. / 1440., 281. / 630., -1983433. / 1935360.});
        p.utg[2] = np[3] * horner({-17. / 480., 37. / 840., 209. / 4480., -5569. / 90720.});
        p.gtu[2] = np[3] * horner({61. / 240., -103. / 140., 15061. / 26880., 167603. / 181440.});
        p.utg[3] = np[4] * horner({-4397. / 161280., 11. / 504., 830251. / 7257600.});
        p.gtu[3] = np[4] * horner({49561. / 161280., -179. / 168., 6601661. / 7257600.});
        p.utg[4] = np[5] * horner({-4583. / 161280., 108847. / 3991680.});
        p.gtu[4] = np[5] * horner({34729. / 80640., -3418889. / 1995840.});
        p.utg[5] = np[6] * (-20648693. / 638668800.);
        p.gtu[5] = np[6] * (212378941. / 319334400.);
        const double z = clenshaw_real(p.cbg, 6, 0.);
        p.Zb = -p.Qn * (z + clenshaw_real(p.gtu, 6, 2 * z));
        return true;
    }
    if (code == 3031 || code == 3413) {
        p.north = code == 3413;
        const double lat_ts = p.north ? 70. * (M_PI / 180.) : (71. * M_PI) / 180.;
        p.lon0 = p.north ? -45. * (M_PI / 180.) : 0.;
        p.e = std::sqrt(e2);
        p.akm1 = std::cos(lat_ts) / tsfn(lat_ts, std::sin(lat_ts), p.e);
        p.akm1 *= kWgsA / std::sqrt(1. - (std::pow(p.e, 2) * std::pow(std::sin(lat_ts), 2)));
        return true;
    }
    if (code == 6933) {
        const double lat_ts = M_PI / 6.;
        p.k0 = std::cos(lat_ts) / std::sqrt(1. - (e2 * std::pow(std::sin(lat_ts), 2)));
        p.e = std::sqrt(e2);
        p.one_es = 1. - e2;
        p.apa[0] = e2 * ((1. / 3.) + (e2 * ((31. / 180.) + (e2 * (517. / 5040.)))));
        p.apa[1] = std::pow(e2, 2) * ((23. / 360.) + (e2 * (251. / 3780.)));
        p.apa[2] = std::pow(e2, 3) * (761. / 45360.);
        p.qp = qsfn(1., p.e, p.one_es);
        return true;
    }
    return false;
}

int proj_fwd(const OrcProj& p, V3 llh, V3& out)
{
    if (p.code == 4326) {
        out = {llh.x * 180.0 / M_PI, llh.y * 180.0 / M_PI, llh.z};
        return 0;
    }
    if (p.code == 4978) {
        out = lonlat_to_xyz(kWgsA, kWgsE2, llh);
        return 0;
    }
    if (p.code > 32600 && p.code <= 32760) {
        const double gauss = clenshaw_real(p.cbg, 6, 2. * llh.y) + llh.y;
        const double lam = llh.x - p.lon0;
        double cn = std::atan2(std::sin(gauss), std::cos(lam) * std::cos(gauss));
        double ce = std::atan2(std::sin(lam) * std::cos(gauss),
                               std::hypot(std::sin(gauss), std::cos(gauss) * std::cos(lam)));
        ce = std::asinh(std::tan(ce));
        double dce = 0.;
        cn += clenshaw_cplx(p.gtu, 6, 2 * cn, 2 * ce, &dce);
        ce += dce;
        if (!(std::fabs(ce) <= 2.623395162778)) return 1;
        out = {(p.Qn * ce * kWgsA) + 500000.,
               (((p.Qn * cn) + p.Zb) * kWgsA) + (p.north ? 0. : 10000000.), llh.z};
        return 0;
    }
    if (p.code == 3031 || p.code == 3413) {
        const double sgn = p.north ? 1. : -1.;
        const double lam = llh.x - p.lon0, phi = llh.y * sgn;
        const double t = p.akm1 * tsfn(phi, std::sin(phi), p.e);
        out = {t * std::sin(lam), -t * std::cos(lam) * sgn, llh.z};
        return 0;
    }
    out = {p.k0 * llh.x * kWgsA, (.5 * kWgsA * qsfn(std::sin(llh.y), p.e, p.one_es)) / p.k0, llh.z};
    return 0;
}

int proj_inv(const OrcProj& p, V3 xyz, V3& llh)
{
    if (p.code == 4326) {
        llh = {xyz.x * M_PI / 180.0, xyz.y * M_PI / 180.0, xyz.z};
        return 0;
    }
    if (p.code == 4978) {
        llh = xyz_to_lonlat(kWgsA, kWgsE2, xyz);
        return 0;
    }
    if (p.code > 32600 && p.code <= 32760) {
        double cn = ((xyz.y - (p.north ? 0. : 10000000.)) / kWgsA - p.Zb) / p.Qn;
        double ce = ((xyz.x - 500000.) / kWgsA) / p.Qn;
        if (!(std::fabs(ce) <= 2.623395162778)) return 1;
        double dce = 0.;
        cn += clenshaw_cplx(p.utg, 6, 2 * cn, 2 * ce, &dce);
        ce = std::atan(std::sinh(ce + dce));
        const double sce = std::sin(ce), cce = std::cos(ce);
        const double lon_s = std::atan2(sce, cce * std::cos(cn));
        const double lat_g = std::atan2(std::sin(cn) * cce, std::hypot(sce, cce * std::cos(cn)));
        llh = {lon_s + p.lon0, clenshaw_real(p.cgb, 6, 2 * lat_g) + lat_g, xyz.z};
        return 0;
    }
    if (p.code == 3031 || p.code == 3413) {
        const double sgn = p.north ? 1. : -1.;
        const double tp = -std::hypot(xyz.x, xyz.y) / p.akm1;
        double prev = (.5 * M_PI) - (2. * std::atan(tp));
        for (int it = 0; it < 8; ++it) {
            const double es = p.e * std::sin(prev);
            const double phi = 0.5 * M_PI + 2. * std::atan(tp * std::pow((1. + es) / (1. - es), -0.5 * p.e));
            if (std::fabs(prev - phi) < 1.e-10) {
                const double lon = (xyz.x == 0. && xyz.y == 0.) ? 0.
                                   : std::atan2(xyz.x, -sgn * xyz.y) + p.lon0;
                llh = {lon, phi * sgn, xyz.z};
                return 0;
            }
            prev = phi;
        }
        return 1;
    }
    const double beta = std::asin((2. * xyz.y * p.k0) / (kWgsA * p.qp));
    llh = {xyz.x / (p.k0 * kWgsA),
           beta + (p.apa[0] * std::sin(2. * beta)) + (p.apa[1] * std::sin(4. * beta)) +
               (p.apa[2] * std::sin(6. * beta)),
           xyz.z};
    return 0;
}

const OrcProj& cached_proj(int code)
{
    thread_local OrcProj p;
    thread_local int have = -1;
    if (have != code) {
        if (!make_proj(code, p)) make_proj(4326, p);
        have = code;
    }
    return p;
}

// geometry/DEMInterpolator.cpp:615-662
double dem_interp_xy(const orc_dem& dem, double x, double y)
{
    double value = dem.ref_height;
    if (!dem.have_raster) return value;
    if (dem.epsg == 4326 && (x > 360 || x < -360)) x = std::fmod(x, 360);
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
    const int irow = static_cast<int>(std::floor(row));
    const int icol = static_cast<int>(std::floor(col));
    if (irow < 2 || irow >= int(dem.length - 1)) return dem.ref_height;
    if (icol < 2 || icol >= int(dem.width - 1)) return dem.ref_height;
    return interp2d<float>(dem.method, col, row, dem.data, dem.width, dem.length);
}

// geometry/DEMInterpolator.cpp:593-610: through the raster's projection
double dem_interp_lonlat(const orc_dem& dem, double lon, double lat)
{
    if (!dem.have_raster) return dem.ref_height;
    V3 xy;
    if (proj_fwd(cached_proj(dem.epsg), V3{lon, lat, 0.0}, xy) != 0) return dem.ref_height;
    return dem_interp_xy(dem, xy.x, xy.y);
}

// ---------------------------------------------------------------- LUT2d
// core/LUT2d.cpp:127-159 (bilinear; out-of-bounds points are clamped)
double lut_eval(const orc_lut2d& l, double y, double x)
{
    if (!l.have_data) return l.ref_value;
    double x_idx = (x - l.xstart) / l.dx;
    double y_idx = (y - l.ystart) / l.dy;
    x_idx = std::min(std::max(x_idx, 0.0), l.width - 1.0);
    y_idx = std::min(std::max(y_idx, 0.0), l.length - 1.0);
    return bilinear_f64(x_idx, y_idx, l.data, l.width);
}

// ---------------------------------------------------------------- Brent
// math/RootFind1dBracket.icc:16-29
inline bool opposite_sign(double a, double b) { return std::signbit(a) ^ std::signbit(b); }
inline int count_bisection_steps(double a, double b, double tol)
{
    return static_cast<int>(std::ceil(std::log2(std::abs((a - b) / tol))));
}

// math/RootFind1dBracket.icc:56-216
template<typename F>
int find_zero_brent(double a, double b, F f, const double tol, double* root)
{
    const double eps = std::numeric_limits<double>::epsilon();
    if (root == nullptr) return ORC_NULL_DEREFERENCE;
    if (tol < 0.) return ORC_INVALID_TOLERANCE;
    double c, d, e, fa, fb, fc, p, q, r, s, tol1;
    fa = f(a);
    if (fa == 0.) { *root = a; return ORC_SUCCESS; }
    fb = f(b);
    if (fb == 0.) { *root = b; return ORC_SUCCESS; }
    if (!opposite_sign(fa, fb)) return ORC_INVALID_INTERVAL;
    c = a;
    fc = fa;
    e = d = b - a;
    tol1 = tol > 0. ? tol : eps;
    const int maxiter = 3 * count_bisection_steps(a, b, tol1);
    for (int i = 0; i < maxiter; ++i) {
        if (std::abs(fc) < std::abs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        tol1 = 2 * eps * std::abs(b) + 0.5 * tol;
        const double xm = 0.5 * (c - b);
        if ((std::abs(xm) <= tol1) || (fb == 0.)) { *root = b; return ORC_SUCCESS; }
        if ((std::abs(e) < tol1) || (std::abs(fa) <= std::abs(fb))) {
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
            if (p > 0.) q = -q; else p = -p;
            s = e;
            e = d;
            if (((2 * p) >= (3 * xm * q - std::abs(tol1 * q))) ||
                    (p >= std::abs(0.5 * s * q))) {
                e = d = xm;
            } else {
                d = p / q;
            }
        }
        a = b;
        fa = fb;
        if (std::abs(d) <= tol1) {
            if (xm <= 0.) b = b - tol1; else b = b + tol1;
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
    return ORC_FAILED_TO_CONVERGE;
}

// ---------------------------------------------------------------- geometry
// geometry/detail/Rdr2Geo.icc:172-242
int rdr2geo_bracket(double aztime, double slant_range, double doppler,
                    const orc_orbit& orbit, const orc_dem& dem, double a, double e2,
                    double wavelength, int side, const orc_r2g_params& params, V3* xyz)
{
    V3 radar, velocity;
    // Orbit::interpolate with the default border mode throws OutOfRange.
    int st = orbit_interp(orbit, aztime, ORC_BORDER_ERROR, &radar, &velocity);
    if (st == ORC_ORBIT_INTERP_DOMAIN_ERROR) return ORC_OUT_OF_RANGE;
    if (st != ORC_SUCCESS) return st;
    const double speed = norm(velocity);
    const V3 along{velocity.x / speed, velocity.y / speed, velocity.z / speed};
    const V3 right = normalized(cross(along, radar));
    const V3 down = cross(along, right);
    const V3 horizontal = (side == ORC_LOOK_RIGHT) ? right : -right;
    const double sin_squint = doppler * wavelength / (2 * speed);
    const double cos_squint = std::sqrt(1.0 - sin_squint * sin_squint);
    const V3 center = radar + (sin_squint * slant_range) * along;
    const double radius = cos_squint * slant_range;
    auto get_xyz = [&](double look) {
        const double hs = radius * std::sin(look);
        const double dc = radius * std::cos(look);
        return V3{center.x + hs * horizontal.x + dc * down.x,
                  center.y + hs * horizontal.y + dc * down.y,
                  center.z + hs * horizontal.z + dc * down.z};
    };
    auto dh = [&](double look) {
        V3 llh = xyz_to_lonlat(a, e2, get_xyz(look));
        return llh.z - dem_interp_lonlat(dem, llh.x, llh.y);
    };
    const double tol_look = params.tol_height / radius;
    double look = 0.0;
    const int err = find_zero_brent(params.look_min, params.look_max, dh, tol_look, &look);
    if (err != ORC_SUCCESS) return err;
    *xyz = get_xyz(look);
    return ORC_SUCCESS;
}

// geometry/detail/Geo2Rdr.icc:182-238
int geo2rdr_bracket(V3 x, const orc_orbit& orbit, const orc_lut2d& doppler,
                    double wavelength, int side, const orc_g2r_params& params,
                    double* aztime, double* range)
{
    double t0, t1;
    if (params.have_time_start) t0 = params.time_start;
    else t0 = doppler.have_data ? std::max(ostart(orbit), doppler.ystart) : ostart(orbit);
    if (params.have_time_end) t1 = params.time_end;
    else {
        if (doppler.have_data) {
            const double dop_end = doppler.ystart + doppler.dy * (doppler.length - 1);
            t1 = std::min(oend(orbit), dop_end);
        } else {
            t1 = oend(orbit);
        }
    }
    V3 xp, v, r;
    auto doppler_error = [&](double t) {
        orbit_interp(orbit, t, ORC_BORDER_FILLNAN, &xp, &v);
        r = x - xp;
        const double rnorm = norm(r);
        double fd = lut_eval(doppler, t, rnorm);
        return 2.0 / wavelength * dot(v, r) / rnorm - fd;
    };
    const int err = find_zero_brent(t0, t1, doppler_error, params.tol_aztime, aztime);
    if (err != ORC_SUCCESS) return err;
    orbit_interp(orbit, *aztime, ORC_BORDER_FILLNAN, &xp, &v);
    r = x - xp;
    *range = norm(r);
    if ((side == ORC_LOOK_RIGHT) ^ (dot(cross(r, v), xp) > 0)) return ORC_WRONG_LOOK_SIDE;
    return ORC_SUCCESS;
}

// focus/BistaticDelay.icc:9-17
inline double bistatic_delay(V3 p, V3 v, V3 x)
{
    V3 r = x - p;
    return 2. * (dot(r, v) - kC * norm(r)) / (sqnorm(v) - (kC * kC));
}

// focus/DryTroposphereModel.icc:9-29
double dry_tropo_tsx(V3 p, V3 llh, double a, double e2)
{
    const double zpd = 2.3, H = 6000.;
    V3 x = lonlat_to_xyz(a, e2, llh);
    V3 r_hat = normalized(p - x);
    V3 n_hat = normalized(nvector(llh.x, llh.y));
    double cos_theta = dot(r_hat, n_hat);
    return 2. * zpd * std::exp(-llh.z / H) / (kC * cos_theta);
}

// ---------------------------------------------------------------- kernels
// math/Sinc.icc:68-91
double sinc(double t)
{
    const double eps1 = 0.1490116119384765625e-7;
    const double eps2 = 0.0001220703125;
    double x = M_PI * std::abs(t);
    if (x < eps2) {
        double out = 1.;
        if (x > eps1) out -= x * x / 6.;
        return out;
    }
    return std::sin(x) / x;
}

// core/Kernels.icc:29-52 (KnabKernel<double>)
double knab(double t, double halfwidth, double bandwidth)
{
    const double c = M_PI * halfwidth * (1.0 - bandwidth);
    const double tf = t / halfwidth;
    std::complex<double> y = std::sqrt(std::complex<double>(1.0 - tf * tf));
    double window = std::real(std::cosh(c * y) / std::cosh(c));
    return window * sinc(t);
}

// core/Kernels.icc:113-135 (table fill; T = float or double)
template<typename T>
void tabulate_knab(double width, double bandwidth, int n, T* table)
{
    const double hw = std::fabs(width / 2.0);
    const double dx = hw / (n - 1.0);
    for (int i = 0; i < n; ++i) table[i] = static_cast<T>(knab(i * dx, hw, bandwidth));
}

// core/Kernels.icc:137-154 (TabulatedKernel<float>::operator())
float tab_eval_f32(const orc_kernel& k, double x)
{
    const double dx = k.halfwidth / (k.n - 1.0);
    const float one_dx = static_cast<float>(1.0 / dx);
    double ax = std::abs(x);
    if (ax > k.halfwidth) return 0.f;
    double axn = ax * one_dx;
    int i = static_cast<int>(std::floor(axn));
    i = std::min(i, k.n - 2);
    const float* t = k.table;
    return static_cast<float>(t[i] + (axn - i) * (t[i + 1] - t[i]));
}

// Same with T = double (TabulatedKernel<double>).
double tab_eval_f64(const orc_kernel& k, double x)
{
    const double dx = k.halfwidth / (k.n - 1.0);
    const double one_dx = 1.0 / dx;
    double ax = std::abs(x);
    if (ax > k.halfwidth) return 0.;
    double axn = ax * one_dx;
    int i = static_cast<int>(std::floor(axn));
    i = std::min(i, k.n - 2);
    const double* t = k.table64;
    return t[i] + (axn - i) * (t[i + 1] - t[i]);
}

// core/detail/Interp1d.h:23-40 (coeffs), 55-84 (zero-padded window),
// 91-106 (inner product, complex<float>, sequential order).
inline std::complex<float> interp1d_ref(const orc_kernel& k, const std::complex<float>* x,
                                        int64_t n, double t)
{
    const int width = static_cast<int>(std::ceil(2.0 * k.halfwidth));
    long i0 = (width % 2 == 0) ? static_cast<long>(std::ceil(t))
                               : static_cast<long>(std::round(t));
    const long low = i0 - width / 2;
    std::complex<float> sum(0.f, 0.f);
    for (int i = 0; i < width; ++i) {
        double ti = (i + low) - t;
        float c = tab_eval_f32(k, ti);
        long j = low + i;
        std::complex<float> v = (j >= 0 && j < n) ? x[j] : std::complex<float>(0.f, 0.f);
        sum += std::complex<float>(c * v.real(), c * v.imag());
    }
    return sum;
}

// All-FP64 variant: double table, complex<double> data path.
inline std::complex<double> interp1d_f64(const orc_kernel& k, const std::complex<float>* x,
                                         int64_t n, double t)
{
    const int width = static_cast<int>(std::ceil(2.0 * k.halfwidth));
    long i0 = (width % 2 == 0) ? static_cast<long>(std::ceil(t))
                               : static_cast<long>(std::round(t));
    const long low = i0 - width / 2;
    std::complex<double> sum(0., 0.);
    for (int i = 0; i < width; ++i) {
        double ti = (i + low) - t;
        double c = tab_eval_f64(k, ti);
        long j = low + i;
        std::complex<double> v = (j >= 0 && j < n) ? std::complex<double>(x[j])
                                                   : std::complex<double>(0., 0.);
        sum += c * v;
    }
    return sum;
}

} // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

double orc_sinc(double t) { return sinc(t); }
double orc_knab(double t, double hw, double bw) { return knab(t, hw, bw); }
void orc_tabulate_knab_f32(double w, double bw, int n, float* t) { tabulate_knab(w, bw, n, t); }
void orc_tabulate_knab_f64(double w, double bw, int n, double* t) { tabulate_knab(w, bw, n, t); }
float orc_tab_eval_f32(const orc_kernel* k, double x) { return tab_eval_f32(*k, x); }
double orc_tab_eval_f64(const orc_kernel* k, double x) { return tab_eval_f64(*k, x); }

void orc_interp1d_c64(const orc_kernel* k, const float* data, int64_t n, double t, float* out2)
{
    auto z = interp1d_ref(*k, reinterpret_cast<const std::complex<float>*>(data), n, t);
    out2[0] = z.real();
    out2[1] = z.imag();
}

int orc_linspace_search(double first, double spacing, int size, double val)
{
    return linspace_search(first, spacing, size, val);
}

int orc_orbit_interpolate(const orc_orbit* orbit, double t, int border, double* pos, double* vel)
{
    V3 p{kNaN, kNaN, kNaN}, v{kNaN, kNaN, kNaN};
    int st = orbit_interp(*orbit, t, border, &p, &v);
    store3(p, pos);
    store3(v, vel);
    return st;
}

void orc_lonlat_to_xyz(double a, double e2, const double* llh, double* xyz)
{
    store3(lonlat_to_xyz(a, e2, load3(llh)), xyz);
}

void orc_xyz_to_lonlat(double a, double e2, const double* xyz, double* llh)
{
    store3(xyz_to_lonlat(a, e2, load3(xyz)), llh);
}

double orc_dem_interpolate_lonlat(const orc_dem* dem, double lon, double lat)
{
    return dem_interp_lonlat(*dem, lon, lat);
}

double orc_dem_interpolate_xy(const orc_dem* dem, double x, double y)
{
    return dem_interp_xy(*dem, x, y);
}

double orc_lut2d_eval(const orc_lut2d* lut, double y, double x) { return lut_eval(*lut, y, x); }

int orc_proj_forward(int epsg, const double* llh, double* xyz)
{
    OrcProj p;
    if (!make_proj(epsg, p)) return -1;
    V3 out{};
    const int rc = proj_fwd(p, load3(llh), out);
    if (rc == 0) store3(out, xyz);
    return rc;
}

int orc_proj_inverse(int epsg, const double* xyz, double* llh)
{
    OrcProj p;
    if (!make_proj(epsg, p)) return -1;
    V3 out{};
    const int rc = proj_inv(p, load3(xyz), out);
    if (rc == 0) store3(out, llh);
    return rc;
}

float orc_interp2d_f32(int method, double x, double y, const float* z, int64_t width,
                       int64_t length)
{
    return interp2d<float>(method, x, y, z, width, length);
}

double orc_interp2d_f64(int method, double x, double y, const double* z, int64_t width,
                        int64_t length)
{
    return interp2d<double>(method, x, y, z, width, length);
}

double orc_bistatic_delay(const double* p, const double* v, const double* x)
{
    return bistatic_delay(load3(p), load3(v), load3(x));
}

double orc_dry_tropo_tsx(const double* p, const double* llh)
{
    return dry_tropo_tsx(load3(p), load3(llh), kWgsA, kWgsE2);
}

// Test functions of tests/cxx/isce3/math/root_find1d.cpp:160-198.
int orc_brent_testfn(int fid, double a, double b, double tol, double* root)
{
    auto f1 = [](double x) { return std::cos(x); };
    auto f2 = [](double x) { return -100.0 * x * std::exp(-2.0 * x); };
    auto f3 = [](double x) {
        if (std::abs(x) < 3.8e-4) return 0.0;
        return x / std::exp(std::pow(x, -2));
    };
    auto f4 = [](double x) { return std::pow(x, 9.0); };
    auto f5 = [](double x) {
        if (std::abs(x) < 1e-3) return std::numeric_limits<double>::quiet_NaN();
        return std::pow(x, 3.0);
    };
    switch (fid) {
    case 1: return find_zero_brent(a, b, f1, tol, root);
    case 2: return find_zero_brent(a, b, f2, tol, root);
    case 3: return find_zero_brent(a, b, f3, tol, root);
    case 4: return find_zero_brent(a, b, f4, tol, root);
    case 5: return find_zero_brent(a, b, f5, tol, root);
    default: return ORC_ARG_ERROR;
    }
}

int orc_rdr2geo_bracket(double aztime, double slant_range, double doppler,
                        const orc_orbit* orbit, const orc_dem* dem, double wavelength,
                        int side, const orc_r2g_params* p, double* xyz)
{
    V3 x{kNaN, kNaN, kNaN};
    int st = rdr2geo_bracket(aztime, slant_range, doppler, *orbit, *dem, kWgsA, kWgsE2,
                             wavelength, side, *p, &x);
    store3(x, xyz);
    return st;
}

int orc_geo2rdr_bracket(const double* x, const orc_orbit* orbit, const orc_lut2d* doppler,
                        double wavelength, int side, const orc_g2r_params* p,
                        double* aztime, double* range)
{
    return geo2rdr_bracket(load3(x), *orbit, *doppler, wavelength, side, *p, aztime, range);
}

// focus/Backproject.cpp:66-212 (orchestration) and 30-63 (sumCoherent).
int orc_backproject(const orc_grid* out_grid, const orc_orbit* out_orbit,
                    const orc_lut2d* out_doppler,
                    const orc_grid* in_grid, const orc_orbit* in_orbit,
                    const orc_lut2d* in_doppler,
                    const orc_dem* dem, double fc, double ds,
                    const orc_kernel* kernel, int tropo,
                    const orc_r2g_params* r2g, const orc_g2r_params* g2r,
                    const float* in_f, float* out_f, float* height,
                    int mode, const int64_t* subset, int64_t n_subset,
                    orc_pixel_geom* diag, int nthreads)
{
    if (tropo != ORC_TROPO_NODELAY && tropo != ORC_TROPO_TSX) {
        g_err = "unexpected dry troposphere model";
        return ORC_ARG_ERROR;
    }
    if (dem->have_raster) {
        OrcProj probe;
        if (!make_proj(dem->epsg, probe)) {
            g_err = "Unknown EPSG code (in factory): " + std::to_string(dem->epsg);
            return ORC_ARG_ERROR;
        }
        if (dem->method < ORC_DEM_SINC || dem->method > ORC_DEM_BIQUINTIC) {
            g_err = "unsupported DEM interpolation method";
            return ORC_ARG_ERROR;
        }
    }
    if (mode == ORC_MODE_FP64 && kernel->table64 == nullptr) {
        g_err = "FP64 mode needs a double kernel table";
        return ORC_ARG_ERROR;
    }
    const auto* in = reinterpret_cast<const std::complex<float>*>(in_f);
    auto* out = reinterpret_cast<std::complex<float>*>(out_f);

    // input grid: Linspace(sensing_start, 1/prf, length) and (r0, dr, width)
    // (container/RadarGeometry.icc:28-59)
    const double in_t0 = in_grid->sensing_start;
    const double in_dt = 1.0 / in_grid->prf;
    const int na = static_cast<int>(in_grid->length);
    const double out_t0 = out_grid->sensing_start;
    const double out_dt = 1.0 / out_grid->prf;
    const int64_t out_w = out_grid->width;
    const int64_t out_l = out_grid->length;

    // platform position & velocity at each pulse (Backproject.cpp:100-106)
    std::vector<V3> pos(na), vel(na);
    for (int i = 0; i < na; ++i) {
        double t = in_t0 + i * in_dt;
        int st = orbit_interp(*in_orbit, t, ORC_BORDER_ERROR, &pos[i], &vel[i]);
        if (st != ORC_SUCCESS) {
            g_err = "orbit interpolation failed for input pulse (out of range)";
            return ORC_OUT_OF_RANGE;
        }
    }
    // range sampling window (Backproject.cpp:108-112)
    const double swst = 2. * in_grid->starting_range / kC;
    const double dtau = 2. * in_grid->range_pixel_spacing / kC;
    const int64_t nr = in_grid->width;
    const double wvl = kC / fc;

    const int64_t npix = subset ? n_subset : out_l * out_w;
    bool all_converged = true;
    bool out_of_range = false;
    const float fnan = std::numeric_limits<float>::quiet_NaN();

#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(&& : all_converged) reduction(|| : out_of_range)
    for (int64_t ip = 0; ip < npix; ++ip) {
        const int64_t flat = subset ? subset[ip] : ip;
        const int64_t j = flat / out_w;
        const int64_t i = flat % out_w;
        orc_pixel_geom g;
        std::memset(&g, 0, sizeof(g));
        V3 x{kNaN, kNaN, kNaN}, llh;
        {
            double t = out_t0 + j * out_dt;
            double r = out_grid->starting_range + i * out_grid->range_pixel_spacing;
            double fD = lut_eval(*out_doppler, t, r);
            int st = rdr2geo_bracket(t, r, fD, *out_orbit, *dem, kWgsA, kWgsE2, wvl,
                                     out_grid->look_side, *r2g, &x);
            if (st == ORC_OUT_OF_RANGE) { out_of_range = true; continue; }
            llh = xyz_to_lonlat(kWgsA, kWgsE2, x);
            if (height) height[ip] = static_cast<float>(llh.z);
            if (st != ORC_SUCCESS) {
                all_converged = false;
                out[ip] = {fnan, fnan};
                if (height) height[ip] = fnan;
                g.status = 1;
                if (diag) diag[ip] = g;
                continue;
            }
        }
        double t, r;
        {
            int st = geo2rdr_bracket(x, *in_orbit, *in_doppler, wvl, in_grid->look_side, *g2r, &t, &r);
            if (st != ORC_SUCCESS) {
                all_converged = false;
                out[ip] = {fnan, fnan};
                g.status = 2;
                store3(x, g.x);
                if (diag) diag[ip] = g;
                continue;
            }
        }
        V3 p, v;
        if (orbit_interp(*in_orbit, t, ORC_BORDER_ERROR, &p, &v) != ORC_SUCCESS) {
            out_of_range = true;
            continue;
        }
        double l = wvl * r * (norm(p) / norm(x)) / (2. * ds);
        double cpi = l / norm(v);
        double tstart = t - 0.5 * cpi;
        double tstop = t + 0.5 * cpi;
        int kstart = static_cast<int>(std::floor((tstart - in_t0) / in_dt));
        int kstop = static_cast<int>(std::ceil((tstop - in_t0) / in_dt));
        kstart = std::max(kstart, 0);
        kstop = std::min(kstop, na);
        double tau_atm = 0.;
        if (tropo == ORC_TROPO_TSX) tau_atm = dry_tropo_tsx(p, llh, kWgsA, kWgsE2);

        std::complex<double> sum(0., 0.);
        if (mode == ORC_MODE_REFERENCE) {
            for (int k = kstart; k < kstop; ++k) {
                double tau = tau_atm + bistatic_delay(pos[k], vel[k], x);
                const std::complex<float>* line = &in[size_t(k) * nr];
                double u = (tau - swst) / dtau;
                std::complex<double> s(interp1d_ref(*kernel, line, nr, u));
                double phi = 2. * M_PI * fc * tau;
                s *= std::complex<double>(std::cos(phi), std::sin(phi));
                sum += s;
            }
        } else {
            for (int k = kstart; k < kstop; ++k) {
                double tau = tau_atm + bistatic_delay(pos[k], vel[k], x);
                const std::complex<float>* line = &in[size_t(k) * nr];
                double u = (tau - swst) / dtau;
                std::complex<double> s = interp1d_f64(*kernel, line, nr, u);
                double phi = 2. * M_PI * fc * tau;
                s *= std::complex<double>(std::cos(phi), std::sin(phi));
                sum += s;
            }
        }
        out[ip] = std::complex<float>(sum);
        if (diag) {
            store3(x, g.x);
            g.tau_atm = tau_atm;
            g.kstart = kstart;
            g.kstop = kstop;
            g.status = 0;
            diag[ip] = g;
        }
    }
    if (out_of_range) {
        g_err = "orbit interpolation out of range for an output pixel";
        return ORC_OUT_OF_RANGE;
    }
    return all_converged ? ORC_SUCCESS : ORC_FAILED_TO_CONVERGE;
}

} // extern "C"
