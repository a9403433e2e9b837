/*
 * Thin C-ABI driver over the reference's own header-only building blocks --
 * TEST INFRASTRUCTURE ONLY.
 *
 * Compiled by oracle/Makefile against the reference headers where they lie
 * (isce3 cxx/isce3/...), together with the reference's except/Error.cpp,
 * into oracle/_ref/libisce3_ref_blocks.so.  Only the pieces of the hot path
 * that compile from their own few sources are exposed here:
 *   - math/RootFind1dBracket.icc  find_zero_brent
 *   - core/Kernels.icc            KnabKernel<double>, TabulatedKernel<float|double>
 *   - core/Interp1d.icc           interp1d on complex<float> data
 *   - core/Linspace.icc           Linspace::search
 *   - math/Sinc.icc               sinc
 * Everything that needs Eigen (orbit, ellipsoid, DEM, delay, backproject
 * itself) cannot be built here; the oracle restatement covers it.
 */
#include <isce3/error/ErrorCode.h>
#include <isce3/core/Interp1d.h>
#include <isce3/core/Kernels.h>
#include <isce3/core/Linspace.h>
#include <isce3/math/RootFind1dBracket.h>
#include <isce3/math/Sinc.h>

#include <complex>
#include <cstdint>
#include <memory>

extern "C" {

typedef double (*ref_fn)(double, void*);

int ref_find_zero_brent(double a, double b, ref_fn f, void* ctx, double tol, double* root)
{
    auto g = [&](double x) { return f(x, ctx); };
    return static_cast<int>(isce3::math::find_zero_brent(a, b, g, tol, root));
}

double ref_sinc(double t) { return isce3::math::sinc<double>(t); }

double ref_knab(double width, double bandwidth, double x)
{
    isce3::core::KnabKernel<double> k(width, bandwidth);
    return k(x);
}

void ref_tab_table_f32(double width, double bandwidth, int n, float* out)
{
    isce3::core::KnabKernel<double> k(width, bandwidth);
    isce3::core::TabulatedKernel<float> t(k, n);
    const auto& tab = t.table();
    for (int i = 0; i < n; ++i) out[i] = tab[i];
}

void ref_tab_table_f64(double width, double bandwidth, int n, double* out)
{
    isce3::core::KnabKernel<double> k(width, bandwidth);
    isce3::core::TabulatedKernel<double> t(k, n);
    const auto& tab = t.table();
    for (int i = 0; i < n; ++i) out[i] = tab[i];
}

void ref_tab_eval_f32(double width, double bandwidth, int n, const double* x, int64_t m,
                      float* out)
{
    isce3::core::KnabKernel<double> k(width, bandwidth);
    isce3::core::TabulatedKernel<float> t(k, n);
    for (int64_t i = 0; i < m; ++i) out[i] = t(x[i]);
}

/* interp1d<float, complex<float>> at m points t[] on one line of n samples. */
void ref_interp1d_c64(double width, double bandwidth, int n_table, const float* data,
                      int64_t n, const double* t, int64_t m, float* out)
{
    isce3::core::KnabKernel<double> k(width, bandwidth);
    isce3::core::TabulatedKernel<float> tab(k, n_table);
    auto* x = reinterpret_cast<const std::complex<float>*>(data);
    auto* o = reinterpret_cast<std::complex<float>*>(out);
    for (int64_t i = 0; i < m; ++i)
        o[i] = isce3::core::interp1d(tab, x, static_cast<size_t>(n), 1, t[i]);
}

int ref_linspace_search(double first, double spacing, int size, double val)
{
    isce3::core::Linspace<double> ls(first, spacing, size);
    return ls.search(val);
}

} // extern "C"
