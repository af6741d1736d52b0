// Host build of the backend's FP64 friction pow and boundary sine (ptha_amd/csrc/swh_mathx.h)
// for the CPU checks of tests/test_mathx.py.
//
//  * shared library (g++ -O2 -ffp-contract=off -shared -fPIC): mx_pow / mx_sin over arrays;
//  * sweep program (-DSWEEP_MAIN, -fopenmp, -lquadmath): N seeded arguments per function in
//    the ranges the configurations reach (depths 1e-5..1e4 m log-uniform with the Manning
//    exponent -7/3 and the Chezy exponent -2; sine arguments in [-100, 100]), each result
//    compared with glibc and with the quad-precision value rounded to double (the
//    correctly rounded result). Prints one JSON line of counts.
//
//     g++ -O2 -ffp-contract=off -fopenmp -DSWEEP_MAIN tests/mathx_sweep.cpp -lquadmath -o sweep
//     ./sweep 100000000
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../ptha_amd/csrc/swh_mathx.h"

extern "C" void mx_pow(int n, const double *x, const double *y, double *out)
{
    for (int k = 0; k < n; k++) out[k] = swhm::pow_acc(x[k], y[k]);
}

extern "C" void mx_sin(int n, const double *t, double *out)
{
    for (int k = 0; k < n; k++) out[k] = swhm::sin_acc(t[k]);
}

#ifdef SWEEP_MAIN
#include <quadmath.h>

static inline uint64_t mix(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

int main(int argc, char **argv)
{
    const long long n = argc > 1 ? atoll(argv[1]) : 1000000;
    long long p_gl = 0, p_cr = 0, gl_cr = 0, c_gl = 0, c_cr = 0, s_gl = 0, s_cr = 0, sgl_cr = 0;
    const double lo = std::log(1e-5), hi = std::log(1e4), y = -7.0 / 3.0;
#pragma omp parallel for reduction(+ : p_gl, p_cr, gl_cr, c_gl, c_cr, s_gl, s_cr, sgl_cr) schedule(static)
    for (long long k = 0; k < n; k++) {
        const double u = (double)(mix(3 * k) >> 11) * 0x1.0p-53;
        const double x = std::exp(lo + u * (hi - lo));
        const double a = swhm::pow_acc(x, y), g = std::pow(x, y), c = (double)powq((__float128)x, (__float128)y);
        p_gl += a != g; p_cr += a != c; gl_cr += g != c;
        const double a2 = swhm::pow_acc(x, -2.0), g2 = std::pow(x, -2.0);
        const double c2 = (double)(1 / ((__float128)x * (__float128)x));
        c_gl += a2 != g2; c_cr += a2 != c2;
        const double t = (double)(mix(3 * k + 1) >> 11) * 0x1.0p-53 * 200.0 - 100.0;
        const double sa = swhm::sin_acc(t), sg = std::sin(t), sc = (double)sinq((__float128)t);
        s_gl += sa != sg; s_cr += sa != sc; sgl_cr += sg != sc;
    }
    printf("{\"n\": %lld, \"pow_m7o3\": {\"differs_from_glibc\": %lld, \"differs_from_correct_rounding\": %lld, "
           "\"glibc_differs_from_correct_rounding\": %lld}, \"pow_m2\": {\"differs_from_glibc\": %lld, "
           "\"differs_from_correct_rounding\": %lld}, \"sin\": {\"differs_from_glibc\": %lld, "
           "\"differs_from_correct_rounding\": %lld, \"glibc_differs_from_correct_rounding\": %lld}}\n",
           n, p_gl, p_cr, gl_cr, c_gl, c_cr, s_gl, s_cr, sgl_cr);
    return 0;
}
#endif
