// boys_table.hpp -- host construction of the device Boys table.
//
// Row i holds exp(-x_i) and F_0..F_14(x_i) at x_i = i/20, i = 0..800.  Each
// F_m(x) = e^-x sum_k (2x)^k / ((2m+1)(2m+3)...(2m+2k+1)) is summed in long
// double until the terms fall below 1e-22 of the sum (all terms positive,
// no cancellation).
#pragma once
#include <cmath>
#include <vector>

#include "device_common.hpp"

inline long double dsscf_boys_series(int m, long double x)
{
    long double s = 1.0L, a = 1.0L;
    for (int k = 1; k < 100000; ++k) {
        a *= 2.0L * x / (long double)(2 * m + 2 * k + 1);
        s += a;
        if (a < 1e-22L * s) break;
    }
    return expl(-x) * s / (long double)(2 * m + 1);
}

// The 801-row Taylor table, then the interval table of DSSCF_BPOLY_*
// (Taylor coefficients (-1)^k F_{m+k}(c)/k! about each interval centre c).
inline void dsscf_boys_table(std::vector<double>& tab)
{
    tab.assign((size_t)DSSCF_BPOLY_OFFSET + (size_t)(DSSCF_BOYS_MMAX + 1) * DSSCF_BPOLY_M_STRIDE, 0.0);
    for (int m = 0; m <= DSSCF_BOYS_MMAX; ++m)
        for (int i = 0; i < DSSCF_BPOLY_ROWS; ++i) {
            const long double c = 0.5L * i + 0.25L;
            double* t = &tab[(size_t)DSSCF_BPOLY_OFFSET + (size_t)m * DSSCF_BPOLY_M_STRIDE];
            auto at = [&](int q) -> double& { return t[2 * ((size_t)(q >> 1) * DSSCF_BPOLY_ROWS + i) + (q & 1)]; };
            at(0) = (double)expl(-c);
            long double fact = 1.0L;
            for (int k = 0; k < DSSCF_BPOLY_TERMS; ++k) {
                if (k) fact *= (long double)k;
                const long double a = dsscf_boys_series(m + k, c) / fact;
                at(1 + k) = (double)((k & 1) ? -a : a);
            }
        }
    for (int i = 0; i < DSSCF_BOYS_ROWS; ++i) {
        const long double x = (long double)i / 20.0L;
        const long double ex = expl(-x);
        double* row = &tab[(size_t)i * DSSCF_BOYS_STRIDE];
        row[0] = (double)ex;
        for (int m = 0; m <= DSSCF_BOYS_MMAX; ++m) {
            long double s = 1.0L, a = 1.0L;
            for (int k = 1; k < 100000; ++k) {
                a *= 2.0L * x / (long double)(2 * m + 2 * k + 1);
                s += a;
                if (a < 1e-22L * s) break;
            }
            row[1 + m] = (double)(ex * s / (long double)(2 * m + 1));
        }
    }
}
