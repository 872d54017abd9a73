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

inline void dsscf_boys_table(std::vector<double>& tab)
{
    tab.assign((size_t)DSSCF_BOYS_ROWS * DSSCF_BOYS_STRIDE, 0.0);
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
