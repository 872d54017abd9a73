// device_common.hpp -- data layout and Boys function for the gfx950 ERI kernels.
//
// HBM layout (built once per geometry by the host, read every Fock build):
//   ShellPair[n_pair]   64 B records: centre A of the first shell, A-B, the
//                        primitive-pair range and the two shells' first AOs.
//   PrimPair[n_prim]    64 B records, one per primitive pair of a shell pair:
//                        p = a+b, 1/p, P = (aA+bB)/p and
//                        Kp = w_a w_b exp(-ab/p |AB|^2) / p.
//   Boys table          801 rows x 16 doubles (one 128 B line per row):
//                        [exp(-x_i), F_0(x_i) .. F_14(x_i)], x_i = i/20.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DSSCF_BOYS_H 0.05
#define DSSCF_BOYS_INV_H 20.0
#define DSSCF_BOYS_TMAX 40.0
#define DSSCF_BOYS_ROWS 801
#define DSSCF_BOYS_STRIDE 16
#define DSSCF_BOYS_MMAX 14

// Boys-function modes: the reference's truncated series, or the accurate
// table/asymptotic evaluation.
enum { BOYS_REFERENCE = 0, BOYS_ACCURATE = 1 };

struct __align__(16) ShellPair {
    double A[3];     // centre of the first shell (higher l)
    double AB[3];    // A - B
    int prim_off;    // first PrimPair record
    int nprim;       // number of PrimPair records
    int ao_a;        // first AO of the first shell
    int ao_b;        // first AO of the second shell
};
static_assert(sizeof(ShellPair) == 64, "ShellPair must be one 64 B record");

struct __align__(16) PrimPair {
    double p, inv_p;
    double P[3];
    double Kp;
    double pad[2];
};
static_assert(sizeof(PrimPair) == 64, "PrimPair must be one 64 B record");

// One context's share of a fused Fock-build launch (one class): its quartet
// list and the buffers the quartets read and accumulate into.  blk0 is the
// first block of the launch that belongs to this segment.
struct JKSeg {
    const uint2* q;
    const ShellPair* pairs;
    const PrimPair* pp;
    const double* Dt;
    const double* Da;
    double* Jh;
    double* Kh;
    int n;
    int N;
    int blk0;
    int pad;
};

// One context's matrices for the fused prepare/fold kernels; row0 is its
// first row in the launch's stacked row range.
struct MatSeg {
    const double* Ptp;  // packed-U total density
    const double* Pap;  // packed-U alpha density
    double* Dt;         // full N x N
    double* Da;
    double* Jh;         // half-accumulated J, K (full N x N)
    double* Kh;
    double* Jp;         // packed-U outputs
    double* Kp;
    int N;
    int row0;
};

__device__ __forceinline__ int find_segment_by_block(const int* blk0, int nseg, int b)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}


// Reference Boys series (GaussianIntegral.cpp:209-223): asymptotic form for
// x > 1.1m+25 (with flo::pow's multiplication order, MathUtil.hpp:57-67),
// otherwise sum a_i = a_{i-1} 2x/(2m+2i+1) until the first a_i <= 1e-9.
__device__ __forceinline__ double flo_pow_dev(double x, unsigned p)
{
    double stack[8];
    int n = 0;
    double r;
    for (;;) {
        if (p == 1) { r = x; break; }
        if (p == 0) { r = 1.0; break; }
        double y = x;
        unsigned l = 2;
        for (; l < p; l *= 2) y = y * y;
        stack[n++] = y;
        p -= l / 2;
    }
    while (n) r = stack[--n] * r;
    return r;
}

__device__ __forceinline__ double boys_reference_m(double x, unsigned m)
{
    if (x > (double)m * 1.1 + 25.0) {
        int df = 1;
        for (int i = 2 * (int)m - 1; i > 0; i -= 2) df *= i;
        const double two_m1 = (double)(1u << (m + 1));
        return (double)df / two_m1 * sqrt(3.14159265358979 / flo_pow_dev(x, 2 * m + 1));
    }
    double F = 1.0, a = 1.0;
    for (unsigned i = 1; a > 0.000000001; ++i) {
        a = 2.0 * x / (double)(2 * m + 2 * i + 1) * a;
        F += a;
    }
    return exp(-x) / (double)(2 * m + 1) * F;
}

// Accurate F_0..F_L: 7-term Taylor expansion of F_L about the nearest grid
// point (|dx| <= 0.025, truncation < 2e-15 relative), e^-T from the tabulated
// e^-x_i times an 8-term series, then the stable downward recursion
// F_m = (2T F_{m+1} + e^-T)/(2m+1).  T >= 40: F_0 = sqrt(pi/T)/2 and upward
// recursion (neglected terms below e^-40).
template <int L>
__device__ __forceinline__ void boys_accurate(double T, const double* __restrict__ tab, double* F)
{
    if (T < DSSCF_BOYS_TMAX) {
        const int i = (int)(T * DSSCF_BOYS_INV_H + 0.5);
        const double nd = i * DSSCF_BOYS_H - T;  // -(T - x_i)
        const double* row = tab + i * DSSCF_BOYS_STRIDE;
        double s = row[1 + L + 6];
        s = fma(s, nd * (1.0 / 6.0), row[1 + L + 5]);
        s = fma(s, nd * (1.0 / 5.0), row[1 + L + 4]);
        s = fma(s, nd * (1.0 / 4.0), row[1 + L + 3]);
        s = fma(s, nd * (1.0 / 3.0), row[1 + L + 2]);
        s = fma(s, nd * (1.0 / 2.0), row[1 + L + 1]);
        s = fma(s, nd, row[1 + L]);
        F[L] = s;
        if (L > 0) {
            double e = 1.0 / 40320.0;
            e = fma(e, nd, 1.0 / 5040.0);
            e = fma(e, nd, 1.0 / 720.0);
            e = fma(e, nd, 1.0 / 120.0);
            e = fma(e, nd, 1.0 / 24.0);
            e = fma(e, nd, 1.0 / 6.0);
            e = fma(e, nd, 0.5);
            e = fma(e, nd, 1.0);
            e = fma(e, nd, 1.0);
            const double eT = row[0] * e;
            const double T2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fma(T2, F[m + 1], eT) * (1.0 / (2 * m + 1));
        }
    } else {
        const double inv_T = 1.0 / T;
        F[0] = 0.88622692545275801365 * sqrt(inv_T);
#pragma unroll
        for (int m = 1; m <= L; ++m) F[m] = F[m - 1] * ((m - 0.5) * inv_T);
    }
}

template <int BM, int L>
__device__ __forceinline__ void boys(double T, const double* __restrict__ tab, double* F)
{
    if (BM == BOYS_ACCURATE) {
        boys_accurate<L>(T, tab, F);
    } else {
#pragma unroll
        for (int m = 0; m <= L; ++m) F[m] = boys_reference_m(T, (unsigned)m);
    }
}

template <typename R>
struct PrimQuartet {
    R PA[3], WP[3], QC[3], WQ[3];
    R o2p, o2q, o2pq, rp, rq;
};

// Per-primitive-quartet quantities (the per-primitive block of
// GaussianIntegral.cpp:577-595 in HGP form): W-P, W-Q, 1/2p, 1/2q,
// 1/2(p+q), rho/p, rho/q and prefactor-scaled Boys values.
template <int BM, typename R, int L>
__device__ __forceinline__ void prim_setup(const PrimPair& P, const PrimPair& Q, const ShellPair& bra,
                                           const ShellPair& ket, const double* __restrict__ btab,
                                           PrimQuartet<R>& g, R* Fm)
{
    const double p = P.p, q = Q.p;
    const double inv_pq = 1.0 / (p + q);
    const double PQx = P.P[0] - Q.P[0], PQy = P.P[1] - Q.P[1], PQz = P.P[2] - Q.P[2];
    const double rho = p * q * inv_pq;
    const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    const double pref = P.Kp * Q.Kp * sqrt(inv_pq);
    double F[L + 1];
    boys<BM, L>(T, btab, F);
#pragma unroll
    for (int m = 0; m <= L; ++m) Fm[m] = (R)(F[m] * pref);
    const double qf = q * inv_pq, pf = p * inv_pq;
    g.PA[0] = (R)(P.P[0] - bra.A[0]);
    g.PA[1] = (R)(P.P[1] - bra.A[1]);
    g.PA[2] = (R)(P.P[2] - bra.A[2]);
    g.QC[0] = (R)(Q.P[0] - ket.A[0]);
    g.QC[1] = (R)(Q.P[1] - ket.A[1]);
    g.QC[2] = (R)(Q.P[2] - ket.A[2]);
    g.WP[0] = (R)(-qf * PQx);
    g.WP[1] = (R)(-qf * PQy);
    g.WP[2] = (R)(-qf * PQz);
    g.WQ[0] = (R)(pf * PQx);
    g.WQ[1] = (R)(pf * PQy);
    g.WQ[2] = (R)(pf * PQz);
    g.o2p = (R)(0.5 * P.inv_p);
    g.o2q = (R)(0.5 * Q.inv_p);
    g.o2pq = (R)(0.5 * inv_pq);
    g.rp = (R)qf;
    g.rq = (R)pf;
}

template <int la, int lb, int lc, int ld>
struct ClassInfo;
template <int la, int lb, int lc, int ld>
struct WaveTab;
