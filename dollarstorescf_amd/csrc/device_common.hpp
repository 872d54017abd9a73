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
// table/asymptotic evaluation.  BOYS_ACCURATE_LDS is the accurate mode
// evaluated from a block's LDS copy of the interval table below (device
// internal; the results agree with BOYS_ACCURATE to ~1e-16 relative).  In
// prim_setup it also means that the bra primitive record holds P - A in
// place of P (the J/K kernels stage their bra's records that way, which
// takes the P - A subtraction out of the primitive-quartet loop).
enum { BOYS_REFERENCE = 0, BOYS_ACCURATE = 1, BOYS_ACCURATE_LDS = 2 };

// Interval table (after the 801-row table in the same buffer): for each
// m = 0..DSSCF_BOYS_MMAX, DSSCF_BPOLY_ROWS intervals [c-1/4, c+1/4) with
// c = (i + 1/2)/2 covering [0, 40).  Coefficient q of an interval: q = 0 is
// e^-c, q = 1..12 the Taylor coefficients (-1)^k F_{m+k}(c)/k!, k = q-1 =
// 0..11 (remainder below 5e-17 relative for |T-c| <= 1/4), q = 13 padding.
// Stored as pairs (q = 2j, 2j+1) in 16 B cells, pair-major:
// cell (j, i) at doubles 2 (j DSSCF_BPOLY_ROWS + i), so one 16 B load
// fetches two coefficients and lanes reading pair j of different intervals
// hit consecutive LDS addresses.
#define DSSCF_BPOLY_ROWS 80
#define DSSCF_BPOLY_COEF 14
#define DSSCF_BPOLY_TERMS 12
#define DSSCF_BPOLY_OFFSET (DSSCF_BOYS_ROWS * DSSCF_BOYS_STRIDE)
#define DSSCF_BPOLY_M_STRIDE (DSSCF_BPOLY_ROWS * DSSCF_BPOLY_COEF)

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
    double e;    // screening factor |Kp| / p^(1/4); a shell pair's records are sorted by it, largest first
    double alpha;  // exponent of the first shell's primitive (p - alpha is the second's)
};
static_assert(sizeof(PrimPair) == 64, "PrimPair must be one 64 B record");

// A bra primitive pair as the FP32 J/K lists stage it in LDS: converted once
// per block, with P - A in place of P.
struct __align__(16) PrimPairF {
    float p, inv_p;
    float PA[3];
    float Kp;
    float pad[2];
};
static_assert(sizeof(PrimPairF) == 32, "PrimPairF must be one 32 B record");

// One context's share of a fused Fock-build launch (one class): its quartet
// list and the buffers the quartets read and accumulate into.  blk0 is the
// first block of the launch that belongs to this segment.
struct JKSeg {
    const uint2* q;            // flat (bra, ket) list: wave-cooperative classes
    const int4* u;             // work units (bra pair, first ket, ket count): thread-per-quartet classes
    const unsigned* k;         // ket pair indices referenced by the units
    const ShellPair* pairs;
    const PrimPair* pp;
    const double* Dt;
    const double* Da;
    double* Jh;
    double* Kh;
    int n;                     // quartets (flat list) or units
    int N;
    int blk0;
    int nslice;                // copies of Jh/Kh (power of two, <= kJKSlicesMax)
    // J_cd slot records (thread-per-quartet classes, or null: atomics into
    // Jh): quartet t of work unit u is slot us[u] + t, and stores its ket
    // block's nc * nd J_cd terms at js + jd[slot].  The records are laid out
    // by ket AO block, so k_jcd sums each block's records from one
    // contiguous range afterwards.
    double* js;
    const int* us;
    const int* jd;
    // unrestricted (open-shell) contexts: the beta density and the beta
    // half-accumulated K (nslice copies like Kh), else null.  K[Db] is
    // digested from the same integrals as K[Da].
    const double* Db;
    double* Kbh;
};

// One context's J_cd reduction (k_jcd): one row per ket AO block; ri =
// (c0, d0, nc, nd), rr = (first double in js, records); a record is nc * nd
// doubles.
struct JcdSeg {
    const int2* rr;
    const int4* ri;
    const double* js;
    double* Jh;  // copy 0 of the half-accumulated J
    int N;
    int nrow;
    int blk0;
    int pad;
};

// One context's matrices for the fused prepare/fold kernels; row0 is its
// first row in the launch's stacked row range.
// J and K are accumulated into up to kJKSlicesMax copies of Jh/Kh (per
// context, a power of two), which the fold sums.  Every J_cd element takes an
// atomic add from each bra it meets, and adds to one address serialize; the
// copies divide the adds per address, and a copy is picked by the XCD the
// wavefront runs on (then by a block-index bit), so that each copy's lines
// stay with few L2s.  The atomics stay device-scope: any block-to-copy map
// gives the same sums.
constexpr int kJKSlicesMax = 16;

__device__ __forceinline__ int jk_slice(int nslice)
{
    // s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4): the XCD this wavefront runs on
    const int xcc = (int)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 7;
    return (xcc + 8 * (((int)blockIdx.x >> 3) & 1)) & (nslice - 1);
}

struct MatSeg {
    const double* Ptp;  // packed-U total density
    const double* Pap;  // packed-U alpha density
    double* Dt;         // full N x N
    double* Da;
    double* Jh;         // half-accumulated J, K (nslice x full N x N)
    double* Kh;
    double* Jp;         // packed-U outputs
    double* Kp;
    int N;
    int row0;
    int nslice;
    int pad;
    // unrestricted contexts: packed-U beta density, its full copy, the beta
    // half-accumulated K and the packed-U K[Pb] output (else all null)
    const double* Pbp;
    double* Db;
    double* Kbh;
    double* Kbp;
};

// Read-only global-memory view of a pointer that arrived through memory (a
// segment table), so loads compile to global_load rather than flat_load.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using GPtr = const __attribute__((address_space(1))) T*;
#else
template <typename T>
using GPtr = const T*;
#endif
template <typename T>
__device__ __forceinline__ GPtr<T> gbl(const T* p)
{
    return (GPtr<T>)p;
}
// Writable global-memory view (stores compile to global_store).
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using GOut = __attribute__((address_space(1))) T*;
#else
template <typename T>
using GOut = T*;
#endif
template <typename T>
__device__ __forceinline__ GOut<T> gout(T* p)
{
    return (GOut<T>)p;
}
// Whole record (a multiple of 8 bytes) through global loads.
template <typename T>
__device__ __forceinline__ T ldg(const T* p)
{
    static_assert(sizeof(T) % 8 == 0, "record size");
    const GPtr<double> q = gbl((const double*)p);
    double tmp[sizeof(T) / 8];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 8); ++i) tmp[i] = q[i];
    T out;
    __builtin_memcpy(&out, tmp, sizeof(T));
    return out;
}

// Bra primitive pairs staged in LDS per block: a shell pair may have at most
// this many primitive pairs (checked by dsscf_set_basis).
constexpr int kBraStage = 128;

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

// 1/sqrt(x) for x > 0: the hardware estimate (v_rsq_f64, relative error
// below 2^-24 on gfx950, measured over 2^20 arguments in [2^-20, 2^20])
// refined by one third-order (Halley) step: with e = 1 - x r^2,
// r <- r (1 + e/2 + 3e^2/8), which leaves about e^3 ~ 2^-72 before rounding
// (two Newton steps r <- r (3 - x r^2)/2 cost two more instructions).
__device__ __forceinline__ double rsqrt_nr(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * r, r, 1.0);
    return fma(r * e, fma(e, 0.375, 0.5), r);
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
        const double r = rsqrt_nr(T);
        const double inv_T = r * r;
        F[0] = 0.88622692545275801365 * r;
#pragma unroll
        for (int m = 1; m <= L; ++m) F[m] = F[m - 1] * ((m - 0.5) * inv_T);
    }
}

// Accurate F_0..F_L from an LDS copy of the interval table of m = L
// (DSSCF_BPOLY_COEF x DSSCF_BPOLY_ROWS, coefficient-major): a degree-11
// Taylor polynomial of F_L about the interval centre, e^-T = e^-c e^-u from
// a degree-12 series in u = T - c (|u| <= 1/4), then the downward recursion;
// T >= 40 as boys_accurate.
template <int L>
__device__ __forceinline__ void boys_interval(double T, const double* __restrict__ tab, double* F)
{
    static_assert(DSSCF_BPOLY_TERMS == 12 && DSSCF_BPOLY_COEF == 14, "pair layout below");
    if (T < DSSCF_BOYS_TMAX) {
        const int i = (int)(T * 2.0);
        const double u = T - (0.5 * i + 0.25);
        const double2* c = reinterpret_cast<const double2*>(tab) + i;
        const double2 p0 = c[0], p1 = c[DSSCF_BPOLY_ROWS], p2 = c[2 * DSSCF_BPOLY_ROWS],
                      p3 = c[3 * DSSCF_BPOLY_ROWS], p4 = c[4 * DSSCF_BPOLY_ROWS], p5 = c[5 * DSSCF_BPOLY_ROWS],
                      p6 = c[6 * DSSCF_BPOLY_ROWS];
        double s = p6.x;  // a_11
        s = fma(s, u, p5.y);
        s = fma(s, u, p5.x);
        s = fma(s, u, p4.y);
        s = fma(s, u, p4.x);
        s = fma(s, u, p3.y);
        s = fma(s, u, p3.x);
        s = fma(s, u, p2.y);
        s = fma(s, u, p2.x);
        s = fma(s, u, p1.y);
        s = fma(s, u, p1.x);
        s = fma(s, u, p0.y);  // a_0
        F[L] = s;
        if (L > 0) {
            double e = 1.0 / 479001600.0;  // sum_k (-u)^k / k!, k <= 12
            e = fma(e, -u, 1.0 / 39916800.0);
            e = fma(e, -u, 1.0 / 3628800.0);
            e = fma(e, -u, 1.0 / 362880.0);
            e = fma(e, -u, 1.0 / 40320.0);
            e = fma(e, -u, 1.0 / 5040.0);
            e = fma(e, -u, 1.0 / 720.0);
            e = fma(e, -u, 1.0 / 120.0);
            e = fma(e, -u, 1.0 / 24.0);
            e = fma(e, -u, 1.0 / 6.0);
            e = fma(e, -u, 0.5);
            e = fma(e, -u, 1.0);
            e = fma(e, -u, 1.0);
            const double eT = p0.x * e;
            const double T2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fma(T2, F[m + 1], eT) * (1.0 / (2 * m + 1));
        }
    } else {
        const double r = rsqrt_nr(T);
        const double inv_T = r * r;
        F[0] = 0.88622692545275801365 * r;
#pragma unroll
        for (int m = 1; m <= L; ++m) F[m] = F[m - 1] * ((m - 0.5) * inv_T);
    }
}

template <int BM, int L>
__device__ __forceinline__ void boys(double T, const double* __restrict__ tab, double* F)
{
    if (BM == BOYS_ACCURATE) {
        boys_accurate<L>(T, tab, F);
    } else if (BM == BOYS_ACCURATE_LDS) {
        boys_interval<L>(T, tab, F);
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
// FP32 form for low-bound quartets: setup, the small-T Boys interpolation
// (Taylor polynomial of F_L and e^-T in float) and the values handed to the
// VRR in FP32; the downward recursion and the large-T Boys branch run in
// FP64.
// FP32 interval table in LDS (k_jk stages it from the FP64 interval table of
// m = L): per interval 8 floats, e^-c and the Taylor coefficients a_0..a_6
// of F_L about c; |u| <= 1/4 leaves a truncation below 1.2e-8 relative,
// under FP32 rounding.
#define DSSCF_BPOLY32_COEF 8

// FP32 lists: the downward recursion F_m = (2T F_{m+1} + e^-T)/(2m+1) from
// the FP32 F_L and e^-T runs in FP64 (as the large-T branch does), and only
// its results are rounded to FP32 for the VRR.
template <int L>
__device__ __forceinline__ void boys_down_f64(double T, double eT, float* Fm)
{
    double F = (double)Fm[L];
    const double T2 = 2.0 * T;
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
        F = fma(T2, F, eT) * (1.0 / (2 * m + 1));
        Fm[m] = (float)F;
    }
}

template <int L, bool kLds>
__device__ __forceinline__ void prim_setup_f32(const PrimPair& P, const PrimPair& Q, const ShellPair& bra,
                                               const ShellPair& ket, const double* __restrict__ btab,
                                               PrimQuartet<float>& g, float* Fm)
{
    // kLds: btab points at the block's FP32 interval table and P.P holds P - A
    const float p = (float)P.p, q = (float)Q.p;
    const float inv_pq = __builtin_amdgcn_rcpf(p + q);
    const float PQx = kLds ? (float)(P.P[0] - (Q.P[0] - bra.A[0])) : (float)(P.P[0] - Q.P[0]);
    const float PQy = kLds ? (float)(P.P[1] - (Q.P[1] - bra.A[1])) : (float)(P.P[1] - Q.P[1]);
    const float PQz = kLds ? (float)(P.P[2] - (Q.P[2] - bra.A[2])) : (float)(P.P[2] - Q.P[2]);
    const float rho = p * q * inv_pq;
    const float T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    const float pref = (float)P.Kp * (float)Q.Kp * __builtin_sqrtf(inv_pq);
    if (kLds && T < (float)DSSCF_BOYS_TMAX) {
        const int i = (int)(T * 2.0f);
        const float u = T - (0.5f * (float)i + 0.25f);
        const float4* c = reinterpret_cast<const float4*>(btab) + 2 * i;
        const float4 c0 = c[0], c1 = c[1];  // e^-c, a_0, a_1, a_2 | a_3 .. a_6
        float s = c1.w;
        s = fmaf(s, u, c1.z);
        s = fmaf(s, u, c1.y);
        s = fmaf(s, u, c1.x);
        s = fmaf(s, u, c0.w);
        s = fmaf(s, u, c0.z);
        s = fmaf(s, u, c0.y);
        Fm[L] = s;
        if (L > 0) {
            float e = 1.0f / 720.0f;  // sum_k (-u)^k / k!, k <= 6
            e = fmaf(e, -u, 1.0f / 120.0f);
            e = fmaf(e, -u, 1.0f / 24.0f);
            e = fmaf(e, -u, 1.0f / 6.0f);
            e = fmaf(e, -u, 0.5f);
            e = fmaf(e, -u, 1.0f);
            e = fmaf(e, -u, 1.0f);
            boys_down_f64<L>((double)T, (double)(c0.x * e), Fm);
        }
    } else if (!kLds && T < (float)DSSCF_BOYS_TMAX) {
        const int i = (int)(T * (float)DSSCF_BOYS_INV_H + 0.5f);
        const float nd = (float)i * (float)DSSCF_BOYS_H - T;
        const double* row = btab + i * DSSCF_BOYS_STRIDE;
        float s = (float)row[1 + L + 4];
        s = fmaf(s, nd * 0.25f, (float)row[1 + L + 3]);
        s = fmaf(s, nd * (1.0f / 3.0f), (float)row[1 + L + 2]);
        s = fmaf(s, nd * 0.5f, (float)row[1 + L + 1]);
        s = fmaf(s, nd, (float)row[1 + L]);
        Fm[L] = s;
        if (L > 0) {
            float e = 1.0f / 24.0f;
            e = fmaf(e, nd, 1.0f / 6.0f);
            e = fmaf(e, nd, 0.5f);
            e = fmaf(e, nd, 1.0f);
            e = fmaf(e, nd, 1.0f);
            boys_down_f64<L>((double)T, row[0] * (double)e, Fm);
        }
    } else {
        double F[L + 1];
        const double inv_T = 1.0 / (double)T;
        F[0] = 0.88622692545275801365 * sqrt(inv_T);
#pragma unroll
        for (int m = 1; m <= L; ++m) F[m] = F[m - 1] * ((m - 0.5) * inv_T);
#pragma unroll
        for (int m = 0; m <= L; ++m) Fm[m] = (float)F[m];
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) Fm[m] *= pref;
    const float qf = q * inv_pq, pf = p * inv_pq;
    g.PA[0] = kLds ? (float)P.P[0] : (float)(P.P[0] - bra.A[0]);
    g.PA[1] = kLds ? (float)P.P[1] : (float)(P.P[1] - bra.A[1]);
    g.PA[2] = kLds ? (float)P.P[2] : (float)(P.P[2] - bra.A[2]);
    g.QC[0] = (float)(Q.P[0] - ket.A[0]);
    g.QC[1] = (float)(Q.P[1] - ket.A[1]);
    g.QC[2] = (float)(Q.P[2] - ket.A[2]);
    g.WP[0] = -qf * PQx;
    g.WP[1] = -qf * PQy;
    g.WP[2] = -qf * PQz;
    g.WQ[0] = pf * PQx;
    g.WQ[1] = pf * PQy;
    g.WQ[2] = pf * PQz;
    g.o2p = 0.5f * (float)P.inv_p;
    g.o2q = 0.5f * (float)Q.inv_p;
    g.o2pq = 0.5f * inv_pq;
    g.rp = qf;
    g.rq = pf;
}

// FP32 lists with their bra records staged as PrimPairF (P - A, converted
// once per block) and the interval table as floats in LDS: as
// prim_setup_f32<L, true>, without the per-quartet conversions of the bra
// record (the ket record's are invariant in the bra loop).
template <int BM, typename R, int L>
__device__ __forceinline__ void prim_setup(const PrimPairF& P, const PrimPair& Q, const ShellPair& bra,
                                           const ShellPair& ket, const double* __restrict__ btab,
                                           PrimQuartet<float>& g, float* Fm)
{
    static_assert(BM == BOYS_ACCURATE_LDS && sizeof(R) == 4, "PrimPairF records are staged for FP32 LDS lists");
    const float p = P.p, q = (float)Q.p;
    const float inv_pq = __builtin_amdgcn_rcpf(p + q);
    const float QAx = (float)(Q.P[0] - bra.A[0]), QAy = (float)(Q.P[1] - bra.A[1]), QAz = (float)(Q.P[2] - bra.A[2]);
    const float PQx = P.PA[0] - QAx, PQy = P.PA[1] - QAy, PQz = P.PA[2] - QAz;
    const float rho = p * q * inv_pq;
    const float T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    const float pref = P.Kp * (float)Q.Kp * __builtin_sqrtf(inv_pq);
    if (T < (float)DSSCF_BOYS_TMAX) {
        const int i = (int)(T * 2.0f);
        const float u = T - (0.5f * (float)i + 0.25f);
        const float4* c = reinterpret_cast<const float4*>(btab) + 2 * i;
        const float4 c0 = c[0], c1 = c[1];  // e^-c, a_0, a_1, a_2 | a_3 .. a_6
        float s = c1.w;
        s = fmaf(s, u, c1.z);
        s = fmaf(s, u, c1.y);
        s = fmaf(s, u, c1.x);
        s = fmaf(s, u, c0.w);
        s = fmaf(s, u, c0.z);
        s = fmaf(s, u, c0.y);
        Fm[L] = s;
        if (L > 0) {
            float e = 1.0f / 720.0f;  // sum_k (-u)^k / k!, k <= 6
            e = fmaf(e, -u, 1.0f / 120.0f);
            e = fmaf(e, -u, 1.0f / 24.0f);
            e = fmaf(e, -u, 1.0f / 6.0f);
            e = fmaf(e, -u, 0.5f);
            e = fmaf(e, -u, 1.0f);
            e = fmaf(e, -u, 1.0f);
            boys_down_f64<L>((double)T, (double)(c0.x * e), Fm);
        }
    } else {
        double F[L + 1];
        const double inv_T = 1.0 / (double)T;
        F[0] = 0.88622692545275801365 * sqrt(inv_T);
#pragma unroll
        for (int m = 1; m <= L; ++m) F[m] = F[m - 1] * ((m - 0.5) * inv_T);
#pragma unroll
        for (int m = 0; m <= L; ++m) Fm[m] = (float)F[m];
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) Fm[m] *= pref;
    const float qf = q * inv_pq, pf = p * inv_pq;
    g.PA[0] = P.PA[0];
    g.PA[1] = P.PA[1];
    g.PA[2] = P.PA[2];
    g.QC[0] = (float)(Q.P[0] - ket.A[0]);
    g.QC[1] = (float)(Q.P[1] - ket.A[1]);
    g.QC[2] = (float)(Q.P[2] - ket.A[2]);
    g.WP[0] = -qf * PQx;
    g.WP[1] = -qf * PQy;
    g.WP[2] = -qf * PQz;
    g.WQ[0] = pf * PQx;
    g.WQ[1] = pf * PQy;
    g.WQ[2] = pf * PQz;
    g.o2p = 0.5f * P.inv_p;
    g.o2q = 0.5f * (float)Q.inv_p;
    g.o2pq = 0.5f * inv_pq;
    g.rp = qf;
    g.rq = pf;
}

template <int BM, typename R, int L>
__device__ __forceinline__ void prim_setup(const PrimPair& P, const PrimPair& Q, const ShellPair& bra,
                                           const ShellPair& ket, const double* __restrict__ btab,
                                           PrimQuartet<R>& g, R* Fm)
{
    if constexpr (sizeof(R) == 4) {
        prim_setup_f32<L, BM == BOYS_ACCURATE_LDS>(P, Q, bra, ket, btab, g, Fm);
        return;
    } else {
    const double p = P.p, q = Q.p;
    const double rs = rsqrt_nr(p + q);
    const double inv_pq = rs * rs;
    constexpr bool kPA = BM == BOYS_ACCURATE_LDS;  // P.P holds P - A
    // P - Q = (P - A) - (Q - A): Q - A is invariant in the bra-primitive loop
    const double PQx = kPA ? P.P[0] - (Q.P[0] - bra.A[0]) : P.P[0] - Q.P[0];
    const double PQy = kPA ? P.P[1] - (Q.P[1] - bra.A[1]) : P.P[1] - Q.P[1];
    const double PQz = kPA ? P.P[2] - (Q.P[2] - bra.A[2]) : P.P[2] - Q.P[2];
    const double rho = p * q * inv_pq;
    const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    const double pref = P.Kp * Q.Kp * rs;
    double F[L + 1];
    boys<BM, L>(T, btab, F);
#pragma unroll
    for (int m = 0; m <= L; ++m) Fm[m] = (R)(F[m] * pref);
    const double qf = q * inv_pq, pf = p * inv_pq;
    g.PA[0] = (R)(kPA ? P.P[0] : P.P[0] - bra.A[0]);
    g.PA[1] = (R)(kPA ? P.P[1] : P.P[1] - bra.A[1]);
    g.PA[2] = (R)(kPA ? P.P[2] : P.P[2] - bra.A[2]);
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
}

// C = 1: the class's d shells are Cartesian functions (6D) instead of real
// solid harmonics; only classes with a d shell have that form.
template <int la, int lb, int lc, int ld, int C = 0>
struct ClassInfo;
template <int la, int lb, int lc, int ld, int C = 0>
struct WaveTab;
