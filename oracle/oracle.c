/*
 * oracle.c -- CPU restatement of DollarStoreSCF's integral and J/K path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this file's library, and only as the
 * checker (or the timed CPU baseline).  The product path never calls it.
 *
 * Every function restates one reference function; the citation names the
 * file:line of DollarStoreSCF it follows.  The AO description passed in is
 * built by oracle/oracle.py (itself a restatement of GTO.cpp:15-94 and
 * SCFBasis.cpp:120-140).
 *
 * Pinning (see tests/test_oracle.py):
 *   orc_boys_ref       == reference boysFunctionExact (oracle/_ref) bit-for-bit,
 *                         F_m(0)=1/(2m+1), F_0 = sqrt(pi) erf(sqrt T)/(2 sqrt T)
 *   orc_boys_accurate  vs mpmath fixture (tests/golden/boys_mpmath.json)
 *   orc_eri            vs reference electronRepulsionIntegral (oracle/_ref and
 *                         tests/golden fixtures), (ss|ss) closed form,
 *                         8-fold symmetry, translation/rotation invariance
 *   orc_overlap/kinetic/nuclear vs reference (oracle/_ref, fixtures), S_ii,
 *                         closed-form s-type integrals
 *   orc_coulomb/exchange vs direct contraction of the full tensor in numpy
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#define ORC_PI 3.14159265358979 /* src/chem/Constants.hpp:5, src/lalib/MathUtil.hpp:8 */

/* ---------------------------------------------------------------- basis -- */
typedef struct {
    int n_ao;
    const double* center;    /* 3*n_ao, bohr */
    const int* l;            /* n_ao */
    const int* prim_off;     /* n_ao */
    const int* nprim;        /* n_ao */
    const double* zeta;      /* primitive exponents */
    const double* weight;    /* normalized primitive weights (GTO.cpp:90-93) */
    const int* term_off;     /* n_ao */
    const int* nterm;        /* n_ao */
    const double* term_w;    /* Cartesian term weights (GTO.cpp:15-88) */
    const int* term_xyz;     /* 3 per term */
} orc_basis;

/* ------------------------------------------------------------- helpers -- */
/* flo::pow(T x, uint p), src/lalib/MathUtil.hpp:57-67: repeated squaring with
 * the reference's exact multiplication order (kept alike for bit parity of the
 * asymptotic Boys branch). */
static double flo_pow(double x, unsigned p)
{
    if (p == 1) return x;
    if (p == 0) return 1.0;
    double y = x;
    unsigned l = 2;
    for (; l < p; l *= 2) y = y * y;
    return y * flo_pow(x, p - l / 2);
}

static int int_pow(int x, unsigned p)
{
    if (p == 1) return x;
    if (p == 0) return 1;
    int y = x;
    unsigned l = 2;
    for (; l < p; l *= 2) y = y * y;
    return y * int_pow(x, p - l / 2);
}

static int int_double_factorial(int x)
{
    int r = 1;
    for (int i = x; i > 0; i -= 2) r *= i;
    return r;
}

/* ---------------------------------------------------------------- Boys -- */
/* boysFunctionExact, src/chem/GaussianIntegral.cpp:209-223.
 * Asymptotic branch for x > 1.1m+25, otherwise the series
 * a_i = a_{i-1} * 2x/(2m+2i+1), stopping after the first a_i <= 1e-9. */
double orc_boys_ref(double x, unsigned m)
{
    if (x > (double)m * 1.1 + 25.0) {
        return (double)int_double_factorial(2 * (int)m - 1) / (double)int_pow(2, m + 1)
               * sqrt(ORC_PI / flo_pow(x, 2 * m + 1));
    }
    double F = 1.0, a = 1.0;
    for (int i = 1; a > 0.000000001; ++i) {
        a = 2.0 * x / (double)(2 * m + 2 * (unsigned)i + 1) * a;
        F += a;
    }
    return exp(-x) / (double)(2 * m + 1) * F;
}

/* Accurate F_m(x): the same series in long double run to convergence
 * (x < 100), else the asymptotic form whose neglected term is below
 * e^-100.  Not a reference function: the accuracy yardstick for the
 * accurate device mode. */
double orc_boys_accurate(double x, unsigned m)
{
    if (x < 100.0) {
        long double X = x, F = 1.0L, a = 1.0L;
        for (int i = 1; i < 100000; ++i) {
            a = a * 2.0L * X / (long double)(2 * m + 2 * (unsigned)i + 1);
            F += a;
            if (a < 1e-22L * F) break;
        }
        return (double)(expl(-X) / (long double)(2 * m + 1) * F);
    }
    long double X = x, df = 1.0L;
    for (int i = 2 * (int)m - 1; i > 0; i -= 2) df *= i;
    return (double)(df / powl(2.0L, m + 1) * sqrtl(3.14159265358979323846264338327950288L / powl(X, 2 * m + 1)));
}

typedef double (*boys_fn)(double, unsigned);
static boys_fn pick_boys(int mode) { return mode == 0 ? orc_boys_ref : orc_boys_accurate; }

/* ------------------------------------------------ 1-D Obara-Saika ERI pass -- */
/* OSRepulsion, src/chem/GaussianIntegral.cpp:334-559.  One Cartesian axis of
 * the four-centre OS recursion with the auxiliary index N carried through
 * buf (length Nt+i+j+k+l+1 in, Nt+1 out).  Indices are raised in the
 * reference's order: i, then j, then k, then l; all coupling terms of the
 * OS scheme are kept (no HRR).  Own loop structure, same recursion. */
#define OS_MAXL 4
#define OS_MAXN 24
static void os_repulsion_1d(double* buf, double A, double B, double C, double D,
                            double P, double Q, double eta, double zeta,
                            int i, int j, int k, int l, int Nt)
{
    const double rho = eta * zeta / (eta + zeta);
    const double PA = P - A, PB = P - B, QC = Q - C, QD = Q - D;
    const double WP = -rho / eta * (P - Q), WQ = -rho / zeta * (Q - P);
    const double o2e = 0.5 / eta, o2z = 0.5 / zeta, o2ez = 0.5 / (eta + zeta);
    const double re = rho / eta, rz = rho / zeta;
    const int Ntot = Nt + i + j + k + l;
    /* X[i'][j'][k'][l'][N] */
    static __thread double X[OS_MAXL + 1][OS_MAXL + 1][OS_MAXL + 1][OS_MAXL + 1][OS_MAXN];
#define XX(a, b, c, d, n) X[a][b][c][d][n]
    for (int n = 0; n <= Ntot; ++n) XX(0, 0, 0, 0, n) = buf[n];
    for (int d = 0; d <= l; ++d)
    for (int c = 0; c <= k; ++c)
    for (int b = 0; b <= j; ++b)
    for (int a = 0; a <= i; ++a) {
        if (a + b + c + d == 0) continue;
        const int nmax = Ntot - (a + b + c + d);
        for (int n = 0; n <= nmax; ++n) {
            double v;
            if (d > 0) {
                v = QD * XX(a, b, c, d - 1, n) + WQ * XX(a, b, c, d - 1, n + 1);
                if (c > 0) v += c * o2z * (XX(a, b, c - 1, d - 1, n) - rz * XX(a, b, c - 1, d - 1, n + 1));
                if (d > 1) v += (d - 1) * o2z * (XX(a, b, c, d - 2, n) - rz * XX(a, b, c, d - 2, n + 1));
                if (a > 0) v += a * o2ez * XX(a - 1, b, c, d - 1, n + 1);
                if (b > 0) v += b * o2ez * XX(a, b - 1, c, d - 1, n + 1);
            } else if (c > 0) {
                v = QC * XX(a, b, c - 1, 0, n) + WQ * XX(a, b, c - 1, 0, n + 1);
                if (c > 1) v += (c - 1) * o2z * (XX(a, b, c - 2, 0, n) - rz * XX(a, b, c - 2, 0, n + 1));
                if (a > 0) v += a * o2ez * XX(a - 1, b, c - 1, 0, n + 1);
                if (b > 0) v += b * o2ez * XX(a, b - 1, c - 1, 0, n + 1);
            } else if (b > 0) {
                v = PB * XX(a, b - 1, 0, 0, n) + WP * XX(a, b - 1, 0, 0, n + 1);
                if (a > 0) v += a * o2e * (XX(a - 1, b - 1, 0, 0, n) - re * XX(a - 1, b - 1, 0, 0, n + 1));
                if (b > 1) v += (b - 1) * o2e * (XX(a, b - 2, 0, 0, n) - re * XX(a, b - 2, 0, 0, n + 1));
            } else {
                v = PA * XX(a - 1, 0, 0, 0, n) + WP * XX(a - 1, 0, 0, 0, n + 1);
                if (a > 1) v += (a - 1) * o2e * (XX(a - 2, 0, 0, 0, n) - re * XX(a - 2, 0, 0, 0, n + 1));
            }
            XX(a, b, c, d, n) = v;
        }
    }
    for (int n = 0; n <= Nt; ++n) buf[n] = XX(i, j, k, l, n);
#undef XX
}

/* electronRepulsionIntegral, src/chem/GaussianIntegral.cpp:561-642.
 * Returns the chemists' (ij|kl): i,j on electron 1 (centres A,B), k,l on
 * electron 2 (centres C,D).  (The reference's own argument order is
 * (mu, sigma, nu, tau) -> (mu nu|sigma tau).)  mode 0: reference Boys;
 * mode 1: accurate Boys. */
double orc_eri(const orc_basis* B, int i, int j, int k, int l, int mode)
{
    boys_fn boys = pick_boys(mode);
    const double* A = B->center + 3 * i;
    const double* Bc = B->center + 3 * j;
    const double* C = B->center + 3 * k;
    const double* D = B->center + 3 * l;
    const int L = B->l[i] + B->l[j] + B->l[k] + B->l[l];
    double F[OS_MAXN], buf[OS_MAXN];
    double integral = 0.0;
    const double AB2 = (A[0] - Bc[0]) * (A[0] - Bc[0]) + (A[1] - Bc[1]) * (A[1] - Bc[1]) + (A[2] - Bc[2]) * (A[2] - Bc[2]);
    const double CD2 = (C[0] - D[0]) * (C[0] - D[0]) + (C[1] - D[1]) * (C[1] - D[1]) + (C[2] - D[2]) * (C[2] - D[2]);
    const double pref0 = 2.0 * pow(ORC_PI, 2.5);
    for (int ip = 0; ip < B->nprim[i]; ++ip)
    for (int jp = 0; jp < B->nprim[j]; ++jp)
    for (int kp = 0; kp < B->nprim[k]; ++kp)
    for (int lp = 0; lp < B->nprim[l]; ++lp) {
        const double al = B->zeta[B->prim_off[i] + ip], be = B->zeta[B->prim_off[j] + jp];
        const double ga = B->zeta[B->prim_off[k] + kp], de = B->zeta[B->prim_off[l] + lp];
        const double eta = al + be, zeta = ga + de;
        double P[3], Q[3];
        for (int x = 0; x < 3; ++x) {
            P[x] = (al * A[x] + be * Bc[x]) / eta;
            Q[x] = (ga * C[x] + de * D[x]) / zeta;
        }
        const double g1 = exp(-al * be / eta * AB2);
        const double g2 = exp(-ga * de / zeta * CD2);
        const double RPQ2 = (P[0] - Q[0]) * (P[0] - Q[0]) + (P[1] - Q[1]) * (P[1] - Q[1]) + (P[2] - Q[2]) * (P[2] - Q[2]);
        const double pw = B->weight[B->prim_off[i] + ip] * B->weight[B->prim_off[j] + jp]
                        * B->weight[B->prim_off[k] + kp] * B->weight[B->prim_off[l] + lp];
        for (int m = 0; m <= L; ++m) F[m] = boys(eta * zeta * RPQ2 / (eta + zeta), (unsigned)m);
        for (int ti = 0; ti < B->nterm[i]; ++ti)
        for (int tj = 0; tj < B->nterm[j]; ++tj)
        for (int tk = 0; tk < B->nterm[k]; ++tk)
        for (int tl = 0; tl < B->nterm[l]; ++tl) {
            const int* ei = B->term_xyz + 3 * (B->term_off[i] + ti);
            const int* ej = B->term_xyz + 3 * (B->term_off[j] + tj);
            const int* ek = B->term_xyz + 3 * (B->term_off[k] + tk);
            const int* el = B->term_xyz + 3 * (B->term_off[l] + tl);
            const int N1 = ei[2] + ej[2] + ek[2] + el[2];
            const int N2 = N1 + ei[1] + ej[1] + ek[1] + el[1];
            const double aw = B->term_w[B->term_off[i] + ti] * B->term_w[B->term_off[j] + tj]
                            * B->term_w[B->term_off[k] + tk] * B->term_w[B->term_off[l] + tl];
            memcpy(buf, F, sizeof(double) * (L + 1));
            os_repulsion_1d(buf, A[0], Bc[0], C[0], D[0], P[0], Q[0], eta, zeta, ei[0], ej[0], ek[0], el[0], N2);
            os_repulsion_1d(buf, A[1], Bc[1], C[1], D[1], P[1], Q[1], eta, zeta, ei[1], ej[1], ek[1], el[1], N1);
            os_repulsion_1d(buf, A[2], Bc[2], C[2], D[2], P[2], Q[2], eta, zeta, ei[2], ej[2], ek[2], el[2], 0);
            integral += pref0 / (eta * zeta * sqrt(eta + zeta)) * g1 * g2 * pw * aw * buf[0];
        }
    }
    return integral;
}

void orc_eri_list(const orc_basis* B, const int* q, long n, int mode, double* out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long t = 0; t < n; ++t) out[t] = orc_eri(B, q[4 * t], q[4 * t + 1], q[4 * t + 2], q[4 * t + 3], mode);
}

/* ----------------------------------------------- packed ERI tensor + J/K -- */
/* assignExactRepulsionTensor, src/chem/scf/ExactCoulomb.cpp:41-72, including
 * its screen (ii|jj)(kk|ll) > 1e-16 (ExactCoulomb.cpp:11,63) and its packed
 * index T - (D-m)(D-m+1)/2 + n - m (ExactCoulomb.cpp:58-61). */
static inline size_t tensor_index(size_t D, size_t T, size_t m, size_t n)
{
    return T - (D - m) * (D - m + 1) / 2 + n - m;
}

/* Threads of the OpenMP loops below (the reference itself is single-threaded:
 * its CPU baseline runs them with one). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_build_tensor(const orc_basis* B, int mode, double* tensor)
{
    const int N = B->n_ao;
    const size_t D = (size_t)N * (N + 1) / 2, T = D * (D + 1) / 2;
    double* diag = (double*)malloc(sizeof(double) * (size_t)N * N);
    for (int a = 0; a < N; ++a)
        for (int b = 0; b <= a; ++b) {
            diag[a * N + b] = diag[b * N + a] = orc_eri(B, a, a, b, b, mode);
        }
    memset(tensor, 0, sizeof(double) * T);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k < N; ++k)
                for (int l = 0; l <= k; ++l) {
                    size_t m = (size_t)j + (size_t)i * (i + 1) / 2;
                    size_t n = (size_t)l + (size_t)k * (k + 1) / 2;
                    if (m > n) continue;
                    if (diag[i * N + j] * diag[k * N + l] > 0.00000001 * 0.00000001)
                        tensor[tensor_index(D, T, m, n)] = orc_eri(B, i, j, k, l, mode);
                }
    free(diag);
}

/* sampleRepulsionTensor, ExactCoulomb.cpp:74-101 ((ij|kl) from the packed
 * tensor; the reference takes (i,k,j,l)). */
static inline double sample(const double* t, size_t D, size_t T, unsigned i, unsigned j, unsigned k, unsigned l)
{
    if (j > i) { unsigned x = i; i = j; j = x; }
    if (l > k) { unsigned x = k; k = l; l = x; }
    size_t m = j + (size_t)i * (i + 1) / 2;
    size_t n = l + (size_t)k * (k + 1) / 2;
    if (m > n) { size_t x = m; m = n; n = x; }
    return t[tensor_index(D, T, m, n)];
}

/* packed-U element (y,x): data[max(x,y)(max+1)/2 + min]  (Matrix.ipp:332-351) */
static inline double pk(const double* p, unsigned y, unsigned x)
{
    if (x < y) { unsigned t = x; x = y; y = t; }
    return p[(size_t)x * (x + 1) / 2 + y];
}

/* computeCoulombTerms + getExactCoulombMatrix (SCFSolver.cpp:29-35,
 * ExactCoulomb.cpp:103-112): J_mn = sum_t sum_{s<=t} (2-d_st) P_st (mn|st). */
void orc_coulomb(const double* tensor, int N, const double* Ptot, double* J)
{
    const size_t D = (size_t)N * (N + 1) / 2, T = D * (D + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
    for (int nu = 0; nu < N; ++nu)
        for (int mu = 0; mu <= nu; ++mu) {
            double sum = 0.0;
            for (unsigned tau = 0; tau < (unsigned)N; ++tau)
                for (unsigned sigma = 0; sigma <= tau; ++sigma) {
                    double v = pk(Ptot, sigma, tau) * sample(tensor, D, T, mu, nu, sigma, tau);
                    sum += (tau == sigma) ? v : 2.0 * v;
                }
            J[(size_t)nu * (nu + 1) / 2 + mu] = sum;
        }
}

/* computeExchangeTerms + getExactExchangeMatrix (SCFSolver.cpp:37-43,
 * ExactCoulomb.cpp:114-123): K_mn = sum_s sum_t P_st (ms|tn). */
void orc_exchange(const double* tensor, int N, const double* Pa, double* K)
{
    const size_t D = (size_t)N * (N + 1) / 2, T = D * (D + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
    for (int nu = 0; nu < N; ++nu)
        for (int mu = 0; mu <= nu; ++mu) {
            double sum = 0.0;
            for (unsigned sigma = 0; sigma < (unsigned)N; ++sigma)
                for (unsigned tau = 0; tau < (unsigned)N; ++tau)
                    sum += pk(Pa, sigma, tau) * sample(tensor, D, T, mu, sigma, tau, nu);
            K[(size_t)nu * (nu + 1) / 2 + mu] = sum;
        }
}

/* -------------------------------------------------- one-electron terms -- */
/* OSOverlap, GaussianIntegral.cpp:29-64: 1-D overlap S_ij / S00 by the OS
 * recursion S_{i+1,j} = PA S_ij + (i S_{i-1,j} + j S_{i,j-1})/(2 eta). */
static void os_overlap_table(double S[][12], double S00, double eta, double PA, double PB, int imax, int jmax)
{
    S[0][0] = S00;
    for (int a = 1; a <= imax; ++a)
        S[a][0] = PA * S[a - 1][0] + (a > 1 ? (a - 1) * S[a - 2][0] / (2 * eta) : 0.0);
    for (int b = 1; b <= jmax; ++b)
        for (int a = 0; a <= imax; ++a) {
            double v = PB * S[a][b - 1];
            if (a > 0) v += a * S[a - 1][b - 1] / (2 * eta);
            if (b > 1) v += (b - 1) * S[a][b - 2] / (2 * eta);
            S[a][b] = v;
        }
}

/* overlapIntegral, GaussianIntegral.cpp:66-92 (type 0), kineticEnergyIntegral
 * GaussianIntegral.cpp:159-193 with OSKineticEnergy :98-157 (type 1). */
static double one_e(const orc_basis* B, int i, int j, int type)
{
    const double* A = B->center + 3 * i;
    const double* Bc = B->center + 3 * j;
    double AB2 = 0;
    for (int x = 0; x < 3; ++x) AB2 += (A[x] - Bc[x]) * (A[x] - Bc[x]);
    double integral = 0.0;
    for (int ip = 0; ip < B->nprim[i]; ++ip)
        for (int jp = 0; jp < B->nprim[j]; ++jp) {
            const double al = B->zeta[B->prim_off[i] + ip], be = B->zeta[B->prim_off[j] + jp];
            const double eta = al + be;
            const double g = exp(-al * be / eta * AB2);
            const double pw = B->weight[B->prim_off[i] + ip] * B->weight[B->prim_off[j] + jp];
            const double S00 = sqrt(ORC_PI / eta);
            for (int ti = 0; ti < B->nterm[i]; ++ti)
                for (int tj = 0; tj < B->nterm[j]; ++tj) {
                    const int* ei = B->term_xyz + 3 * (B->term_off[i] + ti);
                    const int* ej = B->term_xyz + 3 * (B->term_off[j] + tj);
                    const double aw = B->term_w[B->term_off[i] + ti] * B->term_w[B->term_off[j] + tj];
                    double Sd[3], Dd[3];
                    for (int x = 0; x < 3; ++x) {
                        const double P = (al * A[x] + be * Bc[x]) / eta;
                        double S[12][12];
                        os_overlap_table(S, S00, eta, P - A[x], P - Bc[x], ei[x] + 2, ej[x]);
                        const int a = ei[x], b = ej[x];
                        Sd[x] = S[a][b];
                        /* d^2/dx^2 on the bra factor (OSKineticEnergy's T_ij). */
                        Dd[x] = 4.0 * al * al * S[a + 2][b] - 2.0 * al * (2 * a + 1) * S[a][b]
                              + (a > 1 ? (double)(a * (a - 1)) * S[a - 2][b] : 0.0);
                    }
                    double v = type == 0 ? Sd[0] * Sd[1] * Sd[2]
                                         : Dd[0] * Sd[1] * Sd[2] + Sd[0] * Dd[1] * Sd[2] + Sd[0] * Sd[1] * Dd[2];
                    integral += g * aw * pw * v;
                }
        }
    return type == 0 ? integral : -0.5 * integral;
}

double orc_overlap(const orc_basis* B, int i, int j) { return one_e(B, i, j, 0); }
double orc_kinetic(const orc_basis* B, int i, int j) { return one_e(B, i, j, 1); }

/* OSNuclearPotential, GaussianIntegral.cpp:225-291: 1-D nuclear-attraction
 * pass with auxiliary index N, Theta_{i+1,j}^N = PA Th^N - PC Th^{N+1}
 * + i/(2eta)(Th_{i-1}^N - Th_{i-1}^{N+1}) + j/(2eta)(...). */
static void os_nuclear_1d(double* buf, double eta, double A, double B, double C, double P, int i, int j, int Nt)
{
    const int Ntot = Nt + i + j;
    double X[8][8][OS_MAXN];
    const double PA = P - A, PB = P - B, PC = P - C, o2e = 0.5 / eta;
    for (int n = 0; n <= Ntot; ++n) X[0][0][n] = buf[n];
    for (int b = 0; b <= j; ++b)
        for (int a = 0; a <= i; ++a) {
            if (a + b == 0) continue;
            for (int n = 0; n <= Ntot - a - b; ++n) {
                double v;
                if (b > 0) {
                    v = PB * X[a][b - 1][n] - PC * X[a][b - 1][n + 1];
                    if (a > 0) v += a * o2e * (X[a - 1][b - 1][n] - X[a - 1][b - 1][n + 1]);
                    if (b > 1) v += (b - 1) * o2e * (X[a][b - 2][n] - X[a][b - 2][n + 1]);
                } else {
                    v = PA * X[a - 1][0][n] - PC * X[a - 1][0][n + 1];
                    if (a > 1) v += (a - 1) * o2e * (X[a - 2][0][n] - X[a - 2][0][n + 1]);
                }
                X[a][b][n] = v;
            }
        }
    for (int n = 0; n <= Nt; ++n) buf[n] = X[i][j][n];
}

/* nuclearPotentialIntegral, GaussianIntegral.cpp:293-332, summed over nuclei
 * with -Z as assignCoreHamiltonian does (SCFBasis.cpp:75-80). */
double orc_nuclear(const orc_basis* B, int i, int j, int natom, const int* Z, const double* xyz, int mode)
{
    boys_fn boys = pick_boys(mode);
    const double* A = B->center + 3 * i;
    const double* Bc = B->center + 3 * j;
    double AB2 = 0;
    for (int x = 0; x < 3; ++x) AB2 += (A[x] - Bc[x]) * (A[x] - Bc[x]);
    double total = 0.0;
    for (int c = 0; c < natom; ++c) {
        const double* Cn = xyz + 3 * c;
        double integral = 0.0;
        for (int ti = 0; ti < B->nterm[i]; ++ti)
            for (int tj = 0; tj < B->nterm[j]; ++tj) {
                const int* ei = B->term_xyz + 3 * (B->term_off[i] + ti);
                const int* ej = B->term_xyz + 3 * (B->term_off[j] + tj);
                const int N1 = ei[2] + ej[2], N2 = N1 + ei[1] + ej[1], N3 = N2 + ei[0] + ej[0];
                const double aw = B->term_w[B->term_off[i] + ti] * B->term_w[B->term_off[j] + tj];
                for (int ip = 0; ip < B->nprim[i]; ++ip)
                    for (int jp = 0; jp < B->nprim[j]; ++jp) {
                        const double al = B->zeta[B->prim_off[i] + ip], be = B->zeta[B->prim_off[j] + jp];
                        const double eta = al + be;
                        double P[3], RPC2 = 0;
                        for (int x = 0; x < 3; ++x) {
                            P[x] = (al * A[x] + be * Bc[x]) / eta;
                            RPC2 += (P[x] - Cn[x]) * (P[x] - Cn[x]);
                        }
                        const double g = exp(-al * be / eta * AB2);
                        const double pw = B->weight[B->prim_off[i] + ip] * B->weight[B->prim_off[j] + jp];
                        double buf[OS_MAXN];
                        for (int m = 0; m <= N3; ++m) buf[m] = boys(eta * RPC2, (unsigned)m);
                        os_nuclear_1d(buf, eta, A[0], Bc[0], Cn[0], P[0], ei[0], ej[0], N2);
                        os_nuclear_1d(buf, eta, A[1], Bc[1], Cn[1], P[1], ei[1], ej[1], N1);
                        os_nuclear_1d(buf, eta, A[2], Bc[2], Cn[2], P[2], ei[2], ej[2], 0);
                        integral += 2.0 * ORC_PI / eta * g * pw * aw * buf[0];
                    }
            }
        total -= (double)Z[c] * integral;
    }
    return total;
}

/* Packed-U S, T, V for all pairs (SCFBasis.cpp:56-83). */
void orc_one_electron(const orc_basis* B, int natom, const int* Z, const double* xyz, int mode,
                      double* S, double* T, double* V)
{
    const int N = B->n_ao;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j) {
            size_t idx = (size_t)i * (i + 1) / 2 + j;
            S[idx] = orc_overlap(B, j, i);
            T[idx] = orc_kinetic(B, j, i);
            V[idx] = orc_nuclear(B, j, i, natom, Z, xyz, mode);
        }
}
