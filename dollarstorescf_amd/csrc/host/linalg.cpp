// linalg.cpp -- see linalg.hpp.
//
// Compiled twice: as is (any x86-64) and with -DLA_AVX2 -mavx2 -mfma; the
// gemm and syev of la:: pick the AVX2/FMA build at run time where the host
// CPU has it.  The other functions exist once.
#include "linalg.hpp"

#include <algorithm>
#include <cmath>

#ifdef LA_AVX2
#define LA_KERNELS la_avx2
#else
#define LA_KERNELS la_base
#endif

namespace LA_KERNELS {
using la::Mat;

void gemm(const Mat& A, const Mat& B, Mat& C, int n)
{
    C.assign((size_t)n * n, 0.0);
    const int bs = 64;
#pragma omp parallel for schedule(static)
    for (int ib = 0; ib < n; ib += bs)
        for (int kb = 0; kb < n; kb += bs)
            for (int i = ib; i < std::min(ib + bs, n); ++i)
                for (int k = kb; k < std::min(kb + bs, n); ++k) {
                    const double a = A[(size_t)i * n + k];
                    const double* b = &B[(size_t)k * n];
                    double* c = &C[(size_t)i * n];
                    for (int j = 0; j < n; ++j) c[j] += a * b[j];
                }
}

// Symmetric eigensolver in two stages (Golub & Van Loan, "Matrix
// Computations", 8.3):
//  1. Householder tridiagonalization T = Q^T A Q: for column k a reflector
//     H = I - 2 v v^T (|v| = 1) maps A[k+1:, k] onto a multiple of e_1;
//     with p = A v and w = 2 (p - (v^T p) v), H A H = A - v w^T - w v^T on the
//     trailing block.  Q = H_0 H_1 ... is accumulated afterwards.
//  2. Implicit symmetric QR steps with the Wilkinson shift on the unreduced
//     blocks of T, each step a chase of Givens rotations (the first from the
//     shifted first column, the rest zeroing the bulge); the rotations are
//     applied to the eigenvector rows Z (row j = eigenvector j), which start
//     as the columns of Q.  Off-diagonals with |e_k| <= eps (|d_k|+|d_k+1|)
//     are set to zero (deflation).
// Built twice, for AVX2/FMA hosts and for any x86-64, the loader picking one
// at run time.
bool syev(const Mat& A0, int n, std::vector<double>& w,
                                                                       Mat& V)
{
    w.assign(n, 0.0);
    V.assign((size_t)n * n, 0.0);
    if (n == 0) return true;
    Mat A = A0;
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    std::vector<double> refl((size_t)n * n, 0.0);  // row k: reflector of step k (entries k+1..n-1)
    std::vector<double> p(n), wv(n);
    for (int k = 0; k + 2 < n; ++k) {
        double nx = 0.0;
        for (int i = k + 1; i < n; ++i) nx += a(i, k) * a(i, k);
        nx = std::sqrt(nx);
        if (nx == 0.0) continue;
        const double alpha = a(k + 1, k) > 0.0 ? -nx : nx;
        double* v = &refl[(size_t)k * n];
        double nv = 0.0;
        for (int i = k + 1; i < n; ++i) {
            v[i] = a(i, k) - (i == k + 1 ? alpha : 0.0);
            nv += v[i] * v[i];
        }
        nv = std::sqrt(nv);
        if (nv == 0.0) continue;
        for (int i = k + 1; i < n; ++i) v[i] /= nv;
        double vp = 0.0;
        for (int i = k + 1; i < n; ++i) {
            double s = 0.0;
            const double* ai = &A[(size_t)i * n];
            for (int j = k + 1; j < n; ++j) s += ai[j] * v[j];
            p[i] = s;
            vp += v[i] * s;
        }
        for (int i = k + 1; i < n; ++i) wv[i] = 2.0 * (p[i] - vp * v[i]);
        for (int i = k + 1; i < n; ++i) {
            double* ai = &A[(size_t)i * n];
            const double vi = v[i], wi = wv[i];
            for (int j = k + 1; j < n; ++j) ai[j] -= vi * wv[j] + wi * v[j];
        }
        a(k + 1, k) = a(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
    // Q = H_0 H_1 ... H_{n-3}, built as Z <- Z H_k from Z = I (each H_k acts on
    // columns k+1..n-1), then transposed so that row j of Z is column j of Q.
    // The QR rotations below act on rows of Z: at the end row j is the
    // eigenvector of d[j].
    Mat& Z = V;
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    for (int k = 0; k + 2 < n; ++k) {
        const double* v = &refl[(size_t)k * n];
        bool any = false;
        for (int i = k + 1; i < n; ++i) any = any || v[i] != 0.0;
        if (!any) continue;
        for (int j = 0; j < n; ++j) {
            double* zj = &Z[(size_t)j * n];
            double s = 0.0;
            for (int i = k + 1; i < n; ++i) s += zj[i] * v[i];
            s *= 2.0;
            if (s != 0.0)
                for (int i = k + 1; i < n; ++i) zj[i] -= s * v[i];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) std::swap(Z[(size_t)i * n + j], Z[(size_t)j * n + i]);

    const double eps = 2.220446049250313e-16;
    auto rotate_rows = [&](int k, double c, double s) {  // z_k, z_k+1 <- G [z_k; z_k+1]
        double* zk = &Z[(size_t)k * n];
        double* zl = &Z[(size_t)(k + 1) * n];
        for (int i = 0; i < n; ++i) {
            const double x = zk[i], y = zl[i];
            zk[i] = c * x + s * y;
            zl[i] = -s * x + c * y;
        }
    };
    bool ok = true;
    int hi = n - 1;
    int iters = 0;
    while (hi > 0) {
        for (int k = 0; k < hi; ++k)
            if (std::fabs(e[k]) <= eps * (std::fabs(d[k]) + std::fabs(d[k + 1]))) e[k] = 0.0;
        if (e[hi - 1] == 0.0) {  // d[hi] converged
            --hi;
            iters = 0;
            continue;
        }
        int lo = hi - 1;
        while (lo > 0 && e[lo - 1] != 0.0) --lo;
        if (++iters > 60) {
            ok = false;
            break;
        }
        // Wilkinson shift from the trailing 2x2 of the block [lo, hi]
        const double dl = (d[hi - 1] - d[hi]) * 0.5, eh = e[hi - 1];
        // (plain square roots: the entries are Fock-matrix sized, far from
        // overflow, and hypot's care costs more than the rest of the chase)
        const double mu = d[hi] - eh * eh / (dl + (dl >= 0.0 ? 1.0 : -1.0) * std::sqrt(dl * dl + eh * eh));
        double x = d[lo] - mu, z = e[lo], bulge = 0.0;
        for (int k = lo; k < hi; ++k) {
            const double r = std::sqrt(x * x + z * z);
            const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
            // T <- G T G^T on rows/cols (k, k+1), G = [[c, s], [-s, c]]
            if (k > lo) e[k - 1] = r;  // the rotation folded the bulge into e[k-1]
            const double dk = d[k], dk1 = d[k + 1], ek = e[k];
            d[k] = c * c * dk + 2.0 * c * s * ek + s * s * dk1;
            d[k + 1] = s * s * dk - 2.0 * c * s * ek + c * c * dk1;
            e[k] = c * s * (dk1 - dk) + (c * c - s * s) * ek;
            if (k + 1 < hi) {
                const double ek1 = e[k + 1];
                bulge = s * ek1;
                e[k + 1] = c * ek1;
            }
            rotate_rows(k, c, s);
            x = e[k];
            z = bulge;
        }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    Mat Vs((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        w[j] = d[order[j]];
        const double* zj = &Z[(size_t)order[j] * n];
        for (int i = 0; i < n; ++i) Vs[(size_t)i * n + j] = zj[i];
    }
    V.swap(Vs);
    return ok;
}

}  // namespace LA_KERNELS

#ifndef LA_AVX2
namespace la_avx2 {
void gemm(const la::Mat& A, const la::Mat& B, la::Mat& C, int n);
bool syev(const la::Mat& A0, int n, std::vector<double>& w, la::Mat& V);
}  // namespace la_avx2

namespace la {

static bool host_has_avx2()
{
    static const bool v = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    return v;
}

void gemm(const Mat& A, const Mat& B, Mat& C, int n)
{
    if (host_has_avx2()) la_avx2::gemm(A, B, C, n);
    else la_base::gemm(A, B, C, n);
}

bool syev(const Mat& A0, int n, std::vector<double>& w, Mat& V)
{
    return host_has_avx2() ? la_avx2::syev(A0, n, w, V) : la_base::syev(A0, n, w, V);
}

void unpack(const double* pk, int n, Mat& M)
{
    M.assign((size_t)n * n, 0.0);
    for (int x = 0; x < n; ++x)
        for (int y = 0; y <= x; ++y) {
            const double v = pk[(size_t)x * (x + 1) / 2 + y];
            M[(size_t)y * n + x] = v;
            M[(size_t)x * n + y] = v;
        }
}

void pack(const Mat& M, int n, double* pk)
{
    for (int x = 0; x < n; ++x)
        for (int y = 0; y <= x; ++y) pk[(size_t)x * (x + 1) / 2 + y] = M[(size_t)y * n + x];
}

void transpose(const Mat& A, Mat& C, int n)
{
    C.resize((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) C[(size_t)j * n + i] = A[(size_t)i * n + j];
}

bool solve(std::vector<double> A, std::vector<double>& b, int n)
{
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (A[(size_t)piv * n + c] == 0.0) return false;
        if (piv != c) {
            for (int k = 0; k < n; ++k) std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < n; ++r) {
            const double f = A[(size_t)r * n + c] / A[(size_t)c * n + c];
            if (f == 0.0) continue;
            for (int k = c; k < n; ++k) A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < n; ++k) s -= A[(size_t)r * n + k] * b[k];
        b[r] = s / A[(size_t)r * n + r];
    }
    return true;
}

}  // namespace la
#endif
