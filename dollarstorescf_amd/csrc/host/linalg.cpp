// linalg.cpp -- see linalg.hpp.
#include "linalg.hpp"

#include <algorithm>
#include <cmath>

namespace la {

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

void transpose(const Mat& A, Mat& C, int n)
{
    C.resize((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) C[(size_t)j * n + i] = A[(size_t)i * n + j];
}

// Cyclic Jacobi: sweep over all (p,q) pairs, zeroing a_pq with the rotation
// t = sgn(theta)/(|theta| + sqrt(theta^2+1)), theta = (a_qq - a_pp)/(2 a_pq),
// c = 1/sqrt(1+t^2), s = t c, until the off-diagonal norm falls below
// 1e-30 of the total.  Rotations accumulate into V.
bool syev(const Mat& A0, int n, std::vector<double>& w, Mat& V)
{
    Mat A = A0;
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    double total = 0.0;
    for (double x : A) total += x * x;
    bool converged = n <= 1;
    for (int sweep = 0; sweep < 100 && !converged; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += 2.0 * a(p, q) * a(p, q);
        if (off <= 1e-30 * total) {
            converged = true;
            break;
        }
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a(p, q);
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * akq;
                    a(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * aqk;
                    a(q, k) = s * apk + c * aqk;
                }
                a(p, q) = a(q, p) = 0.0;
                for (int k = 0; k < n; ++k) {
                    double* vk = &V[(size_t)k * n];
                    const double vp = vk[p], vq = vk[q];
                    vk[p] = c * vp - s * vq;
                    vk[q] = s * vp + c * vq;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
    w.resize(n);
    Mat Vs((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        w[j] = a(order[j], order[j]);
        for (int i = 0; i < n; ++i) Vs[(size_t)i * n + j] = V[(size_t)i * n + order[j]];
    }
    V.swap(Vs);
    return converged;
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
