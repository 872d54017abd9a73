// scf_device.hip -- SCF iteration on the device (see host/scf_device.hpp).
//
// Layout: full matrices are n x n column-major (S, X, F and P are
// symmetric, so their storage is the same either way); packed-U vectors
// hold element (y,x), y <= x, at x(x+1)/2 + y (src/lalib/Matrix.ipp:332-351).
// Iterations run grouped (group_pre / group_post), every stage one launch
// for all jobs of the group (blockIdx.y = job):
//
//   k_g_assemble     G = J - K (packed), F0 = H + G (full), P (full)
//   k_g_gemm x2      S (P F0)                           (Pulay error, SCFDIIS.cpp:19-25)
//   k_g_error        e = M - M^T and F0 into a history slot
//   k_g_dots         <e_new, e_i> for the B matrix      (DIIS.ipp:18-62)
//     (host: slot choice and bordered solve)
//   k_g_combine      F = sum c_i F_i                    (DIIS.ipp:84-100)
//   k_g_gemm x2      F' = X F X                         (SCFSolver.cpp:64-75)
//   k_g_bounds, k_g_x0, k_g_gemm<true> x k
//                    P' = projector onto the nocc lowest orbitals of F'
//                    (purification; the role of Lapack.cpp:12-39 and
//                    SCFSolver.cpp:77-94 without eigenvectors)
//   k_g_gemm x2      P = X P' X
//   k_g_pack         Pa, Ptot = 2 Pa (packed), tr(Pa G)  (SCFSolver.cpp:118-143)
//
// The per-job functions below (one stream and one rocBLAS handle per job:
// dgemm, rocSOLVER dsyevd) remain for the final orbitals (C, eps of the last
// Fock matrix) and for an iteration whose purification does not finish.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host/scf_device.hpp"

namespace scfdev {

constexpr int kMaxHist = kMaxHistory;

struct Coefs {
    double c[kMaxHist];
};

struct Step {
    int device = 0, n = 0, nocc = 0, kmax = 0;
    size_t np = 0, nn = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    double *H = nullptr, *S = nullptr, *X = nullptr;   // packed H; full S, X
    double *Pa = nullptr, *Pt = nullptr;                // packed densities (Fock-build input)
    double *J = nullptr, *K = nullptr, *G = nullptr;    // packed
    double *F0 = nullptr, *P = nullptr, *T1 = nullptr, *T2 = nullptr, *F = nullptr, *C = nullptr;  // full
    double *Fh = nullptr, *Eh = nullptr;                // kmax full matrices each
    double *w = nullptr, *work = nullptr;               // eigenvalues, dsyevd's E
    double *scal = nullptr;                             // [0]: tr(Pa G), [1..kmax]: dots
    double *part = nullptr, *pstate = nullptr;          // purification: tile partial sums, state (two parities each)
    bool purify = true;                                 // grouped iterations: purification (else the eigensolver)
    rocblas_int* info = nullptr;
    double* h_scal = nullptr;                           // pinned host mirror of scal
    size_t blas_bytes = 0;                              // the handle's workspace (sized at creation)
};

// ---------------------------------------------------------------------------
// Grouped steps: one launch per stage for every job of a group, the density
// from trace-correcting purification (Niklasson, Phys. Rev. B 66, 155115
// (2002)) of F' = X F X instead of its eigenvectors.
//
// Purification maps F' to X_0 = (hi I - F') / (hi - lo) (Gershgorin bounds
// lo, hi, so the spectrum lies in [0, 1] with the lowest orbitals near 1)
// and repeats X <- X^2 or X <- 2X - X^2, whichever puts the trace nearer to
// nocc. Every step is a polynomial in F', so X keeps F''s eigenvectors; the
// eigenvalues go to 1 for the nocc lowest orbitals and to 0 for the rest, so
// the limit is C'_occ C'_occ^T, the density of the diagonalization route
// (SCFSolver.cpp:64-94) whenever orbital nocc lies below orbital nocc+1.
// tr(X) - tr(X^2) = sum_i x_i (1 - x_i) measures the distance from that
// limit; two steps after it falls below kPurifyTol every x_i is within
// rounding of 0 or 1 (the pair X^2, 2X - X^2 squares both distances), and
// that iterate is the density.  Sum of occupied eps = tr(P' F').
constexpr int kGemmTile = 32;                 // output tile edge (four waves splitting k, 4 x 4 per lane)
constexpr int kGemmK = 16;                    // k panel depth
constexpr int kPurifyChunk = 16;              // purification launches between completion checks
constexpr int kPurifyFirstCheck = 32;         // launches before the first check
constexpr int kPurifyMax = 192;               // a job still short of its limit then falls back
constexpr double kPurifyTol = 1e-9;
constexpr int kChunks = 16;                   // partial sums per reduction (summed on the host, in order)
// per job results: purification status, then the chunks of tr(P' F'), of tr(Pa G) and of each <e_new, e_i>
constexpr int kResStatus = 0, kResEps = 1, kResPG = kResEps + kChunks, kResDots = kResPG + kChunks;
constexpr int kResStride = kResDots + kMaxHist * kChunks;

enum GemmOp { OP_PF0, OP_ST1, OP_XF, OP_T1X, OP_XPP, OP_T2X };

struct GJob {
    int n, nocc, slot, count, purify;
    double *H, *S, *X, *Pa, *Pt, *J, *K, *G, *F0, *P, *T1, *T2, *F, *C, *Fh, *Eh;
    double *part, *st;
    double* coef;  // count coefficients (slot >= 0)
    double* res;   // kResStride values
};

struct Group {
    int device = 0, cap = 0;
    hipStream_t stream = nullptr;
    GJob *d_jobs = nullptr, *h_jobs = nullptr;   // job table (pinned host staging)
    double *d_res = nullptr, *h_res = nullptr;   // cap x kResStride
    double *d_coef = nullptr, *h_coef = nullptr; // cap x kMaxHist
};

namespace {

__device__ __forceinline__ size_t pk(int i, int j)
{
    const int x = i > j ? i : j, y = i > j ? j : i;
    return (size_t)x * (x + 1) / 2 + y;
}

__global__ void k_assemble(int n, const double* __restrict__ H, const double* __restrict__ J,
                           const double* __restrict__ K, const double* __restrict__ Pa, double* __restrict__ G,
                           double* __restrict__ F0, double* __restrict__ P)
{
    const size_t nn = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(t / n), i = (int)(t - (size_t)j * n);  // column-major (i, j)
        const size_t k = pk(i, j);
        const double g = J[k] - K[k];
        F0[t] = H[k] + g;
        P[t] = Pa[k];
        if (i <= j) G[k] = g;
    }
}

__global__ void k_antisym(int n, const double* __restrict__ M, double* __restrict__ E)
{
    const size_t nn = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(t / n), i = (int)(t - (size_t)j * n);
        E[t] = M[t] - M[(size_t)i * n + j];
    }
}

template <int NT>
__device__ double block_sum(double v)
{
    __shared__ double part[NT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; ++w) s += part[w];
    return s;
}

// block b: <e_new, e_b>; one block per history slot
template <int NT>
__global__ void __launch_bounds__(NT) k_dots(size_t nn, const double* __restrict__ enew,
                                             const double* __restrict__ hist, double* __restrict__ out)
{
    const double* e = hist + (size_t)blockIdx.x * nn;
    double s = 0.0;
    for (size_t t = threadIdx.x; t < nn; t += NT) s += enew[t] * e[t];
    s = block_sum<NT>(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ void k_combine(size_t nn, const double* __restrict__ hist, Coefs c, int count, double* __restrict__ F)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < count; ++i) s += c.c[i] * hist[(size_t)i * nn + t];
        F[t] = s;
    }
}

// Pa, Ptot from the full P; tr(Pa G) = sum_x<=y (2 - d_xy) Pa_xy G_xy
template <int NT>
__global__ void __launch_bounds__(NT) k_pack_density(int n, const double* __restrict__ P, const double* __restrict__ G,
                                                     double* __restrict__ Pa, double* __restrict__ Pt,
                                                     double* __restrict__ tr)
{
    const size_t np = (size_t)n * (n + 1) / 2;
    double s = 0.0;
    for (size_t k = threadIdx.x; k < np; k += NT) {
        // invert k = x(x+1)/2 + y
        int x = (int)((__builtin_sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
        while ((size_t)(x + 1) * (x + 2) / 2 <= k) ++x;
        while ((size_t)x * (x + 1) / 2 > k) --x;
        const int y = (int)(k - (size_t)x * (x + 1) / 2);
        const double p = P[(size_t)x * n + y];
        Pa[k] = p;
        Pt[k] = 2.0 * p;
        s += (x == y ? 1.0 : 2.0) * p * G[k];
    }
    s = block_sum<NT>(s);
    if (threadIdx.x == 0) tr[0] = s;
}

// ---- grouped kernels: blockIdx.y is the job, blockIdx.x strides its work ----

// a and b summed over the block, the sums in every thread
template <int NT>
__device__ void block_sum2(double& a, double& b)
{
    __shared__ double sa[NT / 64], sb[NT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    __syncthreads();  // earlier readers of sa, sb are done
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = a;
        sb[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    a = b = 0.0;
    for (int w = 0; w < NT / 64; ++w) {
        a += sa[w];
        b += sb[w];
    }
}

// lo = block minimum, hi = block maximum, in every thread
template <int NT>
__device__ void block_minmax(double& lo, double& hi)
{
    __shared__ double sl[NT / 64], sh[NT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off));
        hi = fmax(hi, __shfl_xor(hi, off));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sl[threadIdx.x >> 6] = lo;
        sh[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    for (int w = 0; w < NT / 64; ++w) {
        lo = fmin(lo, sl[w]);
        hi = fmax(hi, sh[w]);
    }
}

#define GJOB_FOR(t, count) \
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < (count); t += (size_t)gridDim.x * blockDim.x)

// G = J - K (packed), F0 = H + G, P = Pa (full)
__global__ void k_g_assemble(const GJob* __restrict__ jobs)
{
    const GJob& q = jobs[blockIdx.y];
    const int n = q.n;
    GJOB_FOR(t, (size_t)n * n)
    {
        const int j = (int)(t / n), i = (int)(t - (size_t)j * n);
        const size_t k = pk(i, j);
        const double g = q.J[k] - q.K[k];
        q.F0[t] = q.H[k] + g;
        q.P[t] = q.Pa[k];
        if (i <= j) q.G[k] = g;
    }
}

// Pulay error e = M - M^T (M = S P F0 in T2) and F0 into history slot
__global__ void k_g_error(const GJob* __restrict__ jobs)
{
    const GJob& q = jobs[blockIdx.y];
    if (q.slot < 0) return;
    const int n = q.n;
    const size_t nn = (size_t)n * n;
    double* e = q.Eh + (size_t)q.slot * nn;
    double* f = q.Fh + (size_t)q.slot * nn;
    GJOB_FOR(t, nn)
    {
        const int j = (int)(t / n), i = (int)(t - (size_t)j * n);
        e[t] = q.T2[t] - q.T2[(size_t)i * n + j];
        f[t] = q.F0[t];
    }
}

// elements [c L, (c+1) L) of a job's n*n, L = ceil(n*n / kChunks): chunk c of a reduction
__device__ __forceinline__ void chunk_range(size_t nn, int c, size_t* b, size_t* e)
{
    const size_t L = (nn + kChunks - 1) / kChunks;
    *b = std::min(nn, (size_t)c * L);
    *e = std::min(nn, *b + L);
}

// <e_new, e_i>: block x = i kChunks + c adds chunk c; the host sums the chunks in order
__global__ void __launch_bounds__(256) k_g_dots(const GJob* __restrict__ jobs)
{
    const GJob& q = jobs[blockIdx.y];
    const int i = blockIdx.x / kChunks, c = blockIdx.x % kChunks;
    if (q.slot < 0 || i >= q.count) return;
    const size_t nn = (size_t)q.n * q.n;
    const double* en = q.Eh + (size_t)q.slot * nn;
    const double* ei = q.Eh + (size_t)i * nn;
    size_t b, e;
    chunk_range(nn, c, &b, &e);
    double s = 0.0, u = 0.0;
    for (size_t t = b + threadIdx.x; t < e; t += 256) s += en[t] * ei[t];
    block_sum2<256>(s, u);
    if (threadIdx.x == 0) q.res[kResDots + i * kChunks + c] = s;
}

// F = sum_i c_i F_i (DIIS.ipp:84-100); the first iteration takes F0
__global__ void k_g_combine(const GJob* __restrict__ jobs)
{
    const GJob& q = jobs[blockIdx.y];
    const size_t nn = (size_t)q.n * q.n;
    if (q.slot < 0) {
        GJOB_FOR(t, nn) q.F[t] = q.F0[t];
        return;
    }
    GJOB_FOR(t, nn)
    {
        double s = 0.0;
        for (int i = 0; i < q.count; ++i) s += q.coef[i] * q.Fh[(size_t)i * nn + t];
        q.F[t] = s;
    }
}

// C = A B for n x n column-major operands, one 32 x 32 tile of C per block of
// four waves.  Each wave takes every fourth k panel (depth 16, staged in its
// own LDS region) for a 4 x 4 patch per lane; wave 0 then adds the other
// three waves' sums in a fixed order, so element (i, j) of a product of
// symmetric factors equals element (j, i) bit for bit.
// PURIFY: A = B = X_k (ping-pong buffers T1, C by k), C = X_k^2 or
// 2 X_k - X_k^2 by the trace rule, with the tile's trace and squared
// Frobenius norm as partial sums for the next launch.
template <bool PURIFY>
__global__ void __launch_bounds__(256) k_g_gemm(const GJob* __restrict__ jobs, int op, int k_iter)
{
    const GJob& q = jobs[blockIdx.y];
    const int n = q.n, nt = (n + kGemmTile - 1) / kGemmTile;
    const int tile = blockIdx.x;
    if (tile >= nt * nt) return;
    const double *A = nullptr, *B = nullptr;
    double* C = nullptr;
    bool twice = false;  // PURIFY: 2X - X^2
    if (PURIFY) {
        if (!q.purify) return;
        const double* sin = q.st + 4 * (k_iter & 1);
        double* sout = q.st + 4 * ((k_iter + 1) & 1);
        if (sin[0] != 0.0) {  // finished earlier: carry the state forward
            if (tile == 0 && threadIdx.x == 0)
                for (int c = 0; c < 4; ++c) sout[c] = sin[c];
            return;
        }
        // tr(X_k) and tr(X_k^2) = ||X_k||_F^2 from the previous launch's tiles
        const double* pin = q.part + (size_t)2 * nt * nt * (k_iter & 1);
        double tr = 0.0, sq = 0.0;
        for (int u = threadIdx.x; u < nt * nt; u += 256) {
            tr += pin[2 * u];
            sq += pin[2 * u + 1];
        }
        block_sum2<256>(tr, sq);
        const double delta = tr - sq;  // sum_i x_i (1 - x_i)
        // sin[2] = delta_{k-1}, sin[3] = delta_{k-2}
        const bool done = (k_iter >= 2 && sin[3] < kPurifyTol) || fabs(delta) < 1e-14;
        if (tile == 0 && threadIdx.x == 0) {
            sout[0] = done ? 1.0 : 0.0;
            sout[1] = (double)k_iter;
            sout[2] = delta;
            sout[3] = sin[2];
        }
        if (done) return;
        const double no = (double)q.nocc;
        twice = fabs(sq - no) > fabs(2.0 * tr - sq - no);
        A = B = (k_iter & 1) ? q.C : q.T1;
        C = (k_iter & 1) ? q.T1 : q.C;
    } else {
        switch (op) {
            case OP_PF0:
                if (q.slot < 0) return;
                A = q.P, B = q.F0, C = q.T1;
                break;
            case OP_ST1:
                if (q.slot < 0) return;
                A = q.S, B = q.T1, C = q.T2;
                break;
            case OP_XF: A = q.X, B = q.F, C = q.T1; break;
            case OP_T1X: A = q.T1, B = q.X, C = q.T2; break;
            case OP_XPP: {  // k_iter: the number of purification launches made
                const double* st = q.st + 4 * (k_iter & 1);
                A = q.X, B = ((int)st[1] & 1) ? q.C : q.T1, C = q.T2;
                break;
            }
            default: A = q.T2, B = q.X, C = q.P; break;  // OP_T2X
        }
    }
    constexpr int W = kGemmTile + 2;                       // padded row of a panel
    __shared__ double lds[4 * 2 * kGemmK * W];             // per wave: A panel As[k][i], B panel Bs[k][j]
    const int i0 = (tile % nt) * kGemmTile, j0 = (tile / nt) * kGemmTile;
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, tx = l & 7, ty = l >> 3;
    double* As = lds + wv * 2 * kGemmK * W;
    double* Bs = As + kGemmK * W;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    const int npanel = (n + kGemmK - 1) / kGemmK;
    for (int r = 0; r < (npanel + 3) / 4; ++r) {
        const int p = 4 * r + wv, k0 = p * kGemmK;
        if (p < npanel) {
#pragma unroll
            for (int s = 0; s < kGemmTile * kGemmK / 64; ++s) {
                const int e = l + 64 * s;
                const int ra = e % kGemmTile, ca = e / kGemmTile;  // A: tile row ra, panel column ca
                const int gi = i0 + ra, gk = k0 + ca;
                As[ca * W + ra] = (gi < n && gk < n) ? A[(size_t)gk * n + gi] : 0.0;
                const int cb = e % kGemmK, jb = e / kGemmK;  // B: panel row cb, tile column jb
                const int gkb = k0 + cb, gj = j0 + jb;
                Bs[cb * W + jb] = (gkb < n && gj < n) ? B[(size_t)gj * n + gkb] : 0.0;
            }
        }
        __syncthreads();
        if (p < npanel) {
#pragma unroll
            for (int kk = 0; kk < kGemmK; ++kk) {
                double a[4], b[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    a[u] = As[kk * W + tx * 4 + u];
                    b[u] = Bs[kk * W + ty * 4 + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
            }
        }
        __syncthreads();
    }
    // waves 1-3 leave their sums in LDS (3 x 1024 of its 4 x 2 x 16 x 34 doubles)
    if (wv > 0) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) lds[(wv - 1) * 1024 + (u * 4 + v) * 64 + l] = acc[u][v];
    }
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int idx = (u * 4 + v) * 64 + l;
            acc[u][v] = ((acc[u][v] + lds[idx]) + lds[1024 + idx]) + lds[2048 + idx];
        }
    double tr = 0.0, sq = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int gj = j0 + ty * 4 + v;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int gi = i0 + tx * 4 + u;
            if (gi < n && gj < n) {
                double o = acc[u][v];
                if (PURIFY) {
                    if (twice) o = 2.0 * A[(size_t)gj * n + gi] - o;
                    if (gi == gj) tr += o;
                    sq += o * o;
                }
                C[(size_t)gj * n + gi] = o;
            }
        }
    }
    if (PURIFY) {
        for (int off = 32; off > 0; off >>= 1) {
            tr += __shfl_xor(tr, off);
            sq += __shfl_xor(sq, off);
        }
        if (l == 0) {
            double* pout = q.part + (size_t)2 * nt * nt * ((k_iter + 1) & 1);
            pout[2 * tile] = tr;
            pout[2 * tile + 1] = sq;
        }
    }
}

// F' (in T2) is read from its upper triangle only, B_ij = F'(min(i,j), max(i,j)),
// so that X_0 and every later iterate are exactly symmetric.
__device__ __forceinline__ double upper(const double* Fp, int n, int i, int j)
{
    return i <= j ? Fp[(size_t)j * n + i] : Fp[(size_t)i * n + j];
}

// Gershgorin bounds of B, rows [32 b, 32 b + 32) per block x = b: 8 lanes
// per row; the block's lo, hi into part (the second parity's space, free
// until the first purification step)
__global__ void __launch_bounds__(256) k_g_bounds(const GJob* __restrict__ jobs)
{
    const GJob& q = jobs[blockIdx.y];
    const int n = q.n, nt = (n + kGemmTile - 1) / kGemmTile;
    if ((int)blockIdx.x >= nt) return;
    const int i = blockIdx.x * kGemmTile + (threadIdx.x >> 3), l8 = threadIdx.x & 7;
    double lo = 1e300, hi = -1e300;
    double r = 0.0;
    if (i < n)
        for (int j = l8; j < n; j += 8)
            if (j != i) r += fabs(upper(q.T2, n, i, j));
    r += __shfl_xor(r, 1);
    r += __shfl_xor(r, 2);
    r += __shfl_xor(r, 4);
    if (i < n) {
        const double d = q.T2[(size_t)i * n + i];
        lo = d - r;
        hi = d + r;
    }
    block_minmax<256>(lo, hi);
    if (threadIdx.x == 0) {
        double* bnd = q.part + (size_t)2 * nt * nt;
        bnd[2 * blockIdx.x] = lo;
        bnd[2 * blockIdx.x + 1] = hi;
    }
}

// X_0 = (hi I - B) / (hi - lo), one 32 x 32 tile per block x, into T1, with
// the tile's trace and squared norm as the first partial sums; tile 0 sets a
// fresh purification state (finished at once for nocc = 0 or n: P' = 0, I)
__global__ void __launch_bounds__(256) k_g_x0(const GJob* __restrict__ jobs)
{
    const GJob& q = jobs[blockIdx.y];
    const int n = q.n, nt = (n + kGemmTile - 1) / kGemmTile;
    const int tile = blockIdx.x;
    if (tile >= nt * nt) return;
    const double* bnd = q.part + (size_t)2 * nt * nt;
    double lo = 1e300, hi = -1e300;
    for (int b = 0; b < nt; ++b) {
        lo = fmin(lo, bnd[2 * b]);
        hi = fmax(hi, bnd[2 * b + 1]);
    }
    const double w = hi > lo ? hi - lo : 1.0;
    const int mode = q.nocc <= 0 ? 1 : (q.nocc >= n ? 2 : 0);
    const int i0 = (tile % nt) * kGemmTile, j0 = (tile / nt) * kGemmTile;
    double tr = 0.0, sq = 0.0;
    for (int e = threadIdx.x; e < kGemmTile * kGemmTile; e += 256) {
        const int i = i0 + e % kGemmTile, j = j0 + e / kGemmTile;
        if (i >= n || j >= n) continue;
        double x;
        if (mode == 0)
            x = ((i == j ? hi : 0.0) - upper(q.T2, n, i, j)) / w;
        else
            x = (mode == 2 && i == j) ? 1.0 : 0.0;
        q.T1[(size_t)j * n + i] = x;
        if (i == j) tr += x;
        sq += x * x;
    }
    block_sum2<256>(tr, sq);
    if (threadIdx.x == 0) {
        q.part[2 * tile] = tr;
        q.part[2 * tile + 1] = sq;
        if (tile == 0) {
            q.st[0] = mode ? 1.0 : 0.0;
            q.st[1] = 0.0;
            q.st[2] = q.st[3] = 1e30;
        }
    }
}

// the purification status after `launches` steps: the final iterate's
// index, or -1 while unfinished
__global__ void k_g_purify_status(const GJob* __restrict__ jobs, int m, int launches)
{
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const double* st = jobs[i].st + 4 * (launches & 1);
        jobs[i].res[kResStatus] = jobs[i].purify && st[0] != 0.0 ? st[1] : -1.0;
    }
}

// chunk x of tr(P' F') (the sum of occupied orbital energies); block 0 also
// the purification status
__global__ void __launch_bounds__(256) k_g_purify_end(const GJob* __restrict__ jobs, int launches)
{
    const GJob& q = jobs[blockIdx.y];
    const double* st = q.st + 4 * (launches & 1);
    const bool done = q.purify && st[0] != 0.0;
    const int kf = (int)st[1];
    const double* Pp = (kf & 1) ? q.C : q.T1;
    size_t b, e;
    chunk_range((size_t)q.n * q.n, blockIdx.x, &b, &e);
    double s = 0.0, u = 0.0;
    if (done)
        for (size_t t = b + threadIdx.x; t < e; t += 256) s += Pp[t] * q.T2[t];
    block_sum2<256>(s, u);
    if (threadIdx.x == 0) {
        q.res[kResEps + blockIdx.x] = s;
        if (blockIdx.x == 0) q.res[kResStatus] = done ? (double)kf : -1.0;
    }
}

// Pa, Ptot = 2 Pa (packed) from the full P for columns x = c, c + kChunks, ...
// (block x = c), with chunk c of tr(Pa G) = sum_x<=y (2 - d_xy) Pa_xy G_xy
__global__ void __launch_bounds__(256) k_g_pack(const GJob* __restrict__ jobs)
{
    const GJob& q = jobs[blockIdx.y];
    const int n = q.n;
    double s = 0.0, u = 0.0;
    for (int x = blockIdx.x; x < n; x += kChunks) {
        const size_t base = (size_t)x * (x + 1) / 2;
        for (int y = threadIdx.x; y <= x; y += 256) {
            const double p = q.P[(size_t)x * n + y];
            q.Pa[base + y] = p;
            q.Pt[base + y] = 2.0 * p;
            s += (x == y ? 1.0 : 2.0) * p * q.G[base + y];
        }
    }
    block_sum2<256>(s, u);
    if (threadIdx.x == 0) q.res[kResPG + blockIdx.x] = s;
}

int grid_for(size_t n) { return (int)std::min<size_t>(4096, (n + 255) / 256); }

#define STEP_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            if (err) *err = std::string("device SCF step: ") + #call + ": " + hipGetErrorString(e_); \
            return DSSCF_ERR_HIP;                                                            \
        }                                                                                    \
    } while (0)
#define STEP_BLAS(call)                                                                      \
    do {                                                                                     \
        rocblas_status s_ = (call);                                                          \
        if (s_ != rocblas_status_success) {                                                  \
            if (err) *err = std::string("device SCF step: ") + #call + ": " + rocblas_status_to_string(s_); \
            return DSSCF_ERR_HIP;                                                            \
        }                                                                                    \
    } while (0)

int alloc(double** p, size_t n, std::string* err)
{
    if (hipMalloc((void**)p, sizeof(double) * std::max<size_t>(n, 1)) != hipSuccess) {
        if (err) *err = "device SCF step: out of device memory";
        return DSSCF_ERR_NOMEM;
    }
    return DSSCF_OK;
}

// column-major C = op(A) op(B), all n x n except the inner dimension k
int gemm(Step* s, rocblas_operation ta, rocblas_operation tb, int k, const double* A, const double* B, double* C,
         std::string* err)
{
    const double one = 1.0, zero = 0.0;
    STEP_BLAS(rocblas_dgemm(s->blas, ta, tb, s->n, s->n, k, &one, A, s->n, B, s->n, &zero, C, s->n));
    return DSSCF_OK;
}

// F (full, in s->F) -> eigenvalues in w, C = X C' (F' = X F X = C' eps C'^T)
int eigen(Step* s, std::string* err)
{
    const int n = s->n;
    const rocblas_operation N = rocblas_operation_none;
    if (int rc = gemm(s, N, N, n, s->X, s->F, s->T1, err)) return rc;
    if (int rc = gemm(s, N, N, n, s->T1, s->X, s->T2, err)) return rc;  // F' in T2
    STEP_BLAS(rocsolver_dsyevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n, s->T2, n, s->w, s->work,
                               s->info));
    return gemm(s, N, N, n, s->X, s->T2, s->C, err);
}

// F (full, in s->F) -> eigenpairs, C, new densities, tr(Pa_new G)
int diagonalize(Step* s, std::vector<double>& eps, double* tr_PG, std::string* err)
{
    const int n = s->n;
    const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
    if (int rc = eigen(s, err)) return rc;
    if (s->nocc > 0) {
        if (int rc = gemm(s, N, T, s->nocc, s->C, s->C, s->P, err)) return rc;
    } else {
        STEP_HIP(hipMemsetAsync(s->P, 0, sizeof(double) * s->nn, s->stream));
    }
    k_pack_density<256><<<1, 256, 0, s->stream>>>(n, s->P, s->G, s->Pa, s->Pt, s->scal);
    STEP_HIP(hipGetLastError());
    eps.resize(n);
    STEP_HIP(hipMemcpyAsync(eps.data(), s->w, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
    STEP_HIP(hipMemcpyAsync(s->h_scal, s->scal, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    rocblas_int info = 0;
    STEP_HIP(hipMemcpyAsync(&info, s->info, sizeof(info), hipMemcpyDeviceToHost, s->stream));
    STEP_HIP(hipStreamSynchronize(s->stream));
    if (info != 0) {
        if (err) *err = "device SCF step: dsyevd did not converge (info " + std::to_string(info) + ")";
        return DSSCF_ERR_NOCONV;
    }
    *tr_PG = s->h_scal[0];
    return DSSCF_OK;
}

}  // namespace

// purification partial sums: two parities of (trace, squared norm) per 32 x 32 tile
static size_t part_doubles(int n)
{
    const size_t nt = ((size_t)n + kGemmTile - 1) / kGemmTile;
    return 4 * nt * nt;
}

int create(Step** out, int device, int n, int nocc, int kmax, const double* Hp, const double* S, const double* X,
           std::string* err)
{
    *out = nullptr;
    if (n <= 0 || nocc < 0 || nocc > n || kmax < 1 || kmax > kMaxHist) {
        if (err) *err = "device SCF step: bad sizes (DIIS history 1.." + std::to_string(kMaxHist) + ")";
        return DSSCF_ERR_ARG;
    }
    Step* s = new Step();
    s->device = device;
    s->n = n;
    s->nocc = nocc;
    s->kmax = kmax;
    s->np = (size_t)n * (n + 1) / 2;
    s->nn = (size_t)n * n;
    int rc = DSSCF_OK;
    auto fail = [&](int code) {
        destroy(s);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        if (err) *err = "device SCF step: cannot create a stream on device " + std::to_string(device);
        return fail(DSSCF_ERR_HIP);
    }
    if (rocblas_create_handle(&s->blas) != rocblas_status_success ||
        rocblas_set_stream(s->blas, s->stream) != rocblas_status_success) {
        if (err) *err = "device SCF step: rocblas_create_handle failed";
        return fail(DSSCF_ERR_HIP);
    }
    // size the handle's workspace for this n now (dsyevd is the only call
    // that takes one), so that the step's device footprint is fixed and counted
    {
        size_t ws = 0;
        bool ok = rocblas_start_device_memory_size_query(s->blas) == rocblas_status_success;
        // in a size query the call only reports what it would need (its status says so)
        if (ok) (void)rocsolver_dsyevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n, nullptr, n, nullptr,
                                       nullptr, nullptr);
        ok = ok && rocblas_stop_device_memory_size_query(s->blas, &ws) == rocblas_status_success &&
             rocblas_set_device_memory_size(s->blas, ws) == rocblas_status_success;
        if (!ok) {
            if (err) *err = "device SCF step: cannot size the rocBLAS workspace";
            return fail(DSSCF_ERR_NOMEM);
        }
        s->blas_bytes = ws;
    }
    for (auto pr : {std::make_pair(&s->H, s->np), std::make_pair(&s->S, s->nn), std::make_pair(&s->X, s->nn),
                    std::make_pair(&s->Pa, s->np), std::make_pair(&s->Pt, s->np), std::make_pair(&s->J, s->np),
                    std::make_pair(&s->K, s->np), std::make_pair(&s->G, s->np), std::make_pair(&s->F0, s->nn),
                    std::make_pair(&s->P, s->nn), std::make_pair(&s->T1, s->nn), std::make_pair(&s->T2, s->nn),
                    std::make_pair(&s->F, s->nn), std::make_pair(&s->C, s->nn),
                    std::make_pair(&s->Fh, s->nn * kmax), std::make_pair(&s->Eh, s->nn * kmax),
                    std::make_pair(&s->w, (size_t)n), std::make_pair(&s->work, (size_t)n),
                    std::make_pair(&s->scal, (size_t)kmax + 1), std::make_pair(&s->part, part_doubles(n)),
                    std::make_pair(&s->pstate, (size_t)8)})
        if ((rc = alloc(pr.first, pr.second, err))) return fail(rc);
    if (hipMalloc((void**)&s->info, sizeof(rocblas_int)) != hipSuccess ||
        hipHostMalloc((void**)&s->h_scal, sizeof(double) * (kmax + 1)) != hipSuccess) {
        if (err) *err = "device SCF step: out of memory";
        return fail(DSSCF_ERR_NOMEM);
    }
    if (hipMemcpyAsync(s->H, Hp, sizeof(double) * s->np, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->S, S, sizeof(double) * s->nn, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->X, X, sizeof(double) * s->nn, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemsetAsync(s->Pa, 0, sizeof(double) * s->np, s->stream) != hipSuccess ||
        hipMemsetAsync(s->Pt, 0, sizeof(double) * s->np, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
        if (err) *err = "device SCF step: upload failed";
        return fail(DSSCF_ERR_HIP);
    }
    *out = s;
    return DSSCF_OK;
}

void destroy(Step* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (double* p : {s->H, s->S, s->X, s->Pa, s->Pt, s->J, s->K, s->G, s->F0, s->P, s->T1, s->T2, s->F, s->C, s->Fh,
                      s->Eh, s->w, s->work, s->scal, s->part, s->pstate})
        if (p) (void)hipFree(p);
    if (s->info) (void)hipFree(s->info);
    if (s->h_scal) (void)hipHostFree(s->h_scal);
    if (s->blas) rocblas_destroy_handle(s->blas);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int set_density(Step* s, const double* Pa, std::string* err)
{
    STEP_HIP(hipSetDevice(s->device));
    std::vector<double> Pt(s->np);
    for (size_t i = 0; i < s->np; ++i) Pt[i] = 2.0 * Pa[i];
    STEP_HIP(hipMemcpyAsync(s->Pa, Pa, sizeof(double) * s->np, hipMemcpyHostToDevice, s->stream));
    STEP_HIP(hipMemcpyAsync(s->Pt, Pt.data(), sizeof(double) * s->np, hipMemcpyHostToDevice, s->stream));
    STEP_HIP(hipStreamSynchronize(s->stream));
    return DSSCF_OK;
}

int fock_build(Step* s, dsscf_ctx* ctx, std::string* err)
{
    STEP_HIP(hipSetDevice(s->device));
    STEP_HIP(hipStreamSynchronize(s->stream));  // densities written on this stream
    if (int rc = dsscf_jk_device(ctx, s->Pt, s->Pa, s->J, s->K)) {  // synchronizes the context's stream
        if (err) *err = dsscf_last_error();
        return rc;
    }
    return DSSCF_OK;
}

int assemble(Step* s, int slot, int count, double* dots, double* rms, std::string* err)
{
    STEP_HIP(hipSetDevice(s->device));
    const int n = s->n;
    k_assemble<<<grid_for(s->nn), 256, 0, s->stream>>>(n, s->H, s->J, s->K, s->Pa, s->G, s->F0, s->P);
    STEP_HIP(hipGetLastError());
    if (slot < 0) return DSSCF_OK;
    if (slot >= s->kmax || count < 1 || count > s->kmax) {
        if (err) *err = "device SCF step: DIIS slot out of range";
        return DSSCF_ERR_ARG;
    }
    const rocblas_operation N = rocblas_operation_none;
    if (int rc = gemm(s, N, N, n, s->P, s->F0, s->T1, err)) return rc;   // P F0
    if (int rc = gemm(s, N, N, n, s->S, s->T1, s->T2, err)) return rc;   // S P F0
    double* e = s->Eh + (size_t)slot * s->nn;
    k_antisym<<<grid_for(s->nn), 256, 0, s->stream>>>(n, s->T2, e);
    STEP_HIP(hipGetLastError());
    STEP_HIP(hipMemcpyAsync(s->Fh + (size_t)slot * s->nn, s->F0, sizeof(double) * s->nn, hipMemcpyDeviceToDevice,
                            s->stream));
    k_dots<1024><<<count, 1024, 0, s->stream>>>(s->nn, e, s->Eh, s->scal + 1);
    STEP_HIP(hipGetLastError());
    STEP_HIP(hipMemcpyAsync(s->h_scal + 1, s->scal + 1, sizeof(double) * count, hipMemcpyDeviceToHost, s->stream));
    STEP_HIP(hipStreamSynchronize(s->stream));
    for (int i = 0; i < count; ++i) dots[i] = s->h_scal[1 + i];
    *rms = dots[slot];
    return DSSCF_OK;
}

int extrapolate(Step* s, const double* coef, int count, std::vector<double>& eps, double* tr_PG, std::string* err)
{
    STEP_HIP(hipSetDevice(s->device));
    if (count < 1 || count > s->kmax) {
        if (err) *err = "device SCF step: DIIS count out of range";
        return DSSCF_ERR_ARG;
    }
    Coefs c{};
    for (int i = 0; i < count; ++i) c.c[i] = coef[i];
    k_combine<<<grid_for(s->nn), 256, 0, s->stream>>>(s->nn, s->Fh, c, count, s->F);
    STEP_HIP(hipGetLastError());
    return diagonalize(s, eps, tr_PG, err);
}

int diagonalize_direct(Step* s, std::vector<double>& eps, double* tr_PG, std::string* err)
{
    STEP_HIP(hipSetDevice(s->device));
    STEP_HIP(hipMemcpyAsync(s->F, s->F0, sizeof(double) * s->nn, hipMemcpyDeviceToDevice, s->stream));
    return diagonalize(s, eps, tr_PG, err);
}

int download(Step* s, double* Pa, double* Pt, double* J, double* K, double* G, double* C, std::string* err)
{
    STEP_HIP(hipSetDevice(s->device));
    const size_t b = sizeof(double) * s->np;
    if (Pa) STEP_HIP(hipMemcpyAsync(Pa, s->Pa, b, hipMemcpyDeviceToHost, s->stream));
    if (Pt) STEP_HIP(hipMemcpyAsync(Pt, s->Pt, b, hipMemcpyDeviceToHost, s->stream));
    if (J) STEP_HIP(hipMemcpyAsync(J, s->J, b, hipMemcpyDeviceToHost, s->stream));
    if (K) STEP_HIP(hipMemcpyAsync(K, s->K, b, hipMemcpyDeviceToHost, s->stream));
    if (G) STEP_HIP(hipMemcpyAsync(G, s->G, b, hipMemcpyDeviceToHost, s->stream));
    std::vector<double> Cc;
    if (C) {
        Cc.resize(s->nn);
        STEP_HIP(hipMemcpyAsync(Cc.data(), s->C, sizeof(double) * s->nn, hipMemcpyDeviceToHost, s->stream));
    }
    STEP_HIP(hipStreamSynchronize(s->stream));
    if (C) {
        const int n = s->n;  // column-major -> row-major
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] = Cc[(size_t)j * n + i];
    }
    return DSSCF_OK;
}

// The step's own matrices.
static long long matrix_bytes(int n, int kmax)
{
    const size_t np = (size_t)n * (n + 1) / 2, nn = (size_t)n * n;
    return (long long)(sizeof(double) *
                       (6 * np + (8 + 2 * (size_t)kmax) * nn + 3 * (size_t)n + kmax + 9 + part_doubles(n)));
}

long long device_bytes_for(int n, int kmax)
{
    // plus an upper estimate of dsyevd's workspace in the rocBLAS handle
    // (divide-and-conquer: about 2n^2 + 6n doubles; the handle's own 1 MiB)
    const size_t nn = (size_t)n * n;
    return matrix_bytes(n, kmax) + (long long)(sizeof(double) * (3 * nn + 16 * (size_t)n)) + (1 << 20);
}

long long device_bytes(const Step* s)
{
    if (!s) return 0;
    return matrix_bytes(s->n, s->kmax) + (long long)s->blas_bytes;
}

void buffers(Step* s, const double** Pt, const double** Pa, double** J, double** K)
{
    *Pt = s->Pt;
    *Pa = s->Pa;
    *J = s->J;
    *K = s->K;
}

void set_purify(Step* s, bool on) { s->purify = on; }

int orbitals(Step* s, std::vector<double>& eps, std::string* err)
{
    STEP_HIP(hipSetDevice(s->device));
    if (int rc = eigen(s, err)) return rc;
    eps.resize(s->n);
    rocblas_int info = 0;
    STEP_HIP(hipMemcpyAsync(eps.data(), s->w, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->stream));
    STEP_HIP(hipMemcpyAsync(&info, s->info, sizeof(info), hipMemcpyDeviceToHost, s->stream));
    STEP_HIP(hipStreamSynchronize(s->stream));
    if (info != 0) {
        if (err) *err = "device SCF step: dsyevd did not converge (info " + std::to_string(info) + ")";
        return DSSCF_ERR_NOCONV;
    }
    return DSSCF_OK;
}

// ------------------------------------------------------------ grouped steps --
int group_create(Group** out, int device, std::string* err)
{
    *out = nullptr;
    Group* g = new Group();
    g->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        if (err) *err = "device SCF step: cannot create a stream on device " + std::to_string(device);
        return DSSCF_ERR_HIP;
    }
    *out = g;
    return DSSCF_OK;
}

static void group_free_tables(Group* g)
{
    if (g->d_jobs) (void)hipFree(g->d_jobs);
    if (g->d_res) (void)hipFree(g->d_res);
    if (g->d_coef) (void)hipFree(g->d_coef);
    if (g->h_jobs) (void)hipHostFree(g->h_jobs);
    if (g->h_res) (void)hipHostFree(g->h_res);
    if (g->h_coef) (void)hipHostFree(g->h_coef);
    g->d_jobs = g->h_jobs = nullptr;
    g->d_res = g->h_res = g->d_coef = g->h_coef = nullptr;
    g->cap = 0;
}

void group_destroy(Group* g)
{
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    group_free_tables(g);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

// tables for m jobs (the stream is idle when this is called)
static int group_reserve(Group* g, int m, std::string* err)
{
    if (m <= g->cap) return DSSCF_OK;
    group_free_tables(g);
    const int cap = std::max(m, 2 * g->cap);
    if (hipMalloc((void**)&g->d_jobs, sizeof(GJob) * cap) != hipSuccess ||
        hipMalloc((void**)&g->d_res, sizeof(double) * kResStride * cap) != hipSuccess ||
        hipMalloc((void**)&g->d_coef, sizeof(double) * kMaxHist * cap) != hipSuccess ||
        hipHostMalloc((void**)&g->h_jobs, sizeof(GJob) * cap) != hipSuccess ||
        hipHostMalloc((void**)&g->h_res, sizeof(double) * kResStride * cap) != hipSuccess ||
        hipHostMalloc((void**)&g->h_coef, sizeof(double) * kMaxHist * cap) != hipSuccess) {
        group_free_tables(g);
        if (err) *err = "device SCF step: out of memory for the group tables";
        return DSSCF_ERR_NOMEM;
    }
    g->cap = cap;
    return DSSCF_OK;
}

static int group_gemm(Group* g, int m, int ntmax, int op, std::string* err)
{
    k_g_gemm<false><<<dim3(ntmax * ntmax, m), 256, 0, g->stream>>>(g->d_jobs, op, 0);
    STEP_HIP(hipGetLastError());
    return DSSCF_OK;
}

static int group_dims(Step* const* steps, int m, int* nmax)
{
    int n = 1;
    for (int i = 0; i < m; ++i) n = std::max(n, steps[i]->n);
    *nmax = n;
    return (n + kGemmTile - 1) / kGemmTile;
}

int group_pre(Group* g, Step* const* steps, int m, const int* slot, const int* count, double* dots, std::string* err)
{
    if (m <= 0) return DSSCF_OK;
    STEP_HIP(hipSetDevice(g->device));
    for (int i = 0; i < m; ++i) {
        const Step* s = steps[i];
        if (s->device != g->device || (slot[i] >= 0 && (slot[i] >= s->kmax || count[i] < 1 || count[i] > s->kmax))) {
            if (err) *err = "device SCF step: DIIS slot out of range or step on another device";
            return DSSCF_ERR_ARG;
        }
    }
    if (int rc = group_reserve(g, m, err)) return rc;
    bool any = false;
    for (int i = 0; i < m; ++i) {
        Step* s = steps[i];
        GJob& o = g->h_jobs[i];
        o.n = s->n;
        o.nocc = s->nocc;
        o.slot = slot[i];
        o.count = slot[i] >= 0 ? count[i] : 0;
        o.purify = s->purify ? 1 : 0;
        o.H = s->H, o.S = s->S, o.X = s->X, o.Pa = s->Pa, o.Pt = s->Pt, o.J = s->J, o.K = s->K, o.G = s->G;
        o.F0 = s->F0, o.P = s->P, o.T1 = s->T1, o.T2 = s->T2, o.F = s->F, o.C = s->C, o.Fh = s->Fh, o.Eh = s->Eh;
        o.part = s->part;
        o.st = s->pstate;
        o.coef = g->d_coef + (size_t)kMaxHist * i;
        o.res = g->d_res + (size_t)kResStride * i;
        any = any || slot[i] >= 0;
    }
    STEP_HIP(hipMemcpyAsync(g->d_jobs, g->h_jobs, sizeof(GJob) * m, hipMemcpyHostToDevice, g->stream));
    int nmax = 1;
    const int ntmax = group_dims(steps, m, &nmax);
    const int chunks = grid_for((size_t)nmax * nmax);
    k_g_assemble<<<dim3(chunks, m), 256, 0, g->stream>>>(g->d_jobs);
    STEP_HIP(hipGetLastError());
    if (!any) return DSSCF_OK;
    if (int rc = group_gemm(g, m, ntmax, OP_PF0, err)) return rc;  // P F0
    if (int rc = group_gemm(g, m, ntmax, OP_ST1, err)) return rc;  // S P F0
    k_g_error<<<dim3(chunks, m), 256, 0, g->stream>>>(g->d_jobs);
    STEP_HIP(hipGetLastError());
    k_g_dots<<<dim3(kMaxHist * kChunks, m), 256, 0, g->stream>>>(g->d_jobs);
    STEP_HIP(hipGetLastError());
    STEP_HIP(hipMemcpyAsync(g->h_res, g->d_res, sizeof(double) * kResStride * m, hipMemcpyDeviceToHost, g->stream));
    STEP_HIP(hipStreamSynchronize(g->stream));
    for (int i = 0; i < m; ++i)
        if (slot[i] >= 0)
            for (int c = 0; c < count[i]; ++c) {
                const double* r = g->h_res + (size_t)kResStride * i + kResDots + (size_t)c * kChunks;
                double d = 0.0;
                for (int u = 0; u < kChunks; ++u) d += r[u];
                dots[(size_t)kMaxHistory * i + c] = d;
            }
    return DSSCF_OK;
}

int group_post(Group* g, Step* const* steps, int m, const double* coef, double* sum_eps, double* tr_PG, int* rcs,
               int* psteps, std::string* err)
{
    if (m <= 0) return DSSCF_OK;
    STEP_HIP(hipSetDevice(g->device));
    for (int i = 0; i < m; ++i) {
        const GJob& o = g->h_jobs[i];
        for (int c = 0; c < kMaxHist; ++c)
            g->h_coef[(size_t)kMaxHist * i + c] = o.slot >= 0 && c < o.count ? coef[(size_t)kMaxHistory * i + c] : 0.0;
    }
    STEP_HIP(hipMemcpyAsync(g->d_coef, g->h_coef, sizeof(double) * kMaxHist * m, hipMemcpyHostToDevice, g->stream));
    int nmax = 1;
    const int ntmax = group_dims(steps, m, &nmax);
    const int chunks = grid_for((size_t)nmax * nmax);
    k_g_combine<<<dim3(chunks, m), 256, 0, g->stream>>>(g->d_jobs);
    STEP_HIP(hipGetLastError());
    if (int rc = group_gemm(g, m, ntmax, OP_XF, err)) return rc;
    if (int rc = group_gemm(g, m, ntmax, OP_T1X, err)) return rc;  // F' = X F X in T2
    k_g_bounds<<<dim3(ntmax, m), 256, 0, g->stream>>>(g->d_jobs);
    STEP_HIP(hipGetLastError());
    k_g_x0<<<dim3(ntmax * ntmax, m), 256, 0, g->stream>>>(g->d_jobs);
    STEP_HIP(hipGetLastError());
    // purification steps in chunks; after the first kPurifyFirstCheck, every
    // chunk ends with a look at the jobs' states (finished jobs' blocks exit at
    // once, so a chunk costs only what the unfinished jobs need)
    int launches = 0;
    for (;;) {
        for (int c = 0; c < kPurifyChunk; ++c, ++launches) {
            k_g_gemm<true><<<dim3(ntmax * ntmax, m), 256, 0, g->stream>>>(g->d_jobs, 0, launches);
            STEP_HIP(hipGetLastError());
        }
        if (launches < kPurifyFirstCheck) continue;
        if (launches >= kPurifyMax) break;
        k_g_purify_status<<<1, 64, 0, g->stream>>>(g->d_jobs, m, launches);
        STEP_HIP(hipGetLastError());
        STEP_HIP(hipMemcpyAsync(g->h_res, g->d_res, sizeof(double) * kResStride * m, hipMemcpyDeviceToHost,
                                g->stream));
        STEP_HIP(hipStreamSynchronize(g->stream));
        bool all = true;
        for (int i = 0; i < m && all; ++i)
            all = g->h_res[(size_t)kResStride * i + kResStatus] >= 0.0 || !g->h_jobs[i].purify;
        if (all) break;
    }
    k_g_purify_end<<<dim3(kChunks, m), 256, 0, g->stream>>>(g->d_jobs, launches);
    STEP_HIP(hipGetLastError());
    k_g_gemm<false><<<dim3(ntmax * ntmax, m), 256, 0, g->stream>>>(g->d_jobs, OP_XPP, launches);
    STEP_HIP(hipGetLastError());
    if (int rc = group_gemm(g, m, ntmax, OP_T2X, err)) return rc;  // P = X P' X
    k_g_pack<<<dim3(kChunks, m), 256, 0, g->stream>>>(g->d_jobs);
    STEP_HIP(hipGetLastError());
    STEP_HIP(hipMemcpyAsync(g->h_res, g->d_res, sizeof(double) * kResStride * m, hipMemcpyDeviceToHost, g->stream));
    STEP_HIP(hipStreamSynchronize(g->stream));
    for (int i = 0; i < m; ++i) {
        const double* r = g->h_res + (size_t)kResStride * i;
        rcs[i] = DSSCF_OK;
        if (psteps) psteps[i] = (int)r[kResStatus];
        if (r[kResStatus] >= 0.0) {
            double se = 0.0, pg = 0.0;
            for (int u = 0; u < kChunks; ++u) {
                se += r[kResEps + u];
                pg += r[kResPG + u];
            }
            sum_eps[i] = se;
            tr_PG[i] = pg;
            continue;
        }
        // purification did not reach its limit (orbitals nocc and nocc+1 too
        // close) or is off for this job: the eigensolver route on its own stream
        Step* s = steps[i];
        const GJob& o = g->h_jobs[i];
        std::vector<double> eps;
        std::string e;
        rcs[i] = o.slot < 0 ? diagonalize_direct(s, eps, &tr_PG[i], &e)
                            : extrapolate(s, coef + (size_t)kMaxHistory * i, o.count, eps, &tr_PG[i], &e);
        double se = 0.0;
        for (int c = 0; c < s->nocc && c < (int)eps.size(); ++c) se += eps[c];
        sum_eps[i] = se;
        if (rcs[i] && err && err->empty()) *err = e;
    }
    return DSSCF_OK;
}

}  // namespace scfdev
