// eri_kernels.hip -- gfx950 kernels of the direct-SCF Fock build.
//
// One thread per contracted shell quartet (thread-per-quartet): the quartet
// lists are grouped by class (la lb|lc ld) and, inside a class, by the bra
// and ket primitive-pair counts, so the 64 lanes of a wavefront run the same
// primitive loop trip counts.  Per primitive quartet: Boys F_m(T), HGP/OS
// VRR in registers, accumulation of [e0|f0]; per quartet: HRR, transform to
// the reference's AO functions, J/K digestion with FP64 atomics into
// half-matrices that a final kernel folds (J = Jh + Jh^T, K = Kh + Kh^T) into
// packed-U output.
#include "device_common.hpp"
#include "kernels.hpp"

#include "eri_classes.inc"

template <int la, int lb, int lc, int ld, int BM, typename R>
struct Quartet;

#define DSSCF_QUARTET_SPECIALIZE(a, b, c, d, name)                                                    \
    template <int BM, typename R>                                                                     \
    struct Quartet<a, b, c, d, BM, R> {                                                               \
        static __device__ __forceinline__ void run(const ShellPair& bra, const ShellPair& ket,        \
                                                   const PrimPair* __restrict__ pp,                   \
                                                   const double* __restrict__ btab, double* out)      \
        {                                                                                             \
            shell_quartet_##name<BM, R>(bra, ket, pp, btab, out);                                     \
        }                                                                                             \
    };
DSSCF_FOR_EACH_CLASS(DSSCF_QUARTET_SPECIALIZE)

__device__ __forceinline__ void atomic_add_f64(double* p, double v)
{
    // global_atomic_add_f64 (no CAS loop)
    unsafeAtomicAdd(p, v);
}

// ----------------------------------------------------------------- J/K --
// J_mn = sum_ls (mn|ls) Dt_ls, K_mn = sum_ls (ml|ns) Da_ls (ExactCoulomb.cpp:
// 103-123).  A unique quartet (ab|cd) with weight g = 1/((1+d_ab)(1+d_cd)
// (1+d_ab,cd)) stands for its 8 permutation images; half of them go into
// Jh/Kh and the other half are the transposes added by the fold kernel.
template <int la, int lb, int lc, int ld, int BM>
__global__ void __launch_bounds__(256) k_jk(const uint2* __restrict__ qlist, int nq,
                                            const ShellPair* __restrict__ pairs,
                                            const PrimPair* __restrict__ pp,
                                            const double* __restrict__ btab,
                                            const double* __restrict__ Dt, const double* __restrict__ Da,
                                            double* __restrict__ Jh, double* __restrict__ Kh, int N,
                                            double scale)
{
    typedef ClassInfo<la, lb, lc, ld> CI;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const uint2 q = qlist[t];
    const ShellPair bra = pairs[q.x];
    const ShellPair ket = pairs[q.y];
    double v[CI::NAO];
    Quartet<la, lb, lc, ld, BM, double>::run(bra, ket, pp, btab, v);

    double g = scale;
    if (bra.ao_a == bra.ao_b) g *= 0.5;
    if (ket.ao_a == ket.ao_b) g *= 0.5;
    if (q.x == q.y) g *= 0.5;
    const int a0 = bra.ao_a, b0 = bra.ao_b, c0 = ket.ao_a, d0 = ket.ao_b;

    // Coulomb: Jh[a][b] += 2g sum_cd v D_cd ; Jh[c][d] += 2g sum_ab v D_ab
    {
        double dcd[NC * ND], dab[NA * NB];
#pragma unroll
        for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int l = 0; l < ND; ++l) dcd[k * ND + l] = Dt[(size_t)(c0 + k) * N + d0 + l];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) dab[i * NB + j] = Dt[(size_t)(a0 + i) * N + b0 + j];
        const double g2 = 2.0 * g;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                double s = 0.0;
#pragma unroll
                for (int kl = 0; kl < NC * ND; ++kl) s += v[(i * NB + j) * NC * ND + kl] * dcd[kl];
                atomic_add_f64(&Jh[(size_t)(a0 + i) * N + b0 + j], g2 * s);
            }
#pragma unroll
        for (int kl = 0; kl < NC * ND; ++kl) {
            double s = 0.0;
#pragma unroll
            for (int ij = 0; ij < NA * NB; ++ij) s += v[ij * NC * ND + kl] * dab[ij];
            atomic_add_f64(&Jh[(size_t)(c0 + kl / ND) * N + d0 + kl % ND], g2 * s);
        }
    }
    // Exchange: Kh[a][c] += g v D_bd, Kh[b][c] += g v D_ad, Kh[a][d] += g v D_bc, Kh[b][d] += g v D_ac
    {
        double dbd[NB * ND], dad[NA * ND], dbc[NB * NC], dac[NA * NC];
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int l = 0; l < ND; ++l) dbd[j * ND + l] = Da[(size_t)(b0 + j) * N + d0 + l];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int l = 0; l < ND; ++l) dad[i * ND + l] = Da[(size_t)(a0 + i) * N + d0 + l];
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int k = 0; k < NC; ++k) dbc[j * NC + k] = Da[(size_t)(b0 + j) * N + c0 + k];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int k = 0; k < NC; ++k) dac[i * NC + k] = Da[(size_t)(a0 + i) * N + c0 + k];
#define VIDX(i, j, k, l) ((((i) * NB + (j)) * NC + (k)) * ND + (l))
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int l = 0; l < ND; ++l) s += v[VIDX(i, j, k, l)] * dbd[j * ND + l];
                atomic_add_f64(&Kh[(size_t)(a0 + i) * N + c0 + k], g * s);
            }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int l = 0; l < ND; ++l) s += v[VIDX(i, j, k, l)] * dad[i * ND + l];
                atomic_add_f64(&Kh[(size_t)(b0 + j) * N + c0 + k], g * s);
            }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int k = 0; k < NC; ++k) s += v[VIDX(i, j, k, l)] * dbc[j * NC + k];
                atomic_add_f64(&Kh[(size_t)(a0 + i) * N + d0 + l], g * s);
            }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int k = 0; k < NC; ++k) s += v[VIDX(i, j, k, l)] * dac[i * NC + k];
                atomic_add_f64(&Kh[(size_t)(b0 + j) * N + d0 + l], g * s);
            }
#undef VIDX
    }
}

// --------------------------------------------------------- ERI batch ---
// Whole AO blocks of listed shell quartets (parity path for the unit API
// flo::electronRepulsionIntegral, GaussianIntegral.cpp:561).
template <int la, int lb, int lc, int ld, int BM>
__global__ void __launch_bounds__(256) k_eri_block(const int4* __restrict__ items, int n,
                                                   const ShellPair* __restrict__ pairs,
                                                   const PrimPair* __restrict__ pp,
                                                   const double* __restrict__ btab, double* __restrict__ out,
                                                   double scale)
{
    typedef ClassInfo<la, lb, lc, ld> CI;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int4 it = items[t];
    const ShellPair bra = pairs[it.x];
    const ShellPair ket = pairs[it.y];
    double v[CI::NAO];
    Quartet<la, lb, lc, ld, BM, double>::run(bra, ket, pp, btab, v);
    double* o = out + (size_t)it.z;
#pragma unroll
    for (int i = 0; i < CI::NAO; ++i) o[i] = scale * v[i];
}

// --------------------------------------------------------- Schwarz -----
// Q_ab = sqrt(max_{mu in a, nu in b} |(mu nu|mu nu)|) for each listed pair.
template <int la, int lb, int BM>
__global__ void __launch_bounds__(256) k_schwarz(const int* __restrict__ plist, int n,
                                                 const ShellPair* __restrict__ pairs,
                                                 const PrimPair* __restrict__ pp,
                                                 const double* __restrict__ btab, double* __restrict__ Q,
                                                 double scale)
{
    typedef ClassInfo<la, lb, la, lb> CI;
    constexpr int NA = CI::NA, NB = CI::NB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int pidx = plist[t];
    const ShellPair pr = pairs[pidx];
    double v[CI::NAO];
    Quartet<la, lb, la, lb, BM, double>::run(pr, pr, pp, btab, v);
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) mx = fmax(mx, fabs(scale * v[(i * NB + j) * NA * NB + i * NB + j]));
    Q[pidx] = sqrt(mx);
}

// ------------------------------------------------------ helper kernels --
__global__ void k_fill(double* __restrict__ p, size_t n, double v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// packed-U (y<=x at x(x+1)/2+y, Matrix.ipp:332-351) -> full row-major
__global__ void k_unpack(const double* __restrict__ pk, double* __restrict__ full, int N)
{
    const int x = blockIdx.y;
    for (int y = threadIdx.x + blockIdx.x * blockDim.x; y <= x; y += gridDim.x * blockDim.x) {
        const double v = pk[(size_t)x * (x + 1) / 2 + y];
        full[(size_t)y * N + x] = v;
        full[(size_t)x * N + y] = v;
    }
}

// J = Jh + Jh^T, K = Kh + Kh^T -> packed-U
__global__ void k_fold(const double* __restrict__ Jh, const double* __restrict__ Kh, double* __restrict__ J,
                       double* __restrict__ K, int N)
{
    const int x = blockIdx.y;
    for (int y = threadIdx.x + blockIdx.x * blockDim.x; y <= x; y += gridDim.x * blockDim.x) {
        const size_t o = (size_t)x * (x + 1) / 2 + y;
        J[o] = Jh[(size_t)y * N + x] + Jh[(size_t)x * N + y];
        K[o] = Kh[(size_t)y * N + x] + Kh[(size_t)x * N + y];
    }
}

// -------------------------------------------------------- host launch ---
void dsscf_upload_xform(const double* x, int n, hipStream_t s)
{
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_ao_xform), x, sizeof(double) * n, 0, hipMemcpyHostToDevice, s);
}

int dk_xform_size() { return DSSCF_XFORM_SIZE; }
int dk_max_l() { return DSSCF_MAXL; }

namespace {
struct ClassDesc {
    int la, lb, lc, ld;
};
const ClassDesc kClasses[] = {
#define DSSCF_CLASS_DESC(a, b, c, d, name) {a, b, c, d},
    DSSCF_FOR_EACH_CLASS(DSSCF_CLASS_DESC)
#undef DSSCF_CLASS_DESC
};
}  // namespace

int dk_n_class() { return (int)(sizeof(kClasses) / sizeof(kClasses[0])); }

void dk_class_l(int c, int* l)
{
    l[0] = kClasses[c].la;
    l[1] = kClasses[c].lb;
    l[2] = kClasses[c].lc;
    l[3] = kClasses[c].ld;
}

hipError_t dsscf_launch_jk(int cls, int boys_mode, const uint2* q, int nq, const ShellPair* pairs,
                           const PrimPair* pp, const double* btab, const double* Dt, const double* Da,
                           double* Jh, double* Kh, int N, double scale, hipStream_t s)
{
    if (nq <= 0) return hipSuccess;
    const dim3 blk(256), grd((nq + 255) / 256);
    int c = 0;
#define DSSCF_LAUNCH_JK(a, b, cc, d, name)                                                                  \
    if (c++ == cls) {                                                                                       \
        if (boys_mode == BOYS_ACCURATE)                                                                     \
            hipLaunchKernelGGL((k_jk<a, b, cc, d, BOYS_ACCURATE>), grd, blk, 0, s, q, nq, pairs, pp, btab, \
                               Dt, Da, Jh, Kh, N, scale);                                                   \
        else                                                                                                \
            hipLaunchKernelGGL((k_jk<a, b, cc, d, BOYS_REFERENCE>), grd, blk, 0, s, q, nq, pairs, pp, btab,\
                               Dt, Da, Jh, Kh, N, scale);                                                   \
        return hipGetLastError();                                                                           \
    }
    DSSCF_FOR_EACH_CLASS(DSSCF_LAUNCH_JK)
#undef DSSCF_LAUNCH_JK
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_eri_block(int cls, int boys_mode, const int4* items, int n, const ShellPair* pairs,
                                  const PrimPair* pp, const double* btab, double* out, double scale,
                                  hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const dim3 blk(256), grd((n + 255) / 256);
    int c = 0;
#define DSSCF_LAUNCH_BLK(a, b, cc, d, name)                                                                       \
    if (c++ == cls) {                                                                                             \
        if (boys_mode == BOYS_ACCURATE)                                                                           \
            hipLaunchKernelGGL((k_eri_block<a, b, cc, d, BOYS_ACCURATE>), grd, blk, 0, s, items, n, pairs, pp,   \
                               btab, out, scale);                                                                 \
        else                                                                                                      \
            hipLaunchKernelGGL((k_eri_block<a, b, cc, d, BOYS_REFERENCE>), grd, blk, 0, s, items, n, pairs, pp,  \
                               btab, out, scale);                                                                 \
        return hipGetLastError();                                                                                 \
    }
    DSSCF_FOR_EACH_CLASS(DSSCF_LAUNCH_BLK)
#undef DSSCF_LAUNCH_BLK
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_schwarz(int la, int lb, int boys_mode, const int* plist, int n, const ShellPair* pairs,
                                const PrimPair* pp, const double* btab, double* Q, double scale, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const dim3 blk(256), grd((n + 255) / 256);
#define DSSCF_LAUNCH_SCH(a, b)                                                                                  \
    if (la == a && lb == b) {                                                                                   \
        if (boys_mode == BOYS_ACCURATE)                                                                         \
            hipLaunchKernelGGL((k_schwarz<a, b, BOYS_ACCURATE>), grd, blk, 0, s, plist, n, pairs, pp, btab, Q, \
                               scale);                                                                          \
        else                                                                                                    \
            hipLaunchKernelGGL((k_schwarz<a, b, BOYS_REFERENCE>), grd, blk, 0, s, plist, n, pairs, pp, btab,   \
                               Q, scale);                                                                       \
        return hipGetLastError();                                                                               \
    }
    DSSCF_LAUNCH_SCH(0, 0)
    DSSCF_LAUNCH_SCH(1, 0)
    DSSCF_LAUNCH_SCH(1, 1)
#if DSSCF_MAXL >= 2
    DSSCF_LAUNCH_SCH(2, 0)
    DSSCF_LAUNCH_SCH(2, 1)
    DSSCF_LAUNCH_SCH(2, 2)
#endif
#undef DSSCF_LAUNCH_SCH
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_fill(double* p, size_t n, double v, hipStream_t s)
{
    const int grd = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_fill, dim3(grd > 0 ? grd : 1), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t dsscf_launch_unpack(const double* pk, double* full, int N, hipStream_t s)
{
    hipLaunchKernelGGL(k_unpack, dim3((N + 255) / 256, N), dim3(256), 0, s, pk, full, N);
    return hipGetLastError();
}

hipError_t dsscf_launch_fold(const double* Jh, const double* Kh, double* J, double* K, int N, hipStream_t s)
{
    hipLaunchKernelGGL(k_fold, dim3((N + 255) / 256, N), dim3(256), 0, s, Jh, Kh, J, K, N);
    return hipGetLastError();
}
