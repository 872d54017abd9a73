// grad_kernels.hip -- ERI part of the RHF nuclear gradient on gfx950 (the
// role of scf::addExactTwoElectronGradient, src/chem/scf/ExactCoulomb.cpp:
// 125-195, over flo::electronRepulsionGradient, GaussianIntegral.cpp:
// 1052-1459), for s and p shells.
//
//   dE/dX = sum over unique shell quartets g * sum_{abcd} Gamma_abcd d(ab|cd)/dX
//   Gamma_abcd = 16 Pa_ab Pa_cd - 4 (Pa_ac Pa_bd + Pa_ad Pa_bc)
//
// with g = 1/((1+d_ab)(1+d_cd)(1+d_ab,cd)) standing for the quartet's eight
// permutation images (the same weight as the J/K kernels).  Work units are
// the context's J/K units (one bra shell pair per block, one ket per
// thread); per quartet the AO-block Gamma is moved to the Cartesian block,
// the generated shell_grad_* evaluates the A, B and C derivatives contracted
// with it, D follows from translational invariance, and the four centres'
// contributions go into a per-block LDS copy of the atom gradient that is
// flushed once per block.
#include "device_common.hpp"
#include "kernels.hpp"

#include "grad_classes.inc"

template <int la, int lb, int lc, int ld, int BM>
struct GradQuartet;
#define DSSCF_GRAD_SPEC(a, b, c, d, name)                                                                   \
    template <int BM>                                                                                     \
    struct GradQuartet<a, b, c, d, BM> {                                                                  \
        static __device__ __forceinline__ void run(const ShellPair& bra, const ShellPair& ket,            \
                                                   const PrimPair* bp, const PrimPair* kp,                 \
                                                   const double* btab, const double* Gc, double* g9)       \
        {                                                                                                 \
            shell_grad_##name<BM>(bra, ket, bp, kp, btab, Gc, g9);                                        \
        }                                                                                                 \
        static __device__ __forceinline__ void density(const double* xf, const double* Gs, double* Gc)    \
        {                                                                                                 \
            grad_density_##name(xf, Gs, Gc);                                                              \
        }                                                                                                 \
    };
DSSCF_FOR_EACH_GRAD_CLASS(DSSCF_GRAD_SPEC)
#undef DSSCF_GRAD_SPEC

template <int la, int lb, int lc, int ld, int BM>
__global__ void __launch_bounds__(128) k_grad(const JKSeg* __restrict__ segs, const int* __restrict__ seg_blk0,
                                              int nseg, const double* __restrict__ btab,
                                              const double* __restrict__ xf, const int* __restrict__ ao_atom,
                                              double* __restrict__ grad, int natom, double scale)
{
    constexpr int NA = 2 * la + 1, NB = 2 * lb + 1, NC = 2 * lc + 1, ND = 2 * ld + 1;
    constexpr int NCART = (la + 1) * (la + 2) / 2 * (lb + 1) * (lb + 2) / 2 * (lc + 1) * (lc + 2) / 2 * (ld + 1) *
                          (ld + 2) / 2;
    extern __shared__ double gsh[];  // [3 * natom]
    const int si = find_segment_by_block(seg_blk0, nseg, (int)blockIdx.x);
    const JKSeg& S = segs[si];
    const int4 u = ldg(S.u + ((int)blockIdx.x - S.blk0));
    const ShellPair bra = ldg(S.pairs + u.x);
    // accurate mode: as the J/K kernels, the Boys function from the block's LDS
    // copy of the interval table of m = L + 1 and the bra's primitive records
    // in LDS with P - A in place of P
    constexpr bool kLds = BM == BOYS_ACCURATE;
    constexpr int BMq = kLds ? BOYS_ACCURATE_LDS : BM;
    constexpr int LG = la + lb + lc + ld + 1;
    __shared__ __align__(16) double bpoly[kLds ? DSSCF_BPOLY_M_STRIDE : 1];
    __shared__ PrimPair braP[kLds ? kBraStage : 1];
    if (kLds) {
        const GPtr<double> src = gbl(btab + DSSCF_BPOLY_OFFSET + LG * DSSCF_BPOLY_M_STRIDE);
#pragma unroll 4
        for (int i = threadIdx.x; i < DSSCF_BPOLY_M_STRIDE; i += blockDim.x) bpoly[i] = src[i];
        for (int i = threadIdx.x; i < bra.nprim; i += blockDim.x) {
            PrimPair r = ldg(S.pp + bra.prim_off + i);
            for (int k = 0; k < 3; ++k) r.P[k] -= bra.A[k];
            braP[i] = r;
        }
    }
    const double* __restrict__ tq = kLds ? bpoly : btab;
    const PrimPair* __restrict__ bp = kLds ? braP : S.pp + bra.prim_off;
    for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x) gsh[i] = 0.0;
    __syncthreads();
    const int a0 = bra.ao_a, b0 = bra.ao_b, N = S.N;
    const GPtr<double> D = gbl(S.Da);
    const int atA = ao_atom[a0], atB = ao_atom[b0];
    const double gbra = scale * (a0 == b0 ? 0.5 : 1.0);
    for (int t = threadIdx.x; t < u.z; t += blockDim.x) {
        const ShellPair ket = ldg(S.pairs + S.k[u.y + t]);
        const int c0 = ket.ao_a, d0 = ket.ao_b;
        const bool same = c0 == a0 && d0 == b0;
        const double g = gbra * (c0 == d0 ? 0.5 : 1.0) * (same ? 0.5 : 1.0);
        double Gs[NA * NB * NC * ND];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int k = 0; k < NC; ++k)
#pragma unroll
                    for (int l = 0; l < ND; ++l) {
                        const size_t a = a0 + i, b = b0 + j, c = c0 + k, d = d0 + l;
                        Gs[((i * NB + j) * NC + k) * ND + l] =
                            g * (16.0 * D[a * N + b] * D[c * N + d] - 4.0 * (D[a * N + c] * D[b * N + d] +
                                                                             D[a * N + d] * D[b * N + c]));
                    }
        double Gc[NCART];
        GradQuartet<la, lb, lc, ld, BMq>::density(xf, Gs, Gc);
        double g9[9];
        GradQuartet<la, lb, lc, ld, BMq>::run(bra, ket, bp, S.pp + ket.prim_off, tq, Gc, g9);
        const int atC = ao_atom[c0], atD = ao_atom[d0];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double gA = g9[x], gB = g9[3 + x], gC = g9[6 + x];
            atomicAdd(&gsh[3 * atA + x], gA);
            atomicAdd(&gsh[3 * atB + x], gB);
            atomicAdd(&gsh[3 * atC + x], gC);
            atomicAdd(&gsh[3 * atD + x], -(gA + gB + gC));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x) {
        const double x = gsh[i];
        if (x != 0.0) unsafeAtomicAdd(&grad[i], x);
    }
}

hipError_t dsscf_launch_grad(const int* l4, int boys_mode, const JKSeg* segs, const int* blk0, int nseg, int nblocks,
                             const double* btab, const double* xform, const int* ao_atom, double* grad, int natom,
                             double scale, hipStream_t s)
{
    if (nblocks <= 0) return hipSuccess;
    const dim3 blk(128), grd(nblocks);
    const size_t lds = sizeof(double) * 3 * (size_t)natom;
#define DSSCF_GRAD_GO(a, b, c, d, name)                                                                       \
    if (l4[0] == a && l4[1] == b && l4[2] == c && l4[3] == d) {                                               \
        if (boys_mode == BOYS_ACCURATE)                                                                       \
            hipLaunchKernelGGL((k_grad<a, b, c, d, BOYS_ACCURATE>), grd, blk, lds, s, segs, blk0, nseg, btab, xform, \
                               ao_atom, grad, natom, scale);                                                  \
        else                                                                                                  \
            hipLaunchKernelGGL((k_grad<a, b, c, d, BOYS_REFERENCE>), grd, blk, lds, s, segs, blk0, nseg, btab,   \
                               xform, ao_atom, grad, natom, scale);                                           \
        return hipGetLastError();                                                                             \
    }
    DSSCF_FOR_EACH_GRAD_CLASS(DSSCF_GRAD_GO)
#undef DSSCF_GRAD_GO
    return hipErrorInvalidValue;
}
