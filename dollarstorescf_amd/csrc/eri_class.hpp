// eri_class.hpp -- gfx950 kernels of the direct-SCF Fock build, per class:
// included by one generated translation unit per shell-quartet class, which
// instantiates ClassLaunch for that class (the classes compile in parallel).
//
// One thread per contracted shell quartet (thread-per-quartet): the quartet
// lists are grouped by class (la lb|lc ld) and, inside a class, by the bra
// and ket primitive-pair counts, so the 64 lanes of a wavefront run the same
// primitive loop trip counts.  Per primitive quartet: Boys F_m(T), HGP/OS
// VRR in registers, accumulation of [e0|f0]; per quartet: HRR, transform to
// the reference's AO functions, J/K digestion with FP64 atomics into
// half-matrices that a final kernel folds (J = Jh + Jh^T, K = Kh + Kh^T) into
// packed-U output.
#pragma once
#include "device_common.hpp"
#include "launch.hpp"

#include <type_traits>

#include "eri_classes.inc"


template <int la, int lb, int lc, int ld, int BM, typename R, int C = 0>
struct Quartet;
#define DSSCF_QSPEC_1(a, b, c, d, name, C, fn)
#define DSSCF_QSPEC_0(a, b, c, d, name, C, fn)                                                        \
    template <int BM, typename R>                                                                     \
    struct Quartet<a, b, c, d, BM, R, C> {                                                            \
        template <typename BP>                                                                        \
        static __device__ __forceinline__ void run(const ShellPair& bra, const ShellPair& ket,        \
                                                   const BP* __restrict__ bra_pp,                     \
                                                   const PrimPair* __restrict__ ket_pp,               \
                                                   const double* __restrict__ btab, double* out,      \
                                                   int gb = 1, int rb = 0, int gk = 1, int rk = 0)    \
        {                                                                                             \
            fn<BM, R, BP>(bra, ket, bra_pp, ket_pp, btab, out, gb, rb, gk, rk);                       \
        }                                                                                             \
    };
#define DSSCF_QUARTET_SPECIALIZE(a, b, c, d, name, wave) \
    DSSCF_QSPEC_##wave(a, b, c, d, name, 0, shell_quartet_##name)
DSSCF_FOR_EACH_CLASS(DSSCF_QUARTET_SPECIALIZE)
#define DSSCF_QUARTET_SPECIALIZE_CART(a, b, c, d, name, wave) \
    DSSCF_QSPEC_##wave(a, b, c, d, name, 1, shell_quartet_##name##_cart)
DSSCF_FOR_EACH_D_CLASS(DSSCF_QUARTET_SPECIALIZE_CART)

#include "wave_kernel.hpp"

__device__ __forceinline__ void atomic_add_f64(double* p, double v)
{
    // global_atomic_add_f64 (no CAS loop)
    unsafeAtomicAdd(p, v);
}

// ----------------------------------------------------------------- J/K --
// J_mn = sum_ls (mn|ls) Dt_ls, K_mn = sum_ls (ml|ns) Da_ls (ExactCoulomb.cpp:
// 103-123).  A unique quartet (ab|cd) with weight g = 1/((1+d_ab)(1+d_cd)
// (1+d_ab,cd)) stands for its 8 permutation images; half of them go into
// Jh/Kh and the other half are the transposes added by the fold kernel:
//   Jh[a][b] += 2g sum_cd v D_cd   Jh[c][d] += 2g sum_ab v D_ab
//   Kh[a][c] += g v D_bd   Kh[b][c] += g v D_ad   Kh[a][d] += g v D_bc   Kh[b][d] += g v D_ac
//
// Work unit = one bra shell pair and a chunk of its ket pairs (bra-major):
// the bra record and its primitive pairs are uniform across the block; the
// Jh[a][b] sums are reduced across the block, the Kh rows of shells a and b
// are accumulated in LDS and flushed once per unit; only Jh[c][d] leaves
// per quartet: as a plain store into the quartet's slot record, which k_jcd
// sums per element after all classes (an atomic add per quartet and element
// serializes at the few hundred addresses a small molecule's J has), or,
// for contexts too large for the slot buffer, as an atomic add.  A unit's
// fourth field gives the lanes per quartet, gb (bra primitives) | gk (ket
// primitives) << 8: a group of gb*gk lanes splits a primitive-heavy
// quartet's primitive loops, and the group's first lane digests the summed
// integrals.
//
// Block b takes units b, b + gridDim, b + 2 gridDim, ... (the list is
// heaviest first; launches use one block per unit, measured as fast as a
// grid of two resident blocks per CU that loop, and faster than blocks
// taking units from an atomic counter, whose same-address adds serialize).
//
// Dynamic LDS: the Kh rows [nks * (NA + NB) * maxN doubles, rounded to an
// even count; nks = 2 when a segment of the launch is unrestricted, whose
// blocks keep the K[Db] rows after the K[Da] rows], then the bra's primitive
// records (at most maxBra of them, sized by the launch rather than for the
// largest possible shell pair, so that more blocks fit on a CU).
template <int la, int lb, int lc, int ld, int BM, typename R, int C>
__global__ void __launch_bounds__(256) k_jk(const JKSeg* __restrict__ segs, const int* __restrict__ seg_blk0,
                                            int nseg, int nunits, int maxN, int nks,
                                            const double* __restrict__ btab, double scale)
{
    typedef ClassInfo<la, lb, lc, ld, C> CI;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND;
    // bra primitive records: FP64 lists PrimPair, FP32 lists (accurate mode) PrimPairF
    constexpr bool kF32Rec = sizeof(R) == 4 && BM == BOYS_ACCURATE;
    typedef typename std::conditional<kF32Rec, PrimPairF, PrimPair>::type BPT;
    extern __shared__ __align__(16) double lds_dyn[];
    double* krow = lds_dyn;  // Kh rows of shell a, then shell b
    BPT* braP = reinterpret_cast<BPT*>(lds_dyn + ((nks * (NA + NB) * maxN + 1) & ~1));
    // FP64 accurate mode: the Boys function from the block's LDS copy of the
    // interval table of m = L (a per-lane gather from the 801-row table
    // touches up to 64 cache lines per load)
    // (FP32 lists: a float copy of its e^-c and first seven coefficients)
    constexpr bool kLds = BM == BOYS_ACCURATE;
    constexpr bool k64 = sizeof(R) == 8;
    constexpr int BMq = kLds ? BOYS_ACCURATE_LDS : BM;
    __shared__ __align__(16) double bpoly[kLds ? (k64 ? DSSCF_BPOLY_M_STRIDE : DSSCF_BPOLY_ROWS * DSSCF_BPOLY32_COEF / 2) : 1];
    if (kLds) {
        const GPtr<double> src = gbl(btab + DSSCF_BPOLY_OFFSET + CI::L * DSSCF_BPOLY_M_STRIDE);
        // four loads in flight per thread
        if (k64) {
#pragma unroll 4
            for (int i = threadIdx.x; i < DSSCF_BPOLY_M_STRIDE; i += blockDim.x) bpoly[i] = src[i];
        } else {
            // coefficient q of interval r sits at double 2 ((q / 2) ROWS + r) + q % 2
            float* f = reinterpret_cast<float*>(bpoly);
#pragma unroll 4
            for (int i = threadIdx.x; i < DSSCF_BPOLY_ROWS * DSSCF_BPOLY32_COEF; i += blockDim.x) {
                const int r = i / DSSCF_BPOLY32_COEF, q = i % DSSCF_BPOLY32_COEF;
                f[i] = (float)src[2 * ((q >> 1) * DSSCF_BPOLY_ROWS + r) + (q & 1)];
            }
        }
    }
    const double* __restrict__ tq = kLds ? bpoly : btab;

    for (int vb = (int)blockIdx.x; vb < nunits; vb += (int)gridDim.x) {
        const int si = find_segment_by_block(seg_blk0, nseg, vb);
        const JKSeg& S = segs[si];
        const int4 u = ldg(S.u + (vb - S.blk0));  // bra pair, first ket, ket count
        const ShellPair* __restrict__ pairs = S.pairs;
        const PrimPair* __restrict__ pp = S.pp;
        const GPtr<double> Dt = gbl(S.Dt);
        const GPtr<double> Da = gbl(S.Da);  // read-only in this kernel
        const int N = S.N;
        const size_t slice = (size_t)jk_slice(S.nslice) * N * N;
        double* __restrict__ Jh = S.Jh + slice;
        double* __restrict__ Kh = S.Kh + slice;
        // unrestricted segment: K[Db] from the same integrals into the rows
        // after the K[Da] rows (uniform over the block)
        const bool beta = S.Db != nullptr;
        const int nrow = (beta ? 2 : 1) * (NA + NB);
        const GPtr<unsigned> kets = gbl(S.k);
        // J_cd: per-quartet slot records (summed by k_jcd) or atomics
        const int slot0 = S.js ? gbl(S.us)[vb - S.blk0] : 0;
        const ShellPair bra = ldg(pairs + u.x);
        const int a0 = bra.ao_a, b0 = bra.ao_b;
        for (int i = threadIdx.x; i < nrow * N; i += blockDim.x) krow[i] = 0.0;
        for (int i = threadIdx.x; i < bra.nprim; i += blockDim.x) {
            PrimPair r = ldg(pp + bra.prim_off + i);
            if (kLds)  // the LDS Boys mode's prim_setup takes P - A
                for (int k = 0; k < 3; ++k) r.P[k] -= bra.A[k];
            if constexpr (kF32Rec) {
                PrimPairF f;
                f.p = (float)r.p;
                f.inv_p = (float)r.inv_p;
                for (int k = 0; k < 3; ++k) f.PA[k] = (float)r.P[k];
                f.Kp = (float)r.Kp;
                f.pad[0] = f.pad[1] = 0.0f;
                braP[i] = f;
            } else {
                braP[i] = r;
            }
        }
        double dab[NA * NB];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) dab[i * NB + j] = Dt[(size_t)(a0 + i) * N + b0 + j];
        double jab[NA * NB];
#pragma unroll
        for (int k = 0; k < NA * NB; ++k) jab[k] = 0.0;
        const double gb = scale * (a0 == b0 ? 0.5 : 1.0);
        __syncthreads();

        const int lgb = u.w & 0xff, lgk = (u.w >> 8) & 0xff, gl = lgb * lgk;  // lanes per quartet (power of two)
        const int lane = (int)threadIdx.x & (gl - 1);
        const int rb = lane & (lgb - 1), rk = lane / lgb;
        for (int t = (int)threadIdx.x / gl; t < u.z; t += (int)blockDim.x / gl) {
            const int ki = (int)kets[u.y + t];
            const ShellPair ket = ldg(pairs + ki);
            const int c0 = ket.ao_a, d0 = ket.ao_b;
            double v[CI::NAO];
            Quartet<la, lb, lc, ld, BMq, R, C>::run(bra, ket, braP, S.pp + ket.prim_off, tq, v, lgb, rb, lgk, rk);
            if (lane != 0) continue;
            // density elements of the digestion: Dt[c][d], Da[b][d], Da[a][d], Da[b][c], Da[a][c]
            // (loading them before the primitive loops, to overlap their latency,
            // was measured slower: the registers held across the loops cost occupancy)
            double dcd[NC * ND], dbd[NB * ND], dad[NA * ND], dbc[NB * NC], dac[NA * NC];
#pragma unroll
            for (int k = 0; k < NC; ++k)
#pragma unroll
                for (int l = 0; l < ND; ++l) dcd[k * ND + l] = Dt[(size_t)(c0 + k) * N + d0 + l];
#pragma unroll
            for (int l = 0; l < ND; ++l) {
#pragma unroll
                for (int j = 0; j < NB; ++j) dbd[j * ND + l] = Da[(size_t)(b0 + j) * N + d0 + l];
#pragma unroll
                for (int i = 0; i < NA; ++i) dad[i * ND + l] = Da[(size_t)(a0 + i) * N + d0 + l];
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) {
#pragma unroll
                for (int j = 0; j < NB; ++j) dbc[j * NC + k] = Da[(size_t)(b0 + j) * N + c0 + k];
#pragma unroll
                for (int i = 0; i < NA; ++i) dac[i * NC + k] = Da[(size_t)(a0 + i) * N + c0 + k];
            }
            // (ab|ab): the bra and ket (sub-)pairs belong to one shell pair
            const bool same = ket.ao_a == a0 && ket.ao_b == b0;
            const double g = gb * (c0 == d0 ? 0.5 : 1.0) * (same ? 0.5 : 1.0);
#define VIDX(i, j, k, l) ((((i) * NB + (j)) * NC + (k)) * ND + (l))
            // Coulomb
            const GOut<double> js = S.js ? gout(S.js) + gbl(S.jd)[slot0 + t] : nullptr;
#pragma unroll
            for (int k = 0; k < NC; ++k)
#pragma unroll
                for (int l = 0; l < ND; ++l) {
                    const double dk = dcd[k * ND + l];
                    double s = 0.0;
#pragma unroll
                    for (int ij = 0; ij < NA * NB; ++ij) {
                        const double x = v[ij * NC * ND + k * ND + l];
                        jab[ij] += g * x * dk;
                        s += x * dab[ij];
                    }
                    if (js)
                        js[k * ND + l] = 2.0 * g * s;
                    else
                        atomic_add_f64(&Jh[(size_t)(c0 + k) * N + d0 + l], 2.0 * g * s);
                }
            // Exchange into the LDS rows: K[Da] (rows 0 .. NA+NB), then, for an
            // unrestricted segment, K[Db] from the same integrals (rows NA+NB ..)
            auto exchange = [&](double* __restrict__ kr) {
#pragma unroll
                for (int k = 0; k < NC; ++k) {
#pragma unroll
                    for (int i = 0; i < NA; ++i) {  // Kh[a][c] += g sum_bd v D_bd
                        double s = 0.0;
#pragma unroll
                        for (int j = 0; j < NB; ++j)
#pragma unroll
                            for (int l = 0; l < ND; ++l) s += v[VIDX(i, j, k, l)] * dbd[j * ND + l];
                        atomicAdd(&kr[i * N + c0 + k], g * s);
                    }
#pragma unroll
                    for (int j = 0; j < NB; ++j) {  // Kh[b][c] += g sum_ad v D_ad
                        double s = 0.0;
#pragma unroll
                        for (int i = 0; i < NA; ++i)
#pragma unroll
                            for (int l = 0; l < ND; ++l) s += v[VIDX(i, j, k, l)] * dad[i * ND + l];
                        atomicAdd(&kr[(NA + j) * N + c0 + k], g * s);
                    }
                }
#pragma unroll
                for (int l = 0; l < ND; ++l) {
#pragma unroll
                    for (int i = 0; i < NA; ++i) {  // Kh[a][d] += g sum_bc v D_bc
                        double s = 0.0;
#pragma unroll
                        for (int j = 0; j < NB; ++j)
#pragma unroll
                            for (int k = 0; k < NC; ++k) s += v[VIDX(i, j, k, l)] * dbc[j * NC + k];
                        atomicAdd(&kr[i * N + d0 + l], g * s);
                    }
#pragma unroll
                    for (int j = 0; j < NB; ++j) {  // Kh[b][d] += g sum_ac v D_ac
                        double s = 0.0;
#pragma unroll
                        for (int i = 0; i < NA; ++i)
#pragma unroll
                            for (int k = 0; k < NC; ++k) s += v[VIDX(i, j, k, l)] * dac[i * NC + k];
                        atomicAdd(&kr[(NA + j) * N + d0 + l], g * s);
                    }
                }
            };
            exchange(krow);
            if (beta) {
                const GPtr<double> Db = gbl(S.Db);
#pragma unroll
                for (int l = 0; l < ND; ++l) {
#pragma unroll
                    for (int j = 0; j < NB; ++j) dbd[j * ND + l] = Db[(size_t)(b0 + j) * N + d0 + l];
#pragma unroll
                    for (int i = 0; i < NA; ++i) dad[i * ND + l] = Db[(size_t)(a0 + i) * N + d0 + l];
                }
#pragma unroll
                for (int k = 0; k < NC; ++k) {
#pragma unroll
                    for (int j = 0; j < NB; ++j) dbc[j * NC + k] = Db[(size_t)(b0 + j) * N + c0 + k];
#pragma unroll
                    for (int i = 0; i < NA; ++i) dac[i * NC + k] = Db[(size_t)(a0 + i) * N + c0 + k];
                }
                exchange(krow + (NA + NB) * N);
            }
#undef VIDX
        }
        // Jh[a][b]: wavefront reduction, one atomic per wavefront and element
#pragma unroll
        for (int k = 0; k < NA * NB; ++k) {
            double x = jab[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
            if ((threadIdx.x & 63) == 0 && x != 0.0)
                atomic_add_f64(&Jh[(size_t)(a0 + k / NB) * N + b0 + k % NB], 2.0 * x);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nrow * N; i += blockDim.x) {
            const double x = krow[i];
            if (x != 0.0) {
                int r = i / N;
                const int col = i % N;
                double* __restrict__ K = Kh;
                if (r >= NA + NB) {  // K[Db] rows of an unrestricted segment
                    r -= NA + NB;
                    K = S.Kbh + slice;
                }
                const int row = r < NA ? a0 + r : b0 + (r - NA);
                atomic_add_f64(&K[(size_t)row * N + col], x);
            }
        }
        __syncthreads();  // the flush has read krow before the next unit clears it
    }
}

// --------------------------------------------------------- ERI batch ---
// Whole AO blocks of listed shell quartets (parity path for the unit API
// flo::electronRepulsionIntegral, GaussianIntegral.cpp:561).
template <int la, int lb, int lc, int ld, int BM, int C>
__global__ void __launch_bounds__(256) k_eri_block(const int4* __restrict__ items, int n,
                                                   const ShellPair* __restrict__ pairs,
                                                   const PrimPair* __restrict__ pp,
                                                   const double* __restrict__ btab, double* __restrict__ out,
                                                   double scale)
{
    typedef ClassInfo<la, lb, lc, ld, C> CI;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int4 it = items[t];
    const ShellPair bra = pairs[it.x];
    const ShellPair ket = pairs[it.y];
    double v[CI::NAO];
    Quartet<la, lb, lc, ld, BM, double, C>::run(bra, ket, pp + bra.prim_off, pp + ket.prim_off, btab, v);
    double* o = out + (size_t)it.z;
#pragma unroll
    for (int i = 0; i < CI::NAO; ++i) o[i] = scale * v[i];
}

// --------------------------------------------------------- Schwarz -----
// Q_ab = sqrt(max_{mu in a, nu in b} |(mu nu|mu nu)|) for each listed pair.
template <int la, int lb, int BM, int C>
__global__ void __launch_bounds__(256) k_schwarz(const int* __restrict__ plist, int n,
                                                 const ShellPair* __restrict__ pairs,
                                                 const PrimPair* __restrict__ pp,
                                                 const double* __restrict__ btab, double* __restrict__ Q,
                                                 double scale)
{
    typedef ClassInfo<la, lb, la, lb, C> CI;
    constexpr int NA = CI::NA, NB = CI::NB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int pidx = plist[t];
    const ShellPair pr = pairs[pidx];
    double v[CI::NAO];
    Quartet<la, lb, la, lb, BM, double, C>::run(pr, pr, pp + pr.prim_off, pp + pr.prim_off, btab, v);
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) mx = fmax(mx, fabs(scale * v[(i * NB + j) * NA * NB + i * NB + j]));
    Q[pidx] = sqrt(mx);
}


// ------------------------------------------------------------ launches --
// cart (run time) selects a class's Cartesian-d form (template C = 1), which
// exists for classes with a d shell; the others have one form.
template <int a, int b, int c, int d, int wave>
template <int BM, int MODE, int C>
void ClassLaunch<a, b, c, d, wave>::wave_go(const void* items, int n, const ShellPair* pairs, const PrimPair* pp,
                                            const double* btab, double* out, double scale, hipStream_t s)
{
    if constexpr (wave) {
        constexpr int G = WaveGroup<a, b, c, d, C>::G;
        const int ng = (n + G - 1) / G;
        const int grd = ng < (1 << 20) ? ng : (1 << 20);
        hipLaunchKernelGGL((k_wave<a, b, c, d, BM, MODE, C>), dim3(grd), dim3(64), 0, s, items, n, pairs, pp, btab,
                           nullptr, nullptr, 0, out, scale);
    }
}

template <int a, int b, int c, int d, int wave>
template <int C>
hipError_t ClassLaunch<a, b, c, d, wave>::jk_c(int bm, int fp32, const JKSeg* segs, const int* blk0, int nseg,
                                               int nblocks, int bs, int maxN, int nks, int maxBra,
                                               const double* btab, double scale, hipStream_t s)
{
    if constexpr (wave) {
        if (bm == BOYS_ACCURATE)
            hipLaunchKernelGGL((k_wave<a, b, c, d, BOYS_ACCURATE, WAVE_JK, C>), dim3(nblocks), dim3(64), 0, s, nullptr,
                               0, nullptr, nullptr, btab, segs, blk0, nseg, nullptr, scale);
        else
            hipLaunchKernelGGL((k_wave<a, b, c, d, BOYS_REFERENCE, WAVE_JK, C>), dim3(nblocks), dim3(64), 0, s,
                               nullptr, 0, nullptr, nullptr, btab, segs, blk0, nseg, nullptr, scale);
    } else {
        typedef ClassInfo<a, b, c, d, C> CI;
        const size_t rows = ((size_t)nks * (CI::NA + CI::NB) * maxN + 1) & ~(size_t)1;
        const size_t lds = sizeof(double) * rows + (fp32 ? sizeof(PrimPairF) : sizeof(PrimPair)) * (size_t)maxBra;
        const dim3 blk(bs), grd(nblocks);  // one block per work unit
        if (fp32)
            hipLaunchKernelGGL((k_jk<a, b, c, d, BOYS_ACCURATE, float, C>), grd, blk, lds, s, segs, blk0, nseg,
                               nblocks, maxN, nks, btab, scale);
        else if (bm == BOYS_ACCURATE)
            hipLaunchKernelGGL((k_jk<a, b, c, d, BOYS_ACCURATE, double, C>), grd, blk, lds, s, segs, blk0, nseg,
                               nblocks, maxN, nks, btab, scale);
        else
            hipLaunchKernelGGL((k_jk<a, b, c, d, BOYS_REFERENCE, double, C>), grd, blk, lds, s, segs, blk0, nseg,
                               nblocks, maxN, nks, btab, scale);
    }
    return hipGetLastError();
}

// J/K over nseg context segments: thread-per-quartet classes one block of bs
// threads per work unit (LDS: the Kh rows of the bra's two shells, nks = 2
// sets of them when a segment is unrestricted, and its primitive records),
// wave-cooperative classes one 64-lane block per WaveGroup::G quartets.
template <int a, int b, int c, int d, int wave>
hipError_t ClassLaunch<a, b, c, d, wave>::jk(int bm, int fp32, const JKSeg* segs, const int* blk0, int nseg,
                                             int nblocks, int bs, int maxN, int nks, int maxBra, int cart,
                                             const double* btab, double scale, hipStream_t s)
{
    if constexpr (kHasD) {
        if (cart) return jk_c<1>(bm, fp32, segs, blk0, nseg, nblocks, bs, maxN, nks, maxBra, btab, scale, s);
    }
    return jk_c<0>(bm, fp32, segs, blk0, nseg, nblocks, bs, maxN, nks, maxBra, btab, scale, s);
}

template <int a, int b, int c, int d, int wave>
template <int C>
hipError_t ClassLaunch<a, b, c, d, wave>::block_c(int bm, const int4* items, int n, const ShellPair* pairs,
                                                  const PrimPair* pp, const double* btab, double* out, double scale,
                                                  hipStream_t s)
{
    if constexpr (wave) {
        if (bm == BOYS_ACCURATE)
            wave_go<BOYS_ACCURATE, WAVE_BLOCK, C>(items, n, pairs, pp, btab, out, scale, s);
        else
            wave_go<BOYS_REFERENCE, WAVE_BLOCK, C>(items, n, pairs, pp, btab, out, scale, s);
    } else {
        const dim3 blk(256), grd((n + 255) / 256);
        if (bm == BOYS_ACCURATE)
            hipLaunchKernelGGL((k_eri_block<a, b, c, d, BOYS_ACCURATE, C>), grd, blk, 0, s, items, n, pairs, pp, btab,
                               out, scale);
        else
            hipLaunchKernelGGL((k_eri_block<a, b, c, d, BOYS_REFERENCE, C>), grd, blk, 0, s, items, n, pairs, pp,
                               btab, out, scale);
    }
    return hipGetLastError();
}

template <int a, int b, int c, int d, int wave>
hipError_t ClassLaunch<a, b, c, d, wave>::block(int bm, int cart, const int4* items, int n, const ShellPair* pairs,
                                                const PrimPair* pp, const double* btab, double* out, double scale,
                                                hipStream_t s)
{
    if constexpr (kHasD) {
        if (cart) return block_c<1>(bm, items, n, pairs, pp, btab, out, scale, s);
    }
    return block_c<0>(bm, items, n, pairs, pp, btab, out, scale, s);
}

// Schwarz bounds exist for the diagonal classes (ab|ab) only.
template <int a, int b, int c, int d, int wave>
template <int C>
hipError_t ClassLaunch<a, b, c, d, wave>::schwarz_c(int bm, const int* plist, int n, const ShellPair* pairs,
                                                    const PrimPair* pp, const double* btab, double* Q, double scale,
                                                    hipStream_t s)
{
    if constexpr (a != c || b != d) {
        return hipErrorInvalidValue;
    } else if constexpr (wave) {
        if (bm == BOYS_ACCURATE)
            wave_go<BOYS_ACCURATE, WAVE_SCHWARZ, C>(plist, n, pairs, pp, btab, Q, scale, s);
        else
            wave_go<BOYS_REFERENCE, WAVE_SCHWARZ, C>(plist, n, pairs, pp, btab, Q, scale, s);
        return hipGetLastError();
    } else {
        const dim3 blk(256), grd((n + 255) / 256);
        if (bm == BOYS_ACCURATE)
            hipLaunchKernelGGL((k_schwarz<a, b, BOYS_ACCURATE, C>), grd, blk, 0, s, plist, n, pairs, pp, btab, Q,
                               scale);
        else
            hipLaunchKernelGGL((k_schwarz<a, b, BOYS_REFERENCE, C>), grd, blk, 0, s, plist, n, pairs, pp, btab, Q,
                               scale);
        return hipGetLastError();
    }
}

template <int a, int b, int c, int d, int wave>
hipError_t ClassLaunch<a, b, c, d, wave>::schwarz(int bm, int cart, const int* plist, int n, const ShellPair* pairs,
                                                  const PrimPair* pp, const double* btab, double* Q, double scale,
                                                  hipStream_t s)
{
    if constexpr (kHasD) {
        if (cart) return schwarz_c<1>(bm, plist, n, pairs, pp, btab, Q, scale, s);
    }
    return schwarz_c<0>(bm, plist, n, pairs, pp, btab, Q, scale, s);
}

// This translation unit's copy of the AO transform (each unit has its own
// __constant__ memory).
template <int a, int b, int c, int d, int wave>
hipError_t ClassLaunch<a, b, c, d, wave>::upload_xform(const double* x, int n, hipStream_t s)
{
    return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_ao_xform), x, sizeof(double) * n, 0, hipMemcpyHostToDevice, s);
}

template <int a, int b, int c, int d, int wave>
int ClassLaunch<a, b, c, d, wave>::wave_group(int cart)
{
    if constexpr (wave) {
        if constexpr (kHasD) {
            if (cart) return WaveGroup<a, b, c, d, 1>::G;
        }
        return WaveGroup<a, b, c, d, 0>::G;
    } else {
        return 1;
    }
}
