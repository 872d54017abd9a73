// grad_kernels.hip -- ERI part of the RHF nuclear gradient on gfx950 (the
// role of scf::addExactTwoElectronGradient, src/chem/scf/ExactCoulomb.cpp:
// 125-195, over flo::electronRepulsionGradient, GaussianIntegral.cpp:
// 1052-1459): s and p classes one lane per quartet (k_grad), classes with a
// d shell wave-cooperative (k_gradw).
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

__device__ __forceinline__ double pick3g(int d, double x, double y, double z) { return d == 0 ? x : (d == 1 ? y : z); }

__device__ __forceinline__ int wave_max_int(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

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

// ------------------------------------------------ classes with d shells --
// One 64-lane wavefront per block works on G shell quartets at a time, LG =
// 64/G lanes per quartet, driven by the class's GradWaveTab (gen_grad.py).
// Per quartet: the AO pair density Gamma (weight g as above) is moved to the
// Cartesian block Gc in LDS (four staged index transforms).  Per primitive
// quartet the group evaluates the VRR table level by level in LDS (each lane
// every LG-th entry) and adds the targets into per-lane register slots of
// four accumulator kinds: weight 1, 2 alpha, 2 beta, 2 gamma (alpha, beta:
// the bra primitives' exponents, gamma the ket's first).  Then, per
// derivative set ((a+1x b|, (a-1x b|, (a b+1x|, (a b-1x|, |c+1x d), |c-1x d)),
// the group writes its kind's slots to LDS, runs the set's HRR level by
// level and contracts the Cartesian outputs with Gc into the nine
// derivatives of A, B and C; D is minus their sum (translational
// invariance).  Quartets: one block per work unit (bra fixed, its kets) or,
// for the flat lists of the wave-cooperative J/K classes, one block per
// kGradFlatChunk quartets.  Each block adds into its LDS copy of the
// gradient (dynamic LDS, 3 natom doubles) and flushes it once.
constexpr int kGradFlatChunk = 64;

template <int la, int lb, int lc, int ld, int BM, int C>
__global__ void __launch_bounds__(64) k_gradw(const JKSeg* __restrict__ segs, const int* __restrict__ seg_blk0,
                                              int nseg, int flat, const double* __restrict__ btab,
                                              const double* __restrict__ xf, const int* __restrict__ ao_atom,
                                              double* __restrict__ grad, int natom, double scale)
{
    typedef GradWaveTab<la, lb, lc, ld> W;
    constexpr int G = W::G, LG = W::LG, L1 = W::L1, SLOTS = W::SLOTS;
    constexpr int ncA = (la + 1) * (la + 2) / 2, ncB = (lb + 1) * (lb + 2) / 2;
    constexpr int ncC = (lc + 1) * (lc + 2) / 2, ncD = (ld + 1) * (ld + 2) / 2;
    constexpr int NA = (C && la >= 2) ? ncA : 2 * la + 1, NB = (C && lb >= 2) ? ncB : 2 * lb + 1;
    constexpr int NC = (C && lc >= 2) ? ncC : 2 * lc + 1, ND = (C && ld >= 2) ? ncD : 2 * ld + 1;
    constexpr int NAO = NA * NB * NC * ND;
    // accurate mode: the interval table of m = L1 straight from global memory
    // (prim_setup's LDS-table form, bra records as P - A)
    constexpr bool kInt = BM == BOYS_ACCURATE;
    constexpr int BMq = kInt ? BOYS_ACCURATE_LDS : BM;
    const double* __restrict__ tq = kInt ? btab + DSSCF_BPOLY_OFFSET + L1 * DSSCF_BPOLY_M_STRIDE : btab;
    __shared__ double smem[G * W::LDS];
    extern __shared__ double gsh[];  // [3 * natom]
    const int grp = (int)threadIdx.x / LG, lg = (int)threadIdx.x % LG;
    double* __restrict__ Gc = smem + grp * W::LDS;  // Cartesian pair density of the group's quartet
    double* __restrict__ R = Gc + W::NC4;           // density staging, then the VRR table, then each set's HRR
    const int* __restrict__ vrr = W::vrr();
    const int* __restrict__ vlev = W::vlev();
    const short* __restrict__ tgt = W::tgt();
    for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x) gsh[i] = 0.0;

    const JKSeg& S = segs[find_segment_by_block(seg_blk0, nseg, (int)blockIdx.x)];
    const int b = (int)blockIdx.x - S.blk0;
    int q0, q1, ubra = 0;
    if (flat) {
        q0 = b * kGradFlatChunk;
        q1 = min(S.n, q0 + kGradFlatChunk);
    } else {
        const int4 u = ldg(S.u + b);
        ubra = u.x;
        q0 = u.y;
        q1 = u.y + u.z;
    }
    const GPtr<double> D = gbl(S.Da);
    const int N = S.N;
    for (int qb = q0; qb < q1; qb += G) {
        const int q = qb + grp;
        const bool valid = q < q1;
        int bi = ubra, ki = 0;
        if (flat) {
            const uint2 e = S.q[valid ? q : q0];
            bi = (int)e.x;
            ki = (int)e.y;
        } else {
            ki = (int)S.k[valid ? q : q0];
        }
        const ShellPair bra = ldg(S.pairs + bi);
        const ShellPair ket = ldg(S.pairs + ki);
        const int a0 = bra.ao_a, b0 = bra.ao_b, c0 = ket.ao_a, d0 = ket.ao_b;
        const bool same = c0 == a0 && d0 == b0;
        const double gq = valid ? scale * (a0 == b0 ? 0.5 : 1.0) * (c0 == d0 ? 0.5 : 1.0) * (same ? 0.5 : 1.0) : 0.0;
        __syncthreads();  // the previous quartet is done with Gc and R
        // AO pair density into R[0 .. NAO)
        for (int o = lg; o < NAO; o += LG) {
            const int l = o % ND, k = (o / ND) % NC, j = (o / (ND * NC)) % NB, i = o / (ND * NC * NB);
            const size_t a = a0 + i, bb = b0 + j, c = c0 + k, d = d0 + l;
            R[o] = gq * (16.0 * D[a * N + bb] * D[c * N + d] - 4.0 * (D[a * N + c] * D[bb * N + d] +
                                                                    D[a * N + d] * D[bb * N + c]));
        }
        __syncthreads();
        // AO -> Cartesian, one index at a time (d, c, b, a): T[..][x] = sum_m xf[m][x] T[..][m]
        {
            constexpr int H = W::NC4;  // second staging buffer
            const double* __restrict__ xd = xf + grad_xform_off(ld, C);
            for (int o = lg; o < NA * NB * NC * ncD; o += LG) {
                const int x = o % ncD, r = o / ncD;
                double s = 0.0;
                for (int m = 0; m < ND; ++m) s += xd[m * ncD + x] * R[r * ND + m];
                R[H + o] = s;
            }
            __syncthreads();
            const double* __restrict__ xc = xf + grad_xform_off(lc, C);
            for (int o = lg; o < NA * NB * ncC * ncD; o += LG) {
                const int y = o % ncD, x = (o / ncD) % ncC, r = o / (ncD * ncC);
                double s = 0.0;
                for (int m = 0; m < NC; ++m) s += xc[m * ncC + x] * R[H + (r * NC + m) * ncD + y];
                R[o] = s;
            }
            __syncthreads();
            const double* __restrict__ xb = xf + grad_xform_off(lb, C);
            for (int o = lg; o < NA * ncB * ncC * ncD; o += LG) {
                const int y = o % (ncC * ncD), x = (o / (ncC * ncD)) % ncB, r = o / (ncC * ncD * ncB);
                double s = 0.0;
                for (int m = 0; m < NB; ++m) s += xb[m * ncB + x] * R[(r * NB + m) * ncC * ncD + y];
                R[H + o] = s;
            }
            __syncthreads();
            const double* __restrict__ xa = xf + grad_xform_off(la, C);
            for (int o = lg; o < W::NC4; o += LG) {
                const int y = o % (ncB * ncC * ncD), x = o / (ncB * ncC * ncD);
                double s = 0.0;
                for (int m = 0; m < NA; ++m) s += xa[m * ncA + x] * R[H + m * ncB * ncC * ncD + y];
                Gc[o] = s;
            }
            __syncthreads();
        }
        // primitive quartets, flattened; the loop runs to the wavefront's largest count
        double acc[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) acc[s] = 0.0;
        const int npq = valid ? bra.nprim * ket.nprim : 0;
        const int npq_max = wave_max_int(npq);
        int ib = 0, ik = 0;
        for (int pq = 0; pq < npq_max; ++pq) {
            const bool active = pq < npq;
            double wA = 0.0, wB = 0.0, wC = 0.0;
            PrimQuartet<double> g;
            if (active) {
                PrimPair P = ldg(S.pp + bra.prim_off + ib);
                const PrimPair Q = ldg(S.pp + ket.prim_off + ik);
                if (++ik == ket.nprim) {
                    ik = 0;
                    ++ib;
                }
                wA = 2.0 * P.alpha;
                wB = 2.0 * (P.p - P.alpha);
                wC = 2.0 * Q.alpha;
                if (kInt)
                    for (int k = 0; k < 3; ++k) P.P[k] -= bra.A[k];
                double Fm[L1 + 1];
                prim_setup<BMq, double, L1>(P, Q, bra, ket, tq, g, Fm);
                if (lg == 0) {
                    R[W::ZERO] = 0.0;
#pragma unroll
                    for (int m = 0; m <= L1; ++m) {
                        const int bm = W::base()[m];
                        if (bm >= 0) R[bm] = Fm[m];
                    }
                }
            }
            __syncthreads();
            for (int lev = 0; lev < W::NLEV; ++lev) {
                if (active) {
                    const int e1 = vlev[lev + 1];
                    for (int k = vlev[lev] + lg; k < e1; k += LG) {
                        const int r0 = vrr[3 * k], r1 = vrr[3 * k + 1], r2 = vrr[3 * k + 2];
                        const int s0 = r0 & 0xffff, s1 = (r0 >> 16) & 0xffff;
                        const int s2 = r1 & 0xffff, s3 = (r1 >> 16) & 0xffff, s4 = r2 & 0xffff;
                        const int kind = (r2 >> 16) & 3, dir = (r2 >> 18) & 3;
                        const double n1 = (double)((r2 >> 20) & 15), n2 = (double)((r2 >> 24) & 15);
                        double A, Wc, o2, rr;
                        if (kind == 1) {
                            A = pick3g(dir, g.PA[0], g.PA[1], g.PA[2]);
                            Wc = pick3g(dir, g.WP[0], g.WP[1], g.WP[2]);
                            o2 = g.o2p;
                            rr = g.rp;
                        } else {
                            A = pick3g(dir, g.QC[0], g.QC[1], g.QC[2]);
                            Wc = pick3g(dir, g.WQ[0], g.WQ[1], g.WQ[2]);
                            o2 = g.o2q;
                            rr = g.rq;
                        }
                        double v = A * R[s0] + Wc * R[s1];
                        v += n1 * o2 * (R[s2] - rr * R[s3]);
                        v += n2 * g.o2pq * R[s4];
                        R[W::NBASE + k] = v;
                    }
                }
                __syncthreads();
            }
            if (active) {
#pragma unroll
                for (int s = 0; s < SLOTS; ++s) {
                    const double w = s < W::KS1 ? 1.0 : (s < W::KS2 ? wA : (s < W::KS3 ? wB : wC));
                    acc[s] = fma(w, R[tgt[lg + LG * s]], acc[s]);
                }
            }
            __syncthreads();
        }
        // derivative sets: HRR of the set's kind, contraction with Gc
        double gs[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) gs[c] = 0.0;
        for (int st = 0; st < W::NSET; ++st) {
            const int kind = W::set()[4 * st], nk = W::set()[4 * st + 1];
            const int hl0 = W::set()[4 * st + 2], hl1 = W::set()[4 * st + 3];
            const int k0 = W::ks(kind), k1 = W::ks(kind + 1);
#pragma unroll
            for (int s = 0; s < SLOTS; ++s)
                if (s >= k0 && s < k1) R[(s - k0) * LG + lg] = acc[s];
            __syncthreads();
            const int hbase = hl0 < hl1 ? W::hlev()[2 * hl0] : 0;
            for (int hl = hl0; hl < hl1; ++hl) {
                const int e1 = W::hlev()[2 * hl + 1];
                for (int k = W::hlev()[2 * hl] + lg; k < e1; k += LG) {
                    const int h = W::hrr()[k];
                    const int d = W::hdir()[k];
                    const double c = (d >> 2) ? pick3g(d & 3, ket.AB[0], ket.AB[1], ket.AB[2])
                                              : pick3g(d & 3, bra.AB[0], bra.AB[1], bra.AB[2]);
                    R[nk + (k - hbase)] = R[h & 0xffff] + c * R[(h >> 16) & 0xffff];
                }
                __syncthreads();
            }
#pragma unroll
            for (int comp = 0; comp < 9; ++comp) {
                const int e1 = W::cstart()[10 * st + comp + 1];
                for (int e = W::cstart()[10 * st + comp] + lg; e < e1; e += LG)
                    gs[comp] += (double)W::cec()[e] * R[W::cey()[e]] * Gc[W::ceg()[e]];
            }
            __syncthreads();
        }
        // sum over the group's lanes; its first lane adds the four centres
#pragma unroll
        for (int c = 0; c < 9; ++c)
#pragma unroll
            for (int off = LG / 2; off > 0; off >>= 1) gs[c] += __shfl_xor(gs[c], off);
        if (valid && lg == 0) {
            const int atA = ao_atom[a0], atB = ao_atom[b0], atC = ao_atom[c0], atD = ao_atom[d0];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                atomicAdd(&gsh[3 * atA + x], gs[x]);
                atomicAdd(&gsh[3 * atB + x], gs[3 + x]);
                atomicAdd(&gsh[3 * atC + x], gs[6 + x]);
                atomicAdd(&gsh[3 * atD + x], -(gs[x] + gs[3 + x] + gs[6 + x]));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x) {
        const double x = gsh[i];
        if (x != 0.0) unsafeAtomicAdd(&grad[i], x);
    }
}

template <int la, int lb, int lc, int ld, int BM, int C>
static hipError_t launch_gradw(int flat, const JKSeg* segs, const int* blk0, int nseg, int nblocks,
                               const double* btab, const double* xform, const int* ao_atom, double* grad, int natom,
                               double scale, hipStream_t s)
{
    const size_t dyn = sizeof(double) * 3 * (size_t)natom;
    const size_t stat = sizeof(double) * (size_t)GradWaveTab<la, lb, lc, ld>::G * GradWaveTab<la, lb, lc, ld>::LDS;
    if (stat + dyn > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_gradw<la, lb, lc, ld, BM, C>), dim3(nblocks), dim3(64), dyn, s, segs, blk0, nseg, flat, btab,
                       xform, ao_atom, grad, natom, scale);
    return hipGetLastError();
}

int dsscf_grad_flat_chunk() { return kGradFlatChunk; }

int dsscf_grad_is_wave(const int* l4)
{
#define DSSCF_GRADW_IS(a, b, c, d, name) \
    if (l4[0] == a && l4[1] == b && l4[2] == c && l4[3] == d) return 1;
    DSSCF_FOR_EACH_GRAD_WAVE_CLASS(DSSCF_GRADW_IS)
#undef DSSCF_GRADW_IS
    return 0;
}

hipError_t dsscf_launch_gradw(const int* l4, int boys_mode, int cart, int flat, const JKSeg* segs, const int* blk0,
                              int nseg, int nblocks, const double* btab, const double* xform, const int* ao_atom,
                              double* grad, int natom, double scale, hipStream_t s)
{
    if (nblocks <= 0) return hipSuccess;
#define DSSCF_GRADW_GO(a, b, c, d, name)                                                                          \
    if (l4[0] == a && l4[1] == b && l4[2] == c && l4[3] == d) {                                                   \
        if (boys_mode == BOYS_ACCURATE)                                                                           \
            return cart ? launch_gradw<a, b, c, d, BOYS_ACCURATE, 1>(flat, segs, blk0, nseg, nblocks, btab, xform, \
                                                                      ao_atom, grad, natom, scale, s)             \
                        : launch_gradw<a, b, c, d, BOYS_ACCURATE, 0>(flat, segs, blk0, nseg, nblocks, btab, xform, \
                                                                      ao_atom, grad, natom, scale, s);            \
        return cart ? launch_gradw<a, b, c, d, BOYS_REFERENCE, 1>(flat, segs, blk0, nseg, nblocks, btab, xform,    \
                                                                   ao_atom, grad, natom, scale, s)                \
                    : launch_gradw<a, b, c, d, BOYS_REFERENCE, 0>(flat, segs, blk0, nseg, nblocks, btab, xform,    \
                                                                   ao_atom, grad, natom, scale, s);               \
    }
    DSSCF_FOR_EACH_GRAD_WAVE_CLASS(DSSCF_GRADW_GO)
#undef DSSCF_GRADW_GO
    return hipErrorInvalidValue;
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
