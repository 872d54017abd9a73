// wave_kernel.hpp -- wave-cooperative shell-quartet kernel for the large
// classes (included by eri_kernels.hip after the generated tables).
//
// One 64-lane wavefront works on G shell quartets at once, 64/G lanes (a
// group) per quartet, each group with its own LDS table.  Per primitive
// quartet a group walks the generated VRR table level by level, each lane
// evaluating every (64/G)-th entry of a level into LDS (entries of one level
// only read lower levels); the lanes then add the [e0|f0] targets into
// per-lane register accumulators.  The groups run the same level structure
// in step (the primitive loop runs to the largest count in the wavefront,
// groups with fewer primitive quartets idle), so one barrier serves them all.
// After the primitive loops the accumulators go back to LDS, the bra and ket
// horizontal recursions run level by level, the Cartesian block is
// transformed to the reference's AO functions, and the result is digested
// into J/K (each output element owned by one lane: one atomic per element and
// quartet), written out, or reduced to a Schwarz bound.
#pragma once

enum { WAVE_JK = 0, WAVE_BLOCK = 1, WAVE_SCHWARZ = 2 };

__device__ __forceinline__ double pick3(int d, double x, double y, double z) { return d == 0 ? x : (d == 1 ? y : z); }

// Quartets per wavefront: the largest power of two up to 8 whose G LDS
// tables fit in DSSCF_WAVE_LDS_BUDGET bytes.
#ifndef DSSCF_WAVE_LDS_BUDGET
#define DSSCF_WAVE_LDS_BUDGET 40960
#endif
template <int la, int lb, int lc, int ld, int C = 0>
struct WaveGroup {
    static constexpr int bytes = (int)sizeof(double) * WaveTab<la, lb, lc, ld, C>::LDS;
    static constexpr int G = 8 * bytes <= DSSCF_WAVE_LDS_BUDGET   ? 8
                             : 4 * bytes <= DSSCF_WAVE_LDS_BUDGET ? 4
                             : 2 * bytes <= DSSCF_WAVE_LDS_BUDGET ? 2
                                                                  : 1;
};

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// MODE WAVE_JK: block b covers quartets G*(b - blk0) .. +G-1 of the segment
// its index falls in (items, n, pairs, pp unused).  WAVE_BLOCK / WAVE_SCHWARZ:
// items over one context, grid-stride over groups of G.
template <int la, int lb, int lc, int ld, int BM, int MODE, int C>
__global__ void __launch_bounds__(64) k_wave(const void* __restrict__ items, int n,
                                             const ShellPair* __restrict__ pairs_in,
                                             const PrimPair* __restrict__ pp_in, const double* __restrict__ btab,
                                             const JKSeg* __restrict__ segs, const int* __restrict__ seg_blk0,
                                             int nseg, double* __restrict__ out, double scale)
{
    typedef WaveTab<la, lb, lc, ld, C> W;
    typedef ClassInfo<la, lb, lc, ld, C> CI;
    constexpr int G = WaveGroup<la, lb, lc, ld, C>::G;
    constexpr int LG = 64 / G;
    constexpr int L = W::L;
    constexpr int SLOTS = (W::NT + LG - 1) / LG;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND;
    constexpr int ncA = (la + 1) * (la + 2) / 2, ncB = (lb + 1) * (lb + 2) / 2;
    constexpr int ncC = (lc + 1) * (lc + 2) / 2, ncD = (ld + 1) * (ld + 2) / 2;
    __shared__ double smem[G * W::LDS];
    const int grp = (int)threadIdx.x / LG, lg = (int)threadIdx.x % LG;
    double* __restrict__ sm = smem + grp * W::LDS;
    const int* __restrict__ vrr = W::vrr();
    const int* __restrict__ vlev = W::vlev();
    const short* __restrict__ tgt = W::tgt();
    const int* __restrict__ hrr = W::hrr();
    const char* __restrict__ hdir = W::hdir();
    const int* __restrict__ hlev = W::hlev();
    const short* __restrict__ carty = W::cart();

    const JKSeg* S = nullptr;
    const ShellPair* __restrict__ pairs = pairs_in;
    const PrimPair* __restrict__ pp = pp_in;
    int b0 = blockIdx.x, bstep = gridDim.x, bend = (n + G - 1) / G;
    if (MODE == WAVE_JK) {
        S = &segs[find_segment_by_block(seg_blk0, nseg, (int)blockIdx.x)];
        pairs = S->pairs;
        pp = S->pp;
        n = S->n;
        b0 = (int)blockIdx.x - S->blk0;
        bend = b0 + 1;  // one group of G quartets per block
        bstep = 1;
    }
    for (int bb = b0; bb < bend; bb += bstep) {  // uniform over the block
        const int t = bb * G + grp;
        const bool valid = t < n;
        int bi = 0, ki = 0;
        if (valid) {
            if (MODE == WAVE_JK) {
                const uint2 q = S->q[t];
                bi = (int)q.x;
                ki = (int)q.y;
            } else if (MODE == WAVE_BLOCK) {
                const int4 it = ((const int4*)items)[t];
                bi = it.x;
                ki = it.y;
            } else {
                bi = ki = ((const int*)items)[t];
            }
        }
        const ShellPair bra = pairs[bi];
        const ShellPair ket = pairs[ki];
        double acc[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) acc[s] = 0.0;

        // primitive quartets, flattened; the loop runs to the wavefront's largest count
        const int npq = valid ? bra.nprim * ket.nprim : 0;
        const int npq_max = wave_max(npq);
        int ib = 0, ik = 0;
        for (int pq = 0; pq < npq_max; ++pq) {
            const bool active = pq < npq;
            if (active) {
                const PrimPair P = pp[bra.prim_off + ib];
                const PrimPair Q = pp[ket.prim_off + ik];
                if (++ik == ket.nprim) {
                    ik = 0;
                    ++ib;
                }
                PrimQuartet<double> g;
                double Fm[L + 1];
                prim_setup<BM, double, L>(P, Q, bra, ket, btab, g, Fm);
                if (lg == 0) {
                    sm[W::ZERO] = 0.0;
#pragma unroll
                    for (int m = 0; m <= L; ++m) {
                        const int b = W::base()[m];
                        if (b >= 0) sm[b] = Fm[m];
                    }
                }
                __syncthreads();
                for (int lev = 0; lev < W::NLEV; ++lev) {
                    const int e1 = vlev[lev + 1];
                    for (int k = vlev[lev] + lg; k < e1; k += LG) {
                        const int r0 = vrr[3 * k], r1 = vrr[3 * k + 1], r2 = vrr[3 * k + 2];
                        const int s0 = r0 & 0xffff, s1 = (r0 >> 16) & 0xffff;
                        const int s2 = r1 & 0xffff, s3 = (r1 >> 16) & 0xffff, s4 = r2 & 0xffff;
                        const int kind = (r2 >> 16) & 3, dir = (r2 >> 18) & 3;
                        const double n1 = (double)((r2 >> 20) & 15), n2 = (double)((r2 >> 24) & 15);
                        double A, Wc, o2, rr;
                        if (kind == 1) {
                            A = pick3(dir, g.PA[0], g.PA[1], g.PA[2]);
                            Wc = pick3(dir, g.WP[0], g.WP[1], g.WP[2]);
                            o2 = g.o2p;
                            rr = g.rp;
                        } else {
                            A = pick3(dir, g.QC[0], g.QC[1], g.QC[2]);
                            Wc = pick3(dir, g.WQ[0], g.WQ[1], g.WQ[2]);
                            o2 = g.o2q;
                            rr = g.rq;
                        }
                        double v = A * sm[s0] + Wc * sm[s1];
                        v += n1 * o2 * (sm[s2] - rr * sm[s3]);
                        v += n2 * g.o2pq * sm[s4];
                        sm[W::NBASE + k] = v;
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int s = 0; s < SLOTS; ++s) {
                    const int tt = lg + LG * s;
                    if (tt < W::NT) acc[s] += sm[tgt[tt]];
                }
                __syncthreads();  // (a block is one wavefront: idle groups need no barrier of their own)
            }
        }
        // horizontal recursions over Y = [targets | bra HRR | ket HRR]
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            const int tt = lg + LG * s;
            if (tt < W::NT) sm[tt] = acc[s];
        }
        __syncthreads();
        for (int hl = 0; hl < W::NHLEV; ++hl) {
            const int e1 = hlev[2 * hl + 1];
            for (int k = hlev[2 * hl] + lg; k < e1; k += LG) {
                const int h = hrr[k];
                const int d = hdir[k];
                const double c = (d >> 2) ? pick3(d & 3, ket.AB[0], ket.AB[1], ket.AB[2])
                                          : pick3(d & 3, bra.AB[0], bra.AB[1], bra.AB[2]);
                sm[W::NT + k] = sm[h & 0xffff] + c * sm[(h >> 16) & 0xffff];
            }
            __syncthreads();
        }
        // Cartesian -> AO, one index at a time (d, c, b, a): each stage output
        // sums the at most three Cartesian components of one AO.  The AO index
        // m is the outer, wavefront-uniform loop, so its terms (weight and
        // Cartesian position) come through scalar loads; the lanes split the
        // remaining indices.  Stages 1 and 3 write after the HRR outputs,
        // stages 2 and 4 from the table's start (the HRR outputs are dead once
        // stage 1 has read them).
        {
            double* __restrict__ R1 = sm + W::NY;
            double* __restrict__ R2 = sm;
            for (int md = 0; md < ND; ++md) {  // T1[ca][cb][cc][md]
                const int pat = ao_pat_row(ld, md, C), nt = ao_pat_n[pat];
                const int c0 = ao_pat_c[3 * pat], c1 = ao_pat_c[3 * pat + 1], c2 = ao_pat_c[3 * pat + 2];
                const double w0 = c_ao_xform[xform_off(ld, C) + md * ncD + c0];
                const double w1 = nt > 1 ? c_ao_xform[xform_off(ld, C) + md * ncD + c1] : 0.0;
                const double w2 = nt > 2 ? c_ao_xform[xform_off(ld, C) + md * ncD + c2] : 0.0;
                for (int r = lg; r < ncA * ncB * ncC; r += LG) {
                    const short* cy = carty + r * ncD;
                    double sum = w0 * sm[cy[c0]];
                    if (nt > 1) sum += w1 * sm[cy[c1]];
                    if (nt > 2) sum += w2 * sm[cy[c2]];
                    R1[r * ND + md] = sum;
                }
            }
            __syncthreads();
            for (int mc = 0; mc < NC; ++mc) {  // T2[ca][cb][mc][md]
                const int pat = ao_pat_row(lc, mc, C), nt = ao_pat_n[pat];
                const int c0 = ao_pat_c[3 * pat], c1 = ao_pat_c[3 * pat + 1], c2 = ao_pat_c[3 * pat + 2];
                const double w0 = c_ao_xform[xform_off(lc, C) + mc * ncC + c0];
                const double w1 = nt > 1 ? c_ao_xform[xform_off(lc, C) + mc * ncC + c1] : 0.0;
                const double w2 = nt > 2 ? c_ao_xform[xform_off(lc, C) + mc * ncC + c2] : 0.0;
                for (int o = lg; o < ncA * ncB * ND; o += LG) {
                    const int md = o % ND, r = o / ND;
                    const double* in = R1 + r * ncC * ND + md;
                    double sum = w0 * in[c0 * ND];
                    if (nt > 1) sum += w1 * in[c1 * ND];
                    if (nt > 2) sum += w2 * in[c2 * ND];
                    R2[(r * NC + mc) * ND + md] = sum;
                }
            }
            __syncthreads();
            for (int mb = 0; mb < NB; ++mb) {  // T3[ca][mb][mc][md]
                const int pat = ao_pat_row(lb, mb, C), nt = ao_pat_n[pat];
                const int c0 = ao_pat_c[3 * pat], c1 = ao_pat_c[3 * pat + 1], c2 = ao_pat_c[3 * pat + 2];
                const double w0 = c_ao_xform[xform_off(lb, C) + mb * ncB + c0];
                const double w1 = nt > 1 ? c_ao_xform[xform_off(lb, C) + mb * ncB + c1] : 0.0;
                const double w2 = nt > 2 ? c_ao_xform[xform_off(lb, C) + mb * ncB + c2] : 0.0;
                for (int o = lg; o < ncA * NC * ND; o += LG) {
                    const int q = o % (NC * ND), ca = o / (NC * ND);
                    const double* in = R2 + ca * ncB * NC * ND + q;
                    double sum = w0 * in[c0 * NC * ND];
                    if (nt > 1) sum += w1 * in[c1 * NC * ND];
                    if (nt > 2) sum += w2 * in[c2 * NC * ND];
                    R1[(ca * NB + mb) * NC * ND + q] = sum;
                }
            }
            __syncthreads();
            for (int ma = 0; ma < NA; ++ma) {  // V[ma][mb][mc][md]
                const int pat = ao_pat_row(la, ma, C), nt = ao_pat_n[pat];
                const int c0 = ao_pat_c[3 * pat], c1 = ao_pat_c[3 * pat + 1], c2 = ao_pat_c[3 * pat + 2];
                const double w0 = c_ao_xform[xform_off(la, C) + ma * ncA + c0];
                const double w1 = nt > 1 ? c_ao_xform[xform_off(la, C) + ma * ncA + c1] : 0.0;
                const double w2 = nt > 2 ? c_ao_xform[xform_off(la, C) + ma * ncA + c2] : 0.0;
                for (int q = lg; q < NB * NC * ND; q += LG) {
                    const double* in = R1 + q;
                    double sum = w0 * in[c0 * NB * NC * ND];
                    if (nt > 1) sum += w1 * in[c1 * NB * NC * ND];
                    if (nt > 2) sum += w2 * in[c2 * NB * NC * ND];
                    R2[ma * NB * NC * ND + q] = scale * sum;
                }
            }
        }
        __syncthreads();
        const double* V = sm;
#define VX(i, j, k, l) V[(((i) * NB + (j)) * NC + (k)) * ND + (l)]
        if (!valid) {
        } else if (MODE == WAVE_BLOCK) {
            const int4 it = ((const int4*)items)[t];
            for (int o = lg; o < W::NAO; o += LG) out[(size_t)it.z + o] = V[o];
        } else if (MODE == WAVE_SCHWARZ) {
            if (lg == 0) {
                double mx = 0.0;
                for (int i = 0; i < NA; ++i)
                    for (int j = 0; j < NB; ++j) mx = fmax(mx, fabs(VX(i, j, i, j)));
                out[bi] = sqrt(mx);
            }
        } else {
            const double* __restrict__ Dt = S->Dt;
            const int N = S->N;
            const size_t slice = (size_t)jk_slice(S->nslice) * N * N;
            double* __restrict__ Jh = S->Jh + slice;
            double g = 1.0;
            if (bra.ao_a == bra.ao_b) g *= 0.5;
            if (ket.ao_a == ket.ao_b) g *= 0.5;
            if (bi == ki) g *= 0.5;
            const int a0 = bra.ao_a, b0 = bra.ao_b, c0 = ket.ao_a, d0 = ket.ao_b;
            constexpr int T0 = NA * NB, T1 = T0 + NC * ND, T2 = T1 + NA * NC, T3 = T2 + NB * NC, T4 = T3 + NA * ND,
                          T5 = T4 + NB * ND;
            // an unrestricted segment digests the exchange terms twice: K[Da]
            // (outputs T1..T5), then K[Db] into Kbh (T5..T5+T5-T1)
            const int Tend = S->Db ? T5 + (T5 - T1) : T5;
            for (int o0 = lg; o0 < Tend; o0 += LG) {
                double s = 0.0;
                const bool bx = o0 >= T5;
                const int o = bx ? o0 - (T5 - T1) : o0;
                const double* __restrict__ Dx = bx ? S->Db : S->Da;
                double* __restrict__ Kh = (bx ? S->Kbh : S->Kh) + slice;
                if (o < T0) {  // Jh[a][b] += 2g sum_cd v Dt_cd
                    const int i = o / NB, j = o % NB;
                    for (int k = 0; k < NC; ++k)
                        for (int l = 0; l < ND; ++l) s += VX(i, j, k, l) * Dt[(size_t)(c0 + k) * N + d0 + l];
                    unsafeAtomicAdd(&Jh[(size_t)(a0 + i) * N + b0 + j], 2.0 * g * s);
                } else if (o < T1) {  // Jh[c][d] += 2g sum_ab v Dt_ab
                    const int k = (o - T0) / ND, l = (o - T0) % ND;
                    for (int i = 0; i < NA; ++i)
                        for (int j = 0; j < NB; ++j) s += VX(i, j, k, l) * Dt[(size_t)(a0 + i) * N + b0 + j];
                    unsafeAtomicAdd(&Jh[(size_t)(c0 + k) * N + d0 + l], 2.0 * g * s);
                } else if (o < T2) {  // Kh[a][c] += g sum_bd v D_bd
                    const int i = (o - T1) / NC, k = (o - T1) % NC;
                    for (int j = 0; j < NB; ++j)
                        for (int l = 0; l < ND; ++l) s += VX(i, j, k, l) * Dx[(size_t)(b0 + j) * N + d0 + l];
                    unsafeAtomicAdd(&Kh[(size_t)(a0 + i) * N + c0 + k], g * s);
                } else if (o < T3) {  // Kh[b][c] += g sum_ad v D_ad
                    const int j = (o - T2) / NC, k = (o - T2) % NC;
                    for (int i = 0; i < NA; ++i)
                        for (int l = 0; l < ND; ++l) s += VX(i, j, k, l) * Dx[(size_t)(a0 + i) * N + d0 + l];
                    unsafeAtomicAdd(&Kh[(size_t)(b0 + j) * N + c0 + k], g * s);
                } else if (o < T4) {  // Kh[a][d] += g sum_bc v D_bc
                    const int i = (o - T3) / ND, l = (o - T3) % ND;
                    for (int j = 0; j < NB; ++j)
                        for (int k = 0; k < NC; ++k) s += VX(i, j, k, l) * Dx[(size_t)(b0 + j) * N + c0 + k];
                    unsafeAtomicAdd(&Kh[(size_t)(a0 + i) * N + d0 + l], g * s);
                } else {  // Kh[b][d] += g sum_ac v D_ac
                    const int j = (o - T4) / ND, l = (o - T4) % ND;
                    for (int i = 0; i < NA; ++i)
                        for (int k = 0; k < NC; ++k) s += VX(i, j, k, l) * Dx[(size_t)(a0 + i) * N + c0 + k];
                    unsafeAtomicAdd(&Kh[(size_t)(b0 + j) * N + d0 + l], g * s);
                }
            }
        }
#undef VX
        __syncthreads();
    }
}
