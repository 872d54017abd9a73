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
//
// The level walks are unrolled at compile time (the level bounds are
// constants of the generated tables): each lane loads its VRR and HRR table
// entries into registers once per block, and reads an entry's coefficients
// (A, W, 1/2zeta, rho/zeta by the entry's side and direction; A-B or C-D)
// from a small per-group LDS block written once per primitive quartet or
// quartet, so no table read waits on global memory inside the walks.
#pragma once
#include <utility>

enum { WAVE_JK = 0, WAVE_BLOCK = 1, WAVE_SCHWARZ = 2 };

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

// Compile-time loop: f(integral_constant<int, 0>) .. f(integral_constant<int, N-1>).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// A lane's share of the VRR and HRR tables: level l's entries are walked by
// the LG lanes of a group in slots of LG entries; slot j of level l holds
// entry start(l) + lg + LG j when that is below the level's end.  The slots of
// all levels are numbered consecutively, so each lane keeps its table entries
// in registers (loaded once per block) instead of reading them per entry.
template <typename W, int LG>
struct WaveSlots {
    static constexpr int vcnt(int l) { return W::VLEV[l + 1] - W::VLEV[l]; }
    static constexpr int vslots(int l) { return (vcnt(l) + LG - 1) / LG; }
    static constexpr int v0(int l)
    {
        int s = 0;
        for (int i = 0; i < l; ++i) s += vslots(i);
        return s;
    }
    static constexpr int NV = v0(W::NLEV);
    static constexpr int hcnt(int l) { return W::HLEV[2 * l + 1] - W::HLEV[2 * l]; }
    static constexpr int hslots(int l) { return (hcnt(l) + LG - 1) / LG; }
    static constexpr int h0(int l)
    {
        int s = 0;
        for (int i = 0; i < l; ++i) s += hslots(i);
        return s;
    }
    static constexpr int NH = h0(W::NHLEV);
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
// At least two wavefronts per SIMD: the register file split in two leaves
// 256 VGPRs a lane (a few spill), which lets a CU hold as many blocks as its
// LDS allows ((dd|dp): six, 0.86 -> 0.60 ms on C20H42/def2-SVP, where one
// wavefront per SIMD held four).
template <int la, int lb, int lc, int ld, int BM, int MODE, int C>
__global__ void __launch_bounds__(64, 2) k_wave(const void* __restrict__ items, int n,
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
    typedef WaveSlots<W, LG> WS;
    __shared__ double smem[G * W::LDS];
    // per group: the VRR coefficients of the current primitive quartet as
    // (A, W) pairs [bra x, y, z, ket x, y, z] and (1/2zeta, rho/zeta) pairs
    // [bra, ket], then 1/2(p+q); the HRR distances A-B (x, y, z) and C-D
    __shared__ __align__(16) double wcoef[G][24];
    const int grp = (int)threadIdx.x / LG, lg = (int)threadIdx.x % LG;
    double* __restrict__ sm = smem + grp * W::LDS;
    double* __restrict__ cf = wcoef[grp];
    // this lane's VRR entries (three words each) and HRR entries (two 13-bit
    // operand indices, then direction and stage in bits 26..28), all levels;
    // the positions in Y of its [e0|f0] targets; the group's copy of the
    // Cartesian-component map of the AO transform
    int vr0[WS::NV > 0 ? WS::NV : 1], vr1[WS::NV > 0 ? WS::NV : 1], vr2[WS::NV > 0 ? WS::NV : 1];
    int hr[WS::NH > 0 ? WS::NH : 1];
    static_assert(W::NY < 8192, "HRR operand indices take 13 bits");
    int tg[SLOTS];
    __shared__ short scart[W::NCART];
    for (int i = (int)threadIdx.x; i < W::NCART; i += (int)blockDim.x) scart[i] = W::cart()[i];
    const short* __restrict__ carty = scart;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        const int tt = lg + LG * s;
        tg[s] = tt < W::NT ? W::tgt()[tt] : W::ZERO;
    }
    {
        const int* __restrict__ vrr = W::vrr();
        const int* __restrict__ hrr = W::hrr();
        const char* __restrict__ hdir = W::hdir();
        static_for<W::NLEV>([&](auto lev) {
            constexpr int l = decltype(lev)::value;
            static_for<WS::vslots(l)>([&](auto jj) {
                constexpr int j = decltype(jj)::value;
                constexpr int s = WS::v0(l) + j;
                // an empty slot reads the level's first entry (never used)
                const int k = W::VLEV[l] + (lg + LG * j < WS::vcnt(l) ? lg + LG * j : 0);
                vr0[s] = vrr[3 * k];
                vr1[s] = vrr[3 * k + 1];
                vr2[s] = vrr[3 * k + 2];
            });
        });
        static_for<W::NHLEV>([&](auto lev) {
            constexpr int l = decltype(lev)::value;
            static_for<WS::hslots(l)>([&](auto jj) {
                constexpr int j = decltype(jj)::value;
                constexpr int s = WS::h0(l) + j;
                const int k = W::HLEV[2 * l] + (lg + LG * j < WS::hcnt(l) ? lg + LG * j : 0);
                const int h = hrr[k];
                hr[s] = (h & 0x1fff) | (((h >> 16) & 0x1fff) << 13) | ((int)hdir[k] << 26);
            });
        });
    }

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
                        const int b = W::BASE[m];
                        if (b >= 0) sm[b] = Fm[m];
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        cf[2 * d] = g.PA[d];
                        cf[2 * d + 1] = g.WP[d];
                        cf[6 + 2 * d] = g.QC[d];
                        cf[6 + 2 * d + 1] = g.WQ[d];
                    }
                    cf[12] = g.o2p;
                    cf[13] = g.rp;
                    cf[14] = g.o2q;
                    cf[15] = g.rq;
                    cf[16] = g.o2pq;
                }
                __syncthreads();
                // the levels and slots unrolled: per entry its coefficients
                // from the group's LDS block (the entry's kind and direction
                // select them), its operands from the levels below
                const double o2pq = cf[16];
                static_for<W::NLEV>([&](auto lev) {
                    constexpr int l = decltype(lev)::value;
                    static_for<WS::vslots(l)>([&](auto jj) {
                        constexpr int j = decltype(jj)::value;
                        constexpr int s = WS::v0(l) + j;
                        if (lg + LG * j < WS::vcnt(l)) {
                            int r0 = vr0[s], r1 = vr1[s], r2 = vr2[s];
                            asm volatile("" : "+v"(r0), "+v"(r1), "+v"(r2));  // decode per use
                            const int s0 = r0 & 0xffff, s1 = (r0 >> 16) & 0xffff;
                            const int s2 = r1 & 0xffff, s3 = (r1 >> 16) & 0xffff, s4 = r2 & 0xffff;
                            const int ket_side = ((r2 >> 16) & 3) != 1, dir = (r2 >> 18) & 3;
                            const double n1 = (double)((r2 >> 20) & 15), n2 = (double)((r2 >> 24) & 15);
                            const double2 aw = reinterpret_cast<const double2*>(cf)[3 * ket_side + dir];
                            const double2 orr = reinterpret_cast<const double2*>(cf)[6 + ket_side];
                            double v = aw.x * sm[s0] + aw.y * sm[s1];
                            v += n1 * orr.x * (sm[s2] - orr.y * sm[s3]);
                            v += n2 * o2pq * sm[s4];
                            sm[W::NBASE + W::VLEV[l] + lg + LG * j] = v;
                        }
                    });
                    __syncthreads();
                });
#pragma unroll
                for (int s = 0; s < SLOTS; ++s) {
                    const int tt = lg + LG * s;
                    if (tt < W::NT) acc[s] += sm[tg[s]];
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
        if (lg == 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                cf[17 + d] = bra.AB[d];
                cf[20 + d] = ket.AB[d];
            }
        }
        __syncthreads();
        static_for<W::NHLEV>([&](auto lev) {
            constexpr int l = decltype(lev)::value;
            static_for<WS::hslots(l)>([&](auto jj) {
                constexpr int j = decltype(jj)::value;
                constexpr int s = WS::h0(l) + j;
                if (lg + LG * j < WS::hcnt(l)) {
                    int h = hr[s];
                    asm volatile("" : "+v"(h));  // decode here, not hoisted out of the quartet loop
                    const int d = (h >> 26) & 7;
                    // direction d & 3 of A-B (stage 0) or C-D (stage 1)
                    const double c = cf[17 + 3 * (d >> 2) + (d & 3)];
                    sm[W::NT + W::HLEV[2 * l] + lg + LG * j] = sm[h & 0x1fff] + c * sm[(h >> 13) & 0x1fff];
                }
            });
            __syncthreads();
        });
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
