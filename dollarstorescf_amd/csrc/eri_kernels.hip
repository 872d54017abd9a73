// eri_kernels.hip -- the Fock build's per-class dispatch (the class kernels
// live in eri_class.hpp, one generated translation unit per class) and the
// per-context prepare / fold kernels.
#include "device_common.hpp"
#include "kernels.hpp"
#include "launch.hpp"

#include "eri_classes.inc"

// ------------------------------------------------------ helper kernels --
__global__ void k_fill(double* __restrict__ p, size_t n, double v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__device__ __forceinline__ int find_segment_by_row(const MatSeg* __restrict__ segs, int nseg, int r)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].row0 <= r) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Per context and row x: packed-U densities (element (y,x), y<=x, at
// x(x+1)/2+y, Matrix.ipp:332-351) -> full row-major Dt, Da (and Db of an
// unrestricted context); zero Jh, Kh (Kbh) rows.
__global__ void k_prep(const MatSeg* __restrict__ segs, int nseg)
{
    const MatSeg& S = segs[find_segment_by_row(segs, nseg, (int)blockIdx.y)];
    const int x = (int)blockIdx.y - S.row0, N = S.N;
    const size_t NN = (size_t)N * N;
    for (int y = threadIdx.x + blockIdx.x * blockDim.x; y < N; y += gridDim.x * blockDim.x) {
        for (int s = 0; s < S.nslice; ++s) {
            S.Jh[s * NN + (size_t)x * N + y] = 0.0;
            S.Kh[s * NN + (size_t)x * N + y] = 0.0;
        }
        if (S.Kbh)
            for (int s = 0; s < S.nslice; ++s) S.Kbh[s * NN + (size_t)x * N + y] = 0.0;
        if (y <= x) {
            const size_t o = (size_t)x * (x + 1) / 2 + y;
            const double t = S.Ptp[o], a = S.Pap[o];
            S.Dt[(size_t)y * N + x] = t;
            S.Dt[(size_t)x * N + y] = t;
            S.Da[(size_t)y * N + x] = a;
            S.Da[(size_t)x * N + y] = a;
            if (S.Pbp) {
                const double b = S.Pbp[o];
                S.Db[(size_t)y * N + x] = b;
                S.Db[(size_t)x * N + y] = b;
            }
        }
    }
}

// J = sum over slices of Jh + Jh^T, likewise K (and K[Db]) -> packed-U, per context.
__global__ void k_fold(const MatSeg* __restrict__ segs, int nseg)
{
    const MatSeg& S = segs[find_segment_by_row(segs, nseg, (int)blockIdx.y)];
    const int x = (int)blockIdx.y - S.row0, N = S.N;
    const size_t NN = (size_t)N * N;
    for (int y = threadIdx.x + blockIdx.x * blockDim.x; y <= x; y += gridDim.x * blockDim.x) {
        const size_t o = (size_t)x * (x + 1) / 2 + y;
        const size_t yx = (size_t)y * N + x, xy = (size_t)x * N + y;
        double j = 0.0, k = 0.0;
        for (int s = 0; s < S.nslice; ++s) {
            j += S.Jh[s * NN + yx] + S.Jh[s * NN + xy];
            k += S.Kh[s * NN + yx] + S.Kh[s * NN + xy];
        }
        S.Jp[o] = j;
        S.Kp[o] = k;
        if (S.Kbp) {
            double kb = 0.0;
            for (int s = 0; s < S.nslice; ++s) kb += S.Kbh[s * NN + yx] + S.Kbh[s * NN + xy];
            S.Kbp[o] = kb;
        }
    }
}

// J_cd terms of the thread-per-quartet classes, stored per quartet in slot
// records by k_jk and laid out by ket AO block, summed per element: one
// 256-thread block per block (row) reads the row's records as one contiguous
// range, thread t on component t % (nc nd); every element belongs to one
// row and the class kernels have finished, so the add into copy 0 of Jh is a
// plain read-modify-write.
__global__ void __launch_bounds__(256) k_jcd(const JcdSeg* __restrict__ segs, const int* __restrict__ blk0, int nseg)
{
    __shared__ double part[256];
    const JcdSeg& S = segs[find_segment_by_block(blk0, nseg, (int)blockIdx.x)];
    const int r = (int)blockIdx.x - S.blk0;
    const int4 ri = S.ri[r];  // c0, d0, nc, nd
    const int2 rr = S.rr[r];  // first double, records
    const int ncnd = ri.z * ri.w;
    const int per = 256 / ncnd;  // records in flight per pass
    const int tid = (int)threadIdx.x;
    const int comp = tid % ncnd, e0 = tid / ncnd;
    double x = 0.0;
    if (e0 < per) {
        const GPtr<double> js = gbl(S.js) + rr.x + comp;
        // four independent loads in flight per thread (the row sums are
        // bound by load latency, not bandwidth)
        double x1 = 0.0, x2 = 0.0, x3 = 0.0;
        int e = e0;
        for (; e + 3 * per < rr.y; e += 4 * per) {
            x += js[(size_t)e * ncnd];
            x1 += js[(size_t)(e + per) * ncnd];
            x2 += js[(size_t)(e + 2 * per) * ncnd];
            x3 += js[(size_t)(e + 3 * per) * ncnd];
        }
        for (; e < rr.y; e += per) x += js[(size_t)e * ncnd];
        x += (x1 + x2) + x3;
    }
    part[tid] = x;
    __syncthreads();
    if (tid < ncnd) {
        double s = 0.0;
        for (int j = 0; j < per; ++j) s += part[tid + j * ncnd];
        const int k = tid / ri.w, l = tid - k * ri.w;
        S.Jh[(size_t)(ri.x + k) * S.N + ri.y + l] += s;
    }
}

// ------------------------------------------------- one-electron terms ---
// S, T and V over the AOs of one shell pair per thread (the pair's first shell
// has the higher l), the roles of OSOverlap / OSKineticEnergy /
// OSNuclearPotential (GaussianIntegral.cpp:29-64, 98-157, 225-291) with the
// reference's PI.  Per primitive pair: 1-D Obara-Saika overlap tables give S
// and, from the bra exponent, T; V runs the vertical recursion
//   [e]^(m) = PA_d [e-1_d]^(m) - PC_d [e-1_d]^(m+1)
//             + (e_d-1)/2p ([e-2_d]^(m) - [e-2_d]^(m+1)),  [0]^(m) = F_m(p PC^2)
// on the first centre up to la+lb for every nucleus, accumulated over
// primitives and nuclei; the transfer to the second shell,
//   (a|b) = sum_k C(b,k) AB^(b-k) (a+k|0)  per direction,
// and the AO transform follow once per pair.  Each packed-U element belongs
// to one pair: plain stores.
namespace one_e {
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int cart_off(int l) { return l * (l + 1) * (l + 2) / 6; }  // components of all lower l
// x-major position of (x, y, l-x-y) within its shell
__device__ __forceinline__ int cart_pos(int l, int x, int y)
{
    const int r = l - x;  // components with this x come after all with a larger x
    return r * (r + 1) / 2 + (r - y);
}
__device__ __forceinline__ void cart_xyz(int l, int k, int& x, int& y, int& z)
{
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= k) ++r;  // k in the block of x = l - r
    x = l - r;
    y = r - (k - r * (r + 1) / 2);
    z = l - x - y;
}
__device__ __forceinline__ double binom(int n, int k) { return (n == 2 && k == 1) ? 2.0 : 1.0; }  // n <= 2
}  // namespace one_e

template <int la, int lb, int BM>
__global__ void __launch_bounds__(64) k_one_electron(const int* __restrict__ plist, int n,
                                                     const ShellPair* __restrict__ pairs,
                                                     const PrimPair* __restrict__ pp,
                                                     const double* __restrict__ atoms, int natom,
                                                     const double* __restrict__ btab,
                                                     const double* __restrict__ xform, int cart,
                                                     double* __restrict__ S, double* __restrict__ T,
                                                     double* __restrict__ V)
{
    using namespace one_e;
    constexpr double kPi = 3.14159265358979;
    constexpr int L = la + lb, NA = ncart(la), NB = ncart(lb);
    constexpr int NE = cart_off(L + 1) - cart_off(la);  // components with la <= |e| <= L
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const ShellPair sp = pairs[plist[t]];
    const double A[3] = {sp.A[0], sp.A[1], sp.A[2]};
    const double B[3] = {sp.A[0] - sp.AB[0], sp.A[1] - sp.AB[1], sp.A[2] - sp.AB[2]};
    double s[NA * NB], k[NA * NB], ve[NE];
    for (int i = 0; i < NA * NB; ++i) s[i] = k[i] = 0.0;
    for (int i = 0; i < NE; ++i) ve[i] = 0.0;
    double th[L + 1][cart_off(L + 1)];
    for (int ip = 0; ip < sp.nprim; ++ip) {
        const PrimPair P = pp[sp.prim_off + ip];
        const double p = P.p, al = P.alpha, o2p = 0.5 * P.inv_p;
        const double K = P.Kp * p;  // w_a w_b exp(-ab/p |AB|^2)
        // 1-D overlaps S1[x][i][j], i <= la+2, j <= lb
        double S1[3][la + 3][lb + 1];
        for (int x = 0; x < 3; ++x) {
            const double PA = P.P[x] - A[x], PB = P.P[x] - B[x];
            S1[x][0][0] = sqrt(kPi * P.inv_p);
            for (int i = 1; i <= la + 2; ++i)
                S1[x][i][0] = PA * S1[x][i - 1][0] + (i > 1 ? (i - 1) * S1[x][i - 2][0] * o2p : 0.0);
            for (int j = 1; j <= lb; ++j)
                for (int i = 0; i <= la + 2; ++i) {
                    double w = PB * S1[x][i][j - 1];
                    if (i > 0) w += i * S1[x][i - 1][j - 1] * o2p;
                    if (j > 1) w += (j - 1) * S1[x][i][j - 2] * o2p;
                    S1[x][i][j] = w;
                }
        }
        for (int ia = 0; ia < NA; ++ia)
            for (int ib = 0; ib < NB; ++ib) {
                int a[3], b[3];
                cart_xyz(la, ia, a[0], a[1], a[2]);
                cart_xyz(lb, ib, b[0], b[1], b[2]);
                double sd[3], dd[3];
                for (int x = 0; x < 3; ++x) {
                    const int i = a[x], j = b[x];
                    sd[x] = S1[x][i][j];
                    dd[x] = 4.0 * al * al * S1[x][i + 2][j] - 2.0 * al * (2 * i + 1) * S1[x][i][j] +
                            (i > 1 ? (double)(i * (i - 1)) * S1[x][i - 2][j] : 0.0);
                }
                s[ia * NB + ib] += K * sd[0] * sd[1] * sd[2];
                k[ia * NB + ib] += -0.5 * K * (dd[0] * sd[1] * sd[2] + sd[0] * dd[1] * sd[2] + sd[0] * sd[1] * dd[2]);
            }
        // nuclear attraction, every nucleus
        const double PA[3] = {P.P[0] - A[0], P.P[1] - A[1], P.P[2] - A[2]};
        for (int c = 0; c < natom; ++c) {
            const double PC[3] = {P.P[0] - atoms[4 * c], P.P[1] - atoms[4 * c + 1], P.P[2] - atoms[4 * c + 2]};
            const double R2 = PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2];
            double F[L + 1];
            boys<BM, L>(p * R2, btab, F);
            for (int m = 0; m <= L; ++m) th[m][0] = F[m];
            for (int l = 1; l <= L; ++l)
                for (int kk = 0; kk < ncart(l); ++kk) {
                    int e[3];
                    cart_xyz(l, kk, e[0], e[1], e[2]);
                    const int d = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
                    int e1[3] = {e[0], e[1], e[2]};
                    e1[d] -= 1;
                    const int i1 = cart_off(l - 1) + cart_pos(l - 1, e1[0], e1[1]);
                    int i2 = -1;
                    if (e1[d] > 0) {
                        int e2[3] = {e1[0], e1[1], e1[2]};
                        e2[d] -= 1;
                        i2 = cart_off(l - 2) + cart_pos(l - 2, e2[0], e2[1]);
                    }
                    const int io = cart_off(l) + kk;
                    for (int m = 0; m <= L - l; ++m) {
                        double w = PA[d] * th[m][i1] - PC[d] * th[m + 1][i1];
                        if (i2 >= 0) w += e1[d] * o2p * (th[m][i2] - th[m + 1][i2]);
                        th[m][io] = w;
                    }
                }
            const double pref = -atoms[4 * c + 3] * 2.0 * kPi * P.inv_p * K;
            for (int i = 0; i < NE; ++i) ve[i] += pref * th[0][cart_off(la) + i];
        }
    }
    // (a|b) from (a+k|0) for V
    const double AB[3] = {sp.AB[0], sp.AB[1], sp.AB[2]};
    double v[NA * NB];
    for (int ia = 0; ia < NA; ++ia)
        for (int ib = 0; ib < NB; ++ib) {
            int a[3], b[3];
            cart_xyz(la, ia, a[0], a[1], a[2]);
            cart_xyz(lb, ib, b[0], b[1], b[2]);
            double w = 0.0;
            for (int kx = 0; kx <= b[0]; ++kx)
                for (int ky = 0; ky <= b[1]; ++ky)
                    for (int kz = 0; kz <= b[2]; ++kz) {
                        const int ex = a[0] + kx, ey = a[1] + ky, ez = a[2] + kz, le = ex + ey + ez;
                        double c = binom(b[0], kx) * binom(b[1], ky) * binom(b[2], kz);
                        for (int q = 0; q < b[0] - kx; ++q) c *= AB[0];
                        for (int q = 0; q < b[1] - ky; ++q) c *= AB[1];
                        for (int q = 0; q < b[2] - kz; ++q) c *= AB[2];
                        w += c * ve[cart_off(le) - cart_off(la) + cart_pos(le, ex, ey)];
                    }
            v[ia * NB + ib] = w;
        }
    // Cartesian -> AO and packed-U stores (cart: d shells as Cartesian functions)
    const double* xa = xform + xform_off(la, cart);
    const double* xb = xform + xform_off(lb, cart);
    const int nfa = cart && la >= 2 ? NA : 2 * la + 1, nfb = cart && lb >= 2 ? NB : 2 * lb + 1;
    for (int ma = 0; ma < nfa; ++ma)
        for (int mb = 0; mb < nfb; ++mb) {
            double ss = 0.0, tt = 0.0, vv = 0.0;
            for (int ia = 0; ia < NA; ++ia) {
                const double wa = xa[ma * NA + ia];
                if (wa == 0.0) continue;
                for (int ib = 0; ib < NB; ++ib) {
                    const double w = wa * xb[mb * NB + ib];
                    ss += w * s[ia * NB + ib];
                    tt += w * k[ia * NB + ib];
                    vv += w * v[ia * NB + ib];
                }
            }
            const int i = sp.ao_a + ma, j = sp.ao_b + mb;
            const size_t o = i >= j ? (size_t)i * (i + 1) / 2 + j : (size_t)j * (j + 1) / 2 + i;
            S[o] = ss;
            T[o] = tt;
            V[o] = vv;
        }
}

// ------------------------------------------ one-electron gradient terms ---
// The one-electron part of the nuclear gradient (Gradient.cpp:61-116): per
// unique shell pair, sum_mn Pt_mn dT_mn/dR - W_mn dS_mn/dR + Pt_mn dV_mn/dR,
// with the reference's AO-pair weights (2 for m != n; a shell paired with
// itself counts m >= n).  One thread per (shell pair, job): job 0 takes the
// kinetic and overlap terms (they cancel within one atom; dB = -dA), job
// 1 + C the attraction to nucleus C (skipped when the pair and C share one
// atom; dC = -(dA + dB)).  Derivatives per primitive pair, d/dA_x <a|O|b> =
// 2 alpha <a+1x|O|b> - a_x <a-1x|O|b> and likewise for B with 2 beta: S and
// T from 1-D Obara-Saika overlaps, V from the vertical recursion on the
// first centre up to la + lb + 1 accumulated unweighted, 2 alpha- and 2
// beta-weighted, then the transfer to the second shell.  The pair densities
// are taken to the Cartesian block once per thread; the contributions go to
// grad with atomic adds.
namespace one_e {
__device__ __forceinline__ double binom3(int n, int k) { return (n == 3 && (k == 1 || k == 2)) ? 3.0 : binom(n, k); }
}  // namespace one_e

template <int la, int lb, int BM>
__global__ void __launch_bounds__(64) k_one_electron_grad(const int* __restrict__ plist, int n,
                                                          const ShellPair* __restrict__ pairs,
                                                          const PrimPair* __restrict__ pp,
                                                          const double* __restrict__ atoms, int natom,
                                                          const double* __restrict__ btab,
                                                          const double* __restrict__ xform,
                                                          const int* __restrict__ ao_atom,
                                                          const double* __restrict__ Pt,
                                                          const double* __restrict__ W, double* __restrict__ grad)
{
    using namespace one_e;
    constexpr double kPi = 3.14159265358979;
    constexpr int NA = ncart(la), NB = ncart(lb), L1 = la + lb + 1;
    constexpr int NFA = 2 * la + 1, NFB = 2 * lb + 1;
    constexpr int E0 = la > 0 ? cart_off(la - 1) : 0, NE = cart_off(L1 + 1) - E0;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)n * (natom + 1)) return;
    const int pi = (int)(t / (natom + 1)), c = (int)(t % (natom + 1)) - 1;
    const ShellPair sp = pairs[plist[pi]];
    const int atA = ao_atom[sp.ao_a], atB = ao_atom[sp.ao_b];
    if (c < 0 ? atA == atB : (atA == atB && c == atA)) return;
    const double A[3] = {sp.A[0], sp.A[1], sp.A[2]};
    const double B[3] = {sp.A[0] - sp.AB[0], sp.A[1] - sp.AB[1], sp.A[2] - sp.AB[2]};
    // AO pair densities -> Cartesian block
    double Dt[NA * NB], Dw[NA * NB];
    for (int i = 0; i < NA * NB; ++i) Dt[i] = Dw[i] = 0.0;
    const double* xa = xform + xform_off(la, 0);
    const double* xb = xform + xform_off(lb, 0);
    const bool same = sp.ao_a == sp.ao_b;
    for (int ma = 0; ma < NFA; ++ma)
        for (int mb = 0; mb < NFB; ++mb) {
            if (same && mb > ma) continue;
            const int i = sp.ao_a + ma, j = sp.ao_b + mb;
            const size_t o = i >= j ? (size_t)i * (i + 1) / 2 + j : (size_t)j * (j + 1) / 2 + i;
            const double w = i == j ? 1.0 : 2.0;
            const double pt = w * Pt[o], pw = c < 0 ? w * W[o] : 0.0;
            for (int ia = 0; ia < NA; ++ia) {
                const double fa = xa[ma * NA + ia];
                if (fa == 0.0) continue;
                for (int ib = 0; ib < NB; ++ib) {
                    const double f = fa * xb[mb * NB + ib];
                    Dt[ia * NB + ib] += f * pt;
                    Dw[ia * NB + ib] += f * pw;
                }
            }
        }
    if (c < 0) {
        // kinetic and overlap terms
        double g[3] = {0.0, 0.0, 0.0};
        for (int ip = 0; ip < sp.nprim; ++ip) {
            const PrimPair P = pp[sp.prim_off + ip];
            const double p = P.p, al = P.alpha, o2p = 0.5 * P.inv_p;
            const double K = P.Kp * p;  // w_a w_b exp(-ab/p |AB|^2)
            double S1[3][la + 4][lb + 1];
            for (int x = 0; x < 3; ++x) {
                const double PA = P.P[x] - A[x], PB = P.P[x] - B[x];
                S1[x][0][0] = sqrt(kPi * P.inv_p);
                for (int i = 1; i <= la + 3; ++i)
                    S1[x][i][0] = PA * S1[x][i - 1][0] + (i > 1 ? (i - 1) * S1[x][i - 2][0] * o2p : 0.0);
                for (int j = 1; j <= lb; ++j)
                    for (int i = 0; i <= la + 3; ++i) {
                        double w = PB * S1[x][i][j - 1];
                        if (i > 0) w += i * S1[x][i - 1][j - 1] * o2p;
                        if (j > 1) w += (j - 1) * S1[x][i][j - 2] * o2p;
                        S1[x][i][j] = w;
                    }
            }
            // S and T of the Cartesian pair (e | b): T applies -1/2 nabla^2 to the bra
            auto st = [&](const int* e, const int* b, double& sv, double& tv) {
                double sd[3], dd[3];
                for (int x = 0; x < 3; ++x) {
                    const int i = e[x], j = b[x];
                    sd[x] = S1[x][i][j];
                    dd[x] = 4.0 * al * al * S1[x][i + 2][j] - 2.0 * al * (2 * i + 1) * S1[x][i][j] +
                            (i > 1 ? (double)(i * (i - 1)) * S1[x][i - 2][j] : 0.0);
                }
                sv = sd[0] * sd[1] * sd[2];
                tv = -0.5 * (dd[0] * sd[1] * sd[2] + sd[0] * dd[1] * sd[2] + sd[0] * sd[1] * dd[2]);
            };
            for (int ia = 0; ia < NA; ++ia)
                for (int ib = 0; ib < NB; ++ib) {
                    const double dt = Dt[ia * NB + ib], dw = Dw[ia * NB + ib];
                    if (dt == 0.0 && dw == 0.0) continue;
                    int a[3], b[3];
                    cart_xyz(la, ia, a[0], a[1], a[2]);
                    cart_xyz(lb, ib, b[0], b[1], b[2]);
                    for (int d = 0; d < 3; ++d) {
                        int e[3] = {a[0], a[1], a[2]};
                        e[d] += 1;
                        double su, tu, sl = 0.0, tl = 0.0;
                        st(e, b, su, tu);
                        if (a[d] > 0) {
                            e[d] -= 2;
                            st(e, b, sl, tl);
                        }
                        const double dS = 2.0 * al * su - a[d] * sl, dT = 2.0 * al * tu - a[d] * tl;
                        g[d] += K * (dt * dT - dw * dS);
                    }
                }
        }
        for (int d = 0; d < 3; ++d) {
            unsafeAtomicAdd(&grad[3 * atA + d], g[d]);
            unsafeAtomicAdd(&grad[3 * atB + d], -g[d]);
        }
        return;
    }
    // attraction to nucleus c
    const double Cc[3] = {atoms[4 * c], atoms[4 * c + 1], atoms[4 * c + 2]};
    const double Zc = atoms[4 * c + 3];
    double U[NE], RA[NE], RB[NE];
    for (int i = 0; i < NE; ++i) U[i] = RA[i] = RB[i] = 0.0;
    double th[L1 + 1][cart_off(L1 + 1)];
    for (int ip = 0; ip < sp.nprim; ++ip) {
        const PrimPair P = pp[sp.prim_off + ip];
        const double p = P.p, al = P.alpha, be = p - al, o2p = 0.5 * P.inv_p;
        const double K = P.Kp * p;
        const double PA[3] = {P.P[0] - A[0], P.P[1] - A[1], P.P[2] - A[2]};
        const double PC[3] = {P.P[0] - Cc[0], P.P[1] - Cc[1], P.P[2] - Cc[2]};
        const double R2 = PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2];
        double F[L1 + 1];
        boys<BM, L1>(p * R2, btab, F);
        for (int m = 0; m <= L1; ++m) th[m][0] = F[m];
        for (int l = 1; l <= L1; ++l)
            for (int kk = 0; kk < ncart(l); ++kk) {
                int e[3];
                cart_xyz(l, kk, e[0], e[1], e[2]);
                const int d = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
                int e1[3] = {e[0], e[1], e[2]};
                e1[d] -= 1;
                const int i1 = cart_off(l - 1) + cart_pos(l - 1, e1[0], e1[1]);
                int i2 = -1;
                if (e1[d] > 0) {
                    int e2[3] = {e1[0], e1[1], e1[2]};
                    e2[d] -= 1;
                    i2 = cart_off(l - 2) + cart_pos(l - 2, e2[0], e2[1]);
                }
                const int io = cart_off(l) + kk;
                for (int m = 0; m <= L1 - l; ++m) {
                    double w = PA[d] * th[m][i1] - PC[d] * th[m + 1][i1];
                    if (i2 >= 0) w += e1[d] * o2p * (th[m][i2] - th[m + 1][i2]);
                    th[m][io] = w;
                }
            }
        const double pref = -Zc * 2.0 * kPi * P.inv_p * K;
        for (int i = 0; i < NE; ++i) {
            const double v = pref * th[0][E0 + i];
            U[i] += v;
            RA[i] += 2.0 * al * v;
            RB[i] += 2.0 * be * v;
        }
    }
    // (e|f) from (e+k|0): sum_k C(f,k) AB^(f-k) X(e+k|0) per direction
    const double AB[3] = {sp.AB[0], sp.AB[1], sp.AB[2]};
    auto xfer = [&](const double* X, const int* e, const int* f) {
        double w = 0.0;
        for (int kx = 0; kx <= f[0]; ++kx)
            for (int ky = 0; ky <= f[1]; ++ky)
                for (int kz = 0; kz <= f[2]; ++kz) {
                    const int ex = e[0] + kx, ey = e[1] + ky, ez = e[2] + kz, le = ex + ey + ez;
                    double cf = binom3(f[0], kx) * binom3(f[1], ky) * binom3(f[2], kz);
                    for (int q = 0; q < f[0] - kx; ++q) cf *= AB[0];
                    for (int q = 0; q < f[1] - ky; ++q) cf *= AB[1];
                    for (int q = 0; q < f[2] - kz; ++q) cf *= AB[2];
                    w += cf * X[cart_off(le) - E0 + cart_pos(le, ex, ey)];
                }
        return w;
    };
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    for (int ia = 0; ia < NA; ++ia)
        for (int ib = 0; ib < NB; ++ib) {
            const double D = Dt[ia * NB + ib];
            if (D == 0.0) continue;
            int a[3], b[3];
            cart_xyz(la, ia, a[0], a[1], a[2]);
            cart_xyz(lb, ib, b[0], b[1], b[2]);
            for (int d = 0; d < 3; ++d) {
                int e[3] = {a[0], a[1], a[2]}, f[3] = {b[0], b[1], b[2]};
                e[d] += 1;
                double vA = xfer(RA, e, b);
                if (a[d] > 0) {
                    e[d] -= 2;
                    vA -= a[d] * xfer(U, e, b);
                }
                f[d] += 1;
                double vB = xfer(RB, a, f);
                if (b[d] > 0) {
                    f[d] -= 2;
                    vB -= b[d] * xfer(U, a, f);
                }
                gA[d] += D * vA;
                gB[d] += D * vB;
            }
        }
    for (int d = 0; d < 3; ++d) {
        unsafeAtomicAdd(&grad[3 * atA + d], gA[d]);
        unsafeAtomicAdd(&grad[3 * atB + d], gB[d]);
        unsafeAtomicAdd(&grad[3 * c + d], -(gA[d] + gB[d]));
    }
}

hipError_t dsscf_launch_one_electron_grad(int la, int lb, int boys_mode, const int* plist, int n,
                                          const ShellPair* pairs, const PrimPair* pp, const double* atoms, int natom,
                                          const double* btab, const double* xform, const int* ao_atom,
                                          const double* Pt, const double* W, double* grad, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const long long nt = (long long)n * (natom + 1);
    const dim3 blk(64), grd((unsigned)((nt + 63) / 64));
#define DSSCF_ONE_EG(A_, B_)                                                                                       \
    if (la == A_ && lb == B_) {                                                                                  \
        if (boys_mode == BOYS_ACCURATE)                                                                          \
            hipLaunchKernelGGL((k_one_electron_grad<A_, B_, BOYS_ACCURATE>), grd, blk, 0, s, plist, n, pairs, pp,  \
                               atoms, natom, btab, xform, ao_atom, Pt, W, grad);                                 \
        else                                                                                                     \
            hipLaunchKernelGGL((k_one_electron_grad<A_, B_, BOYS_REFERENCE>), grd, blk, 0, s, plist, n, pairs, pp, \
                               atoms, natom, btab, xform, ao_atom, Pt, W, grad);                                 \
        return hipGetLastError();                                                                                \
    }
    DSSCF_ONE_EG(0, 0)
    DSSCF_ONE_EG(1, 0)
    DSSCF_ONE_EG(1, 1)
    DSSCF_ONE_EG(2, 0)
    DSSCF_ONE_EG(2, 1)
    DSSCF_ONE_EG(2, 2)
#undef DSSCF_ONE_EG
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_one_electron(int la, int lb, int boys_mode, const int* plist, int n, const ShellPair* pairs,
                                     const PrimPair* pp, const double* atoms, int natom, const double* btab,
                                     const double* xform, int cart, double* S, double* T, double* V, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const dim3 blk(64), grd((n + 63) / 64);
#define DSSCF_ONE_E(A_, B_)                                                                                        \
    if (la == A_ && lb == B_) {                                                                                  \
        if (boys_mode == BOYS_ACCURATE)                                                                          \
            hipLaunchKernelGGL((k_one_electron<A_, B_, BOYS_ACCURATE>), grd, blk, 0, s, plist, n, pairs, pp, atoms, \
                               natom, btab, xform, cart, S, T, V);                                               \
        else                                                                                                     \
            hipLaunchKernelGGL((k_one_electron<A_, B_, BOYS_REFERENCE>), grd, blk, 0, s, plist, n, pairs, pp,     \
                               atoms, natom, btab, xform, cart, S, T, V);                                        \
        return hipGetLastError();                                                                                \
    }
    DSSCF_ONE_E(0, 0)
    DSSCF_ONE_E(1, 0)
    DSSCF_ONE_E(1, 1)
    DSSCF_ONE_E(2, 0)
    DSSCF_ONE_E(2, 1)
    DSSCF_ONE_E(2, 2)
#undef DSSCF_ONE_E
    return hipErrorInvalidValue;
}

// -------------------------------------------------------- host launch ---
void dsscf_upload_xform(const double* x, int n, hipStream_t s)
{
#define DSSCF_UPLOAD(a, b, cc, d, name, wave) (void)ClassLaunch<a, b, cc, d, wave>::upload_xform(x, n, s);
    DSSCF_FOR_EACH_CLASS(DSSCF_UPLOAD)
#undef DSSCF_UPLOAD
}

int dk_xform_size() { return DSSCF_XFORM_SIZE; }
int dk_max_l() { return DSSCF_MAXL; }

namespace {
struct ClassDesc {
    int la, lb, lc, ld;
};
const ClassDesc kClasses[] = {
#define DSSCF_CLASS_DESC(a, b, c, d, name, wave) {a, b, c, d},
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

hipError_t dsscf_launch_jk(int cls, int boys_mode, int fp32, const JKSeg* segs, const int* blk0, int nseg,
                           int nblocks, int block_size, int maxN, int nks, int maxBra, int cart, const double* btab,
                           double scale, hipStream_t s)
{
    if (nseg <= 0 || nblocks <= 0) return hipSuccess;
    int c = 0;
#define DSSCF_LAUNCH_JK(a, b, cc, d, name, wave)                                                                 \
    if (c++ == cls)                                                                                              \
        return ClassLaunch<a, b, cc, d, wave>::jk(boys_mode, fp32, segs, blk0, nseg, nblocks, block_size, maxN,   \
                                                  nks, maxBra, cart, btab, scale, s);
    DSSCF_FOR_EACH_CLASS(DSSCF_LAUNCH_JK)
#undef DSSCF_LAUNCH_JK
    return hipErrorInvalidValue;
}

int dsscf_wave_group(int cls, int cart)
{
    int c = 0;
#define DSSCF_WAVE_G(a, b, cc, d, name, wave) \
    if (c++ == cls) return ClassLaunch<a, b, cc, d, wave>::wave_group(cart);
    DSSCF_FOR_EACH_CLASS(DSSCF_WAVE_G)
#undef DSSCF_WAVE_G
    return 1;
}

int dsscf_class_is_wave(int cls)
{
    int c = 0;
#define DSSCF_WAVE_OF(a, b, cc, d, name, wave) \
    if (c++ == cls) return wave;
    DSSCF_FOR_EACH_CLASS(DSSCF_WAVE_OF)
#undef DSSCF_WAVE_OF
    return 0;
}

hipError_t dsscf_launch_eri_block(int cls, int boys_mode, int cart, const int4* items, int n, const ShellPair* pairs,
                                  const PrimPair* pp, const double* btab, double* out, double scale,
                                  hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int c = 0;
#define DSSCF_LAUNCH_BLK(a, b, cc, d, name, wave)                                                                    \
    if (c++ == cls)                                                                                                 \
        return ClassLaunch<a, b, cc, d, wave>::block(boys_mode, cart, items, n, pairs, pp, btab, out, scale, s);
    DSSCF_FOR_EACH_CLASS(DSSCF_LAUNCH_BLK)
#undef DSSCF_LAUNCH_BLK
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_schwarz(int la, int lb, int boys_mode, int cart, const int* plist, int n,
                                const ShellPair* pairs, const PrimPair* pp, const double* btab, double* Q, double scale,
                                hipStream_t s)
{
    if (n <= 0) return hipSuccess;
#define DSSCF_LAUNCH_SCH(a, b, cc, d, name, wave)                                                             \
    if (a == cc && b == d && la == a && lb == b)                                                                \
        return ClassLaunch<a, b, cc, d, wave>::schwarz(boys_mode, cart, plist, n, pairs, pp, btab, Q, scale, s);
    DSSCF_FOR_EACH_CLASS(DSSCF_LAUNCH_SCH)
#undef DSSCF_LAUNCH_SCH
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_fill(double* p, size_t n, double v, hipStream_t s)
{
    const int grd = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_fill, dim3(grd > 0 ? grd : 1), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t dsscf_launch_prep(const MatSeg* segs, int nseg, int total_rows, int maxN, hipStream_t s)
{
    hipLaunchKernelGGL(k_prep, dim3((maxN + 255) / 256, total_rows), dim3(256), 0, s, segs, nseg);
    return hipGetLastError();
}

hipError_t dsscf_launch_jcd(const JcdSeg* segs, const int* blk0, int nseg, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_jcd, dim3(nblocks), dim3(256), 0, s, segs, blk0, nseg);
    return hipGetLastError();
}

hipError_t dsscf_launch_fold(const MatSeg* segs, int nseg, int total_rows, int maxN, hipStream_t s)
{
    hipLaunchKernelGGL(k_fold, dim3((maxN + 255) / 256, total_rows), dim3(256), 0, s, segs, nseg);
    return hipGetLastError();
}

// ------------------------------------------------ Boys-function probe ---
// F_0..F_L at n arguments in one of the device modes (the ERI kernels'
// own evaluation, for checking against tabulated values): out[k*(L+1)+m].
template <int L, int BM>
__global__ void __launch_bounds__(256) k_boys_eval(const double* __restrict__ T, int n,
                                                   const double* __restrict__ btab, double* __restrict__ out)
{
    __shared__ __align__(16) double tab[BM == BOYS_ACCURATE_LDS ? DSSCF_BPOLY_M_STRIDE : 1];
    const double* t = btab;
    if (BM == BOYS_ACCURATE_LDS) {
        for (int i = threadIdx.x; i < DSSCF_BPOLY_M_STRIDE; i += blockDim.x)
            tab[i] = btab[DSSCF_BPOLY_OFFSET + L * DSSCF_BPOLY_M_STRIDE + i];
        __syncthreads();
        t = tab;
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double f[L + 1];
    boys<BM, L>(T[k], t, f);
#pragma unroll
    for (int m = 0; m <= L; ++m) out[(size_t)k * (L + 1) + m] = f[m];
}

template <int BM>
static hipError_t launch_boys_eval_mode(int L, const double* T, int n, const double* btab, double* out, hipStream_t s)
{
    const dim3 blk(256), grd((n + 255) / 256);
    switch (L) {
#define DSSCF_BOYS_CASE(L_)                                                                   \
    case L_:                                                                                \
        hipLaunchKernelGGL((k_boys_eval<L_, BM>), grd, blk, 0, s, T, n, btab, out);          \
        return hipGetLastError();
        DSSCF_BOYS_CASE(0)
        DSSCF_BOYS_CASE(1)
        DSSCF_BOYS_CASE(2)
        DSSCF_BOYS_CASE(3)
        DSSCF_BOYS_CASE(4)
        DSSCF_BOYS_CASE(5)
        DSSCF_BOYS_CASE(6)
        DSSCF_BOYS_CASE(7)
        DSSCF_BOYS_CASE(8)
#undef DSSCF_BOYS_CASE
    }
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_boys_eval(int mode, int L, const double* T, int n, const double* btab, double* out,
                                  hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (mode == BOYS_REFERENCE) return launch_boys_eval_mode<BOYS_REFERENCE>(L, T, n, btab, out, s);
    if (mode == BOYS_ACCURATE) return launch_boys_eval_mode<BOYS_ACCURATE>(L, T, n, btab, out, s);
    if (mode == BOYS_ACCURATE_LDS) return launch_boys_eval_mode<BOYS_ACCURATE_LDS>(L, T, n, btab, out, s);
    return hipErrorInvalidValue;
}
