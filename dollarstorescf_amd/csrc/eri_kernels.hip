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
// x(x+1)/2+y, Matrix.ipp:332-351) -> full row-major Dt, Da; zero Jh, Kh rows.
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
        if (y <= x) {
            const size_t o = (size_t)x * (x + 1) / 2 + y;
            const double t = S.Ptp[o], a = S.Pap[o];
            S.Dt[(size_t)y * N + x] = t;
            S.Dt[(size_t)x * N + y] = t;
            S.Da[(size_t)y * N + x] = a;
            S.Da[(size_t)x * N + y] = a;
        }
    }
}

// J = sum over slices of Jh + Jh^T, likewise K -> packed-U, per context.
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
    }
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
                           int nblocks, int block_size, int maxN, const double* btab, double scale, hipStream_t s)
{
    if (nseg <= 0 || nblocks <= 0) return hipSuccess;
    int c = 0;
#define DSSCF_LAUNCH_JK(a, b, cc, d, name, wave)                                                                 \
    if (c++ == cls)                                                                                              \
        return ClassLaunch<a, b, cc, d, wave>::jk(boys_mode, fp32, segs, blk0, nseg, nblocks, block_size, maxN, btab, \
                                                  scale, s);
    DSSCF_FOR_EACH_CLASS(DSSCF_LAUNCH_JK)
#undef DSSCF_LAUNCH_JK
    return hipErrorInvalidValue;
}

int dsscf_wave_group(int cls)
{
    int c = 0;
#define DSSCF_WAVE_G(a, b, cc, d, name, wave) \
    if (c++ == cls) return ClassLaunch<a, b, cc, d, wave>::wave_group();
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

hipError_t dsscf_launch_eri_block(int cls, int boys_mode, const int4* items, int n, const ShellPair* pairs,
                                  const PrimPair* pp, const double* btab, double* out, double scale,
                                  hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int c = 0;
#define DSSCF_LAUNCH_BLK(a, b, cc, d, name, wave) \
    if (c++ == cls) return ClassLaunch<a, b, cc, d, wave>::block(boys_mode, items, n, pairs, pp, btab, out, scale, s);
    DSSCF_FOR_EACH_CLASS(DSSCF_LAUNCH_BLK)
#undef DSSCF_LAUNCH_BLK
    return hipErrorInvalidValue;
}

hipError_t dsscf_launch_schwarz(int la, int lb, int boys_mode, const int* plist, int n, const ShellPair* pairs,
                                const PrimPair* pp, const double* btab, double* Q, double scale, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
#define DSSCF_LAUNCH_SCH(a, b, cc, d, name, wave)                                                       \
    if (a == cc && b == d && la == a && lb == b)                                                          \
        return ClassLaunch<a, b, cc, d, wave>::schwarz(boys_mode, plist, n, pairs, pp, btab, Q, scale, s);
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

hipError_t dsscf_launch_fold(const MatSeg* segs, int nseg, int total_rows, int maxN, hipStream_t s)
{
    hipLaunchKernelGGL(k_fold, dim3((maxN + 255) / 256, total_rows), dim3(256), 0, s, segs, nseg);
    return hipGetLastError();
}
