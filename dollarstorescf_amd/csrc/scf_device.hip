// scf_device.hip -- SCF iteration on the device (see host/scf_device.hpp).
//
// Layout: full matrices are n x n column-major (S, X, F and P are
// symmetric, so their storage is the same either way); packed-U vectors
// hold element (y,x), y <= x, at x(x+1)/2 + y (src/lalib/Matrix.ipp:332-351).
// One HIP stream and one rocBLAS handle per job.  Per iteration:
//
//   k_assemble       G = J - K (packed), F0 = H + G (full), P (full)
//   dgemm x2         S (P F0)                           (Pulay error, SCFDIIS.cpp:19-25)
//   k_antisym        e = M - M^T into history slot
//   k_dots           <e_new, e_i> for the B matrix      (DIIS.ipp:18-62)
//   k_combine        F = sum c_i F_i                    (DIIS.ipp:84-100)
//   dgemm x2         F' = X F X                         (SCFSolver.cpp:64-75)
//   dsyevd           F' = C' eps C'^T                   (Lapack.cpp:12-39, dspev_)
//   dgemm            C = X C'
//   dgemm            P = C_occ C_occ^T                  (SCFSolver.cpp:77-94)
//   k_pack_density   Pa, Ptot = 2 Pa (packed), tr(Pa G)  (SCFSolver.cpp:118-143)
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host/scf_device.hpp"

namespace scfdev {

constexpr int kMaxHist = 16;

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
    rocblas_int* info = nullptr;
    double* h_scal = nullptr;                           // pinned host mirror of scal
    size_t blas_bytes = 0;                              // the handle's workspace (sized at creation)
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

// F (full, in s->F) -> eigenpairs, C, new densities, tr(Pa_new G)
int diagonalize(Step* s, std::vector<double>& eps, double* tr_PG, std::string* err)
{
    const int n = s->n;
    const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
    if (int rc = gemm(s, N, N, n, s->X, s->F, s->T1, err)) return rc;
    if (int rc = gemm(s, N, N, n, s->T1, s->X, s->T2, err)) return rc;  // F' in T2
    STEP_BLAS(rocsolver_dsyevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n, s->T2, n, s->w, s->work,
                               s->info));
    if (int rc = gemm(s, N, N, n, s->X, s->T2, s->C, err)) return rc;
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
                    std::make_pair(&s->scal, (size_t)kmax + 1)})
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
                      s->Eh, s->w, s->work, s->scal})
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
    return (long long)(sizeof(double) * (6 * np + (8 + 2 * (size_t)kmax) * nn + 3 * (size_t)n + kmax));
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

}  // namespace scfdev
