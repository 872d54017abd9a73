// dsscf_abi.cpp -- implementation of the C-ABI in include/dsscf.h.
//
// Owns the per-context device buffers, builds the shell-pair and
// primitive-pair records, the Schwarz bounds and the class-sorted quartet
// lists, and drives the per-class kernels of eri_kernels.hip.
#include "../../include/dsscf.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <atomic>
#include <thread>

#include "boys_table.hpp"
#include "kernels.hpp"

#include <mutex>

// ------------------------------------------------------- no-alloc scratch --
// Device memory of dsscf_eri_batch_noalloc (its shell-quartet item list and
// AO-block outputs): a file-scope array, so that the call needs no device
// allocation.  Each GPU the library runs on holds its own copy, at its own
// address, until the process ends.  Host code must not pass `scratch` itself
// to a kernel (on the host the name is only a stand-in): it uses the address
// from hipGetSymbolAddress, taken once per device at context creation.
#define DSSCF_SCRATCH_BYTES ((size_t)16 << 20)
static __device__ __align__(256) unsigned char scratch[DSSCF_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_dev;  // per device: the device address of `scratch`

// Look up `scratch` on `device` (which must be current) once.
int scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize(device + 1, nullptr);
    if (g_scratch_dev[device]) return DSSCF_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return DSSCF_ERR_HIP;
    }
    g_scratch_dev[device] = p;
    return DSSCF_OK;
}

void* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < (int)g_scratch_dev.size() ? g_scratch_dev[device] : nullptr;
}

// Byte offset of the block outputs behind n int4 items (256-byte aligned).
size_t staging_offset_out(size_t n_items) { return (sizeof(int4) * n_items + 255) & ~(size_t)255; }

}  // namespace

namespace {

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(DSSCF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// 2 pi^(5/2) with the reference's PI (GaussianIntegral.cpp:632, Constants.hpp:5).
const double kEriScale = 2.0 * std::pow(3.14159265358979, 2.5);

// Smallest primitive-pair run of a sub-pair.
constexpr int kMinSubPrims = 5;
constexpr int kMaxGroup = 16;  // lanes per quartet, at most
// Kets per work unit that a class schedule may ask for (the default stays at
// 64..512).
constexpr int kMaxKetsPerUnit = 4096;

// Lanes per quartet for a (sub-)quartet of nb x nk primitive quartets: the
// largest power of two G <= kMaxGroup with nb*nk/G >= target, split into
// gb lanes over bra primitives and gk = G/gb over ket primitives (each at
// most its primitive count) so that the busiest lane's share
// ceil(nb/gb)*ceil(nk/gk) is smallest.  Returns gb | gk << 8.
int quartet_group(int nb, int nk, double target)
{
    if (target <= 0.0) return 1 | (1 << 8);
    int G = 1;
    while (G < kMaxGroup && (double)nb * nk / (2.0 * G) >= target) G *= 2;
    int best = 1 | (1 << 8), best_load = nb * nk + 1;
    for (int gb = 1; gb <= G; gb *= 2) {
        const int gk = G / gb;
        if (gb > nb || gk > nk) continue;
        const int load = ((nb + gb - 1) / gb) * ((nk + gk - 1) / gk);
        if (load < best_load) {
            best_load = load;
            best = gb | (gk << 8);
        }
    }
    return best;
}

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
// AOs of an l-shell: 2l+1 real solid harmonics, or with cartesian d (l >= 2)
// the ncart(l) Cartesian functions
int nfun(int l, int cart) { return cart && l >= 2 ? ncart(l) : 2 * l + 1; }

// Algorithmic FLOPs of one primitive quartet of class (la lb|lc ld): setup,
// Boys, Head-Gordon-Pople VRR at 5 FLOPs per entry, contraction into the
// HRR inputs (the model of dollarstorescf_amd/flops.py).
double prim_quartet_flops(int la, int lb, int lc, int ld)
{
    const int L = la + lb + lc + ld;
    long long vrr = -(L + 1), contr = 0;
    for (int e = 0; e <= la + lb; ++e)
        for (int f = 0; f <= lc + ld; ++f) vrr += (long long)ncart(e) * ncart(f) * (L - e - f + 1);
    for (int e = la; e <= la + lb; ++e)
        for (int f = lc; f <= lc + ld; ++f) contr += (long long)ncart(e) * ncart(f);
    return 40.0 + (27 + 3 * L) + 10.0 * vrr + 2.0 * contr;
}
int pair_class_of(int l1, int l2) { return l1 * (l1 + 1) / 2 + l2; }

template <typename T>
int dev_alloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * n);
    if (e != hipSuccess) return fail(DSSCF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return DSSCF_OK;
}

template <typename T>
void dev_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}


// ------------------------------------------------------------------ plan --
// A fused Fock-build launch over one or more contexts on one device: per
// class, a segment table (one JKSeg per context with quartets of the class);
// for the prepare/fold kernels, one MatSeg per context over stacked rows.
constexpr int kStreams = 4;  // streams per fused Fock build (GPU_MAX_HW_QUEUES defaults to 4)

struct Plan {
    int boys_mode = DSSCF_BOYS_ACCURATE;
    int cart = 0;  // Cartesian d shells (every member context agrees)
    int nks = 1;   // exchange sets per block: 2 when a member is unrestricted (K[Pa] and K[Pb])
    const double* btab = nullptr;
    int ncls = 0;
    std::vector<int> cls_nseg, cls_nblocks, cls_off, cls_bs, cls_maxN, cls_maxbra;
    int nmat = 0, total_rows = 0, maxN = 0;
    std::vector<JKSeg> h_segs;
    std::vector<int> h_blk0;
    std::vector<MatSeg> h_mat;
    JKSeg* d_segs = nullptr;
    int* d_blk0 = nullptr;
    MatSeg* d_mat = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool built = false;
    // Class kernels are independent (they only add into J and K), so they run
    // on kStreams streams at once: the latency-bound tails of small classes
    // then overlap the large ones.  cls_stream: stream of each slot (0 = the
    // caller's), balanced by cls_cost; side streams and fork/join events.
    std::vector<double> cls_cost;
    std::vector<int> cls_stream;
    // launch order of the slots on the forked build: highest L first, so that
    // the classes with the longest per-quartet chains start early and their
    // tails overlap the rest (then heaviest first)
    std::vector<int> launch_order;
    hipStream_t side[kStreams - 1] = {};
    hipEvent_t fork = nullptr, join[kStreams - 1] = {};
    std::vector<hipEvent_t> slot_ev;  // start/end of each slot's kernel (timed builds)

    size_t cap_segs = 0, cap_blk0 = 0, cap_mat = 0;  // device table capacities (grow only)
    // J_cd slot reduction (k_jcd) over the contexts that have slot records
    std::vector<JcdSeg> h_jcd;
    std::vector<int> h_jblk0;
    JcdSeg* d_jcd = nullptr;
    int* d_jblk0 = nullptr;
    size_t cap_jcd = 0, cap_jblk0 = 0;
    int jcd_blocks = 0;

    // Forget the launch (graph, built flag); keep tables and streams for the next build.
    void reset()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        built = false;
    }

    void release()
    {
        reset();
        for (int i = 0; i < kStreams - 1; ++i) {
            if (side[i]) (void)hipStreamDestroy(side[i]);
            if (join[i]) (void)hipEventDestroy(join[i]);
            side[i] = nullptr;
            join[i] = nullptr;
        }
        if (fork) (void)hipEventDestroy(fork);
        fork = nullptr;
        for (hipEvent_t e : slot_ev) (void)hipEventDestroy(e);
        slot_ev.clear();
        dev_free(d_segs);
        dev_free(d_blk0);
        dev_free(d_mat);
        dev_free(d_jcd);
        dev_free(d_jblk0);
        cap_segs = cap_blk0 = cap_mat = cap_jcd = cap_jblk0 = 0;
    }
};
// Fisher-Yates shuffle driven by splitmix64 (deterministic for a seed).
template <typename T>
void shuffle_seq(T* v, size_t n, unsigned long long seed)
{
    for (size_t i = n; i > 1; --i) {
        seed += 0x9e3779b97f4a7c15ULL;
        unsigned long long z = seed;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        std::swap(v[i - 1], v[z % i]);
    }
}

}  // namespace

struct dsscf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    dsscf_options opt{DSSCF_BOYS_ACCURATE, 1e-12, 0.0, 0.0};

    // basis
    int n_shell = 0, n_ao = 0;
    std::vector<int> sh_l, sh_nprim, sh_prim_off, sh_ao, ao_shell;
    std::vector<double> sh_center, prim_exp, prim_coef, xform;
    bool have_basis = false;

    // pairs (index a(a+1)/2+b for shells a >= b)
    std::vector<ShellPair> pairs;
    std::vector<PrimPair> prims;
    std::vector<int> pair_first, pair_class;
    std::vector<double> Q;

    // quartet lists
    // per class: FP64 list [q_off, q_off+q_cnt), FP32 list [q_off32, q_off32+q_cnt32)
    std::vector<long long> q_off, q_cnt, q_prim, q_off32, q_cnt32, q_prim32;
    // thread-per-quartet classes: work units (bra pair, first ket, count) into `kets`,
    // per class [u_off, u_off+u_cnt) FP64 and [u_off32, u_off32+u_cnt32) FP32
    std::vector<long long> u_off, u_cnt, u_off32, u_cnt32;
    std::vector<int> u_maxbra, u_maxbra32;  // per class: most primitive pairs of a unit's bra
    // per class: wavefront fill of the work units (dsscf_stats class_lanes etc.)
    std::vector<long long> fill_lanes, fill_slots, fill_lane_trips, fill_wave_trips;
    long long n_unscreened = 0;
    long long n_work_pairs = 0, n_work_items = 0;  // shell pairs + sub-pairs; (sub-)quartets
    int n_slice = 1;                                // copies of Jh/Kh
    long long device_bytes = 0;                     // geometry-dependent device memory after prepare
    bool prepared = false;

    // device
    ShellPair* d_pairs = nullptr;
    PrimPair* d_pp = nullptr;
    double* d_btab = nullptr;
    uint2* d_q = nullptr;
    int4* d_units = nullptr;
    unsigned* d_kets = nullptr;
    double* d_Q = nullptr;
    int* d_plist = nullptr;
    double *d_Dt = nullptr, *d_Da = nullptr, *d_Jh = nullptr, *d_Kh = nullptr;
    double *d_Pt_pk = nullptr, *d_Pa_pk = nullptr, *d_J_pk = nullptr, *d_K_pk = nullptr;
    // unrestricted builds (dsscf_jk_uhf), allocated on first use: the beta
    // density (packed, full), K[Pb] half-accumulated (n_slice copies) and packed
    double *d_Pb_pk = nullptr, *d_Db = nullptr, *d_Kbh = nullptr, *d_Kb_pk = nullptr;
    double* d_xform = nullptr;                   // AO transform (one-electron kernel)
    // J_cd slot records of the thread-per-quartet classes (null: atomics):
    // per unit its first double in d_js; per ket AO block (row) its records
    double* d_js = nullptr;
    int* d_us = nullptr;
    int* d_jd = nullptr;
    int2* d_jrr = nullptr;
    int4* d_jri = nullptr;
    int jcd_rows = 0;
    long long jcd_bytes = 0;
    std::vector<int> pc_start, pc_count;         // pair classes in d_plist

    std::vector<int> block_size;  // per class, thread-per-quartet kernels (autotuned)
    std::vector<int> chunk;       // per class, kets per work unit
    // per-class schedule overrides for prepare (dsscf_set_class_schedule; 0 or
    // empty: the options' value / the default): sub-pair FLOP size,
    // lane-group target, kets per work unit
    std::vector<double> sched_wif, sched_gp;
    std::vector<int> sched_chunk;
    Plan plan;                    // fused launch plan over this context alone
    Plan plan_u;                  // the same for unrestricted builds (K[Pa] and K[Pb])

    void free_geometry()
    {
        plan.release();
        plan_u.release();
        dev_free(d_Pb_pk);
        dev_free(d_Db);
        dev_free(d_Kbh);
        dev_free(d_Kb_pk);
        dev_free(d_pairs);
        dev_free(d_pp);
        dev_free(d_q);
        dev_free(d_units);
        dev_free(d_kets);
        dev_free(d_Q);
        dev_free(d_plist);
        dev_free(d_Dt);
        dev_free(d_Da);
        dev_free(d_Jh);
        dev_free(d_Kh);
        dev_free(d_Pt_pk);
        dev_free(d_Pa_pk);
        dev_free(d_J_pk);
        dev_free(d_K_pk);
        dev_free(d_xform);
        dev_free(d_js);
        dev_free(d_us);
        dev_free(d_jd);
        dev_free(d_jrr);
        dev_free(d_jri);
        jcd_rows = 0;
        jcd_bytes = 0;
        prepared = false;
    }
};

namespace {
// First AO of each shell and the shell of each AO, from the shells' l and
// the cartesian option.
void layout_aos(dsscf_ctx* c)
{
    int nao = 0;
    for (int s = 0; s < c->n_shell; ++s) {
        c->sh_ao[s] = nao;
        nao += nfun(c->sh_l[s], c->opt.cartesian);
    }
    c->n_ao = nao;
    c->ao_shell.resize(nao);
    for (int s = 0; s < c->n_shell; ++s)
        for (int m = 0; m < nfun(c->sh_l[s], c->opt.cartesian); ++m) c->ao_shell[c->sh_ao[s] + m] = s;
}
}  // namespace

extern "C" {

const char* dsscf_last_error(void) { return g_err.c_str(); }
int dsscf_max_l(void) { return dk_max_l(); }
int dsscf_class_count(void) { return dk_n_class(); }
int dsscf_xform_size(void) { return dk_xform_size(); }

int dsscf_class_l(int c, int* l4)
{
    if (c < 0 || c >= dk_n_class() || !l4) return fail(DSSCF_ERR_ARG, "dsscf_class_l: bad class");
    dk_class_l(c, l4);
    return DSSCF_OK;
}

int dsscf_ctx_create(dsscf_ctx** out, int device)
{
    if (!out) return fail(DSSCF_ERR_ARG, "dsscf_ctx_create: null output");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(DSSCF_ERR_ARG, "dsscf_ctx_create: no such device");
    HIP_TRY(hipSetDevice(device));
    // the no-alloc entry points' scratch address on this device (once per device)
    if (int rc = scratch_setup(device)) return rc;
    dsscf_ctx* c = new dsscf_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(DSSCF_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    std::vector<double> tab;
    dsscf_boys_table(tab);
    if (int rc = dev_alloc(&c->d_btab, tab.size())) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    e = hipMemcpy(c->d_btab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dev_free(c->d_btab);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(DSSCF_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    *out = c;
    return DSSCF_OK;
}

int dsscf_ctx_destroy(dsscf_ctx* c)
{
    if (!c) return DSSCF_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->free_geometry();
    dev_free(c->d_btab);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return DSSCF_OK;
}

int dsscf_set_options(dsscf_ctx* c, const dsscf_options* o)
{
    if (!c || !o) return fail(DSSCF_ERR_ARG, "dsscf_set_options: null argument");
    if (o->boys_mode != DSSCF_BOYS_REFERENCE && o->boys_mode != DSSCF_BOYS_ACCURATE)
        return fail(DSSCF_ERR_ARG, "dsscf_set_options: unknown boys_mode");
    if (o->ket_order < 0 || o->ket_order > 2) return fail(DSSCF_ERR_ARG, "dsscf_set_options: ket_order 0..2");
    if (o->cartesian != 0 && o->cartesian != 1) return fail(DSSCF_ERR_ARG, "dsscf_set_options: cartesian 0 or 1");
    if (!(o->schwarz_threshold >= 0.0) || !(o->fp32_threshold >= 0.0) || !(o->prim_threshold >= 0.0))
        return fail(DSSCF_ERR_ARG, "dsscf_set_options: negative threshold");
    if (o->fp32_threshold > 0.0 && o->boys_mode != DSSCF_BOYS_ACCURATE)
        return fail(DSSCF_ERR_ARG, "dsscf_set_options: FP32 quartets need the accurate Boys mode");
    const bool changed = o->boys_mode != c->opt.boys_mode || o->schwarz_threshold != c->opt.schwarz_threshold ||
                         o->fp32_threshold != c->opt.fp32_threshold || o->prim_threshold != c->opt.prim_threshold ||
                         o->work_item_flops != c->opt.work_item_flops || o->group_prims != c->opt.group_prims ||
                         o->jcd_max_bytes != c->opt.jcd_max_bytes || o->ket_order != c->opt.ket_order ||
                         o->cartesian != c->opt.cartesian;
    const bool relayout = o->cartesian != c->opt.cartesian && c->have_basis;
    c->opt = *o;
    if (changed) c->prepared = false;
    if (relayout) layout_aos(c);
    return DSSCF_OK;
}

int dsscf_set_basis(dsscf_ctx* c, int n_shell, const int* shell_l, const int* shell_nprim,
                    const double* shell_center, const double* prim_exp, const double* prim_coef,
                    const double* ao_xform)
{
    if (!c || n_shell <= 0 || !shell_l || !shell_nprim || !shell_center || !prim_exp || !prim_coef || !ao_xform)
        return fail(DSSCF_ERR_ARG, "dsscf_set_basis: null argument or no shells");
    const int maxl = dk_max_l();
    c->sh_l.assign(shell_l, shell_l + n_shell);
    c->sh_nprim.assign(shell_nprim, shell_nprim + n_shell);
    c->sh_center.assign(shell_center, shell_center + 3 * n_shell);
    c->sh_prim_off.resize(n_shell);
    c->sh_ao.resize(n_shell);
    int np = 0;
    for (int s = 0; s < n_shell; ++s) {
        if (shell_l[s] < 0 || shell_l[s] > maxl)
            return fail(DSSCF_ERR_UNSUPPORTED, "dsscf_set_basis: shell angular momentum " + std::to_string(shell_l[s]) +
                                                   " exceeds the built maximum " + std::to_string(maxl));
        if (shell_nprim[s] <= 0 || shell_nprim[s] > 100)
            return fail(DSSCF_ERR_ARG, "dsscf_set_basis: primitive count outside 1-100");
        c->sh_prim_off[s] = np;
        np += shell_nprim[s];
    }
    for (int i = 0; i < np; ++i)
        if (!(prim_exp[i] > 0.0)) return fail(DSSCF_ERR_ARG, "dsscf_set_basis: non-positive exponent");
    int maxp = 0;
    for (int s2 = 0; s2 < n_shell; ++s2) maxp = std::max(maxp, shell_nprim[s2]);
    if (maxp * maxp > kBraStage)
        return fail(DSSCF_ERR_UNSUPPORTED, "dsscf_set_basis: shells with more than " + std::to_string((int)std::sqrt((double)kBraStage)) +
                                               " primitives are not supported (primitive pairs are staged in LDS)");
    c->prim_exp.assign(prim_exp, prim_exp + np);
    c->prim_coef.assign(prim_coef, prim_coef + np);
    c->xform.assign(ao_xform, ao_xform + dk_xform_size());
    c->n_shell = n_shell;
    layout_aos(c);
    c->have_basis = true;
    c->prepared = false;
    return DSSCF_OK;
}

namespace {

// Shell pairs and primitive pairs; the first shell of a pair has the higher l
// (equal l: the higher shell index).
void build_pairs(dsscf_ctx* c)
{
    const int ns = c->n_shell;
    const size_t npair = (size_t)ns * (ns + 1) / 2;
    c->pairs.assign(npair, ShellPair{});
    c->pair_first.assign(npair, 0);
    c->pair_class.assign(npair, 0);
    c->prims.clear();
    for (int a = 0; a < ns; ++a)
        for (int b = 0; b <= a; ++b) {
            const size_t idx = (size_t)a * (a + 1) / 2 + b;
            const int f = c->sh_l[a] >= c->sh_l[b] ? a : b;
            const int s = f == a ? b : a;
            ShellPair& P = c->pairs[idx];
            const double* A = &c->sh_center[3 * f];
            const double* B = &c->sh_center[3 * s];
            double AB2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                P.A[x] = A[x];
                P.AB[x] = A[x] - B[x];
                AB2 += P.AB[x] * P.AB[x];
            }
            P.prim_off = (int)c->prims.size();
            P.ao_a = c->sh_ao[f];
            P.ao_b = c->sh_ao[s];
            for (int i = 0; i < c->sh_nprim[f]; ++i)
                for (int j = 0; j < c->sh_nprim[s]; ++j) {
                    const double al = c->prim_exp[c->sh_prim_off[f] + i];
                    const double be = c->prim_exp[c->sh_prim_off[s] + j];
                    PrimPair pp{};
                    pp.p = al + be;
                    pp.inv_p = 1.0 / pp.p;
                    for (int x = 0; x < 3; ++x) pp.P[x] = (al * A[x] + be * B[x]) * pp.inv_p;
                    pp.Kp = c->prim_coef[c->sh_prim_off[f] + i] * c->prim_coef[c->sh_prim_off[s] + j] *
                            std::exp(-al * be * pp.inv_p * AB2) * pp.inv_p;
                    pp.e = std::fabs(pp.Kp) / std::sqrt(std::sqrt(pp.p));
                    pp.alpha = al;
                    c->prims.push_back(pp);
                }
            P.nprim = (int)c->prims.size() - P.prim_off;
            std::stable_sort(c->prims.begin() + P.prim_off, c->prims.end(),
                             [](const PrimPair& x, const PrimPair& y) { return x.e > y.e; });
            c->pair_first[idx] = f;
            c->pair_class[idx] = pair_class_of(c->sh_l[f], c->sh_l[s]);
        }
}

void shuffle_run(unsigned* v, size_t n, unsigned long long seed) { shuffle_seq(v, n, seed); }
void shuffle_units(int4* v, size_t n, unsigned long long seed) { shuffle_seq(v, n, seed); }

int pair_index(int a, int b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

// Primitive screening: a primitive pair whose e times the largest e of any
// primitive pair stays below pthr bounds every primitive quartet it could
// enter (2 pi^(5/2) e_ab e_cd >= |(ab|cd)| for the s-type prefactor), so it
// is removed from its shell pair's records once, at prepare time.
void drop_primitive_pairs(dsscf_ctx* c, double pthr)
{
    double emax = 0.0;
    for (const PrimPair& p : c->prims) emax = std::max(emax, p.e);
    std::vector<PrimPair> kept;
    kept.reserve(c->prims.size());
    for (ShellPair& P : c->pairs) {
        const int off = (int)kept.size();
        for (int i = 0; i < P.nprim; ++i) {
            const PrimPair& r = c->prims[P.prim_off + i];
            if (r.e * emax >= pthr) kept.push_back(r);
        }
        P.prim_off = off;
        P.nprim = (int)kept.size() - off;
    }
    c->prims.swap(kept);
}


}  // namespace

int dsscf_prepare(dsscf_ctx* c)
{
    if (!c) return fail(DSSCF_ERR_ARG, "dsscf_prepare: null context");
    if (!c->have_basis) return fail(DSSCF_ERR_STATE, "dsscf_prepare: no basis set");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->free_geometry();
    build_pairs(c);
    if (c->opt.prim_threshold > 0.0) drop_primitive_pairs(c, c->opt.prim_threshold / kEriScale);
    const size_t npair = c->pairs.size();
    const int N = c->n_ao;
    if (int rc = dev_alloc(&c->d_pairs, npair)) return rc;
    if (int rc = dev_alloc(&c->d_pp, c->prims.size())) return rc;
    if (int rc = dev_alloc(&c->d_Q, npair)) return rc;
    if (int rc = dev_alloc(&c->d_plist, npair)) return rc;
    const size_t nn = (size_t)N * N, np = (size_t)N * (N + 1) / 2;
    for (double** p : {&c->d_Dt, &c->d_Da})
        if (int rc = dev_alloc(p, nn)) return rc;
    for (double** p : {&c->d_Pt_pk, &c->d_Pa_pk, &c->d_J_pk, &c->d_K_pk})
        if (int rc = dev_alloc(p, np)) return rc;
    HIP_TRY(hipMemcpyAsync(c->d_pairs, c->pairs.data(), sizeof(ShellPair) * npair, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_pp, c->prims.data(), sizeof(PrimPair) * c->prims.size(), hipMemcpyHostToDevice,
                           c->stream));
    dsscf_upload_xform(c->xform.data(), (int)c->xform.size(), c->stream);
    if (int rc = dev_alloc(&c->d_xform, c->xform.size())) return rc;
    HIP_TRY(hipMemcpyAsync(c->d_xform, c->xform.data(), sizeof(double) * c->xform.size(), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipGetLastError());

    // Schwarz bounds, one launch per pair class.
    const int maxl = dk_max_l();
    const int npc = (maxl + 1) * (maxl + 2) / 2;
    std::vector<std::vector<int>> by_class(npc);
    for (size_t p = 0; p < npair; ++p) by_class[c->pair_class[p]].push_back((int)p);
    {
        std::vector<int> plist;
        std::vector<int> start(npc);
        for (int pc = 0; pc < npc; ++pc) {
            start[pc] = (int)plist.size();
            plist.insert(plist.end(), by_class[pc].begin(), by_class[pc].end());
        }
        HIP_TRY(hipMemcpyAsync(c->d_plist, plist.data(), sizeof(int) * plist.size(), hipMemcpyHostToDevice,
                               c->stream));
        c->pc_start = start;
        c->pc_count.assign(npc, 0);
        for (int pc = 0; pc < npc; ++pc) c->pc_count[pc] = (int)by_class[pc].size();
        for (int l1 = 0; l1 <= maxl; ++l1)
            for (int l2 = 0; l2 <= l1; ++l2) {
                const int pc = pair_class_of(l1, l2);
                HIP_TRY(dsscf_launch_schwarz(l1, l2, c->opt.boys_mode, c->opt.cartesian, c->d_plist + start[pc],
                                             (int)by_class[pc].size(), c->d_pairs, c->d_pp, c->d_btab, c->d_Q,
                                             kEriScale, c->stream));
            }
        c->Q.resize(npair);
        HIP_TRY(hipMemcpyAsync(c->Q.data(), c->d_Q, sizeof(double) * npair, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }

    // Sort each pair class by (primitive pairs desc, Q desc) and cut it into
    // groups of equal primitive-pair count.
    for (auto& v : by_class)
        std::sort(v.begin(), v.end(), [&](int x, int y) {
            if (c->pairs[x].nprim != c->pairs[y].nprim) return c->pairs[x].nprim > c->pairs[y].nprim;
            if (c->Q[x] != c->Q[y]) return c->Q[x] > c->Q[y];
            return x < y;
        });
    auto groups = [&](const std::vector<int>& v) {
        std::vector<int> g{0};
        for (size_t i = 1; i < v.size(); ++i)
            if (c->pairs[v[i]].nprim != c->pairs[v[i - 1]].nprim) g.push_back((int)i);
        g.push_back((int)v.size());
        return g;
    };

    const int ncls = dk_n_class();
    const double thr = c->opt.schwarz_threshold;
    const double thr32 = c->opt.fp32_threshold;
    const double item_flops = c->opt.work_item_flops == 0.0 ? DSSCF_WORK_ITEM_FLOPS : c->opt.work_item_flops;
    const double group_prims = c->opt.group_prims == 0.0 ? DSSCF_GROUP_PRIMS : c->opt.group_prims;

    // Work pairs: the shell pairs (same indices), then sub-pairs.  A sub-pair
    // is a shell pair restricted to a run of consecutive primitive pairs;
    // the primitive sum of an ERI splits over such runs, so the sub-quartets
    // of a shell quartet add up to it and digest into J and K independently.
    std::vector<ShellPair> work(c->pairs);
    std::vector<std::vector<std::pair<int, int>>> sub_cache(npair);  // (P, first work index)
    auto sub_pairs = [&](int p, int P, std::vector<unsigned>& out) {
        const ShellPair& sp = c->pairs[p];
        const int m = (sp.nprim + P - 1) / P;
        if (m <= 1) {
            out.push_back((unsigned)p);
            return;
        }
        int first = -1;
        for (const auto& e : sub_cache[p])
            if (e.first == P) first = e.second;
        if (first < 0) {
            first = (int)work.size();
            sub_cache[p].push_back({P, first});
            for (int k = 0; k < m; ++k) {  // m near-equal runs
                const int lo = (int)((long long)sp.nprim * k / m), hi = (int)((long long)sp.nprim * (k + 1) / m);
                ShellPair s = sp;
                s.prim_off = sp.prim_off + lo;
                s.nprim = hi - lo;
                work.push_back(s);
            }
        }
        for (int k = 0; k < m; ++k) out.push_back((unsigned)(first + k));
    };

    c->chunk.assign(ncls, 512);
    for (auto* v : {&c->fill_lanes, &c->fill_slots, &c->fill_lane_trips, &c->fill_wave_trips}) v->assign(ncls, 0);
    c->u_maxbra.assign(ncls, 0);
    c->u_maxbra32.assign(ncls, 0);
    for (auto* v : {&c->q_off, &c->q_cnt, &c->q_prim, &c->q_off32, &c->q_cnt32, &c->q_prim32, &c->u_off, &c->u_cnt,
                    &c->u_off32, &c->u_cnt32})
        v->assign(ncls, 0);
    c->n_unscreened = 0;
    c->n_work_items = 0;

    // Per class: its pair classes, work-unit size and sub-pair run length.
    struct ClassJob {
        int X, Y, cls, P, kChunk;
        bool wave, split;
        double gp;  // lane-group target (primitive quartets per lane)
    };
    std::vector<ClassJob> cj;
    for (int X = 0; X < npc; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            ClassJob j{X, Y, X * (X + 1) / 2 + Y, 0, 0, false, false, group_prims};
            j.wave = dsscf_class_is_wave(j.cls) != 0;
            j.split = thr32 > 0.0 && !j.wave;
            const size_t nx = by_class[X].size(), ny = by_class[Y].size();
            const long long est = X == Y ? (long long)nx * (nx + 1) / 2 : (long long)nx * ny;
            c->n_unscreened += est;
            // kets per work unit: enough units to fill the chip (~2048), 64..512, multiple of 64
            j.kChunk = (int)std::min<long long>(512, std::max<long long>(64, (est / 2048 + 63) / 64 * 64));
            if (!c->sched_chunk.empty() && c->sched_chunk[j.cls] > 0)
                j.kChunk = std::min(kMaxKetsPerUnit, std::max(64, (c->sched_chunk[j.cls] + 63) / 64 * 64));
            if (!c->sched_gp.empty() && c->sched_gp[j.cls] != 0.0) j.gp = c->sched_gp[j.cls];
            c->chunk[j.cls] = j.kChunk;
            // primitive pairs per sub-pair, both sides: P^2 primitive quartets ~ item_flops,
            // and at least kMinSubPrims (shorter runs cost more in per-item digestion and
            // J/K atomics than they gain in balance)
            int l4[4];
            dk_class_l(j.cls, l4);
            const double fq = prim_quartet_flops(l4[0], l4[1], l4[2], l4[3]);
            const double wif = !c->sched_wif.empty() && c->sched_wif[j.cls] != 0.0 ? c->sched_wif[j.cls] : item_flops;
            j.P = wif < 0.0 ? (1 << 30) : std::max(kMinSubPrims, (int)std::sqrt(wif / fq));
            cj.push_back(j);
        }
    // every sub-pair any class will use, created before the parallel part
    for (const ClassJob& j : cj) {
        if (j.wave) continue;
        std::vector<unsigned> tmp;
        for (int pc : {j.X, j.Y})
            for (int pidx : by_class[pc]) {
                tmp.clear();
                sub_pairs(pidx, j.P, tmp);
            }
    }
    auto sub_lookup = [&](int p, int P, std::vector<unsigned>& out) {  // read-only from here on
        const int m = (c->pairs[p].nprim + P - 1) / P;
        if (m <= 1) {
            out.push_back((unsigned)p);
            return;
        }
        int first = -1;
        for (const auto& e : sub_cache[p])
            if (e.first == P) first = e.second;
        for (int k = 0; k < m; ++k) out.push_back((unsigned)(first + k));
    };

    // Tasks: a class's bras cut into ranges of about equal quartet counts; each
    // task builds its own lists, merged below in class and range order.
    struct Task {
        int job, i0, i1;
        std::vector<uint2> all;
        std::vector<int4> units, units32;  // ket offsets local to `kets`
        std::vector<unsigned> kets;
        long long q_cnt = 0, q_cnt32 = 0, q_prim = 0, q_prim32 = 0, items = 0;
    };
    std::vector<Task> tasks;
    for (int jn = 0; jn < (int)cj.size(); ++jn) {
        const ClassJob& j = cj[jn];
        const int nx = (int)by_class[j.X].size();
        const long long ny = (long long)by_class[j.Y].size();
        const long long total = j.X == j.Y ? (long long)nx * (nx + 1) / 2 : (long long)nx * ny;
        const long long per = std::max<long long>(1 << 18, total / 64);
        long long acc = 0;
        int i0 = 0;
        for (int i = 0; i < nx; ++i) {
            acc += j.X == j.Y ? (long long)i + 1 : ny;
            if (acc >= per || i == nx - 1) {
                Task t;
                t.job = jn;
                t.i0 = i0;
                t.i1 = i + 1;
                tasks.push_back(std::move(t));
                i0 = i + 1;
                acc = 0;
            }
        }
    }
    const bool by_distance = c->opt.ket_order == 2 ||
                             (c->opt.ket_order == 0 && (double)c->n_unscreened >= DSSCF_KET_DISTANCE_QUARTETS);
    auto run_task = [&](Task& t) {
        const ClassJob& j = cj[t.job];
        const std::vector<int>& SX = by_class[j.X];
        const std::vector<int>& SY = by_class[j.Y];
        const std::vector<int> GY = groups(SY);
        std::vector<unsigned> k64, k32, bsub, ksub;
        std::vector<std::pair<double, unsigned>> byd;
        for (int i = t.i0; i < t.i1; ++i) {
            const int bi = SX[i];
            k64.clear();
            k32.clear();
            for (size_t gk = 0; gk + 1 < GY.size(); ++gk) {
                const int jend = j.X == j.Y ? std::min<int>(GY[gk + 1], i + 1) : GY[gk + 1];
                for (int jj = GY[gk]; jj < jend; ++jj) {
                    const int kj = SY[jj];
                    const double bound = c->Q[bi] * c->Q[kj];
                    if (thr > 0.0 && bound < thr) break;  // Q sorted desc in a group
                    const long long np = (long long)c->pairs[bi].nprim * c->pairs[kj].nprim;
                    if (j.split && bound < thr32) {
                        k32.push_back((unsigned)kj);
                        t.q_prim32 += np;
                    } else {
                        k64.push_back((unsigned)kj);
                        t.q_prim += np;
                    }
                }
            }
            if (j.wave) {
                for (unsigned kj : k64) t.all.push_back(make_uint2((unsigned)bi, kj));
                continue;
            }
            t.q_cnt += (long long)k64.size();
            t.q_cnt32 += (long long)k32.size();
            bsub.clear();
            sub_lookup(bi, j.P, bsub);
            for (int pass = 0; pass < 2; ++pass) {
                const std::vector<unsigned>& kk = pass ? k32 : k64;
                if (kk.empty()) continue;
                std::vector<int4>& uu = pass ? t.units32 : t.units;
                ksub.clear();
                for (unsigned kj : kk) sub_lookup((int)kj, j.P, ksub);
                // Equal primitive counts side by side: uniform loop trips in a
                // wavefront.  In a large context, inside each run of equal counts
                // by distance from the bra: neighbouring lanes then see similar
                // Boys arguments T, so a wavefront more often takes one branch
                // (T < 40 or the asymptotic form) instead of both (C20H42/def2-SVP
                // build -5%).  Below DSSCF_KET_DISTANCE_QUARTETS the sort costs
                // more host time at preparation than it saves on the device, and a
                // per-bra shuffle inside each run spreads the J_cd atomics of
                // contexts without slot records instead (ket_order chooses).
                if (!by_distance) {
                    std::stable_sort(ksub.begin(), ksub.end(),
                                     [&](unsigned x, unsigned y) { return work[x].nprim > work[y].nprim; });
                    for (size_t r0 = 0; r0 < ksub.size();) {
                        size_t r1 = r0 + 1;
                        while (r1 < ksub.size() && work[ksub[r1]].nprim == work[ksub[r0]].nprim) ++r1;
                        shuffle_run(ksub.data() + r0, r1 - r0,
                                    0x9e3779b97f4a7c15ULL * (unsigned long long)(bi + 1) + pass);
                        r0 = r1;
                    }
                } else {
                    const ShellPair& bp = work[bi];
                    const double bx = bp.A[0] - 0.5 * bp.AB[0], by = bp.A[1] - 0.5 * bp.AB[1],
                                 bz = bp.A[2] - 0.5 * bp.AB[2];
                    byd.clear();
                    for (unsigned k : ksub) {
                        const ShellPair& kp = work[k];
                        const double x = kp.A[0] - 0.5 * kp.AB[0] - bx, y = kp.A[1] - 0.5 * kp.AB[1] - by,
                                     z = kp.A[2] - 0.5 * kp.AB[2] - bz;
                        // more primitives first; the distance is below 2^40 bohr^2
                        byd.emplace_back((double)(1024 - kp.nprim) * 1099511627776.0 + (x * x + y * y + z * z), k);
                    }
                    std::sort(byd.begin(), byd.end());
                    for (size_t q = 0; q < ksub.size(); ++q) ksub[q] = byd[q].second;
                }
                const int k0 = (int)t.kets.size();
                t.kets.insert(t.kets.end(), ksub.begin(), ksub.end());
                for (unsigned sb : bsub) {
                    t.items += (long long)ksub.size();
                    // kets come by primitive count, so lanes per quartet only fall along
                    // the list: a unit keeps one group size and about kChunk lanes
                    for (size_t o = 0; o < ksub.size();) {
                        const int grp = quartet_group(work[sb].nprim, work[ksub[o]].nprim, j.gp);
                        const int G = (grp & 0xff) * (grp >> 8);
                        const size_t lim = std::max<size_t>(1, (size_t)j.kChunk / G);
                        size_t e = o + 1;
                        while (e < ksub.size() && e - o < lim &&
                               quartet_group(work[sb].nprim, work[ksub[e]].nprim, j.gp) == grp)
                            ++e;
                        const int cnt = (int)(e - o);
                        long long cost = 0;  // primitive quartets of the unit
                        for (int q = 0; q < cnt; ++q) cost += work[ksub[o + q]].nprim;
                        cost *= work[sb].nprim;
                        uu.push_back(make_int4((int)sb, k0 + (int)o, cnt, (int)std::min<long long>(cost, 0x7fffffff)));
                        o = e;
                    }
                }
            }
        }
    };
    {
        std::atomic<int> next{0};
        auto worker = [&] {
            for (int k; (k = next.fetch_add(1)) < (int)tasks.size();) run_task(tasks[k]);
        };
        const int nt = std::max(1, std::min<int>((int)std::thread::hardware_concurrency(), 16));
        std::vector<std::thread> th;
        for (int k = 1; k < std::min<int>(nt, (int)tasks.size()); ++k) th.emplace_back(worker);
        worker();
        for (auto& x : th) x.join();
    }

    // merge in class order: per class its FP64 units (heaviest first), then its FP32 units
    std::vector<uint2> all;  // flat lists (wave-cooperative classes)
    std::vector<int4> units, units32;
    std::vector<unsigned> kets;
    auto by_cost = [](const int4& x, const int4& y) { return x.w > y.w; };
    size_t tk = 0;
    for (int jn = 0; jn < (int)cj.size(); ++jn) {
        const int cls = cj[jn].cls;
        c->q_off[cls] = (long long)all.size();
        const long long u0 = (long long)units.size();
        units32.clear();
        for (; tk < tasks.size() && tasks[tk].job == jn; ++tk) {
            Task& t = tasks[tk];
            const int kbase = (int)kets.size();
            if (kets.size() + t.kets.size() > 0x7fffffffULL) return fail(DSSCF_ERR_UNSUPPORTED, "too many quartets");
            kets.insert(kets.end(), t.kets.begin(), t.kets.end());
            for (int4 u : t.units) units.push_back(make_int4(u.x, u.y + kbase, u.z, u.w));
            for (int4 u : t.units32) units32.push_back(make_int4(u.x, u.y + kbase, u.z, u.w));
            all.insert(all.end(), t.all.begin(), t.all.end());
            c->q_cnt[cls] += t.q_cnt;
            c->q_cnt32[cls] += t.q_cnt32;
            c->q_prim[cls] += t.q_prim;
            c->q_prim32[cls] += t.q_prim32;
            c->n_work_items += t.items;
            std::vector<unsigned>().swap(t.kets);
            std::vector<int4>().swap(t.units);
            std::vector<int4>().swap(t.units32);
            std::vector<uint2>().swap(t.all);
        }
        if (cj[jn].wave) c->q_cnt[cls] = (long long)all.size() - c->q_off[cls];
        // heaviest units first, so that the longest blocks start at once; units of
        // equal cost (often chunks of one bra, which flush the same K rows) in
        // shuffled order
        shuffle_units(units.data() + u0, units.size() - (size_t)u0, 0x5851f42d4c957f2dULL + cls);
        shuffle_units(units32.data(), units32.size(), 0x14057b7ef767814fULL + cls);
        std::stable_sort(units.begin() + u0, units.end(), by_cost);
        std::stable_sort(units32.begin(), units32.end(), by_cost);
        // the device reads the lanes per quartet where the sort key was
        auto set_group = [&](int4& u) {
            u.w = quartet_group(work[u.x].nprim, work[kets[u.y]].nprim, cj[jn].gp);
        };
        for (size_t k = (size_t)u0; k < units.size(); ++k) set_group(units[k]);
        for (int4& u : units32) set_group(u);
        // wavefront fill of the lists (diagnostics for dsscf_get_stats, at
        // 64-lane granularity): lanes the units keep busy against the lanes of
        // the wavefronts they occupy, and lane-trips of the primitive loops
        // against the trips the wavefronts run (each runs its longest lane)
        auto fill = [&](const int4& u) {
            const int gb = u.w & 0xff, gk = (u.w >> 8) & 0xff, G = gb * gk;
            const int per_wave = std::max(1, 64 / G);
            const long long tb = (work[u.x].nprim + gb - 1) / gb;
            for (int w0 = 0; w0 < u.z; w0 += per_wave) {
                const int w1 = std::min(u.z, w0 + per_wave);
                long long mx = 0;
                for (int q = w0; q < w1; ++q) {
                    const long long trips = tb * ((work[kets[u.y + q]].nprim + gk - 1) / gk);
                    c->fill_lane_trips[cls] += trips * G;
                    mx = std::max(mx, trips);
                }
                c->fill_wave_trips[cls] += mx * 64;
                c->fill_lanes[cls] += (long long)(w1 - w0) * G;
                c->fill_slots[cls] += 64;
            }
        };
        for (size_t k = (size_t)u0; k < units.size(); ++k) fill(units[k]);
        for (const int4& u : units32) fill(u);
        c->u_maxbra[cls] = c->u_maxbra32[cls] = 0;
        for (size_t k = (size_t)u0; k < units.size(); ++k)
            c->u_maxbra[cls] = std::max(c->u_maxbra[cls], work[units[k].x].nprim);
        for (const int4& u : units32) c->u_maxbra32[cls] = std::max(c->u_maxbra32[cls], work[u.x].nprim);
        c->u_off[cls] = u0;
        c->u_cnt[cls] = (long long)units.size() - u0;
        c->u_off32[cls] = (long long)units.size();
        c->u_cnt32[cls] = (long long)units32.size();
        units.insert(units.end(), units32.begin(), units32.end());
    }
    if (work.size() > 0x7fffffffULL) return fail(DSSCF_ERR_UNSUPPORTED, "too many sub-pairs");
    if (work.size() > npair) {  // the JK kernels index the work pairs; Schwarz and ERI batches the first npair
        dev_free(c->d_pairs);
        if (int rc = dev_alloc(&c->d_pairs, work.size())) return rc;
        HIP_TRY(hipMemcpyAsync(c->d_pairs, work.data(), sizeof(ShellPair) * work.size(), hipMemcpyHostToDevice,
                               c->stream));
    }
    c->n_work_pairs = (long long)work.size();
    {
        // J/K copies: about one per 64 quartets that meet each element of J in a
        // build, as a power of two up to kJKSlicesMax (with the J_cd slot
        // records only the J_ab and exchange-row adds remain atomic: one per
        // 32 had been the better choice when J_cd was atomic too)
        long long items = c->n_work_items;
        for (int k = 0; k < ncls; ++k)
            if (dsscf_class_is_wave(k)) items += c->q_cnt[k];
        const double per_elem = (double)items / std::max<double>(1.0, (double)N * (N + 1) / 2);
        c->n_slice = 1;
        while (c->n_slice < kJKSlicesMax && c->n_slice * 64.0 < per_elem) c->n_slice *= 2;
    }
    for (double** p : {&c->d_Jh, &c->d_Kh})
        if (int rc = dev_alloc(p, nn * c->n_slice)) return rc;
    if (kets.size() > 0xffffffffULL) return fail(DSSCF_ERR_UNSUPPORTED, "too many quartets");
    if (int rc = dev_alloc(&c->d_units, units.size())) return rc;
    if (int rc = dev_alloc(&c->d_kets, kets.size())) return rc;
    if (!units.empty())
        HIP_TRY(hipMemcpyAsync(c->d_units, units.data(), sizeof(int4) * units.size(), hipMemcpyHostToDevice, c->stream));
    if (!kets.empty())
        HIP_TRY(hipMemcpyAsync(c->d_kets, kets.data(), sizeof(unsigned) * kets.size(), hipMemcpyHostToDevice, c->stream));
    {
        // J_cd slot records: a context whose records fit jcd_max_bytes stores
        // each quartet's J_cd terms (k_jk) into a record laid out by ket AO
        // block and sums each block's records per element (k_jcd); a larger
        // one keeps the atomic adds.
        dev_free(c->d_js);
        dev_free(c->d_us);
        dev_free(c->d_jd);
        dev_free(c->d_jrr);
        dev_free(c->d_jri);
        c->jcd_rows = 0;
        c->jcd_bytes = 0;
        std::vector<int> us(units.size(), 0);
        std::vector<int> cls_ncnd(ncls, 0);
        long long nd = 0, ne = 0;
        for (int cls = 0; cls < ncls; ++cls) {
            if (dsscf_class_is_wave(cls)) continue;
            int l4[4];
            dk_class_l(cls, l4);
            cls_ncnd[cls] = nfun(l4[2], c->opt.cartesian) * nfun(l4[3], c->opt.cartesian);
            for (long long ui = c->u_off[cls]; ui < c->u_off32[cls] + c->u_cnt32[cls]; ++ui) {
                us[ui] = (int)std::min<long long>(ne, 0x7fffffff);
                nd += (long long)units[ui].z * cls_ncnd[cls];
                ne += units[ui].z;
            }
        }
        const double bytes = 8.0 * (double)nd + 4.0 * (double)ne + 4.0 * (double)units.size();
        const double cap = c->opt.jcd_max_bytes == 0.0 ? DSSCF_JCD_MAX_BYTES : c->opt.jcd_max_bytes;
        if (ne > 0 && bytes <= cap && nd < 0x7fffffffLL) {
            const int N = c->n_ao;
            std::vector<int> row_of((size_t)N * N, -1), slot_row((size_t)ne), jd((size_t)ne);
            std::vector<int4> ri;
            std::vector<int2> rr;
            size_t slot = 0;
            for (int cls = 0; cls < ncls; ++cls) {  // rows and their record counts
                if (!cls_ncnd[cls]) continue;
                int l4[4];
                dk_class_l(cls, l4);
                for (long long ui = c->u_off[cls]; ui < c->u_off32[cls] + c->u_cnt32[cls]; ++ui) {
                    const int4 u = units[ui];
                    for (int t = 0; t < u.z; ++t, ++slot) {
                        const ShellPair& kp = work[kets[(size_t)u.y + t]];
                        int& r = row_of[(size_t)kp.ao_a * N + kp.ao_b];
                        if (r < 0) {
                            r = (int)ri.size();
                            ri.push_back(make_int4(kp.ao_a, kp.ao_b, nfun(l4[2], c->opt.cartesian),
                                                   nfun(l4[3], c->opt.cartesian)));
                            rr.push_back(make_int2(0, 0));
                        }
                        ++rr[r].y;
                        slot_row[slot] = r;
                    }
                }
            }
            int first = 0;
            for (size_t r = 0; r < rr.size(); ++r) {
                rr[r].x = first;
                first += rr[r].y * ri[r].z * ri[r].w;
            }
            std::vector<int> fill(rr.size(), 0);
            for (size_t q = 0; q < (size_t)ne; ++q) {
                const int r = slot_row[q];
                jd[q] = rr[r].x + fill[r]++ * ri[r].z * ri[r].w;
            }
            c->jcd_rows = (int)ri.size();
            if (int rc = dev_alloc(&c->d_js, (size_t)nd)) return rc;
            if (int rc = dev_alloc(&c->d_us, us.size())) return rc;
            if (int rc = dev_alloc(&c->d_jd, jd.size())) return rc;
            if (int rc = dev_alloc(&c->d_jrr, rr.size())) return rc;
            if (int rc = dev_alloc(&c->d_jri, ri.size())) return rc;
            HIP_TRY(hipMemcpyAsync(c->d_us, us.data(), sizeof(int) * us.size(), hipMemcpyHostToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(c->d_jd, jd.data(), sizeof(int) * jd.size(), hipMemcpyHostToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(c->d_jrr, rr.data(), sizeof(int2) * rr.size(), hipMemcpyHostToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(c->d_jri, ri.data(), sizeof(int4) * ri.size(), hipMemcpyHostToDevice, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
            c->jcd_bytes = (long long)bytes + (long long)((sizeof(int2) + sizeof(int4)) * rr.size());
        }
    }
    if (int rc = dev_alloc(&c->d_q, all.size())) return rc;
    if (!all.empty())
        HIP_TRY(hipMemcpyAsync(c->d_q, all.data(), sizeof(uint2) * all.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if ((int)c->block_size.size() != ncls) c->block_size.assign(ncls, 256);
    c->device_bytes = (long long)(sizeof(ShellPair) * work.size() + sizeof(PrimPair) * c->prims.size() +
                                  (sizeof(double) + sizeof(int)) * npair + sizeof(double) * nn * (2 + 2 * c->n_slice) +
                                  sizeof(double) * np * 4 + sizeof(int4) * units.size() +
                                  sizeof(unsigned) * kets.size() + sizeof(uint2) * all.size() + c->jcd_bytes);
    c->prepared = true;
    return DSSCF_OK;
}

namespace {

// Build `p` over contexts ctxs[0..n): their packed inputs/outputs at
// dPt/dPa/dJ/dK (device), block sizes per class from bs (nullptr: each
// class's default of 256 or the first context's tuned value).  dPb/dKb
// (both null, or arrays whose entry k is null for a restricted member): the
// packed beta density and K[Pb] output of unrestricted members, whose beta
// buffers (d_Db, d_Kbh) must exist.
int plan_build(Plan& p, dsscf_ctx* const* ctxs, int n, const double* const* dPt, const double* const* dPa,
               double* const* dJ, double* const* dK, const int* bs, hipStream_t stream,
               const double* const* dPb = nullptr, double* const* dKb = nullptr)
{
    p.reset();
    const int ncls = dk_n_class();
    p.ncls = ncls;
    p.boys_mode = ctxs[0]->opt.boys_mode;
    p.cart = ctxs[0]->opt.cartesian;
    auto is_u = [&](int k) { return dPb && dKb && dPb[k] && dKb[k]; };
    p.nks = 1;
    for (int k = 0; k < n; ++k)
        if (is_u(k)) p.nks = 2;
    p.btab = ctxs[0]->d_btab;
    const int nslot = 2 * ncls;  // slot 2*cls: FP64 list, 2*cls+1: FP32 list
    p.cls_nseg.assign(nslot, 0);
    p.cls_nblocks.assign(nslot, 0);
    p.cls_off.assign(nslot, 0);
    p.cls_bs.assign(nslot, 256);
    p.cls_maxN.assign(nslot, 0);
    p.cls_maxbra.assign(nslot, 0);
    p.cls_cost.assign(nslot, 0.0);
    p.cls_stream.assign(nslot, 0);
    p.h_segs.clear();
    p.h_blk0.clear();
    p.h_mat.clear();
    for (int slot = 0; slot < nslot; ++slot) {
        const int cls = slot / 2, prec32 = slot % 2;
        const bool wave = dsscf_class_is_wave(cls) != 0;
        int b = wave ? 64 : (bs ? bs[cls] : (ctxs[0]->block_size.empty() ? 256 : ctxs[0]->block_size[cls]));
        if (!wave) {  // no more threads per block than the largest work unit
            int maxchunk = 64;
            for (int k = 0; k < n; ++k) maxchunk = std::max(maxchunk, ctxs[k]->chunk.empty() ? 512 : ctxs[k]->chunk[cls]);
            b = std::min(b, maxchunk);
        }
        p.cls_bs[slot] = b;
        p.cls_off[slot] = (int)p.h_segs.size();
        long long blocks = 0;
        int slot_maxN = 0;
        for (int k = 0; k < n; ++k) {
            dsscf_ctx* c = ctxs[k];
            const long long cnt = wave ? (prec32 ? 0 : c->q_cnt[cls]) : (prec32 ? c->u_cnt32[cls] : c->u_cnt[cls]);
            if (cnt <= 0) continue;
            if (cnt > 0x7fffffffLL) return fail(DSSCF_ERR_UNSUPPORTED, "work list longer than 2^31");
            JKSeg sg{};
            sg.q = c->d_q + c->q_off[cls];
            sg.u = c->d_units + (prec32 ? c->u_off32[cls] : c->u_off[cls]);
            sg.k = c->d_kets;
            slot_maxN = std::max(slot_maxN, c->n_ao);
            if (!wave) p.cls_maxbra[slot] = std::max(p.cls_maxbra[slot], prec32 ? c->u_maxbra32[cls] : c->u_maxbra[cls]);
            sg.pairs = c->d_pairs;
            sg.pp = c->d_pp;
            sg.Dt = c->d_Dt;
            sg.Da = c->d_Da;
            sg.Jh = c->d_Jh;
            sg.Kh = c->d_Kh;
            if (is_u(k)) {
                sg.Db = c->d_Db;
                sg.Kbh = c->d_Kbh;
            }
            sg.nslice = c->n_slice;
            if (!wave && c->d_js) {
                sg.js = c->d_js;
                sg.us = c->d_us + (prec32 ? c->u_off32[cls] : c->u_off[cls]);
                sg.jd = c->d_jd;
            }
            sg.n = (int)cnt;
            sg.N = c->n_ao;
            sg.blk0 = (int)blocks;
            p.cls_cost[slot] += (double)(prec32 ? c->q_prim32[cls] : c->q_prim[cls]);
            // one block per work unit, or per dsscf_wave_group quartets (wave classes)
            blocks += wave ? (cnt + dsscf_wave_group(cls, p.cart) - 1) / dsscf_wave_group(cls, p.cart) : cnt;
            if (blocks > 0x7fffffffLL) return fail(DSSCF_ERR_UNSUPPORTED, "launch grid too large");
            p.h_segs.push_back(sg);
            p.h_blk0.push_back(sg.blk0);
        }
        p.cls_nseg[slot] = (int)p.h_segs.size() - p.cls_off[slot];
        p.cls_nblocks[slot] = (int)blocks;
        p.cls_maxN[slot] = slot_maxN;
        {
            int l4[4];
            dk_class_l(cls, l4);
            p.cls_cost[slot] *= prim_quartet_flops(l4[0], l4[1], l4[2], l4[3]) * (prec32 ? 0.5 : 1.0);
        }
        if (!wave && p.cls_nseg[slot]) {
            int l4[4];
            dk_class_l(cls, l4);
            const size_t lds =
                sizeof(double) * ((size_t)p.nks * (nfun(l4[0], p.cart) + nfun(l4[1], p.cart)) * slot_maxN + 1) +
                sizeof(PrimPair) * (size_t)p.cls_maxbra[slot];
            if (lds > 64 * 1024) return fail(DSSCF_ERR_UNSUPPORTED, "basis too large for the LDS exchange rows");
        }
    }
    {  // slots onto streams, heaviest first onto the least loaded
        std::vector<int> order(nslot);
        for (int i = 0; i < nslot; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return p.cls_cost[x] > p.cls_cost[y]; });
        std::vector<int> lsum(nslot);
        for (int slot = 0; slot < nslot; ++slot) {
            int l4[4];
            dk_class_l(slot / 2, l4);
            lsum[slot] = l4[0] + l4[1] + l4[2] + l4[3];
        }
        p.launch_order = order;
        std::stable_sort(p.launch_order.begin(), p.launch_order.end(), [&](int x, int y) { return lsum[x] > lsum[y]; });
        double load[kStreams] = {};
        for (int slot : order) {
            if (p.cls_nseg[slot] == 0) continue;
            int best = 0;
            for (int k = 1; k < kStreams; ++k)
                if (load[k] < load[best]) best = k;
            p.cls_stream[slot] = best;
            load[best] += p.cls_cost[slot];
        }
    }
    int row = 0, maxN = 0;
    for (int k = 0; k < n; ++k) {
        dsscf_ctx* c = ctxs[k];
        MatSeg m{};
        m.Ptp = dPt[k];
        m.Pap = dPa[k];
        m.Dt = c->d_Dt;
        m.Da = c->d_Da;
        m.Jh = c->d_Jh;
        m.Kh = c->d_Kh;
        m.Jp = dJ[k];
        m.Kp = dK[k];
        if (is_u(k)) {
            m.Pbp = dPb[k];
            m.Db = c->d_Db;
            m.Kbh = c->d_Kbh;
            m.Kbp = dKb[k];
        }
        m.N = c->n_ao;
        m.nslice = c->n_slice;
        m.row0 = row;
        row += c->n_ao;
        maxN = std::max(maxN, c->n_ao);
        p.h_mat.push_back(m);
    }
    p.nmat = n;
    p.total_rows = row;
    p.maxN = maxN;
    p.h_jcd.clear();
    p.h_jblk0.clear();
    p.jcd_blocks = 0;
    for (int k = 0; k < n; ++k) {
        dsscf_ctx* c = ctxs[k];
        if (!c->d_js || c->jcd_rows == 0) continue;
        JcdSeg js{};
        js.rr = c->d_jrr;
        js.ri = c->d_jri;
        js.js = c->d_js;
        js.Jh = c->d_Jh;
        js.N = c->n_ao;
        js.nrow = c->jcd_rows;
        js.blk0 = p.jcd_blocks;
        p.jcd_blocks += c->jcd_rows;
        p.h_jcd.push_back(js);
        p.h_jblk0.push_back(js.blk0);
    }
    auto grow = [](auto** ptr, size_t& cap, size_t need) -> int {
        if (need <= cap && *ptr) return DSSCF_OK;
        dev_free(*ptr);
        cap = 0;
        const size_t n = std::max<size_t>(need + need / 2, 16);
        if (int rc = dev_alloc(ptr, n)) return rc;
        cap = n;
        return DSSCF_OK;
    };
    if (int rc = grow(&p.d_segs, p.cap_segs, p.h_segs.size())) return rc;
    if (int rc = grow(&p.d_blk0, p.cap_blk0, p.h_blk0.size())) return rc;
    if (int rc = grow(&p.d_mat, p.cap_mat, p.h_mat.size())) return rc;
    if (!p.h_jcd.empty()) {
        if (int rc = grow(&p.d_jcd, p.cap_jcd, p.h_jcd.size())) return rc;
        if (int rc = grow(&p.d_jblk0, p.cap_jblk0, p.h_jblk0.size())) return rc;
        HIP_TRY(hipMemcpyAsync(p.d_jcd, p.h_jcd.data(), sizeof(JcdSeg) * p.h_jcd.size(), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(p.d_jblk0, p.h_jblk0.data(), sizeof(int) * p.h_jblk0.size(), hipMemcpyHostToDevice,
                               stream));
    }
    if (!p.h_segs.empty()) {
        HIP_TRY(hipMemcpyAsync(p.d_segs, p.h_segs.data(), sizeof(JKSeg) * p.h_segs.size(), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(p.d_blk0, p.h_blk0.data(), sizeof(int) * p.h_blk0.size(), hipMemcpyHostToDevice, stream));
    }
    HIP_TRY(hipMemcpyAsync(p.d_mat, p.h_mat.data(), sizeof(MatSeg) * p.h_mat.size(), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    p.built = true;
    return DSSCF_OK;
}

int plan_launch_slot(Plan& p, int slot, hipStream_t s)
{
    if (p.cls_nseg[slot] == 0) return DSSCF_OK;
    HIP_TRY(dsscf_launch_jk(slot / 2, p.boys_mode, slot % 2, p.d_segs + p.cls_off[slot], p.d_blk0 + p.cls_off[slot],
                            p.cls_nseg[slot], p.cls_nblocks[slot], p.cls_bs[slot], p.cls_maxN[slot], p.nks,
                            p.cls_maxbra[slot], p.cart, p.btab, kEriScale, s));
    return DSSCF_OK;
}

int plan_launch_class(Plan& p, int cls, hipStream_t s)
{
    if (int rc = plan_launch_slot(p, 2 * cls, s)) return rc;
    return plan_launch_slot(p, 2 * cls + 1, s);
}

// The plan's side streams and fork/join events (created once, outside any
// stream capture).
int plan_streams(Plan& p)
{
    if (p.fork) return DSSCF_OK;
    HIP_TRY(hipEventCreateWithFlags(&p.fork, hipEventDisableTiming));
    for (int i = 0; i < kStreams - 1; ++i) {
        HIP_TRY(hipStreamCreateWithFlags(&p.side[i], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&p.join[i], hipEventDisableTiming));
    }
    return DSSCF_OK;
}

// prepare (unpack densities, zero accumulators), all classes, fold.  With
// ev (per-class timing) every class runs on s in turn; otherwise the classes
// fork onto the plan's streams and join before the fold, and with timed each
// slot's kernel is bracketed by p.slot_ev on the stream it runs on.
int plan_enqueue(Plan& p, hipStream_t s, std::vector<hipEvent_t>* ev, bool timed = false)
{
    HIP_TRY(dsscf_launch_prep(p.d_mat, p.nmat, p.total_rows, p.maxN, s));
    if (ev) {
        HIP_TRY(hipEventRecord((*ev)[0], s));
        for (int cls = 0; cls < p.ncls; ++cls) {
            if (int rc = plan_launch_class(p, cls, s)) return rc;
            HIP_TRY(hipEventRecord((*ev)[cls + 1], s));
        }
    } else {
        if (int rc = plan_streams(p)) return rc;
        bool used[kStreams] = {};
        for (int slot = 0; slot < 2 * p.ncls; ++slot)
            if (p.cls_nseg[slot] > 0) used[p.cls_stream[slot]] = true;
        HIP_TRY(hipEventRecord(p.fork, s));
        for (int i = 1; i < kStreams; ++i)
            if (used[i]) HIP_TRY(hipStreamWaitEvent(p.side[i - 1], p.fork, 0));
        if (timed && p.slot_ev.empty()) {
            p.slot_ev.assign(4 * p.ncls, nullptr);
            for (hipEvent_t& e : p.slot_ev) HIP_TRY(hipEventCreate(&e));
        }
        for (int slot : p.launch_order) {
            const int k = p.cls_stream[slot];
            hipStream_t st = k ? p.side[k - 1] : s;
            if (timed && p.cls_nseg[slot]) HIP_TRY(hipEventRecord(p.slot_ev[2 * slot], st));
            if (int rc = plan_launch_slot(p, slot, st)) return rc;
            if (timed && p.cls_nseg[slot]) HIP_TRY(hipEventRecord(p.slot_ev[2 * slot + 1], st));
        }
        for (int i = 1; i < kStreams; ++i)
            if (used[i]) {
                HIP_TRY(hipEventRecord(p.join[i - 1], p.side[i - 1]));
                HIP_TRY(hipStreamWaitEvent(s, p.join[i - 1], 0));
            }
    }
    if (p.jcd_blocks > 0)
        HIP_TRY(dsscf_launch_jcd(p.d_jcd, p.d_jblk0, (int)p.h_jcd.size(), p.jcd_blocks, s));
    HIP_TRY(dsscf_launch_fold(p.d_mat, p.nmat, p.total_rows, p.maxN, s));
    return DSSCF_OK;
}

// Run one fused Fock build; with use_graph the launch sequence is captured
// once into a hipGraph (kernels only) and replayed.
int plan_run(Plan& p, hipStream_t s, bool use_graph, std::vector<float>* cls_ms)
{
    if (cls_ms) {
        // every event created here is destroyed on every path
        std::vector<hipEvent_t> ev(p.ncls + 1, nullptr);
        hipError_t e = hipSuccess;
        for (auto& x : ev)
            if (e == hipSuccess) e = hipEventCreate(&x);
        int rc = e == hipSuccess ? plan_enqueue(p, s, &ev) : DSSCF_OK;
        if (e == hipSuccess && !rc) {
            e = hipEventSynchronize(ev[p.ncls]);
            cls_ms->assign(p.ncls, 0.0f);
            for (int k = 0; k < p.ncls && e == hipSuccess; ++k) e = hipEventElapsedTime(&(*cls_ms)[k], ev[k], ev[k + 1]);
        }
        for (auto& x : ev)
            if (x) (void)hipEventDestroy(x);
        if (e != hipSuccess) return fail(DSSCF_ERR_HIP, std::string("plan_run (timed): ") + hipGetErrorString(e));
        return rc;
    }
    if (!use_graph) return plan_enqueue(p, s, nullptr);
    if (!p.gexec) {
        if (int rc = plan_streams(p)) return rc;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = plan_enqueue(p, s, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(DSSCF_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        p.graph = g;
        HIP_TRY(hipGraphInstantiate(&p.gexec, p.graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(p.gexec, s));
    return DSSCF_OK;
}

int ctx_plan(dsscf_ctx* c)
{
    if (c->plan.built) return DSSCF_OK;
    const double* pt = c->d_Pt_pk;
    const double* pa = c->d_Pa_pk;
    double* j = c->d_J_pk;
    double* k = c->d_K_pk;
    return plan_build(c->plan, &c, 1, &pt, &pa, &j, &k, nullptr, c->stream);
}

// One Fock build on the context's stream: a replay of the captured build
// (per-class timing runs the launches eagerly instead).
int run_jk(dsscf_ctx* c, std::vector<float>* cls_ms)
{
    if (int rc = ctx_plan(c)) return rc;
    return plan_run(c->plan, c->stream, cls_ms == nullptr, cls_ms);
}

// The beta buffers of an unrestricted build and its plan (the restricted
// plan and its graph stay as they are).
int ctx_plan_u(dsscf_ctx* c)
{
    const size_t N = (size_t)c->n_ao, nn = N * N, np = N * (N + 1) / 2;
    if (!c->d_Db) {
        if (int rc = dev_alloc(&c->d_Pb_pk, np)) return rc;
        if (int rc = dev_alloc(&c->d_Kb_pk, np)) return rc;
        if (int rc = dev_alloc(&c->d_Db, nn)) return rc;
        if (int rc = dev_alloc(&c->d_Kbh, nn * c->n_slice)) return rc;
    }
    if (c->plan_u.built) return DSSCF_OK;
    const double* pt = c->d_Pt_pk;
    const double* pa = c->d_Pa_pk;
    const double* pb = c->d_Pb_pk;
    double* j = c->d_J_pk;
    double* k = c->d_K_pk;
    double* kb = c->d_Kb_pk;
    return plan_build(c->plan_u, &c, 1, &pt, &pa, &j, &k, nullptr, c->stream, &pb, &kb);
}

}  // namespace

int dsscf_jk_uhf(dsscf_ctx* c, const double* Pt, const double* Pa, const double* Pb, double* J, double* Ka,
                 double* Kb)
{
    if (!c || !Pt || !Pa || !Pb || !J || !Ka || !Kb) return fail(DSSCF_ERR_ARG, "dsscf_jk_uhf: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_jk_uhf: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = ctx_plan_u(c)) return rc;
    const size_t np = (size_t)c->n_ao * (c->n_ao + 1) / 2;
    HIP_TRY(hipMemcpyAsync(c->d_Pt_pk, Pt, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_Pa_pk, Pa, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_Pb_pk, Pb, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    if (int rc = plan_run(c->plan_u, c->stream, true, nullptr)) return rc;
    HIP_TRY(hipMemcpyAsync(J, c->d_J_pk, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(Ka, c->d_K_pk, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(Kb, c->d_Kb_pk, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DSSCF_OK;
}

int dsscf_jk_device(dsscf_ctx* c, const double* d_Pt, const double* d_Pa, double* d_J, double* d_K)
{
    if (!c || !d_Pt || !d_Pa || !d_J || !d_K) return fail(DSSCF_ERR_ARG, "dsscf_jk_device: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_jk_device: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    const size_t np = (size_t)c->n_ao * (c->n_ao + 1) / 2;
    HIP_TRY(hipMemcpyAsync(c->d_Pt_pk, d_Pt, sizeof(double) * np, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_Pa_pk, d_Pa, sizeof(double) * np, hipMemcpyDeviceToDevice, c->stream));
    if (int rc = run_jk(c, nullptr)) return rc;
    HIP_TRY(hipMemcpyAsync(d_J, c->d_J_pk, sizeof(double) * np, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_K, c->d_K_pk, sizeof(double) * np, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DSSCF_OK;
}

int dsscf_boys_eval(dsscf_ctx* c, int mode, int L, int n, const double* T, double* F)
{
    if (!c || n < 0 || (n > 0 && (!T || !F))) return fail(DSSCF_ERR_ARG, "dsscf_boys_eval: bad argument");
    if (mode < 0 || mode > 2 || L < 0 || L > 8) return fail(DSSCF_ERR_ARG, "dsscf_boys_eval: mode 0..2, L 0..8");
    if (n == 0) return DSSCF_OK;
    HIP_TRY(hipSetDevice(c->device));
    double *dT = nullptr, *dF = nullptr;
    if (int rc = dev_alloc(&dT, (size_t)n)) return rc;
    if (int rc = dev_alloc(&dF, (size_t)n * (L + 1))) {
        dev_free(dT);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(dT, T, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = dsscf_launch_boys_eval(mode, L, dT, n, c->d_btab, dF, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(F, dF, sizeof(double) * n * (L + 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(dT);
    dev_free(dF);
    if (e != hipSuccess) return fail(DSSCF_ERR_HIP, std::string("dsscf_boys_eval: ") + hipGetErrorString(e));
    return DSSCF_OK;
}

int dsscf_eri_gradient(dsscf_ctx* c, const double* Pa, int natom, const int* ao_atom, double* grad)
{
    if (!c || !Pa || natom <= 0 || !ao_atom || !grad) return fail(DSSCF_ERR_ARG, "dsscf_eri_gradient: bad argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_eri_gradient: call dsscf_prepare first");
    const int N = c->n_ao;
    for (int i = 0; i < N; ++i)
        if (ao_atom[i] < 0 || ao_atom[i] >= natom) return fail(DSSCF_ERR_ARG, "dsscf_eri_gradient: AO atom index");
    if ((size_t)3 * natom * sizeof(double) > 48 * 1024)
        return fail(DSSCF_ERR_UNSUPPORTED, "dsscf_eri_gradient: more than 2048 atoms");
    const int ncls = dk_n_class();
    HIP_TRY(hipSetDevice(c->device));
    const size_t np = (size_t)N * (N + 1) / 2;
    std::vector<double> Pt(np);
    for (size_t i = 0; i < np; ++i) Pt[i] = 2.0 * Pa[i];
    HIP_TRY(hipMemcpyAsync(c->d_Pt_pk, Pt.data(), sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_Pa_pk, Pa, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    // full densities from the packed ones (the prepare kernel of a Fock build)
    MatSeg m{};
    m.Ptp = c->d_Pt_pk;
    m.Pap = c->d_Pa_pk;
    m.Dt = c->d_Dt;
    m.Da = c->d_Da;
    m.Jh = c->d_Jh;
    m.Kh = c->d_Kh;
    m.N = N;
    m.nslice = c->n_slice;
    // per class: up to two segments (the FP64 and the FP32 work units; both
    // run in FP64 here), or the flat quartet list of a wave-cooperative class
    std::vector<JKSeg> segs;
    std::vector<int> blk0, cls_of, nblk;
    for (int cls = 0; cls < ncls; ++cls)
        for (int pass = 0; pass < 2; ++pass) {
            const bool wave = dsscf_class_is_wave(cls) != 0;
            if (wave && pass) continue;
            const long long cnt = wave ? c->q_cnt[cls] : (pass ? c->u_cnt32[cls] : c->u_cnt[cls]);
            if (cnt <= 0) continue;
            JKSeg sg{};
            sg.q = wave ? c->d_q + c->q_off[cls] : nullptr;
            sg.u = c->d_units + (pass ? c->u_off32[cls] : c->u_off[cls]);
            sg.k = c->d_kets;
            sg.pairs = c->d_pairs;
            sg.pp = c->d_pp;
            sg.Dt = c->d_Dt;
            sg.Da = c->d_Da;
            sg.n = (int)cnt;
            sg.N = N;
            sg.blk0 = 0;
            segs.push_back(sg);
            blk0.push_back(0);
            cls_of.push_back(cls);
            nblk.push_back(wave ? (int)((cnt + dsscf_grad_flat_chunk() - 1) / dsscf_grad_flat_chunk()) : (int)cnt);
        }
    MatSeg* d_m = nullptr;
    JKSeg* d_segs = nullptr;
    int *d_blk0 = nullptr, *d_atom = nullptr;
    double* d_grad = nullptr;
    auto cleanup = [&] {
        dev_free(d_m);
        dev_free(d_segs);
        dev_free(d_blk0);
        dev_free(d_atom);
        dev_free(d_grad);
    };
    int rc = dev_alloc(&d_m, 1);
    if (!rc && !segs.empty()) rc = dev_alloc(&d_segs, segs.size());
    if (!rc && !segs.empty()) rc = dev_alloc(&d_blk0, blk0.size());
    if (!rc) rc = dev_alloc(&d_atom, (size_t)N);
    if (!rc) rc = dev_alloc(&d_grad, (size_t)3 * natom);
    if (rc) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_m, &m, sizeof(MatSeg), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !segs.empty())
        e = hipMemcpyAsync(d_segs, segs.data(), sizeof(JKSeg) * segs.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !segs.empty())
        e = hipMemcpyAsync(d_blk0, blk0.data(), sizeof(int) * blk0.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_atom, ao_atom, sizeof(int) * N, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_grad, 0, sizeof(double) * 3 * natom, c->stream);
    if (e == hipSuccess) e = dsscf_launch_prep(d_m, 1, N, N, c->stream);
    for (size_t k = 0; e == hipSuccess && k < segs.size(); ++k) {
        int l4[4];
        dk_class_l(cls_of[k], l4);
        if (!dsscf_grad_is_wave(l4))
            e = dsscf_launch_grad(l4, c->opt.boys_mode, d_segs + k, d_blk0 + k, 1, segs[k].n, c->d_btab, c->d_xform,
                                  d_atom, d_grad, natom, kEriScale, c->stream);
        else
            e = dsscf_launch_gradw(l4, c->opt.boys_mode, c->opt.cartesian, dsscf_class_is_wave(cls_of[k]), d_segs + k,
                                   d_blk0 + k, 1, nblk[k], c->d_btab, c->d_xform, d_atom, d_grad, natom, kEriScale,
                                   c->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(grad, d_grad, sizeof(double) * 3 * natom, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(DSSCF_ERR_HIP, std::string("dsscf_eri_gradient: ") + hipGetErrorString(e));
    return DSSCF_OK;
}

int dsscf_jk(dsscf_ctx* c, const double* Pt, const double* Pa, double* J, double* K)
{
    if (!c || !Pt || !Pa || !J || !K) return fail(DSSCF_ERR_ARG, "dsscf_jk: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_jk: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    const size_t np = (size_t)c->n_ao * (c->n_ao + 1) / 2;
    HIP_TRY(hipMemcpyAsync(c->d_Pt_pk, Pt, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_Pa_pk, Pa, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    if (int rc = run_jk(c, nullptr)) return rc;
    HIP_TRY(hipMemcpyAsync(J, c->d_J_pk, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(K, c->d_K_pk, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DSSCF_OK;
}

int dsscf_set_density(dsscf_ctx* c, const double* Pt, const double* Pa)
{
    if (!c || !Pt || !Pa) return fail(DSSCF_ERR_ARG, "dsscf_set_density: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_set_density: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    const size_t np = (size_t)c->n_ao * (c->n_ao + 1) / 2;
    HIP_TRY(hipMemcpyAsync(c->d_Pt_pk, Pt, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_Pa_pk, Pa, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DSSCF_OK;
}

int dsscf_jk_resident(dsscf_ctx* c)
{
    if (!c) return fail(DSSCF_ERR_ARG, "dsscf_jk_resident: null context");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_jk_resident: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = run_jk(c, nullptr)) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DSSCF_OK;
}

int dsscf_jk_resident_timed(dsscf_ctx* c, double* kernel_ms)
{
    if (!c || !kernel_ms) return fail(DSSCF_ERR_ARG, "dsscf_jk_resident_timed: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_jk_resident_timed: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = ctx_plan(c)) return rc;
    Plan& p = c->plan;
    if (int rc = plan_enqueue(p, c->stream, nullptr, true)) return rc;  // eager: events need it
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int slot = 0; slot < 2 * p.ncls; ++slot) {
        if (!p.cls_nseg[slot]) continue;
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, p.slot_ev[2 * slot], p.slot_ev[2 * slot + 1]));
        kernel_ms[slot / 2] += ms;
    }
    return DSSCF_OK;
}

int dsscf_get_jk(dsscf_ctx* c, double* J, double* K)
{
    if (!c || !J || !K) return fail(DSSCF_ERR_ARG, "dsscf_get_jk: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_get_jk: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    const size_t np = (size_t)c->n_ao * (c->n_ao + 1) / 2;
    HIP_TRY(hipMemcpyAsync(J, c->d_J_pk, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(K, c->d_K_pk, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DSSCF_OK;
}

int dsscf_time_jk(dsscf_ctx* c, int warmup, int iters, double* ms_per_build, double* ms_dominant, int* dom_cls)
{
    if (!c || iters <= 0 || warmup < 0) return fail(DSSCF_ERR_ARG, "dsscf_time_jk: bad argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_time_jk: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    for (int i = 0; i < warmup; ++i)
        if (int rc = run_jk(c, nullptr)) return rc;
    const int ncls = dk_n_class();
    std::vector<double> acc(ncls, 0.0);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i)
        if (int rc = run_jk(c, nullptr)) return rc;
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float total = 0.0f;
    HIP_TRY(hipEventElapsedTime(&total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // Per-class durations from a separate pass with events between launches.
    std::vector<float> ms;
    for (int i = 0; i < iters; ++i) {
        if (int rc = run_jk(c, &ms)) return rc;
        for (int k = 0; k < ncls; ++k) acc[k] += ms[k];
    }
    int best = 0;
    for (int k = 1; k < ncls; ++k)
        if (acc[k] > acc[best]) best = k;
    if (ms_per_build) *ms_per_build = total / iters;
    if (ms_dominant) *ms_dominant = acc[best] / iters;
    if (dom_cls) *dom_cls = best;
    return DSSCF_OK;
}

int dsscf_time_jk_classes(dsscf_ctx* c, int iters, double* ms_per_class)
{
    if (!c || iters <= 0 || !ms_per_class) return fail(DSSCF_ERR_ARG, "dsscf_time_jk_classes: bad argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_time_jk_classes: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    const int ncls = dk_n_class();
    std::vector<double> acc(ncls, 0.0);
    std::vector<float> ms;
    for (int i = 0; i < iters; ++i) {
        if (int rc = run_jk(c, &ms)) return rc;
        for (int k = 0; k < ncls; ++k) acc[k] += ms[k];
    }
    for (int k = 0; k < ncls; ++k) ms_per_class[k] = acc[k] / iters;
    return DSSCF_OK;
}

int dsscf_one_electron(dsscf_ctx* c, int n_atom, const int* Z, const double* xyz, double* S, double* T, double* V)
{
    if (!c || n_atom < 0 || (n_atom > 0 && (!Z || !xyz)) || !S || !T || !V)
        return fail(DSSCF_ERR_ARG, "dsscf_one_electron: bad argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_one_electron: call dsscf_prepare first");
    HIP_TRY(hipSetDevice(c->device));
    const size_t np = (size_t)c->n_ao * (c->n_ao + 1) / 2;
    std::vector<double> at(4 * (size_t)std::max(n_atom, 1));
    for (int i = 0; i < n_atom; ++i) {
        at[4 * i] = xyz[3 * i];
        at[4 * i + 1] = xyz[3 * i + 1];
        at[4 * i + 2] = xyz[3 * i + 2];
        at[4 * i + 3] = (double)Z[i];
    }
    double *d_at = nullptr, *d_out = nullptr;
    if (int rc = dev_alloc(&d_at, at.size())) return rc;
    if (int rc = dev_alloc(&d_out, 3 * np)) {
        dev_free(d_at);
        return rc;
    }
    int rc = DSSCF_OK;
    hipError_t e = hipMemcpyAsync(d_at, at.data(), sizeof(double) * at.size(), hipMemcpyHostToDevice, c->stream);
    const int maxl = dk_max_l();
    for (int l1 = 0; l1 <= maxl && e == hipSuccess; ++l1)
        for (int l2 = 0; l2 <= l1 && e == hipSuccess; ++l2) {
            const int pc = pair_class_of(l1, l2);
            e = dsscf_launch_one_electron(l1, l2, c->opt.boys_mode, c->d_plist + c->pc_start[pc], c->pc_count[pc],
                                          c->d_pairs, c->d_pp, d_at, n_atom, c->d_btab, c->d_xform, c->opt.cartesian,
                                          d_out, d_out + np, d_out + 2 * np, c->stream);
        }
    if (e == hipSuccess) e = hipMemcpyAsync(S, d_out, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T, d_out + np, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(V, d_out + 2 * np, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(DSSCF_ERR_HIP, std::string("dsscf_one_electron: ") + hipGetErrorString(e));
    dev_free(d_at);
    dev_free(d_out);
    return rc;
}

int dsscf_one_electron_gradient(dsscf_ctx* c, int n_atom, const int* Z, const double* xyz, const int* ao_atom,
                                const double* Pt, const double* W, double* grad)
{
    if (!c || n_atom <= 0 || !Z || !xyz || !ao_atom || !Pt || !W || !grad)
        return fail(DSSCF_ERR_ARG, "dsscf_one_electron_gradient: bad argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_one_electron_gradient: call dsscf_prepare first");
    if (c->opt.cartesian) return fail(DSSCF_ERR_UNSUPPORTED, "dsscf_one_electron_gradient: spherical AOs only");
    const int N = c->n_ao;
    for (int i = 0; i < N; ++i)
        if (ao_atom[i] < 0 || ao_atom[i] >= n_atom) return fail(DSSCF_ERR_ARG, "dsscf_one_electron_gradient: AO atom index");
    HIP_TRY(hipSetDevice(c->device));
    const size_t np = (size_t)N * (N + 1) / 2;
    std::vector<double> at(4 * (size_t)n_atom);
    for (int i = 0; i < n_atom; ++i) {
        at[4 * i] = xyz[3 * i];
        at[4 * i + 1] = xyz[3 * i + 1];
        at[4 * i + 2] = xyz[3 * i + 2];
        at[4 * i + 3] = (double)Z[i];
    }
    double *d_at = nullptr, *d_m = nullptr, *d_g = nullptr;
    int* d_atom = nullptr;
    auto cleanup = [&] {
        dev_free(d_at);
        dev_free(d_m);
        dev_free(d_g);
        dev_free(d_atom);
    };
    int rc = dev_alloc(&d_at, at.size());
    if (!rc) rc = dev_alloc(&d_m, 2 * np);
    if (!rc) rc = dev_alloc(&d_g, 3 * (size_t)n_atom);
    if (!rc) rc = dev_alloc(&d_atom, (size_t)N);
    if (rc) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_at, at.data(), sizeof(double) * at.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_m, Pt, sizeof(double) * np, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_m + np, W, sizeof(double) * np, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_atom, ao_atom, sizeof(int) * N, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = dsscf_launch_fill(d_g, 3 * (size_t)n_atom, 0.0, c->stream);
    const int maxl = dk_max_l();
    for (int l1 = 0; l1 <= maxl && e == hipSuccess; ++l1)
        for (int l2 = 0; l2 <= l1 && e == hipSuccess; ++l2) {
            const int pc = pair_class_of(l1, l2);
            e = dsscf_launch_one_electron_grad(l1, l2, c->opt.boys_mode, c->d_plist + c->pc_start[pc], c->pc_count[pc],
                                               c->d_pairs, c->d_pp, d_at, n_atom, c->d_btab, c->d_xform, d_atom, d_m,
                                               d_m + np, d_g, c->stream);
        }
    if (e == hipSuccess) e = hipMemcpyAsync(grad, d_g, sizeof(double) * 3 * n_atom, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(DSSCF_ERR_HIP, std::string("dsscf_one_electron_gradient: ") + hipGetErrorString(e));
    return DSSCF_OK;
}

long long dsscf_ctx_device_bytes(dsscf_ctx* c)
{
    if (!c) return 0;
    return 1024LL * 128 + (c->prepared ? c->device_bytes : 0);  // + Boys table and launch tables
}

int dsscf_schwarz(dsscf_ctx* c, double* Q)
{
    if (!c || !Q) return fail(DSSCF_ERR_ARG, "dsscf_schwarz: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_schwarz: call dsscf_prepare first");
    std::memcpy(Q, c->Q.data(), sizeof(double) * c->Q.size());
    return DSSCF_OK;
}

int dsscf_get_stats(dsscf_ctx* c, dsscf_stats* st)
{
    if (!c || !st) return fail(DSSCF_ERR_ARG, "dsscf_get_stats: null argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_get_stats: call dsscf_prepare first");
    std::memset(st, 0, sizeof(*st));
    st->n_pairs = (long long)c->pairs.size();
    st->n_prim_pairs = (long long)c->prims.size();
    st->n_quartets_unscreened = c->n_unscreened;
    st->n_class = dk_n_class();
    for (int k = 0; k < st->n_class && k < 21; ++k) {
        st->class_quartets[k] = c->q_cnt[k] + c->q_cnt32[k];
        st->class_prim_quartets[k] = c->q_prim[k] + c->q_prim32[k];
        st->class_quartets_fp32[k] = c->q_cnt32[k];
        st->class_prim_quartets_fp32[k] = c->q_prim32[k];
        st->n_quartets += c->q_cnt[k] + c->q_cnt32[k];
        st->n_prim_quartets += c->q_prim[k] + c->q_prim32[k];
        st->n_quartets_fp32 += c->q_cnt32[k];
        if (k < (int)c->fill_lanes.size()) {
            st->class_lanes[k] = c->fill_lanes[k];
            st->class_lane_slots[k] = c->fill_slots[k];
            st->class_lane_trips[k] = c->fill_lane_trips[k];
            st->class_wave_trips[k] = c->fill_wave_trips[k];
        }
    }
    st->n_work_pairs = c->n_work_pairs;
    st->n_work_items = c->n_work_items;
    return DSSCF_OK;
}

}  // extern "C"

namespace {

// dsscf_eri_batch and its no-alloc form: map every request onto (bra pair,
// ket pair, AO component of the block), evaluate each class's unique shell
// quartets as whole AO blocks, pick the requested components.  staging null:
// the class's item list and block outputs go into buffers allocated per
// class; otherwise into staging (at most staging_bytes, checked first).
int eri_batch_impl(dsscf_ctx* c, const int* ijkl, long long n, double* out, void* staging, size_t staging_bytes,
                   const char* who)
{
    const int ncls = dk_n_class();
    struct Req {
        long long t;
        int bra, ket, comp;
    };
    std::vector<std::vector<Req>> per_cls(ncls);
    for (long long t = 0; t < n; ++t) {
        int ao[4];
        for (int k = 0; k < 4; ++k) {
            ao[k] = ijkl[4 * t + k];
            if (ao[k] < 0 || ao[k] >= c->n_ao) return fail(DSSCF_ERR_ARG, std::string(who) + ": AO index out of range");
        }
        int pidx[2], ca[2], cb[2], nb[2];
        for (int h = 0; h < 2; ++h) {
            const int s1 = c->ao_shell[ao[2 * h]], s2 = c->ao_shell[ao[2 * h + 1]];
            const int p = pair_index(s1, s2);
            const int f = c->pair_first[p];
            const int m1 = ao[2 * h] - c->sh_ao[s1], m2 = ao[2 * h + 1] - c->sh_ao[s2];
            // component order inside the pair block is (first shell, second shell)
            if (f == s1) {
                ca[h] = m1;
                cb[h] = m2;
            } else {
                ca[h] = m2;
                cb[h] = m1;
            }
            const int s_second = f == s1 ? s2 : s1;
            nb[h] = nfun(c->sh_l[s_second], c->opt.cartesian);
            pidx[h] = p;
        }
        int X = c->pair_class[pidx[0]], Y = c->pair_class[pidx[1]];
        int bra = 0, ket = 1;
        if (X < Y) {
            std::swap(X, Y);
            bra = 1;
            ket = 0;
        }
        const int cls = X * (X + 1) / 2 + Y;
        const int nket = nb[ket] * nfun(c->sh_l[c->pair_first[pidx[ket]]], c->opt.cartesian);
        const int comp = (ca[bra] * nb[bra] + cb[bra]) * nket + ca[ket] * nb[ket] + cb[ket];
        per_cls[cls].push_back(Req{t, pidx[bra], pidx[ket], comp});
    }
    std::vector<std::vector<long long>> keys(ncls), uniq(ncls);
    std::vector<int> nblock(ncls, 0);
    for (int cls = 0; cls < ncls; ++cls) {
        auto& rq = per_cls[cls];
        if (rq.empty()) continue;
        int l4[4];
        dk_class_l(cls, l4);
        const int ct = c->opt.cartesian;
        nblock[cls] = nfun(l4[0], ct) * nfun(l4[1], ct) * nfun(l4[2], ct) * nfun(l4[3], ct);
        // unique shell quartets of this class
        auto& key = keys[cls];
        key.resize(rq.size());
        for (size_t i = 0; i < rq.size(); ++i) key[i] = (long long)rq[i].bra * (long long)c->pairs.size() + rq[i].ket;
        auto& uk = uniq[cls];
        uk = key;
        std::sort(uk.begin(), uk.end());
        uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
        if (staging) {
            const size_t need = staging_offset_out(uk.size()) + sizeof(double) * uk.size() * nblock[cls];
            if (need > staging_bytes)
                return fail(DSSCF_ERR_ARG, std::string(who) + ": class " + std::to_string(cls) + " needs " +
                                               std::to_string(need) + " bytes of staging for " +
                                               std::to_string(uk.size()) + " shell quartets, more than the " +
                                               std::to_string(staging_bytes) +
                                               "-byte scratch: split the request (or call dsscf_eri_batch)");
        }
    }
    for (int cls = 0; cls < ncls; ++cls) {
        auto& rq = per_cls[cls];
        if (rq.empty()) continue;
        const auto& uk = uniq[cls];
        const int nb = nblock[cls];
        std::vector<int4> items(uk.size());
        for (size_t u = 0; u < uk.size(); ++u)
            items[u] = make_int4((int)(uk[u] / (long long)c->pairs.size()), (int)(uk[u] % (long long)c->pairs.size()),
                                 (int)(u * nb), 0);
        int4* d_items = nullptr;
        double* d_out = nullptr;
        if (staging) {
            d_items = (int4*)staging;
            d_out = (double*)((char*)staging + staging_offset_out(items.size()));
        } else {
            if (int rc = dev_alloc(&d_items, items.size())) return rc;
            if (int rc = dev_alloc(&d_out, items.size() * nb)) {
                dev_free(d_items);
                return rc;
            }
        }
        std::vector<double> blk(items.size() * nb);
        // every block's nb components are written by the kernel: nothing to clear
        hipError_t e = hipMemcpyAsync(d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = dsscf_launch_eri_block(cls, c->opt.boys_mode, c->opt.cartesian, d_items, (int)items.size(), c->d_pairs,
                                       c->d_pp, c->d_btab,
                                       d_out, kEriScale, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(blk.data(), d_out, sizeof(double) * blk.size(), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (!staging) {
            dev_free(d_items);
            dev_free(d_out);
        }
        if (e != hipSuccess) return fail(DSSCF_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
        for (size_t i = 0; i < rq.size(); ++i) {
            const size_t u = std::lower_bound(uk.begin(), uk.end(), keys[cls][i]) - uk.begin();
            out[rq[i].t] = blk[u * nb + rq[i].comp];
        }
    }
    return DSSCF_OK;
}

}  // namespace

extern "C" {

int dsscf_eri_batch(dsscf_ctx* c, const int* ijkl, long long n, double* out)
{
    if (!c || (n > 0 && (!ijkl || !out)) || n < 0) return fail(DSSCF_ERR_ARG, "dsscf_eri_batch: bad argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_eri_batch: call dsscf_prepare first");
    if (n == 0) return DSSCF_OK;
    HIP_TRY(hipSetDevice(c->device));
    return eri_batch_impl(c, ijkl, n, out, nullptr, 0, "dsscf_eri_batch");
}

int dsscf_eri_batch_noalloc(dsscf_ctx* c, const int* ijkl, long long n, double* out)
{
    if (!c || (n > 0 && (!ijkl || !out)) || n < 0) return fail(DSSCF_ERR_ARG, "dsscf_eri_batch_noalloc: bad argument");
    if (!c->prepared) return fail(DSSCF_ERR_STATE, "dsscf_eri_batch_noalloc: call dsscf_prepare first");
    if (n == 0) return DSSCF_OK;
    HIP_TRY(hipSetDevice(c->device));
    void* st = scratch_ptr(c->device);
    if (!st) return fail(DSSCF_ERR_STATE, "dsscf_eri_batch_noalloc: no scratch address for this device");
    return eri_batch_impl(c, ijkl, n, out, st, DSSCF_SCRATCH_BYTES, "dsscf_eri_batch_noalloc");
}

}  // extern "C"

// ----------------------------------------------------------------- batch --
struct dsscf_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<dsscf_ctx*> ctxs;
    std::vector<size_t> off;  // packed offset of each context
    size_t total = 0;         // packed elements over all contexts
    double* d_buf = nullptr;  // [Pt | Pa | J | K], each `total`
    double* h_pin = nullptr;  // pinned staging, same layout
    size_t cap = 0;           // elements of each of the four buffer sections allocated (grow only)
    Plan plan;
    int use_graph = 1;
    std::vector<int> block_size;
    // dsscf_batch_set_device: the members' own device buffers
    bool device_io = false;
    std::vector<const double*> io_pt, io_pa;
    std::vector<double*> io_j, io_k;

    void release_buffers()
    {
        plan.release();
        dev_free(d_buf);
        if (h_pin) (void)hipHostFree(h_pin);
        h_pin = nullptr;
        cap = 0;
    }
};

extern "C" {

int dsscf_batch_create(dsscf_batch** out, int device)
{
    if (!out) return fail(DSSCF_ERR_ARG, "dsscf_batch_create: null output");
    HIP_TRY(hipSetDevice(device));
    dsscf_batch* b = new dsscf_batch();
    b->device = device;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return fail(DSSCF_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    b->block_size.assign(dk_n_class(), 256);
    *out = b;
    return DSSCF_OK;
}

int dsscf_batch_destroy(dsscf_batch* b)
{
    if (!b) return DSSCF_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    b->release_buffers();
    (void)hipStreamDestroy(b->stream);
    delete b;
    return DSSCF_OK;
}

namespace {
int batch_build_plan(dsscf_batch* b)
{
    const int n = (int)b->ctxs.size();
    if (b->device_io)
        return plan_build(b->plan, b->ctxs.data(), n, b->io_pt.data(), b->io_pa.data(), b->io_j.data(),
                          b->io_k.data(), b->block_size.data(), b->stream);
    std::vector<const double*> pt(n), pa(n);
    std::vector<double*> j(n), k(n);
    for (int i = 0; i < n; ++i) {
        pt[i] = b->d_buf + b->off[i];
        pa[i] = b->d_buf + b->total + b->off[i];
        j[i] = b->d_buf + 2 * b->total + b->off[i];
        k[i] = b->d_buf + 3 * b->total + b->off[i];
    }
    return plan_build(b->plan, b->ctxs.data(), n, pt.data(), pa.data(), j.data(), k.data(), b->block_size.data(),
                      b->stream);
}
}  // namespace

namespace {
int batch_check_members(dsscf_batch* b, dsscf_ctx* const* ctxs, int n, const char* who)
{
    if (!b || n <= 0 || !ctxs) return fail(DSSCF_ERR_ARG, std::string(who) + ": bad argument");
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i] || !ctxs[i]->prepared) return fail(DSSCF_ERR_STATE, std::string(who) + ": context not prepared");
        if (ctxs[i]->device != b->device) return fail(DSSCF_ERR_ARG, std::string(who) + ": contexts on another device");
        if (ctxs[i]->opt.boys_mode != ctxs[0]->opt.boys_mode)
            return fail(DSSCF_ERR_ARG, std::string(who) + ": mixed Boys modes in one batch");
        if (ctxs[i]->opt.cartesian != ctxs[0]->opt.cartesian)
            return fail(DSSCF_ERR_ARG, std::string(who) + ": spherical and Cartesian-d contexts in one batch");
        for (int k = 0; k < i; ++k)
            if (ctxs[k] == ctxs[i]) return fail(DSSCF_ERR_ARG, std::string(who) + ": context listed twice");
    }
    return DSSCF_OK;
}
}  // namespace

int dsscf_batch_set(dsscf_batch* b, dsscf_ctx* const* ctxs, int n)
{
    if (int rc = batch_check_members(b, ctxs, n, "dsscf_batch_set")) return rc;
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->plan.reset();  // tables, streams and (if large enough) buffers are reused
    b->device_io = false;
    b->ctxs.assign(ctxs, ctxs + n);
    b->off.resize(n);
    b->total = 0;
    for (int i = 0; i < n; ++i) {
        b->off[i] = b->total;
        b->total += (size_t)ctxs[i]->n_ao * (ctxs[i]->n_ao + 1) / 2;
    }
    if (b->total > b->cap || !b->d_buf) {
        dev_free(b->d_buf);
        if (b->h_pin) (void)hipHostFree(b->h_pin);
        b->h_pin = nullptr;
        b->cap = 0;
        const size_t c = std::max<size_t>(b->total + b->total / 2, 1024);
        if (int rc = dev_alloc(&b->d_buf, 4 * c)) return rc;
        HIP_TRY(hipHostMalloc((void**)&b->h_pin, sizeof(double) * 4 * c, hipHostMallocDefault));
        b->cap = c;
    }
    // the contexts' own streams must be idle: the batch writes their buffers
    for (dsscf_ctx* c : b->ctxs) HIP_TRY(hipStreamSynchronize(c->stream));
    return batch_build_plan(b);
}

int dsscf_batch_set_device(dsscf_batch* b, dsscf_ctx* const* ctxs, int n, const double* const* d_Ptot,
                           const double* const* d_Palpha, double* const* d_J, double* const* d_K)
{
    if (int rc = batch_check_members(b, ctxs, n, "dsscf_batch_set_device")) return rc;
    if (!d_Ptot || !d_Palpha || !d_J || !d_K) return fail(DSSCF_ERR_ARG, "dsscf_batch_set_device: null argument");
    for (int i = 0; i < n; ++i)
        if (!d_Ptot[i] || !d_Palpha[i] || !d_J[i] || !d_K[i])
            return fail(DSSCF_ERR_ARG, "dsscf_batch_set_device: null matrix");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->plan.reset();
    b->device_io = true;
    b->ctxs.assign(ctxs, ctxs + n);
    b->io_pt.assign(d_Ptot, d_Ptot + n);
    b->io_pa.assign(d_Palpha, d_Palpha + n);
    b->io_j.assign(d_J, d_J + n);
    b->io_k.assign(d_K, d_K + n);
    b->off.assign(n, 0);
    b->total = 0;
    for (dsscf_ctx* c : b->ctxs) HIP_TRY(hipStreamSynchronize(c->stream));
    return batch_build_plan(b);
}

int dsscf_batch_jk_device(dsscf_batch* b)
{
    if (!b) return fail(DSSCF_ERR_ARG, "dsscf_batch_jk_device: null batch");
    if (!b->plan.built || !b->device_io)
        return fail(DSSCF_ERR_STATE, "dsscf_batch_jk_device: call dsscf_batch_set_device first");
    HIP_TRY(hipSetDevice(b->device));
    if (int rc = plan_run(b->plan, b->stream, b->use_graph != 0, nullptr)) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return DSSCF_OK;
}

int dsscf_batch_set_graph(dsscf_batch* b, int use_graph)
{
    if (!b) return fail(DSSCF_ERR_ARG, "dsscf_batch_set_graph: null batch");
    b->use_graph = use_graph;
    return DSSCF_OK;
}

int dsscf_batch_jk(dsscf_batch* b, const double* const* Ptot, const double* const* Palpha, double* const* J,
                   double* const* K)
{
    if (!b || !Ptot || !Palpha || !J || !K) return fail(DSSCF_ERR_ARG, "dsscf_batch_jk: null argument");
    if (!b->plan.built || b->device_io) return fail(DSSCF_ERR_STATE, "dsscf_batch_jk: call dsscf_batch_set first");
    HIP_TRY(hipSetDevice(b->device));
    const int n = (int)b->ctxs.size();
    for (int i = 0; i < n; ++i) {
        const size_t np = (size_t)b->ctxs[i]->n_ao * (b->ctxs[i]->n_ao + 1) / 2;
        if (!Ptot[i] || !Palpha[i] || !J[i] || !K[i]) return fail(DSSCF_ERR_ARG, "dsscf_batch_jk: null matrix");
        std::memcpy(b->h_pin + b->off[i], Ptot[i], sizeof(double) * np);
        std::memcpy(b->h_pin + b->total + b->off[i], Palpha[i], sizeof(double) * np);
    }
    HIP_TRY(hipMemcpyAsync(b->d_buf, b->h_pin, sizeof(double) * 2 * b->total, hipMemcpyHostToDevice, b->stream));
    if (int rc = plan_run(b->plan, b->stream, b->use_graph != 0, nullptr)) return rc;
    HIP_TRY(hipMemcpyAsync(b->h_pin + 2 * b->total, b->d_buf + 2 * b->total, sizeof(double) * 2 * b->total,
                           hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (int i = 0; i < n; ++i) {
        const size_t np = (size_t)b->ctxs[i]->n_ao * (b->ctxs[i]->n_ao + 1) / 2;
        std::memcpy(J[i], b->h_pin + 2 * b->total + b->off[i], sizeof(double) * np);
        std::memcpy(K[i], b->h_pin + 3 * b->total + b->off[i], sizeof(double) * np);
    }
    return DSSCF_OK;
}

int dsscf_batch_time(dsscf_batch* b, int warmup, int iters, double* ms_per_build)
{
    if (!b || iters <= 0 || warmup < 0) return fail(DSSCF_ERR_ARG, "dsscf_batch_time: bad argument");
    if (!b->plan.built) return fail(DSSCF_ERR_STATE, "dsscf_batch_time: call dsscf_batch_set first");
    HIP_TRY(hipSetDevice(b->device));
    for (int i = 0; i < warmup; ++i)
        if (int rc = plan_run(b->plan, b->stream, b->use_graph != 0, nullptr)) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, b->stream));
    for (int i = 0; i < iters; ++i)
        if (int rc = plan_run(b->plan, b->stream, b->use_graph != 0, nullptr)) return rc;
    HIP_TRY(hipEventRecord(e1, b->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_build) *ms_per_build = ms / iters;
    return DSSCF_OK;
}

// Project autotuner: for every thread-per-quartet class with work, time the
// class kernel alone at 64, 128 and 256 threads per block (3 launches each,
// HIP events) and keep the fastest; the result goes into the batch's plan
// and, for a single-context batch, into the context.
}  // extern "C"

namespace {
// A pair of timing events, destroyed on every path.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~EventPair()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    int create()
    {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        return DSSCF_OK;
    }
};

// Device time (ms) of `reps` launches of one class's kernels on the batch's
// current plan, after one warm-up launch.
int time_class(dsscf_batch* b, int cls, EventPair& ev, int reps, float* ms)
{
    if (int rc = plan_launch_class(b->plan, cls, b->stream)) return rc;  // warm
    HIP_TRY(hipEventRecord(ev.e0, b->stream));
    for (int r = 0; r < reps; ++r)
        if (int rc = plan_launch_class(b->plan, cls, b->stream)) return rc;
    HIP_TRY(hipEventRecord(ev.e1, b->stream));
    HIP_TRY(hipEventSynchronize(ev.e1));
    HIP_TRY(hipEventElapsedTime(ms, ev.e0, ev.e1));
    return DSSCF_OK;
}

bool class_in_plan(const Plan& p, int cls)
{
    return !dsscf_class_is_wave(cls) && (p.cls_nseg[2 * cls] > 0 || p.cls_nseg[2 * cls + 1] > 0);
}
}  // namespace

extern "C" {

int dsscf_batch_autotune(dsscf_batch* b, int* chosen)
{
    if (!b || !b->plan.built) return fail(DSSCF_ERR_STATE, "dsscf_batch_autotune: call dsscf_batch_set first");
    HIP_TRY(hipSetDevice(b->device));
    const int ncls = dk_n_class();
    const int cand[3] = {64, 128, 256};
    EventPair ev;
    if (int rc = ev.create()) return rc;
    std::vector<int> best = b->block_size;
    for (int cls = 0; cls < ncls; ++cls) {
        if (!class_in_plan(b->plan, cls)) continue;
        float best_ms = 1e30f;
        for (int ci = 0; ci < 3; ++ci) {
            b->block_size[cls] = cand[ci];
            if (int rc = batch_build_plan(b)) return rc;
            float ms = 0.0f;
            if (int rc = time_class(b, cls, ev, 3, &ms)) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best[cls] = cand[ci];
            }
        }
        b->block_size[cls] = best[cls];
    }
    b->block_size = best;
    if (int rc = batch_build_plan(b)) return rc;
    if (b->ctxs.size() == 1) {
        b->ctxs[0]->block_size = best;
        b->ctxs[0]->plan.release();
        b->ctxs[0]->plan_u.release();
    }
    if (chosen)
        for (int k = 0; k < ncls; ++k) chosen[k] = best[k];
    return DSSCF_OK;
}

int dsscf_set_class_schedule(dsscf_ctx* c, const double* work_item_flops, const double* group_prims,
                             const int* kets_per_unit)
{
    if (!c) return fail(DSSCF_ERR_ARG, "dsscf_set_class_schedule: null context");
    const int ncls = dk_n_class();
    if (work_item_flops) c->sched_wif.assign(work_item_flops, work_item_flops + ncls); else c->sched_wif.clear();
    if (group_prims) c->sched_gp.assign(group_prims, group_prims + ncls); else c->sched_gp.clear();
    if (kets_per_unit) c->sched_chunk.assign(kets_per_unit, kets_per_unit + ncls); else c->sched_chunk.clear();
    for (int k = 0; k < (int)c->sched_chunk.size(); ++k)
        if (c->sched_chunk[k] < 0 || c->sched_chunk[k] > kMaxKetsPerUnit)
            return fail(DSSCF_ERR_ARG, "dsscf_set_class_schedule: kets per unit 0..4096");
    c->prepared = false;
    return DSSCF_OK;
}

int dsscf_batch_autotune_schedule(dsscf_batch* b, double* work_item_flops, double* group_prims, int* kets_per_unit,
                                  int* block_size)
{
    if (!b || !b->plan.built) return fail(DSSCF_ERR_STATE, "dsscf_batch_autotune_schedule: call dsscf_batch_set first");
    HIP_TRY(hipSetDevice(b->device));
    const int ncls = dk_n_class();
    // candidates of each schedule parameter; 0 = the context's own default
    const std::vector<double> c_wif = {0.0, 4096.0, 8192.0, 32768.0, 65536.0};
    const std::vector<double> c_gp = {0.0, 4.0, 16.0, -1.0};
    const std::vector<double> c_chunk = {0.0, 64.0, 128.0, 256.0, 512.0};
    std::vector<double> wif(ncls, 0.0), gp(ncls, 0.0);
    std::vector<int> chunk(ncls, 0);
    EventPair ev;
    if (int rc = ev.create()) return rc;
    // apply the schedule to every member and rebuild its lists and the fused plan
    auto apply = [&]() -> int {
        for (dsscf_ctx* c : b->ctxs) {
            c->sched_wif = wif;
            c->sched_gp = gp;
            c->sched_chunk = chunk;
            c->prepared = false;
            if (int rc = dsscf_prepare(c)) return rc;
        }
        return batch_build_plan(b);
    };
    // one parameter at a time (sub-pair size, lane groups, kets per unit),
    // every class timed at each candidate; each class keeps its fastest
    for (int param = 0; param < 3; ++param) {
        const std::vector<double>& cand = param == 0 ? c_wif : (param == 1 ? c_gp : c_chunk);
        std::vector<float> best_ms(ncls, 1e30f);
        std::vector<double> best_v(ncls, 0.0);
        for (double v : cand) {
            for (int k = 0; k < ncls; ++k) {
                if (param == 0) wif[k] = v;
                else if (param == 1) gp[k] = v;
                else chunk[k] = (int)v;
            }
            // the parameters already chosen stay at their per-class choice
            if (int rc = apply()) return rc;
            for (int cls = 0; cls < ncls; ++cls) {
                if (!class_in_plan(b->plan, cls)) continue;
                float ms = 0.0f;
                if (int rc = time_class(b, cls, ev, 3, &ms)) return rc;
                if (ms < best_ms[cls]) {
                    best_ms[cls] = ms;
                    best_v[cls] = v;
                }
            }
        }
        for (int k = 0; k < ncls; ++k) {
            if (param == 0) wif[k] = best_v[k];
            else if (param == 1) gp[k] = best_v[k];
            else chunk[k] = (int)best_v[k];
        }
    }
    if (int rc = apply()) return rc;
    // block sizes for the chosen lists
    if (int rc = dsscf_batch_autotune(b, block_size)) return rc;
    for (int k = 0; k < ncls; ++k) {
        if (work_item_flops) work_item_flops[k] = wif[k];
        if (group_prims) group_prims[k] = gp[k];
        if (kets_per_unit) kets_per_unit[k] = chunk[k];
    }
    return DSSCF_OK;
}

}  // extern "C"
