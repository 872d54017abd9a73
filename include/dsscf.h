/*
 * dsscf.h -- C-ABI of the MI355X direct-SCF Fock-build engine.
 *
 * The hot path of DollarStoreSCF (Obara-Saika ERIs, Boys F_m(T), screening
 * and the J/K contraction into the Fock matrix) behind plain-C entry points:
 * no C++ or torch types, plain pointers and sizes, int status (0 = ok) with
 * dsscf_last_error() for the message.  Matrices cross the boundary in the
 * reference's packed-upper layout: element (y,x), y <= x, at x(x+1)/2 + y
 * (src/lalib/Matrix.ipp:332-351), so SymmetricMatrixNd::getRawData()
 * (src/lalib/Matrix.hpp:56) passes through unchanged.
 *
 * Threading: one HIP stream per context; contexts are independent, one
 * context must not be used from two threads at once.
 */
#ifndef DSSCF_H
#define DSSCF_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DSSCF_OK = 0,
    DSSCF_ERR_ARG = 1,         /* bad argument (null pointer, size, index) */
    DSSCF_ERR_HIP = 2,         /* HIP runtime error */
    DSSCF_ERR_STATE = 3,       /* call out of order (e.g. jk before prepare) */
    DSSCF_ERR_UNSUPPORTED = 4, /* angular momentum or mode not built */
    DSSCF_ERR_NOMEM = 5,       /* device allocation failed */
    DSSCF_ERR_NOCONV = 6       /* SCF iteration cap reached */
};

/* Boys-function evaluation on the device. */
enum {
    DSSCF_BOYS_REFERENCE = 0, /* the reference's truncated series, GaussianIntegral.cpp:209-223 */
    DSSCF_BOYS_ACCURATE = 1   /* tabulated Taylor + downward recursion, ~1e-15 relative */
};

typedef struct dsscf_ctx dsscf_ctx;

typedef struct {
    int boys_mode;            /* DSSCF_BOYS_* */
    double schwarz_threshold; /* drop shell quartets with Q_ab*Q_cd below it; 0 keeps all.
                                 Replaces the (ii|jj)(kk|ll) > 1e-16 screen of
                                 src/chem/scf/ExactCoulomb.cpp:63 */
    double fp32_threshold;    /* quartets with Q_ab*Q_cd below it run their setup, small-T Boys and
                                 VRR in FP32 (contraction, HRR, J/K accumulation stay FP64); 0 = all
                                 FP64.  Needs DSSCF_BOYS_ACCURATE. */
    double prim_threshold;    /* drop primitive pairs with 2 pi^(5/2) e_ab e_max below it, where
                                 e = w_a w_b exp(-ab/p |AB|^2) / p^(5/4) and e_max is the largest e:
                                 they cannot lift any primitive quartet above it (no reference
                                 counterpart: the reference evaluates every primitive); 0 = none */
    double work_item_flops;   /* a shell pair with many primitive pairs is evaluated as sub-pairs over
                                 consecutive ranges of its primitive pairs, so that one thread's
                                 share of a shell quartet stays near this many FLOPs of primitive
                                 work, with runs of at least 5 primitive pairs (the sub-quartets
                                 add up to the shell quartet); 0 = default
                                 (DSSCF_WORK_ITEM_FLOPS), < 0 = never split */
    double group_prims;       /* a (sub-)quartet with many primitive quartets is shared by a group of
                                 up to 16 lanes that split its primitive loops, so that each lane
                                 keeps about this many primitive quartets; its first lane digests
                                 the summed integrals into J and K; 0 = default
                                 (DSSCF_GROUP_PRIMS), < 0 = one lane per quartet */
    double jcd_max_bytes;     /* device memory the context may spend on J_cd slot records: each
                                 quartet then stores its ket block's Coulomb terms, and a reduction
                                 kernel sums them per element, in place of one atomic add per
                                 quartet and element (which serialize where many quartets share a
                                 ket block); a context whose records would exceed it keeps the
                                 atomic adds; 0 = default (DSSCF_JCD_MAX_BYTES), < 0 = never */
    int ket_order;            /* order of a bra's kets of equal primitive count: 0 = by context size
                                 (distance from the bra from DSSCF_KET_DISTANCE_QUARTETS unscreened
                                 quartets on, else shuffled), 1 = shuffled, 2 = by distance; the
                                 results agree to rounding */
    int cartesian;            /* 1: shells with l >= 2 are ncart(l) Cartesian functions x^a y^b z^c
                                 (x-major: xx, xy, xz, yy, yz, zz; the "6D" convention of 6-31G*),
                                 each primitive normalized; 0: the reference's 2l+1 real solid
                                 harmonics (its only form, src/chem/GTO.cpp:15-94).  No reference
                                 counterpart; s and p shells are the same either way */
} dsscf_options;

#define DSSCF_WORK_ITEM_FLOPS 16384.0
#define DSSCF_GROUP_PRIMS 8.0
#define DSSCF_JCD_MAX_BYTES 8589934592.0
#define DSSCF_KET_DISTANCE_QUARTETS 100000000.0

typedef struct {
    long long n_pairs;             /* unique shell pairs */
    long long n_prim_pairs;        /* primitive pairs over all shell pairs */
    long long n_quartets;          /* surviving unique shell quartets */
    long long n_prim_quartets;     /* primitive quartets evaluated (after primitive-pair screening) */
    long long n_quartets_unscreened;
    int n_class;                   /* shell-quartet classes */
    long long class_quartets[21];  /* per class (order of dsscf_class_l) */
    long long class_prim_quartets[21];
    long long n_quartets_fp32;     /* of n_quartets, evaluated in FP32 */
    long long class_quartets_fp32[21];
    long long class_prim_quartets_fp32[21];
    long long n_work_pairs;        /* shell pairs plus the sub-pairs of split ones (work_item_flops) */
    long long n_work_items;        /* (sub-)quartets the thread-per-quartet kernels evaluate */
    /* Wavefront fill of the thread-per-quartet classes' work units, per class
     * (64-lane wavefronts): lanes with a quartet against the lanes of the
     * wavefronts the units occupy, and the primitive-loop trips of those
     * lanes against the trips the wavefronts run (each runs its longest
     * lane).  Diagnostics; no reference counterpart. */
    long long class_lanes[21];
    long long class_lane_slots[21];
    long long class_lane_trips[21];
    long long class_wave_trips[21];
} dsscf_stats;

/* Last error message of the calling thread ("" if none). */
const char* dsscf_last_error(void);

/* Library facts: highest angular momentum built, number of classes, the
 * (la,lb,lc,ld) of class c, and the size of the AO-transform table. */
int dsscf_max_l(void);
int dsscf_class_count(void);
int dsscf_class_l(int c, int* l4);
int dsscf_xform_size(void);

/* Context bound to HIP device `device` with its own stream.  Replaces the
 * namespace-scope state of src/chem/scf/ExactCoulomb.cpp:13-15
 * (repulsion_tensor, diagonal_size, tensor_size). */
int dsscf_ctx_create(dsscf_ctx** ctx, int device);
int dsscf_ctx_destroy(dsscf_ctx* ctx);
int dsscf_set_options(dsscf_ctx* ctx, const dsscf_options* opt);
/* Device memory the context holds (bytes, approximate: its buffers after
 * dsscf_prepare; no reference counterpart). */
long long dsscf_ctx_device_bytes(dsscf_ctx* ctx);

/* Shell data, shells in AO order (atom -> shell in basis-file order, each
 * shell contributing 2l+1 AOs m=-l..l: src/chem/scf/SCFBasis.cpp:120-140;
 * with dsscf_options.cartesian, l >= 2 shells contribute ncart(l) AOs).
 *   shell_l[n_shell], shell_nprim[n_shell], shell_center[3*n_shell] (bohr)
 *   prim_exp[], prim_coef[]  -- the reference's normalized primitive weights
 *                               (GTOPrimitive::weight, src/chem/GTO.cpp:90-93)
 *   ao_xform[dsscf_xform_size()] -- for each l = 0..max_l, a (2l+1) x ncart(l)
 *       row-major matrix: weight of Cartesian monomial c (x-major order) in
 *       AO m (ContractedGaussian::spherical_harmonic, src/chem/GTO.cpp:15-88);
 *       then for each l = 2..max_l an ncart(l) x ncart(l) matrix, the
 *       weights of the Cartesian AOs (diagonal)
 * Replaces the std::vector<ContractedGaussian> basis that
 * assignExactRepulsionTensor reads (ExactCoulomb.cpp:41-72, SCFCommon.hpp:10). */
int dsscf_set_basis(dsscf_ctx* ctx, int n_shell, const int* shell_l, const int* shell_nprim,
                    const double* shell_center, const double* prim_exp, const double* prim_coef,
                    const double* ao_xform);

/* Build pair data, Schwarz bounds and quartet lists on the device.  Replaces
 * scf::assignExactRepulsionTensor (src/chem/scf/ExactCoulomb.cpp:41-72): no
 * tensor is stored, integrals are recomputed per Fock build (direct SCF). */
int dsscf_prepare(dsscf_ctx* ctx);

/* One Fock build: J[Ptot] and K[Palpha] in packed-U form from packed-U
 * densities (host pointers).  J_mn = sum_ls (mn|ls) Ptot_ls replaces
 * getExactCoulombMatrix (ExactCoulomb.cpp:103-112) as called by
 * computeCoulombTerms (SCFSolver.cpp:29-35); K_mn = sum_ls (ml|sn) Pa_ls
 * replaces getExactExchangeMatrix (ExactCoulomb.cpp:114-123) as called by
 * computeExchangeTerms (SCFSolver.cpp:37-43). */
int dsscf_jk(dsscf_ctx* ctx, const double* Ptot, const double* Palpha, double* J, double* K);

/* Unrestricted (open-shell) Fock build: J[Ptot], K[Palpha] and K[Pbeta] from
 * one pass over the integrals (host pointers, packed-U).  Replaces the
 * open-shell sweep of runSCFIteration (SCFSolver.cpp:145-166):
 * computeCoulombTerms for each spin (SCFSolver.cpp:29-35, both read the
 * total density) and computeExchangeTerms(spin) (SCFSolver.cpp:37-43 over
 * getExactExchangeMatrix(mu, nu, spin), ExactCoulomb.cpp:114-123).  The
 * first call allocates the beta buffers (two packed-U arrays and 2 to 17
 * full N x N arrays, as many as the context's J/K copies plus one); a
 * context over the LDS limit of the exchange rows (about
 * N = 400 with d shells) returns DSSCF_ERR_UNSUPPORTED. */
int dsscf_jk_uhf(dsscf_ctx* ctx, const double* Ptot, const double* Palpha, const double* Pbeta, double* J,
                 double* Kalpha, double* Kbeta);

/* As dsscf_jk with device pointers (device-resident density / output). */
int dsscf_jk_device(dsscf_ctx* ctx, const double* d_Ptot, const double* d_Palpha, double* d_J, double* d_K);

/* Device-resident form: upload packed-U densities once, run Fock builds on
 * them (each call synchronizes the context's stream), read J and K back. */
int dsscf_set_density(dsscf_ctx* ctx, const double* Ptot, const double* Palpha);
int dsscf_jk_resident(dsscf_ctx* ctx);
int dsscf_get_jk(dsscf_ctx* ctx, double* J, double* K);
/* dsscf_jk_resident that also adds each class kernel's duration in ms (HIP
 * events on the stream the kernel runs on; the build's launches are made
 * directly rather than replayed from its captured graph; FP64 and FP32 lists
 * summed) to kernel_ms[class] (dsscf_class_count() entries, not cleared). */
int dsscf_jk_resident_timed(dsscf_ctx* ctx, double* kernel_ms);

/* One-electron matrices over the context's AOs (packed-U, N(N+1)/2 each):
 * overlap S, kinetic T and nuclear attraction V of the n_atom nuclei
 * (charges Z, positions xyz in bohr).  Replaces the host integrals of
 * SCFBasis.cpp:56-83 (flo::overlapIntegral, kineticEnergyIntegral,
 * nuclearAttractionIntegral, GaussianIntegral.cpp:29-332).  Needs
 * dsscf_prepare; synchronous. */
int dsscf_one_electron(dsscf_ctx* ctx, int n_atom, const int* Z, const double* xyz, double* S, double* T, double* V);

/* (ij|kl) for n AO quartets ijkl[4*t..4*t+3], written to out[t].  The unit
 * API flo::electronRepulsionIntegral(mu, sigma, nu, tau) = (mu nu|sigma tau)
 * (src/chem/GaussianIntegral.cpp:561-642, GaussianIntegral.hpp:12) maps to
 * ijkl = {mu, nu, sigma, tau}.  No screening is applied. */
int dsscf_eri_batch(dsscf_ctx* ctx, const int* ijkl, long long n, double* out);

/* dsscf_eri_batch without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments
 * and results as dsscf_eri_batch; the item list and AO-block outputs of each
 * class are staged in a fixed 16 MiB device array per GPU instead of
 * per-call allocations.  A call whose largest class needs more staging
 * (16 B per unique shell quartet plus 8 B per AO integral of its blocks)
 * fails with DSSCF_ERR_ARG before any work.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (from several threads) race. */
int dsscf_eri_batch_noalloc(dsscf_ctx* ctx, const int* ijkl, long long n, double* out);

/* F_0..F_L (L <= 8) of the Boys function at n arguments T[], evaluated on the
 * device exactly as the ERI kernels do: mode 0 = the reference series
 * (boysFunctionExact, GaussianIntegral.cpp:209-223), 1 = accurate from the
 * 801-row Taylor table, 2 = accurate from the interval table staged in LDS
 * (the FP64 J/K kernels' form).  F[k*(L+1)+m] receives F_m(T[k]).  A
 * checking aid (no reference counterpart beyond the series); synchronous. */
int dsscf_boys_eval(dsscf_ctx* ctx, int mode, int L, int n, const double* T, double* F);

/* Two-electron part of the RHF nuclear gradient at alpha density Pa
 * (packed-U; Ptot = 2 Pa): grad[3*natom] (x, y, z per atom, Eh/bohr)
 * receives sum over AO quartets of the pair density times d(ij|kl)/dR, the
 * quantity scf::addExactTwoElectronGradient adds (ExactCoulomb.cpp:125-195,
 * over flo::electronRepulsionGradient, GaussianIntegral.cpp:1052-1459);
 * ao_atom[N] gives each AO's atom.  Uses the prepared quartet lists (Schwarz
 * screening as the Fock build; 0 keeps every quartet).  s, p and d shells,
 * spherical or Cartesian d; the class with the most LDS, (dd|dd), leaves
 * room for the block's copy of about 1800 atoms (DSSCF_ERR_HIP beyond).
 * Synchronous. */
int dsscf_eri_gradient(dsscf_ctx* ctx, const double* Pa, int natom, const int* ao_atom, double* grad);

/* One-electron part of the nuclear gradient on the device: grad[3*natom]
 * receives sum_mn Pt_mn dT_mn/dR - W_mn dS_mn/dR + Pt_mn dV_mn/dR for the
 * packed-U total density Pt and energy-weighted density W, the kinetic,
 * overlap and nuclear-attraction terms of scf::getEnergyGradient
 * (Gradient.cpp:61-116, over overlapGradient / kineticEnergyGradient /
 * nuclearPotentialGradient, GaussianIntegral.cpp:707-1050); nuclei Z[natom]
 * at xyz[3*natom] (bohr), ao_atom[N] each AO's atom.  Spherical AOs
 * (DSSCF_ERR_UNSUPPORTED for a Cartesian-d context).  Synchronous. */
int dsscf_one_electron_gradient(dsscf_ctx* ctx, int natom, const int* Z, const double* xyz, const int* ao_atom,
                                const double* Pt, const double* W, double* grad);

/* Schwarz bounds Q_ab of all unique shell pairs (a >= b, index a(a+1)/2+b). */
int dsscf_schwarz(dsscf_ctx* ctx, double* Q);

int dsscf_get_stats(dsscf_ctx* ctx, dsscf_stats* st);

/* Run K Fock builds back to back on the current device density and return
 * the average device time per build in milliseconds (HIP events on the
 * context's stream).  Timing helper for benchmarks. */
int dsscf_time_jk(dsscf_ctx* ctx, int warmup, int iters, double* ms_per_build, double* ms_dominant_kernel,
                  int* dominant_class);

/* Average device time of each class's kernel(s) over `iters` Fock builds
 * (ms_per_class[dsscf_class_count()]). */
int dsscf_time_jk_classes(dsscf_ctx* ctx, int iters, double* ms_per_class);

/* ---------------------------------------------------------------- batches --
 * Fock builds of several prepared contexts (different molecules and bases,
 * one device, one Boys mode) fused into one launch per shell-quartet class:
 * a prepare kernel, one kernel per class over all contexts' quartet lists,
 * a fold kernel (captured once into a hipGraph and replayed when graphs are
 * on).  No reference counterpart: the reference runs one molecule at a time. */
typedef struct dsscf_batch dsscf_batch;
int dsscf_batch_create(dsscf_batch** batch, int device);
int dsscf_batch_destroy(dsscf_batch* batch);
/* Set the member contexts (replaces any previous set). */
int dsscf_batch_set(dsscf_batch* batch, dsscf_ctx* const* ctxs, int n);
int dsscf_batch_set_graph(dsscf_batch* batch, int use_graph);
/* J[Ptot_i], K[Palpha_i] for every member i (packed-U host arrays). */
int dsscf_batch_jk(dsscf_batch* batch, const double* const* Ptot, const double* const* Palpha, double* const* J,
                   double* const* K);
/* Member contexts whose densities and J/K stay in caller-owned device
 * buffers (packed-U, one set per member, alive until the next set call):
 * the fused plan reads d_Ptot[i], d_Palpha[i] and writes d_J[i], d_K[i]
 * directly, with no host copies.  dsscf_batch_jk then refuses; call
 * dsscf_batch_jk_device, which returns when the build has finished.  The
 * caller must have finished its own writes to the densities. */
int dsscf_batch_set_device(dsscf_batch* batch, dsscf_ctx* const* ctxs, int n, const double* const* d_Ptot,
                           const double* const* d_Palpha, double* const* d_J, double* const* d_K);
int dsscf_batch_jk_device(dsscf_batch* batch);
/* Average device time of `iters` fused builds on the resident densities. */
int dsscf_batch_time(dsscf_batch* batch, int warmup, int iters, double* ms_per_build);
/* Choose threads per block for each thread-per-quartet class by timing
 * 64/128/256 on the batch's own quartet lists; chosen[dsscf_class_count()]
 * receives the choice (may be null). */
int dsscf_batch_autotune(dsscf_batch* batch, int* chosen);

/* Per-class schedule of a context's quartet lists, used by its next
 * dsscf_prepare (arrays of dsscf_class_count() entries; a null array or a 0
 * entry keeps the options' value or the default):
 *   work_item_flops -- sub-pair size (dsscf_options.work_item_flops) per class
 *   group_prims     -- lane-group target (dsscf_options.group_prims) per class
 *   kets_per_unit   -- quartets per work unit (one block), 64..4096 (rounded
 *                      up to a multiple of 64) instead of the size-based
 *                      default (64..512); larger values are refused
 * No reference counterpart.  Results agree to rounding for any schedule. */
int dsscf_set_class_schedule(dsscf_ctx* ctx, const double* work_item_flops, const double* group_prims,
                             const int* kets_per_unit);

/* The project's schedule autotuner: for each class of the batch's members,
 * times the class kernels on the members' own lists at several sub-pair
 * sizes, lane-group targets and kets per unit (one parameter at a time,
 * re-preparing the members for each candidate) and keeps per class the
 * fastest of each, then picks threads per block as dsscf_batch_autotune.
 * The members end prepared with the chosen schedule; the arrays
 * (dsscf_class_count() entries each, any may be null) receive it, ready for
 * dsscf_set_class_schedule on other contexts of the same kind of workload. */
int dsscf_batch_autotune_schedule(dsscf_batch* batch, double* work_item_flops, double* group_prims,
                                  int* kets_per_unit, int* block_size);

#ifdef __cplusplus
}
#endif
#endif /* DSSCF_H */
