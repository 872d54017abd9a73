// launch.hpp -- host launch interface of one shell-quartet class (defined in
// eri_class.hpp and instantiated in that class's generated translation unit).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"

template <int a, int b, int c, int d, int wave>
struct ClassLaunch {
    // cart: the class's Cartesian-d form (ignored by classes without a d shell)
    // nks: 2 when a segment is unrestricted (K[Db] rows beside the K[Da] rows), else 1
    static hipError_t jk(int bm, int fp32, const JKSeg* segs, const int* blk0, int nseg, int nblocks, int bs,
                         int maxN, int nks, int maxBra, int cart, const double* btab, double scale, hipStream_t s);
    static hipError_t block(int bm, int cart, const int4* items, int n, const ShellPair* pairs, const PrimPair* pp,
                            const double* btab, double* out, double scale, hipStream_t s);
    static hipError_t schwarz(int bm, int cart, const int* plist, int n, const ShellPair* pairs, const PrimPair* pp,
                              const double* btab, double* Q, double scale, hipStream_t s);
    static hipError_t upload_xform(const double* x, int n, hipStream_t s);
    static int wave_group(int cart);

private:
    static constexpr bool kHasD = a >= 2 || b >= 2 || c >= 2 || d >= 2;
    template <int C>
    static hipError_t jk_c(int bm, int fp32, const JKSeg* segs, const int* blk0, int nseg, int nblocks, int bs,
                           int maxN, int nks, int maxBra, const double* btab, double scale, hipStream_t s);
    template <int C>
    static hipError_t block_c(int bm, const int4* items, int n, const ShellPair* pairs, const PrimPair* pp,
                              const double* btab, double* out, double scale, hipStream_t s);
    template <int C>
    static hipError_t schwarz_c(int bm, const int* plist, int n, const ShellPair* pairs, const PrimPair* pp,
                                const double* btab, double* Q, double scale, hipStream_t s);
    template <int BM, int MODE, int C>
    static void wave_go(const void* items, int n, const ShellPair* pairs, const PrimPair* pp, const double* btab,
                        double* out, double scale, hipStream_t s);
};
