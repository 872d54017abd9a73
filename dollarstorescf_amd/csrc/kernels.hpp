// kernels.hpp -- host-side launch interface of eri_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "device_common.hpp"

int dk_xform_size();
int dk_max_l();
int dk_n_class();
void dk_class_l(int cls, int* l4);
void dsscf_upload_xform(const double* x, int n, hipStream_t s);

// Fused J/K launch of one class over nseg context segments (block_size is
// ignored by the wave-cooperative classes, which use one 64-lane block per
// quartet).
// fp32: the FP32 form of a thread-per-quartet class (always accurate Boys).
// maxN: largest AO count among the segments (sizes the LDS exchange rows);
// maxBra: most primitive pairs of any unit's bra (sizes the LDS bra records);
// nblocks: work units (thread-per-quartet classes) or blocks
// (wave-cooperative classes, which ignore maxBra).
// cart: d shells as Cartesian functions (the classes' Cartesian-d form).
// nks: 2 when a segment is unrestricted (JKSeg::Db set), else 1.
hipError_t dsscf_launch_jk(int cls, int boys_mode, int fp32, const JKSeg* segs, const int* blk0, int nseg,
                           int nblocks, int block_size, int maxN, int nks, int maxBra, int cart, const double* btab,
                           double scale, hipStream_t s);
int dsscf_class_is_wave(int cls);
// Shell quartets per wavefront of a wave-cooperative class (1 for the others).
int dsscf_wave_group(int cls, int cart);
hipError_t dsscf_launch_eri_block(int cls, int boys_mode, int cart, const int4* items, int n, const ShellPair* pairs,
                                  const PrimPair* pp, const double* btab, double* out, double scale,
                                  hipStream_t s);
hipError_t dsscf_launch_schwarz(int la, int lb, int boys_mode, int cart, const int* plist, int n,
                                const ShellPair* pairs, const PrimPair* pp, const double* btab, double* Q, double scale,
                                hipStream_t s);
hipError_t dsscf_launch_fill(double* p, size_t n, double v, hipStream_t s);
// S, T, V (packed-U) of the pairs plist[0..n) of pair class (la, lb), la >= lb;
// atoms: x, y, z, Z per nucleus; xform: the AO transform (device copy).
hipError_t dsscf_launch_one_electron(int la, int lb, int boys_mode, const int* plist, int n, const ShellPair* pairs,
                                     const PrimPair* pp, const double* atoms, int natom, const double* btab,
                                     const double* xform, int cart, double* S, double* T, double* V, hipStream_t s);
// One-electron part of the nuclear gradient over the pairs plist[0..n) of
// pair class (la, lb) (spherical AOs): grad[3*natom] += the kinetic, overlap
// and nuclear-attraction terms for the packed-U total density Pt and
// energy-weighted density W; ao_atom[N] the atom of each AO.
hipError_t dsscf_launch_one_electron_grad(int la, int lb, int boys_mode, const int* plist, int n,
                                          const ShellPair* pairs, const PrimPair* pp, const double* atoms, int natom,
                                          const double* btab, const double* xform, const int* ao_atom,
                                          const double* Pt, const double* W, double* grad, hipStream_t s);
hipError_t dsscf_launch_prep(const MatSeg* segs, int nseg, int total_rows, int maxN, hipStream_t s);
hipError_t dsscf_launch_fold(const MatSeg* segs, int nseg, int total_rows, int maxN, hipStream_t s);
// Sum the J_cd slot records into copy 0 of Jh: nblocks = rows over all segments.
hipError_t dsscf_launch_jcd(const JcdSeg* segs, const int* blk0, int nseg, int nblocks, hipStream_t s);
// F_0..F_L (L <= 8) at n arguments in device Boys mode `mode` (BOYS_*):
// out[k*(L+1)+m].
hipError_t dsscf_launch_boys_eval(int mode, int L, const double* T, int n, const double* btab, double* out,
                                  hipStream_t s);
// ERI part of the nuclear gradient for one s/p class (l4) over nseg context
// segments of J/K work units: grad[3*natom] += its contribution (see
// grad_kernels.hip); xform the AO transform, ao_atom[N] the atom of each AO.
// The same for a class with a d shell (wave-cooperative, one 64-lane block
// per work unit, or with flat != 0 per dsscf_grad_flat_chunk() quartets of
// the segments' flat lists); cart: the Cartesian-d form.  Fails with
// hipErrorInvalidValue when the class's LDS tables and the 3 natom doubles
// of the block's gradient copy exceed the 160 KiB of a workgroup.
hipError_t dsscf_launch_gradw(const int* l4, int boys_mode, int cart, int flat, const JKSeg* segs, const int* blk0,
                              int nseg, int nblocks, const double* btab, const double* xform, const int* ao_atom,
                              double* grad, int natom, double scale, hipStream_t s);
int dsscf_grad_flat_chunk();
// 1 if class l4's gradient runs in the wave-cooperative kernel (every class
// with a d shell and the s/p classes with many accumulators), else 0.
int dsscf_grad_is_wave(const int* l4);
hipError_t dsscf_launch_grad(const int* l4, int boys_mode, const JKSeg* segs, const int* blk0, int nseg, int nblocks,
                             const double* btab, const double* xform, const int* ao_atom, double* grad, int natom,
                             double scale, hipStream_t s);
