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

hipError_t dsscf_launch_jk(int cls, int boys_mode, const uint2* q, int nq, const ShellPair* pairs,
                           const PrimPair* pp, const double* btab, const double* Dt, const double* Da,
                           double* Jh, double* Kh, int N, double scale, hipStream_t s);
hipError_t dsscf_launch_eri_block(int cls, int boys_mode, const int4* items, int n, const ShellPair* pairs,
                                  const PrimPair* pp, const double* btab, double* out, double scale,
                                  hipStream_t s);
hipError_t dsscf_launch_schwarz(int la, int lb, int boys_mode, const int* plist, int n, const ShellPair* pairs,
                                const PrimPair* pp, const double* btab, double* Q, double scale, hipStream_t s);
hipError_t dsscf_launch_fill(double* p, size_t n, double v, hipStream_t s);
hipError_t dsscf_launch_unpack(const double* pk, double* full, int N, hipStream_t s);
hipError_t dsscf_launch_fold(const double* Jh, const double* Kh, double* J, double* K, int N, hipStream_t s);
