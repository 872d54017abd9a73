// linalg.hpp -- host linear algebra for the SCF driver (the roles of
// DollarStoreSCF's src/lalib: packed-U symmetric storage, dense products,
// symmetric eigensolver in place of LAPACK dspev_, symmetric solve in place
// of dspsv_).  Own implementations; dense matrices are row-major N x N.
#pragma once
#include <vector>

namespace la {

typedef std::vector<double> Mat;  // row-major dense

// packed-U (element (y,x), y<=x at x(x+1)/2+y; Matrix.ipp:332-351) <-> dense
void unpack(const double* pk, int n, Mat& M);
void pack(const Mat& M, int n, double* pk);

// C = A * B (n x n), blocked loops
void gemm(const Mat& A, const Mat& B, Mat& C, int n);
// C = A^T
void transpose(const Mat& A, Mat& C, int n);

// Eigen-decomposition of a symmetric matrix (Householder tridiagonalization,
// implicit QR with Wilkinson shifts).  Eigenvalues ascending in w,
// eigenvectors in the columns of V (dspev_('V','U') semantics,
// src/lalib/Lapack.cpp:12-39).  Returns false if the QR steps do not converge.
bool syev(const Mat& A, int n, std::vector<double>& w, Mat& V);

// Solve A x = b for a small dense system by Gaussian elimination with
// partial pivoting (the role of dspsv_, src/lalib/Lapack.cpp:145-171).
// Returns false if A is singular.
bool solve(std::vector<double> A, std::vector<double>& b, int n);

}  // namespace la
