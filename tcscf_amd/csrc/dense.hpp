// FP64 dense eigen-solvers for the small (nb <= a few hundred) SCF matrices.
#pragma once

#include <complex>
#include <vector>

namespace tc::dense {

using cplx = std::complex<double>;

// F x = lambda S x with F symmetric, S symmetric positive definite (the role of
// LAPACK hegvd in the reference).  Solved per symmetry block (connected
// components of the nonzero pattern of F and S).  evals ascending; C (n x n,
// row-major) holds S-orthonormal eigenvectors as columns.  false if S is not SPD.
bool gen_sym_eig(int n, const double* F, const double* S, std::vector<double>& evals, std::vector<double>& C);

// F x = lambda S x with F general real, S SPD (the role of LAPACK ggev).  Per
// symmetry block; evals unsorted (block order), X columns are the right
// eigenvectors.  false if S is not SPD or the QR iteration fails.
bool gen_nonsym_eig(int n, const double* F, const double* S, std::vector<cplx>& evals, std::vector<cplx>& X);

// Modified Gram-Schmidt in the S metric (src/orthogonalization.hpp:253-271):
// vector 0 gets the metric norm, later vectors the Euclidean norm when
// euclid_after_first (the reference's behaviour), else the metric norm.
// vecs: k x n row-major.
void mgs(int k, int n, std::vector<cplx>& vecs, const double* S, bool euclid_after_first);

}  // namespace tc::dense
