// FP64 dense eigen-solvers for the small (nb <= a few hundred) SCF matrices.
#pragma once

#include <complex>
#include <deque>
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

// Pulay's direct inversion in the iterative subspace (Pulay, Chem. Phys. Lett.
// 73, 393 (1980); J. Comput. Chem. 3, 556 (1982)): keeps the last `depth`
// (Fock, error) pairs and replaces F by sum_i c_i F_i with c minimising
// |sum_i c_i e_i| subject to sum_i c_i = 1.  Vectors are flat (all spins).
class Diis {
 public:
  explicit Diis(int depth) : depth_(depth) {}
  // Adds (F, e) and overwrites F with the extrapolation (F is kept as is while
  // fewer than two pairs are stored or the subspace system is singular).
  void extrapolate(std::vector<double>& F, const std::vector<double>& e);

 private:
  int depth_;
  std::deque<std::vector<double>> f_, e_;
  std::deque<std::deque<double>> b_;  // b_[i][j] = <e_i, e_j> for j <= i (the stored pairs)
};

// DIIS error of a Hermitian Fock operator, per spin: F D S - S D F.
void commutator_error(int nspin, int nb, const double* F, const double* D, const double* S, std::vector<double>& e);
// DIIS error of a non-Hermitian (TC) Fock operator, per spin: F D - S D F D,
// which vanishes when the occupied S-orthonormal orbitals are right
// eigenvectors of F.
void right_eigen_error(int nspin, int nb, const double* F, const double* D, const double* S, std::vector<double>& e);

}  // namespace tc::dense
