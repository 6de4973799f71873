// FP64 dense eigen-solvers: Cholesky reduction to a standard problem, then
//   symmetric:     Householder tridiagonalisation + implicit QL with Wilkinson
//                  shifts (EISPACK tql2 / Numerical Recipes tqli),
//   nonsymmetric:  Householder Hessenberg reduction + complex Schur form by
//                  single-shift QR (Golub & Van Loan 7.5) with accumulated
//                  transformations, eigenvectors by back-substitution on T.
#include "dense.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace tc::dense {
namespace {

bool cholesky(int n, const std::vector<double>& S, std::vector<double>& L) {
  L.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = S[j * n + j];
    for (int k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0)) return false;
    const double ljj = std::sqrt(d);
    L[j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double s = S[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / ljj;
    }
  }
  return true;
}

// A <- L^{-1} A L^{-T}
void to_standard(int n, const std::vector<double>& L, std::vector<double>& A) {
  for (int c = 0; c < n; ++c)  // A <- L^{-1} A
    for (int i = 0; i < n; ++i) {
      double s = A[i * n + c];
      for (int k = 0; k < i; ++k) s -= L[i * n + k] * A[k * n + c];
      A[i * n + c] = s / L[i * n + i];
    }
  for (int r = 0; r < n; ++r)  // A <- A L^{-T}
    for (int i = 0; i < n; ++i) {
      double s = A[r * n + i];
      for (int k = 0; k < i; ++k) s -= A[r * n + k] * L[i * n + k];
      A[r * n + i] = s / L[i * n + i];
    }
}

template <class T>
void solve_lt(int n, const std::vector<double>& L, const T* y, T* x) {  // x = L^{-T} y
  for (int i = n - 1; i >= 0; --i) {
    T s = y[i];
    for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * x[k];
    x[i] = s / L[i * n + i];
  }
}

// Symmetric A (n x n): eigenvalues d ascending, eigenvectors as columns of Z.
bool sym_eig(int n, std::vector<double> A, std::vector<double>& d, std::vector<double>& Z) {
  Z.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Z[i * n + i] = 1;
  // Householder tridiagonalisation, A <- P A P, Z <- Z P.
  for (int k = 0; k + 2 < n; ++k) {
    double nx = 0;
    for (int i = k + 1; i < n; ++i) nx += A[i * n + k] * A[i * n + k];
    nx = std::sqrt(nx);
    if (nx == 0) continue;
    const double x0 = A[(k + 1) * n + k];
    const double alpha = x0 > 0 ? -nx : nx;
    std::vector<double> v(n, 0.0);
    for (int i = k + 1; i < n; ++i) v[i] = A[i * n + k];
    v[k + 1] -= alpha;
    double vv = 0;
    for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
    if (vv == 0) continue;
    const double beta = 2.0 / vv;
    // p = beta A v, K = beta/2 v.p, q = p - K v, A -= v q^T + q v^T
    std::vector<double> p(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int j = k + 1; j < n; ++j) s += A[i * n + j] * v[j];
      p[i] = beta * s;
    }
    double K = 0;
    for (int i = k + 1; i < n; ++i) K += v[i] * p[i];
    K *= 0.5 * beta;
    for (int i = 0; i < n; ++i) p[i] -= K * v[i];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) A[i * n + j] -= v[i] * p[j] + p[i] * v[j];
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int j = k + 1; j < n; ++j) s += Z[i * n + j] * v[j];
      s *= beta;
      for (int j = k + 1; j < n; ++j) Z[i * n + j] -= s * v[j];
    }
  }
  d.assign(n, 0.0);
  std::vector<double> e(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = A[i * n + i];
  for (int i = 0; i + 1 < n; ++i) e[i] = A[(i + 1) * n + i];
  // Implicit QL with Wilkinson shifts.
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-17 * dd) break;
      }
      if (m != l) {
        if (iter++ == 100) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
        double s = 1, c = 1, p = 0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = Z[k * n + i + 1];
            Z[k * n + i + 1] = s * Z[k * n + i] + c * f;
            Z[k * n + i] = c * Z[k * n + i] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  // sort ascending
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> d2(n), Z2((size_t)n * n);
  for (int c = 0; c < n; ++c) {
    d2[c] = d[idx[c]];
    for (int r = 0; r < n; ++r) Z2[r * n + c] = Z[r * n + idx[c]];
  }
  d.swap(d2);
  Z.swap(Z2);
  return true;
}

// Complex Schur form of a real matrix: A = Q T Q^H; returns eigenvalues
// (diag T) and eigenvectors Q y with y from T.
bool nonsym_eig(int n, const std::vector<double>& A0, std::vector<cplx>& ev, std::vector<cplx>& X) {
  std::vector<cplx> H((size_t)n * n), Q((size_t)n * n, 0.0);
  for (int i = 0; i < n * n; ++i) H[i] = A0[i];
  for (int i = 0; i < n; ++i) Q[i * n + i] = 1;
  for (int k = 0; k + 2 < n; ++k) {  // Householder to Hessenberg
    double nx = 0;
    for (int i = k + 1; i < n; ++i) nx += std::norm(H[i * n + k]);
    nx = std::sqrt(nx);
    if (nx == 0) continue;
    const cplx x0 = H[(k + 1) * n + k];
    const cplx ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cplx(1);
    std::vector<cplx> v(n, 0.0);
    for (int i = k + 1; i < n; ++i) v[i] = H[i * n + k];
    v[k + 1] += ph * nx;
    double vv = 0;
    for (int i = k + 1; i < n; ++i) vv += std::norm(v[i]);
    if (vv == 0) continue;
    const double beta = 2.0 / vv;
    for (int c = 0; c < n; ++c) {  // H <- P H
      cplx s = 0;
      for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * H[i * n + c];
      s *= beta;
      for (int i = k + 1; i < n; ++i) H[i * n + c] -= v[i] * s;
    }
    for (int r = 0; r < n; ++r) {  // H <- H P, Q <- Q P
      cplx s = 0, t = 0;
      for (int i = k + 1; i < n; ++i) {
        s += H[r * n + i] * v[i];
        t += Q[r * n + i] * v[i];
      }
      s *= beta;
      t *= beta;
      for (int i = k + 1; i < n; ++i) {
        H[r * n + i] -= s * std::conj(v[i]);
        Q[r * n + i] -= t * std::conj(v[i]);
      }
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j + 1 < i; ++j) H[i * n + j] = 0;
  int hi = n - 1, iter = 0;
  while (hi > 0) {
    int lo = hi;
    for (; lo > 0; --lo) {
      const double s = std::abs(H[(lo - 1) * n + lo - 1]) + std::abs(H[lo * n + lo]);
      if (std::abs(H[lo * n + lo - 1]) <= 1e-16 * (s > 0 ? s : 1.0)) {
        H[lo * n + lo - 1] = 0;
        break;
      }
    }
    if (lo == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++iter > 2000) return false;
    const cplx a = H[(hi - 1) * n + hi - 1], b = H[(hi - 1) * n + hi], c = H[hi * n + hi - 1], d = H[hi * n + hi];
    const cplx half_tr = 0.5 * (a + d);
    const cplx disc = std::sqrt(half_tr * half_tr - (a * d - b * c));
    const cplx l1 = half_tr + disc, l2 = half_tr - disc;
    cplx mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
    if (iter % 10 == 0) mu += 0.75 * std::abs(H[hi * n + hi - 1]);
    for (int i = lo; i <= hi; ++i) H[i * n + i] -= mu;
    std::vector<cplx> gc(hi - lo), gs(hi - lo);
    for (int k = lo; k < hi; ++k) {
      const cplx x = H[k * n + k], y = H[(k + 1) * n + k];
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      const cplx cc = r > 0 ? x / r : cplx(1), ss = r > 0 ? y / r : cplx(0);
      gc[k - lo] = cc;
      gs[k - lo] = ss;
      for (int col = k; col < n; ++col) {
        const cplx h1 = H[k * n + col], h2 = H[(k + 1) * n + col];
        H[k * n + col] = std::conj(cc) * h1 + std::conj(ss) * h2;
        H[(k + 1) * n + col] = -ss * h1 + cc * h2;
      }
    }
    for (int k = lo; k < hi; ++k) {
      const cplx cc = gc[k - lo], ss = gs[k - lo];
      for (int row = 0; row <= std::min(k + 2, hi); ++row) {
        const cplx h1 = H[row * n + k], h2 = H[row * n + k + 1];
        H[row * n + k] = h1 * cc + h2 * ss;
        H[row * n + k + 1] = -h1 * std::conj(ss) + h2 * std::conj(cc);
      }
      for (int row = 0; row < n; ++row) {
        const cplx q1 = Q[row * n + k], q2 = Q[row * n + k + 1];
        Q[row * n + k] = q1 * cc + q2 * ss;
        Q[row * n + k + 1] = -q1 * std::conj(ss) + q2 * std::conj(cc);
      }
    }
    for (int i = lo; i <= hi; ++i) H[i * n + i] += mu;
  }
  ev.resize(n);
  for (int i = 0; i < n; ++i) ev[i] = H[i * n + i];
  double tnorm = 0;
  for (int i = 0; i < n * n; ++i) tnorm = std::max(tnorm, std::abs(H[i]));
  const double small = 1e-15 * (tnorm > 0 ? tnorm : 1.0);
  X.assign((size_t)n * n, 0.0);
  std::vector<cplx> y(n);
  for (int j = 0; j < n; ++j) {
    std::fill(y.begin(), y.end(), cplx(0));
    y[j] = 1;
    for (int i = j - 1; i >= 0; --i) {
      cplx s = 0;
      for (int k = i + 1; k <= j; ++k) s += H[i * n + k] * y[k];
      cplx den = H[i * n + i] - H[j * n + j];
      if (std::abs(den) < small) den = small;
      y[i] = -s / den;
    }
    double nrm = 0;
    for (int r = 0; r < n; ++r) {
      cplx s = 0;
      for (int k = 0; k <= j; ++k) s += Q[r * n + k] * y[k];
      X[r * n + j] = s;
      nrm += std::norm(s);
    }
    nrm = std::sqrt(nrm);
    for (int r = 0; r < n; ++r) X[r * n + j] /= nrm;
  }
  return true;
}

// Symmetry blocks: connected components of the nonzero pattern of F and S.
std::vector<std::vector<int>> symmetry_blocks(int n, const double* F, const double* S) {
  std::vector<int> label(n, -1);
  std::vector<std::vector<int>> blocks;
  for (int s = 0; s < n; ++s) {
    if (label[s] >= 0) continue;
    const int id = (int)blocks.size();
    std::vector<int> members{s}, todo{s};
    label[s] = id;
    while (!todo.empty()) {
      const int u = todo.back();
      todo.pop_back();
      for (int v = 0; v < n; ++v)
        if (label[v] < 0 && (F[u * n + v] != 0 || F[v * n + u] != 0 || S[u * n + v] != 0)) {
          label[v] = id;
          members.push_back(v);
          todo.push_back(v);
        }
    }
    std::sort(members.begin(), members.end());
    blocks.push_back(members);
  }
  return blocks;
}

}  // namespace

bool gen_sym_eig(int n, const double* F, const double* S, std::vector<double>& evals, std::vector<double>& C) {
  struct Pair {
    double e;
    std::vector<double> v;
  };
  std::vector<Pair> pairs;
  for (const auto& blk : symmetry_blocks(n, F, S)) {
    const int m = (int)blk.size();
    std::vector<double> A((size_t)m * m), Sb((size_t)m * m), L, d, Z;
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) {
        A[a * m + b] = 0.5 * (F[blk[a] * n + blk[b]] + F[blk[b] * n + blk[a]]);
        Sb[a * m + b] = S[blk[a] * n + blk[b]];
      }
    if (!cholesky(m, Sb, L)) return false;
    to_standard(m, L, A);
    for (int a = 0; a < m; ++a)
      for (int b = a + 1; b < m; ++b) A[a * m + b] = A[b * m + a] = 0.5 * (A[a * m + b] + A[b * m + a]);
    if (!sym_eig(m, A, d, Z)) return false;
    std::vector<double> y(m), x(m);
    for (int c = 0; c < m; ++c) {
      for (int r = 0; r < m; ++r) y[r] = Z[r * m + c];
      solve_lt(m, L, y.data(), x.data());
      Pair p{d[c], std::vector<double>(n, 0.0)};
      for (int r = 0; r < m; ++r) p.v[blk[r]] = x[r];
      pairs.push_back(std::move(p));
    }
  }
  std::stable_sort(pairs.begin(), pairs.end(), [](const Pair& a, const Pair& b) { return a.e < b.e; });
  evals.resize(n);
  C.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c) {
    evals[c] = pairs[c].e;
    for (int r = 0; r < n; ++r) C[r * n + c] = pairs[c].v[r];
  }
  return true;
}

bool gen_nonsym_eig(int n, const double* F, const double* S, std::vector<cplx>& evals, std::vector<cplx>& X) {
  evals.clear();
  std::vector<std::vector<cplx>> vecs;
  for (const auto& blk : symmetry_blocks(n, F, S)) {
    const int m = (int)blk.size();
    std::vector<double> A((size_t)m * m), Sb((size_t)m * m), L;
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) {
        A[a * m + b] = F[blk[a] * n + blk[b]];
        Sb[a * m + b] = S[blk[a] * n + blk[b]];
      }
    if (!cholesky(m, Sb, L)) return false;
    to_standard(m, L, A);
    std::vector<cplx> ev, Y;
    if (!nonsym_eig(m, A, ev, Y)) return false;
    std::vector<cplx> y(m), x(m);
    for (int c = 0; c < m; ++c) {
      for (int r = 0; r < m; ++r) y[r] = Y[r * m + c];
      solve_lt(m, L, y.data(), x.data());
      std::vector<cplx> full(n, 0.0);
      for (int r = 0; r < m; ++r) full[blk[r]] = x[r];
      evals.push_back(ev[c]);
      vecs.push_back(std::move(full));
    }
  }
  X.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) X[r * n + c] = vecs[c][r];
  return true;
}

void mgs(int k, int n, std::vector<cplx>& v, const double* S, bool euclid_after_first) {
  auto metric_ip = [&](int a, int b) {
    cplx d = 0;
    for (int i = 0; i < n; ++i) {
      cplx row = 0;
      for (int j = 0; j < n; ++j) row += S[i * n + j] * v[b * n + j];
      d += std::conj(v[a * n + i]) * row;
    }
    return d;
  };
  auto scale = [&](int a, double f) {
    for (int i = 0; i < n; ++i) v[a * n + i] *= f;
  };
  if (k == 0) return;
  scale(0, 1.0 / std::sqrt(metric_ip(0, 0).real()));
  for (int i = 1; i < k; ++i) {
    for (int j = i; j < k; ++j) {
      const cplx s = metric_ip(i - 1, j);
      for (int q = 0; q < n; ++q) v[j * n + q] -= s * v[(i - 1) * n + q];
    }
    double nv;
    if (euclid_after_first) {
      nv = 0;
      for (int q = 0; q < n; ++q) nv += std::norm(v[i * n + q]);
      nv = std::sqrt(nv);
    } else {
      nv = std::sqrt(metric_ip(i, i).real());
    }
    scale(i, 1.0 / nv);
  }
}

void Diis::extrapolate(std::vector<double>& F, const std::vector<double>& e) {
  // the new error's inner products with the stored ones (and itself); the
  // older entries of B are kept from earlier calls
  std::deque<double> row;
  for (size_t j = 0; j <= e_.size(); ++j) {
    const std::vector<double>& ej = j < e_.size() ? e_[j] : e;
    double t = 0;
    for (size_t q = 0; q < e.size(); ++q) t += e[q] * ej[q];
    row.push_back(t);
  }
  f_.push_back(F);
  e_.push_back(e);
  b_.push_back(std::move(row));
  auto drop_oldest = [&] {
    f_.pop_front();
    e_.pop_front();
    b_.pop_front();
    for (auto& r : b_) r.pop_front();
  };
  while ((int)f_.size() > depth_) drop_oldest();
  while (f_.size() >= 2) {
    const int m = (int)f_.size(), n = m + 1;
    // [B 1; 1^T 0] [c; lambda] = [0; 1], B_ij = <e_i, e_j>
    std::vector<double> M((size_t)n * n, 0.0), x(n, 0.0);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j <= i; ++j) M[i * n + j] = M[j * n + i] = b_[i][j];
    for (int i = 0; i < m; ++i) M[i * n + m] = M[m * n + i] = 1.0;
    x[m] = 1.0;
    // Gaussian elimination with partial pivoting, relative pivot threshold
    double scale = 0;
    for (int i = 0; i < m; ++i) scale = std::max(scale, std::fabs(M[i * n + i]));
    bool ok = true;
    for (int c = 0; c < n && ok; ++c) {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (std::fabs(M[r * n + c]) > std::fabs(M[piv * n + c])) piv = r;
      if (!(std::fabs(M[piv * n + c]) > 1e-14 * std::max(scale, 1e-300))) {
        ok = false;
        break;
      }
      if (piv != c) {
        for (int k = 0; k < n; ++k) std::swap(M[c * n + k], M[piv * n + k]);
        std::swap(x[c], x[piv]);
      }
      for (int r = c + 1; r < n; ++r) {
        const double f = M[r * n + c] / M[c * n + c];
        for (int k = c; k < n; ++k) M[r * n + k] -= f * M[c * n + k];
        x[r] -= f * x[c];
      }
    }
    if (ok) {
      for (int c = n - 1; c >= 0; --c) {
        double t = x[c];
        for (int k = c + 1; k < n; ++k) t -= M[c * n + k] * x[k];
        x[c] = t / M[c * n + c];
      }
      std::fill(F.begin(), F.end(), 0.0);
      for (int i = 0; i < m; ++i)
        for (size_t q = 0; q < F.size(); ++q) F[q] += x[i] * f_[i][q];
      return;
    }
    drop_oldest();  // singular subspace: drop the oldest pair and retry
  }
}

namespace {
// C = A B for n x n row-major.  Row i of C accumulates A[i][k] B[k][:] in k
// order (each element's sum in the order of a dot product over k, so the
// result is that of the i-j-k loop), with the inner loop over contiguous j.
void matmul(int n, const double* A, const double* B, double* C) {
  for (int i = 0; i < n; ++i) {
    double* c = C + (int64_t)i * n;
    for (int j = 0; j < n; ++j) c[j] = 0;
    for (int k = 0; k < n; ++k) {
      const double a = A[(int64_t)i * n + k];
      const double* b = B + (int64_t)k * n;
      for (int j = 0; j < n; ++j) c[j] += a * b[j];
    }
  }
}

// Spin s's F and D equal spin 0's bit for bit (a closed shell whose two spin
// operators coincide): its error vector is spin 0's.
bool same_as_spin0(int s, int64_t nb2, const double* F, const double* D) {
  return s > 0 && std::memcmp(F, F + s * nb2, sizeof(double) * nb2) == 0 &&
         std::memcmp(D, D + s * nb2, sizeof(double) * nb2) == 0;
}
}  // namespace

void commutator_error(int nspin, int nb, const double* F, const double* D, const double* S, std::vector<double>& e) {
  const int64_t nb2 = (int64_t)nb * nb;
  e.assign(nspin * nb2, 0.0);
  std::vector<double> FD(nb2), FDS(nb2), SD(nb2), SDF(nb2);
  for (int s = 0; s < nspin; ++s) {
    if (same_as_spin0(s, nb2, F, D)) {
      std::copy(e.begin(), e.begin() + nb2, e.begin() + s * nb2);
      continue;
    }
    matmul(nb, F + s * nb2, D + s * nb2, FD.data());
    matmul(nb, FD.data(), S, FDS.data());
    matmul(nb, S, D + s * nb2, SD.data());
    matmul(nb, SD.data(), F + s * nb2, SDF.data());
    for (int64_t q = 0; q < nb2; ++q) e[s * nb2 + q] = FDS[q] - SDF[q];
  }
}

void right_eigen_error(int nspin, int nb, const double* F, const double* D, const double* S, std::vector<double>& e) {
  const int64_t nb2 = (int64_t)nb * nb;
  e.assign(nspin * nb2, 0.0);
  std::vector<double> FD(nb2), SD(nb2), SDFD(nb2);
  for (int s = 0; s < nspin; ++s) {
    if (same_as_spin0(s, nb2, F, D)) {
      std::copy(e.begin(), e.begin() + nb2, e.begin() + s * nb2);
      continue;
    }
    matmul(nb, F + s * nb2, D + s * nb2, FD.data());
    matmul(nb, S, D + s * nb2, SD.data());
    matmul(nb, SD.data(), FD.data(), SDFD.data());
    for (int64_t q = 0; q < nb2; ++q) e[s * nb2 + q] = FD[q] - SDFD[q];
  }
}

}  // namespace tc::dense
