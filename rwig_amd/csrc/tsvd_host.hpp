// Host-side small dense linear algebra of the truncated SVD (pure C++, no
// device code): the leading eigenvectors of G = T^T T by subspace iteration
// with Rayleigh-Ritz, where the caller supplies the product Z = G Q (on the
// device, two GEMMs over T). Everything here is O(d * b^2) or O(b^3) with b
// the small block width; the O(R * d * b) work stays in the caller's product.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace rwig {
namespace tsvd_host {

// Uniform values in [-1, 1) from an integer key (splitmix64 finaliser, Steele,
// Lea & Flood, "Fast splittable pseudorandom number generators", 2014): a
// fixed start block, so that results do not depend on any global RNG state.
inline double hash_uniform(uint64_t key) {
  uint64_t z = key + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline double dot(const double *x, const double *y, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

// Eigen-decomposition of the symmetric b x b matrix H (column-major; destroyed)
// by cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations, 4th ed.,
// sec. 8.5.2-8.5.3): H = W diag(lam) W^T, returned with lam in descending
// order and W's columns to match.
inline void jacobi_eig(std::vector<double> &H, int b, std::vector<double> &W,
                       std::vector<double> &lam) {
  W.assign((size_t)b * b, 0.0);
  for (int i = 0; i < b; ++i) W[i + (size_t)i * b] = 1.0;
  auto h = [&](int i, int j) -> double & { return H[i + (size_t)j * b]; };
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int q = 0; q < b; ++q) {
      diag += h(q, q) * h(q, q);
      for (int p = 0; p < q; ++p) off += h(p, q) * h(p, q);
    }
    if (!(off > 1e-32 * diag)) break;
    for (int q = 1; q < b; ++q)
      for (int p = 0; p < q; ++p) {
        const double apq = h(p, q);
        if (apq == 0.0) continue;
        // rotation J (J_pp = c, J_pq = s, J_qp = -s, J_qq = c) with
        // (J^T H J)_pq = 0: t the smaller root of t^2 + 2 tau t - 1 = 0
        const double tau = (h(q, q) - h(p, p)) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < b; ++k) {  // H J
          const double hp = h(k, p), hq = h(k, q);
          h(k, p) = c * hp - s * hq;
          h(k, q) = s * hp + c * hq;
        }
        for (int k = 0; k < b; ++k) {  // J^T (H J)
          const double hp = h(p, k), hq = h(q, k);
          h(p, k) = c * hp - s * hq;
          h(q, k) = s * hp + c * hq;
        }
        for (int k = 0; k < b; ++k) {  // W J
          double &wp = W[k + (size_t)p * b], &wq = W[k + (size_t)q * b];
          const double a = wp, e = wq;
          wp = c * a - s * e;
          wq = s * a + c * e;
        }
      }
  }
  std::vector<int> order(b);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return h(x, x) > h(y, y); });
  std::vector<double> W2((size_t)b * b);
  lam.resize(b);
  for (int i = 0; i < b; ++i) {
    lam[i] = h(order[i], order[i]);
    std::copy(W.begin() + (size_t)order[i] * b, W.begin() + (size_t)(order[i] + 1) * b,
              W2.begin() + (size_t)i * b);
  }
  W.swap(W2);
}

// Orthonormalises the b columns of Z (d x b, column-major, in place) by
// modified Gram-Schmidt applied twice. A column that loses all but 1e-10 of
// its norm to the earlier ones (rank deficiency) is replaced by a hashed
// vector orthogonalised the same way, so the result always has b orthonormal
// columns (b <= d).
inline void orthonormalise(std::vector<double> &Z, int d, int b, uint64_t salt) {
  for (int j = 0; j < b; ++j) {
    double *z = &Z[(size_t)j * d];
    for (int attempt = 0;; ++attempt) {
      const double n0 = std::sqrt(dot(z, z, d));
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < j; ++i) {
          const double *q = &Z[(size_t)i * d];
          const double r = dot(q, z, d);
          for (int k = 0; k < d; ++k) z[k] -= r * q[k];
        }
      const double n1 = std::sqrt(dot(z, z, d));
      if (n0 > 0.0 && n1 > 1e-10 * n0) {
        for (int k = 0; k < d; ++k) z[k] /= n1;
        break;
      }
      for (int k = 0; k < d; ++k)
        z[k] = hash_uniform((salt * 1315423911ull + (uint64_t)j * 2654435761ull + attempt) *
                                (uint64_t)(d + 1) + (uint64_t)k);
    }
  }
}

struct EigResult {
  std::vector<double> V;    // d x kk: leading eigenvectors of G
  std::vector<double> lam;  // kk Ritz values, descending
  int iters = 0;
  double residual = 0.0;    // max_i ||G v_i - lam_i v_i|| / lam_0 at the last step
};

// Leading kk eigenpairs of the d x d symmetric positive semidefinite G,
// through apply(Q, Z): Z = G Q for a d x b column-major block Q. Block width
// b = min(d, max(kk + 16, 32)) (the product streams T once per step whatever
// the width, so a wide block costs little and converges faster); with b = d
// the first step is exact (Q spans everything). Stops when the Ritz residuals
// of the kk leading pairs reach 1e-13 of the largest eigenvalue, when they
// stop decreasing near the rounding floor, or after max_iter steps.
template <class Apply>
EigResult leading_eigvecs(int d, int kk, Apply apply, int max_iter = 500) {
  const int b = std::min(d, std::max(kk + 16, 32));
  std::vector<double> Q((size_t)d * b, 0.0), Z((size_t)d * b), H((size_t)b * b), W, lam;
  if (b == d) {
    for (int i = 0; i < d; ++i) Q[i + (size_t)i * d] = 1.0;
  } else {
    for (size_t i = 0; i < Q.size(); ++i) Q[i] = hash_uniform(i);
    orthonormalise(Q, d, b, 1);
  }
  EigResult out;
  std::vector<double> hist;
  std::vector<double> QW((size_t)d * b), ZW((size_t)d * b);
  for (int it = 1;; ++it) {
    apply(Q, Z);
    for (int j = 0; j < b; ++j)
      for (int i = 0; i <= j; ++i) {
        const double v = 0.5 * (dot(&Q[(size_t)i * d], &Z[(size_t)j * d], d) +
                                dot(&Q[(size_t)j * d], &Z[(size_t)i * d], d));
        H[i + (size_t)j * b] = v;
        H[j + (size_t)i * b] = v;
      }
    jacobi_eig(H, b, W, lam);
    // Ritz vectors Q W, their images Z W, and the residuals of the leading kk
    for (int j = 0; j < b; ++j)
      for (int k = 0; k < d; ++k) {
        double a = 0.0, c = 0.0;
        for (int i = 0; i < b; ++i) {
          const double w = W[i + (size_t)j * b];
          a += Q[k + (size_t)i * d] * w;
          c += Z[k + (size_t)i * d] * w;
        }
        QW[k + (size_t)j * d] = a;
        ZW[k + (size_t)j * d] = c;
      }
    double res = 0.0;
    for (int j = 0; j < kk; ++j) {
      double s = 0.0;
      for (int k = 0; k < d; ++k) {
        const double r = ZW[k + (size_t)j * d] - lam[j] * QW[k + (size_t)j * d];
        s += r * r;
      }
      res = std::max(res, std::sqrt(s));
    }
    res = lam[0] > 0.0 ? res / lam[0] : 0.0;
    hist.push_back(res);
    bool stop = res <= 1e-13 || it >= max_iter;
    if (!stop && it >= 6 && res < 1e-9) {  // at the rounding floor: no progress in 5 steps
      const double best = *std::min_element(hist.end() - 6, hist.end() - 1);
      stop = res >= best;
    }
    if (stop) {
      out.V.assign(QW.begin(), QW.begin() + (size_t)d * kk);
      out.lam.assign(lam.begin(), lam.begin() + kk);
      out.iters = it;
      out.residual = res;
      return out;
    }
    Q = ZW;
    orthonormalise(Q, d, b, (uint64_t)it + 1);
  }
}

}  // namespace tsvd_host
}  // namespace rwig
