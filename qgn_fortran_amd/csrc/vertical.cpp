// Vertical structure of the layered model: the LAPACK DGEEV step of
// Initialize (QGN_Module.f90:84-148) done without LAPACK.
//
// The stretching matrix L is tridiagonal with H_k L(k,k+1) = H_{k+1} L(k+1,k),
// so A = D^{1/2} L D^{-1/2} (D = diag H) is symmetric.  Its eigenpairs come
// from cyclic Jacobi rotations (accurate to a few ulps for the small nz of a
// layered model); the eigenvectors of L are D^{-1/2} w.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "qgn.h"

namespace {

void jacobi_eig(int n, std::vector<double>& A, std::vector<double>& V) {
  // A row-major n x n symmetric (destroyed: diagonal holds eigenvalues);
  // V columns are eigenvectors.
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double a = A[(size_t)i * n + j];
        if (i != j) off += a * a;
        else diag += a * a;
      }
    if (off <= 1e-36 * diag || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // rotate rows/cols p, q
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
}

}  // namespace

extern "C" int qgn_vertical_modes(int nz, double Htot, const double* H_in, const double* S, const double* uBar,
                                  double* H_out, double* L_out, double* Modes, double* EVals, double* qyBar) {
  if (nz < 2 || !H_in || !S || !H_out || !Modes || !EVals) return QGN_EINVAL;
  std::vector<double> H(nz);
  double sum = 0.0;
  for (int k = 0; k < nz; ++k) sum += H_in[k];
  if (!(sum > 0)) return QGN_EINVAL;
  for (int k = 0; k < nz; ++k) H[k] = Htot * H_in[k] / sum;  // QGN_Module.f90:87
  // L, column-major L(k,kk) at [kk*nz + k]   (QGN_Module.f90:102-113)
  std::vector<double> L((size_t)nz * nz, 0.0);
  auto at = [&](int r, int c) -> double& { return L[(size_t)c * nz + r]; };
  at(0, 1) = 2.0 * S[1] / (H[0] * (H[0] + H[1]));
  at(0, 0) = -at(0, 1);
  for (int k = 1; k < nz - 1; ++k) {
    at(k, k - 1) = 2.0 * S[k] / (H[k] * (H[k] + H[k - 1]));
    at(k, k + 1) = 2.0 * S[k + 1] / (H[k] * (H[k] + H[k + 1]));
    at(k, k) = -(at(k, k - 1) + at(k, k + 1));
  }
  at(nz - 1, nz - 2) = 2.0 * S[nz - 1] / (H[nz - 1] * (H[nz - 1] + H[nz - 2]));
  at(nz - 1, nz - 1) = -at(nz - 1, nz - 2);
  // symmetric A = D^1/2 L D^-1/2, row-major
  std::vector<double> A((size_t)nz * nz, 0.0), V;
  for (int r = 0; r < nz; ++r)
    for (int c = 0; c < nz; ++c) A[(size_t)r * nz + c] = std::sqrt(H[r] / H[c]) * at(r, c);
  for (int r = 0; r < nz; ++r)  // exact symmetry before the rotations
    for (int c = r + 1; c < nz; ++c) {
      const double m = 0.5 * (A[(size_t)r * nz + c] + A[(size_t)c * nz + r]);
      A[(size_t)r * nz + c] = A[(size_t)c * nz + r] = m;
    }
  jacobi_eig(nz, A, V);
  std::vector<int> idx(nz);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(),
            [&](int a, int b) { return A[(size_t)a * nz + a] > A[(size_t)b * nz + b]; });
  for (int m = 0; m < nz; ++m) {
    const int src = idx[m];
    EVals[m] = A[(size_t)src * nz + src];
    std::vector<double> v(nz);
    for (int k = 0; k < nz; ++k) v[k] = V[(size_t)k * nz + src] / std::sqrt(H[k]);
    double nrm = 0.0;
    for (int k = 0; k < nz; ++k) nrm += H[k] * v[k] * v[k];
    const double sc = std::copysign(1.0, v[0]) / std::sqrt(nrm / Htot);  // QGN_Module.f90:125-127
    for (int k = 0; k < nz; ++k) Modes[(size_t)m * nz + k] = v[k] * sc;
  }
  for (int k = 0; k < nz; ++k) H_out[k] = H[k];
  if (L_out)
    for (size_t i = 0; i < L.size(); ++i) L_out[i] = L[i];
  if (qyBar && uBar)
    for (int k = 0; k < nz; ++k) {  // qyBar = matmul(L, -uBar)  (QGN_Module.f90:148)
      double s = 0.0;
      for (int kk = 0; kk < nz; ++kk) s += at(k, kk) * (-uBar[kk]);
      qyBar[k] = s;
    }
  return QGN_OK;
}
