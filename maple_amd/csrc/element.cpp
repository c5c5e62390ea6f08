// Host-side element matrices of the periodic voxel discretization.
//
// The same quantities as the reference's hexahedron / hexahedron_scalar
// (_utils.py:49-109, 116-196): 3x3x3 Gauss-Legendre on a cube of side
// h = 1/r, B rows [xx, yy, zz, xy, yz, xz] with engineering shear, node
// order (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1),
// X0 from Ke[f,f] X0[f] = Fe[f] with DOFs {0,1,2,4,5,11} (elastic) or node 0
// (thermal) fixed.  Also the 8 child Galerkin matrices G_c = P_c^T Ke P_c of
// the first coarse level (EBE_GMG.py:202-224, 261, 284-294).
#include "element.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

void shape_grad(double xi, double eta, double zeta, double half, double g[3][8]) {
  for (int a = 0; a < 8; ++a) {
    const double sx = 2.0 * kCorner[a][0] - 1.0, sy = 2.0 * kCorner[a][1] - 1.0, sz = 2.0 * kCorner[a][2] - 1.0;
    const double fx = 1.0 + sx * xi, fy = 1.0 + sy * eta, fz = 1.0 + sz * zeta;
    g[0][a] = sx * fy * fz / 8.0 / half;
    g[1][a] = fx * sy * fz / 8.0 / half;
    g[2][a] = fx * fy * sz / 8.0 / half;
  }
}

// Solve A X = B in place (A n x n, B n x m, row-major), Gaussian elimination with partial pivoting.
bool solve_dense(int n, int m, std::vector<double>& A, std::vector<double>& B) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[piv * n + k])) piv = i;
    if (A[piv * n + k] == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
      for (int j = 0; j < m; ++j) std::swap(B[k * m + j], B[piv * m + j]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      if (f == 0.0) continue;
      for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      for (int j = 0; j < m; ++j) B[i * m + j] -= f * B[k * m + j];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    for (int j = 0; j < m; ++j) {
      double s = B[k * m + j];
      for (int i = k + 1; i < n; ++i) s -= A[k * n + i] * B[i * m + j];
      B[k * m + j] = s / A[k * n + k];
    }
  }
  return true;
}

void fixed_solve(int nd, int ncase, const double* Ke, const double* Fe, const std::vector<int>& fixed, double* X0) {
  const int ne = 8 * nd;
  std::vector<int> freeidx;
  for (int i = 0; i < ne; ++i) {
    bool f = false;
    for (int k : fixed) f |= (k == i);
    if (!f) freeidx.push_back(i);
  }
  const int nf = (int)freeidx.size();
  std::vector<double> A(nf * nf), B(nf * ncase);
  for (int i = 0; i < nf; ++i) {
    for (int j = 0; j < nf; ++j) A[i * nf + j] = Ke[freeidx[i] * ne + freeidx[j]];
    for (int j = 0; j < ncase; ++j) B[i * ncase + j] = Fe[freeidx[i] * ncase + j];
  }
  solve_dense(nf, ncase, A, B);
  std::memset(X0, 0, sizeof(double) * ne * ncase);
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < ncase; ++j) X0[freeidx[i] * ncase + j] = B[i * ncase + j];
}

}  // namespace

void hex8_elastic(int r, double E, double nu, double* Ke, double* Fe, double* X0) {
  const double lam = nu / (1.0 + nu) / (1.0 - 2.0 * nu) * E;
  const double mu = 1.0 / (2.0 * (1.0 + nu)) * E;
  double C[6][6] = {};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = lam;
  for (int i = 0; i < 3; ++i) C[i][i] = lam + 2.0 * mu;
  for (int i = 3; i < 6; ++i) C[i][i] = mu;
  const double half = 0.5 / r;
  const double p[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)};
  const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  std::memset(Ke, 0, sizeof(double) * 576);
  std::memset(Fe, 0, sizeof(double) * 144);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        double g[3][8];
        shape_grad(p[i], p[j], p[k], half, g);
        double B[6][24] = {};
        for (int a = 0; a < 8; ++a) {
          const int c = 3 * a;
          B[0][c] = g[0][a];
          B[1][c + 1] = g[1][a];
          B[2][c + 2] = g[2][a];
          B[3][c] = g[1][a];
          B[3][c + 1] = g[0][a];
          B[4][c + 1] = g[2][a];
          B[4][c + 2] = g[1][a];
          B[5][c] = g[2][a];
          B[5][c + 2] = g[0][a];
        }
        const double wt = w[i] * w[j] * w[k] * half * half * half;
        double CB[6][24];
        for (int s = 0; s < 6; ++s)
          for (int c = 0; c < 24; ++c) {
            double t = 0.0;
            for (int q = 0; q < 6; ++q) t += C[s][q] * B[q][c];
            CB[s][c] = t;
          }
        for (int a = 0; a < 24; ++a) {
          for (int b = 0; b < 24; ++b) {
            double t = 0.0;
            for (int s = 0; s < 6; ++s) t += B[s][a] * CB[s][b];
            Ke[a * 24 + b] += wt * t;
          }
          for (int s = 0; s < 6; ++s) {
            double t = 0.0;
            for (int q = 0; q < 6; ++q) t += B[q][a] * C[q][s];
            Fe[a * 6 + s] += wt * t;
          }
        }
      }
  fixed_solve(3, 6, Ke, Fe, {0, 1, 2, 4, 5, 11}, X0);
}

void hex8_thermal(int r, double kappa, double* Ke, double* Fe, double* X0) {
  const double half = 0.5 / r;
  const double p[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)};
  const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  std::memset(Ke, 0, sizeof(double) * 64);
  std::memset(Fe, 0, sizeof(double) * 24);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        double g[3][8];
        shape_grad(p[i], p[j], p[k], half, g);
        const double wt = w[i] * w[j] * w[k] * half * half * half * kappa;
        for (int a = 0; a < 8; ++a) {
          for (int b = 0; b < 8; ++b) Ke[a * 8 + b] += wt * (g[0][a] * g[0][b] + g[1][a] * g[1][b] + g[2][a] * g[2][b]);
          for (int d = 0; d < 3; ++d) Fe[a * 3 + d] += wt * g[d][a];
        }
      }
  fixed_solve(1, 3, Ke, Fe, {0}, X0);
}

void child_galerkin(int nd, const double* Ke, double* G) {
  const int ne = 8 * nd;
  std::vector<double> P(ne * ne), T(ne * ne);
  for (int c = 0; c < 8; ++c) {
    const int off[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
    std::fill(P.begin(), P.end(), 0.0);
    for (int a = 0; a < 8; ++a) {
      double pos[3];
      for (int d = 0; d < 3; ++d) pos[d] = 0.5 * (off[d] + kCorner[a][d]);
      for (int b = 0; b < 8; ++b) {
        double wb = 1.0;
        for (int d = 0; d < 3; ++d) wb *= 1.0 - std::fabs(pos[d] - kCorner[b][d]);
        for (int d = 0; d < nd; ++d) P[(a * nd + d) * ne + b * nd + d] = wb;
      }
    }
    // T = Ke P, G_c = P^T T
    for (int i = 0; i < ne; ++i)
      for (int j = 0; j < ne; ++j) {
        double t = 0.0;
        for (int k = 0; k < ne; ++k) t += Ke[i * ne + k] * P[k * ne + j];
        T[i * ne + j] = t;
      }
    double* Gc = G + (size_t)c * ne * ne;
    for (int i = 0; i < ne; ++i)
      for (int j = 0; j < ne; ++j) {
        double t = 0.0;
        for (int k = 0; k < ne; ++k) t += P[k * ne + i] * T[k * ne + j];
        Gc[i * ne + j] = t;
      }
  }
}
