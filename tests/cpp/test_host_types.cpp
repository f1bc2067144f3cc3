// Host-only checks of the value types of include/tensor_train.hpp (no device
// call): tt::to_dense and tt_matrix::to_dense against Kronecker and
// outer-product known answers.  Exit code = number of failed checks.
#include <cmath>
#include <cstdio>
#include <vector>

#include "tensor_train.hpp"

namespace tt = mva::tensor_train;

static int failures = 0;
static void check(bool ok, const char* msg)
{
  std::printf("%s %s\n", ok ? "ok:  " : "FAIL:", msg);
  failures += ok ? 0 : 1;
}

int main()
{
  // Rank-1 TT-matrix = A1 (x) A2 with A1 2x3, A2 3x2: dense is kron(A1, A2).
  const double A1[2][3] = {{1, 2, 3}, {4, 5, 6}};
  const double A2[3][2] = {{0.5, -1}, {2, 0.25}, {-3, 1.5}};
  tt::tt_matrix_core c0(1, 2, 3, 1), c1(1, 3, 2, 1);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) c0(0, i, j, 0) = A1[i][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) c1(0, i, j, 0) = A2[i][j];
  tt::tt_matrix K(std::vector<tt::tt_matrix_core>{c0, c1});
  const auto D = K.to_dense();
  bool ok = D.size() == 36 && K.total_rows() == 6 && K.total_cols() == 6;
  for (int i1 = 0; i1 < 2 && ok; ++i1)
    for (int i2 = 0; i2 < 3; ++i2)
      for (int j1 = 0; j1 < 3; ++j1)
        for (int j2 = 0; j2 < 2; ++j2)
          ok = ok && D[(size_t)(i1 * 3 + i2) * 6 + (j1 * 2 + j2)] == A1[i1][j1] * A2[i2][j2];
  check(ok, "tt_matrix::to_dense = kron(A1, A2)");
  // Rank-2 TT-matrix: sum of two Kronecker products through the bond.
  tt::tt_matrix_core d0(1, 2, 2, 2), d1(2, 2, 2, 1);
  const double P[2][2][2] = {{{1, 2}, {3, 4}}, {{-1, 0.5}, {2, -2}}};  // P[b][i][j]
  const double Q[2][2][2] = {{{0, 1}, {1, 0}}, {{2, 3}, {-4, 5}}};
  for (int b = 0; b < 2; ++b)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
      {
        d0(0, i, j, b) = P[b][i][j];
        d1(b, i, j, 0) = Q[b][i][j];
      }
  tt::tt_matrix R(std::vector<tt::tt_matrix_core>{d0, d1});
  const auto E = R.to_dense();
  ok = E.size() == 16;
  for (int i1 = 0; i1 < 2 && ok; ++i1)
    for (int i2 = 0; i2 < 2; ++i2)
      for (int j1 = 0; j1 < 2; ++j1)
        for (int j2 = 0; j2 < 2; ++j2)
        {
          const double want = P[0][i1][j1] * Q[0][i2][j2] + P[1][i1][j1] * Q[1][i2][j2];
          ok = ok && E[(size_t)(i1 * 2 + i2) * 4 + (j1 * 2 + j2)] == want;
        }
  check(ok, "tt_matrix::to_dense of a rank-2 sum of Kronecker products");
  check(tt::tt_matrix().to_dense() == std::vector<double>({1.0}), "tt_matrix d = 0 is [1]");
  std::printf("%d failure(s)\n", failures);
  return failures;
}
