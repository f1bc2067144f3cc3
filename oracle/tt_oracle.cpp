// TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
//
// FP64 CPU oracle for TT rounding and MPO apply-and-round.  It restates the
// algorithm of bdaszuta/tensor_train (header-only C++ over Eigen) step by step
// in plain C++ with no third-party library, so that the HIP path can be
// checked against it on identical seeded inputs.  Only tests/, the smoke()
// check and the CPU-baseline leg of bench.py load it.
//
// Eigen (the reference's un-vendored arithmetic dependency, libs/ext_eigen,
// API >= 3.4.90) is absent, so its HouseholderQR and BDCSVD are restated from
// their published algorithms:
//   * thin QR: Householder with the LAPACK dlarfg sign convention
//     (beta = -sign(alpha) * ||x||, tau = (beta - alpha) / beta), the same
//     convention Eigen's makeHouseholder uses; Q formed by backward
//     accumulation of the reflectors on the identity (Golub & Van Loan 5.1.6).
//   * thin SVD: one-sided (Hestenes) Jacobi, Demmel & Veselic 1992, on the
//     orientation with min(rows, cols) columns; singular values sorted
//     descending.  BDCSVD and Jacobi give the same U, s, V up to rounding and
//     up to sign/rotation inside clusters of equal singular values.
//
// Every function cites the reference file:line it restates.  Parity pins:
// qr_thin / svd_thin against numpy (LAPACK) golden vectors in tests/golden,
// truncate_eps_rank against hand-written known answers, and the TT-level
// functions against closed forms and brute-force dense products on small
// inputs (tests/test_oracle.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <atomic>
#include <random>
#include <thread>
#include <vector>

namespace ora
{

// ---------------------------------------------------------------------------
// Containers.  A core is (r, n, rp) row-major: linear = (a*n + i)*rp + b
// (include/core/detail/storage_backend.hpp:237-253).  A TT-matrix core
// (r, m, n, rp) is byte-identical to (r, m*n, rp)
// (include/core/detail/matrix_pack.hpp:31-55), so one container serves both;
// `m` and `nn` keep the row / column mode sizes for matrices.
// ---------------------------------------------------------------------------
struct core
{
  int r = 1, n = 1, rp = 1;
  std::vector<double> v;
  core() = default;
  core(int r_, int n_, int rp_) : r(r_), n(n_), rp(rp_), v((size_t)r_ * n_ * rp_, 0.0) {}
  size_t size() const { return v.size(); }
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
};

struct tt
{
  std::vector<core> c;
  std::vector<int> m, nn;  // non-empty only for TT-matrices
  int d() const { return (int)c.size(); }
  bool is_matrix() const { return !m.empty(); }
  std::vector<int> shape() const
  {
    std::vector<int> s;
    for (auto& x : c) s.push_back(x.n);
    return s;
  }
};

[[noreturn]] static void die(const char* msg)
{
  std::fprintf(stderr, "oracle: %s\n", msg);
  std::abort();
}

// Row-major dense matrix helper.
struct mat
{
  int rows = 0, cols = 0;
  std::vector<double> v;
  mat() = default;
  mat(int r_, int c_) : rows(r_), cols(c_), v((size_t)r_ * c_, 0.0) {}
  double& operator()(int i, int j) { return v[(size_t)i * cols + j]; }
  double operator()(int i, int j) const { return v[(size_t)i * cols + j]; }
};

// C (M x N) = A (M x K) * B (K x N); all row-major with leading dims.
static void gemm(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C,
                 int ldc)
{
  for (int i = 0; i < M; ++i)
  {
    double* ci = C + (size_t)i * ldc;
    for (int j = 0; j < N; ++j) ci[j] = 0.0;
    const double* ai = A + (size_t)i * lda;
    for (int k = 0; k < K; ++k)
    {
      const double a = ai[k];
      if (a == 0.0) continue;
      const double* bk = B + (size_t)k * ldb;
      for (int j = 0; j < N; ++j) ci[j] += a * bk[j];
    }
  }
}

// ---------------------------------------------------------------------------
// Thin Householder QR of a row-major (rows x cols) matrix.
// Restates detail::qr_thin (include/core/types/tt_eigen_bridge.hpp:42-60):
// Q (rows x k) with orthonormal columns, R (k x cols) upper trapezoidal,
// k = min(rows, cols).
// ---------------------------------------------------------------------------
static void qr_thin(const double* M, int rows, int cols, mat& Q, mat& R)
{
  if (rows <= 0 || cols <= 0)
  {
    Q = mat();
    R = mat();
    return;
  }
  const int k = std::min(rows, cols);
  // Work column-major so each Householder column is contiguous.
  std::vector<double> A((size_t)rows * cols);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) A[(size_t)j * rows + i] = M[(size_t)i * cols + j];
  std::vector<double> tau(k, 0.0);
  for (int j = 0; j < k; ++j)
  {
    double* x = &A[(size_t)j * rows + j];
    const int len = rows - j;
    double alpha = x[0];
    double tail = 0.0;
    for (int i = 1; i < len; ++i) tail += x[i] * x[i];
    if (tail <= std::numeric_limits<double>::min())
    {
      // dlarfg / makeHouseholder: no reflection needed.
      tau[j] = 0.0;
      for (int i = 1; i < len; ++i) x[i] = 0.0;
      continue;
    }
    double beta = std::sqrt(alpha * alpha + tail);
    if (alpha >= 0.0) beta = -beta;
    tau[j] = (beta - alpha) / beta;
    const double scal = 1.0 / (alpha - beta);
    for (int i = 1; i < len; ++i) x[i] *= scal;
    x[0] = beta;
    // Apply H = I - tau v v^T (v[0] = 1) to the trailing columns.
    for (int c = j + 1; c < cols; ++c)
    {
      double* y = &A[(size_t)c * rows + j];
      double w = y[0];
      for (int i = 1; i < len; ++i) w += x[i] * y[i];
      w *= tau[j];
      y[0] -= w;
      for (int i = 1; i < len; ++i) y[i] -= w * x[i];
    }
  }
  R = mat(k, cols);
  for (int i = 0; i < k; ++i)
    for (int j = i; j < cols; ++j) R(i, j) = A[(size_t)j * rows + i];
  // Q = H_0 ... H_{k-1} I[:, :k], accumulated backward (column-major work).
  std::vector<double> Qc((size_t)rows * k, 0.0);
  for (int j = 0; j < k; ++j) Qc[(size_t)j * rows + j] = 1.0;
  for (int j = k - 1; j >= 0; --j)
  {
    const double* x = &A[(size_t)j * rows + j];
    const int len = rows - j;
    if (tau[j] == 0.0) continue;
    for (int c = j; c < k; ++c)
    {
      double* y = &Qc[(size_t)c * rows + j];
      double w = y[0];
      for (int i = 1; i < len; ++i) w += x[i] * y[i];
      w *= tau[j];
      y[0] -= w;
      for (int i = 1; i < len; ++i) y[i] -= w * x[i];
    }
  }
  Q = mat(rows, k);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < k; ++j) Q(i, j) = Qc[(size_t)j * rows + i];
}

// One-sided Jacobi on the columns of X (p x q, column-major, columns
// contiguous), accumulating the rotations in J (q x q, column-major).
// On return the columns of X are mutually orthogonal and X_in * J = X_out.
// T is the accumulation type: the rotations are applied to X and J in T, so
// with long double (x87 extended, u = 2^-64) the product X_out J^T returns
// X_in to far below FP64 rounding however many rotations ran; the stopping
// rule is the FP64 one in either case.
template <typename T>
static void jacobi_columns(std::vector<T>& X, int p, int q, std::vector<T>& J)
{
  J.assign((size_t)q * q, T(0));
  for (int i = 0; i < q; ++i) J[(size_t)i * q + i] = T(1);
  // Rotate while |x_i . x_j| > tol * ||x_i|| ||x_j||, tol = max(1e-15, sqrt(p) u)
  // (the dot product itself carries ~sqrt(p) u relative rounding).
  const T tol = (T)std::max(1.0e-15, std::sqrt((double)p) * 2.220446049250313e-16);
  for (int sweep = 0; sweep < 80; ++sweep)
  {
    bool rotated = false;
    for (int i = 0; i < q - 1; ++i)
      for (int j = i + 1; j < q; ++j)
      {
        T* xi = &X[(size_t)i * p];
        T* xj = &X[(size_t)j * p];
        T a = 0, b = 0, g = 0;
        for (int t = 0; t < p; ++t)
        {
          a += xi[t] * xi[t];
          b += xj[t] * xj[t];
          g += xi[t] * xj[t];
        }
        // No rotation when the pair is orthogonal to tol, a squared norm
        // underflowed to zero, or the angle rounds to zero (|zeta| > 1e150
        // uses the asymptotic tangent 1 / (2 zeta)).
        using std::fabs;
        using std::sqrt;
        if (a == T(0) || b == T(0) || !(fabs(g) > tol * sqrt(a) * sqrt(b))) continue;
        const T zeta = (b - a) / (T(2) * g);
        const T az = fabs(zeta);
        const T t = (zeta >= T(0) ? T(1) : T(-1)) / (az > T(1.0e150) ? T(2) * az : az + sqrt(T(1) + zeta * zeta));
        const T cs = T(1) / sqrt(T(1) + t * t);
        const T sn = cs * t;
        if (sn == T(0)) continue;
        rotated = true;
        for (int u = 0; u < p; ++u)
        {
          const T x1 = xi[u], x2 = xj[u];
          xi[u] = cs * x1 - sn * x2;
          xj[u] = sn * x1 + cs * x2;
        }
        T* ji = &J[(size_t)i * q];
        T* jj = &J[(size_t)j * q];
        for (int u = 0; u < q; ++u)
        {
          const T y1 = ji[u], y2 = jj[u];
          ji[u] = cs * y1 - sn * y2;
          jj[u] = sn * y1 + cs * y2;
        }
      }
    if (!rotated) break;
  }
}

// Extended-precision accumulation of the SVD (on by default; the FP64 form
// is what a timing of the reference's arithmetic runs).
static bool g_svd_extended = true;

template <typename T>
static void svd_thin_t(const double* M, int rows, int cols, mat& U, std::vector<double>& s, mat& Vt)
{
  const int k = std::min(rows, cols);
  std::vector<T> X, J;
  std::vector<T> norms(k);
  U = mat(rows, k);
  Vt = mat(k, cols);
  s.assign(k, 0.0);
  std::vector<int> perm(k);
  std::iota(perm.begin(), perm.end(), 0);
  if (rows <= cols)
  {
    // X = M^T (cols x rows) column-major == M row-major: columns = rows of M.
    X.assign(M, M + (size_t)rows * cols);
    jacobi_columns(X, cols, rows, J);
    for (int i = 0; i < k; ++i)
    {
      T nn = 0;
      for (int t = 0; t < cols; ++t) nn += X[(size_t)i * cols + t] * X[(size_t)i * cols + t];
      norms[i] = std::sqrt(nn);
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return norms[a] > norms[b]; });
    for (int o = 0; o < k; ++o)
    {
      const int i = perm[o];
      s[o] = (double)norms[i];
      for (int t = 0; t < rows; ++t) U(t, o) = (double)J[(size_t)i * rows + t];
      const T inv = norms[i] > T(0) ? T(1) / norms[i] : T(0);
      for (int t = 0; t < cols; ++t) Vt(o, t) = (double)(X[(size_t)i * cols + t] * inv);
    }
  }
  else
  {
    // X = M (rows x cols) column-major: columns = columns of M.
    X.assign((size_t)rows * cols, T(0));
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) X[(size_t)j * rows + i] = M[(size_t)i * cols + j];
    jacobi_columns(X, rows, cols, J);
    for (int i = 0; i < k; ++i)
    {
      T nn = 0;
      for (int t = 0; t < rows; ++t) nn += X[(size_t)i * rows + t] * X[(size_t)i * rows + t];
      norms[i] = std::sqrt(nn);
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return norms[a] > norms[b]; });
    for (int o = 0; o < k; ++o)
    {
      const int i = perm[o];
      s[o] = (double)norms[i];
      const T inv = norms[i] > T(0) ? T(1) / norms[i] : T(0);
      for (int t = 0; t < rows; ++t) U(t, o) = (double)(X[(size_t)i * rows + t] * inv);
      for (int t = 0; t < cols; ++t) Vt(o, t) = (double)J[(size_t)i * cols + t];
    }
  }
}

// ---------------------------------------------------------------------------
// Thin SVD of a row-major (rows x cols) matrix: U (rows x k), s (k,
// descending), Vt (k x cols), k = min(rows, cols).  Restates the contract of
// detail::svd_thin (include/core/types/tt_eigen_bridge.hpp:67-104).
// ---------------------------------------------------------------------------
static void svd_thin(const double* M, int rows, int cols, mat& U, std::vector<double>& s, mat& Vt)
{
  if (rows <= 0 || cols <= 0)
  {
    U = mat();
    s.clear();
    Vt = mat();
    return;
  }
  if (g_svd_extended)
    svd_thin_t<long double>(M, rows, cols, U, s, Vt);
  else
    svd_thin_t<double>(M, rows, cols, U, s, Vt);
}

// Restates detail::truncate_eps_rank (tt_eigen_bridge.hpp:172-197): walk
// from the smallest singular value, accumulating the squared tail while it
// stays <= delta^2 (strict '>' stops); at least 1; capped by max_rank.
static int truncate_eps_rank(const double* s, int n, double delta, int max_rank)
{
  if (n == 0) return 1;
  const double thresh = delta * delta;
  double tail2 = 0.0;
  int r_new = n;
  for (int i = n - 1; i >= 0; --i)
  {
    const double next = tail2 + s[i] * s[i];
    if (next > thresh) break;
    tail2 = next;
    r_new = i;
  }
  if (r_new < 1) r_new = 1;
  if (max_rank > 0 && r_new > max_rank) r_new = max_rank;
  return r_new;
}

// Restates frob_norm_buffer (include/core/detail/shape_check.hpp:164-172):
// sequential sum of squares.
static double frob_norm_buffer(const double* p, size_t n)
{
  double s2 = 0.0;
  for (size_t i = 0; i < n; ++i) s2 += p[i] * p[i];
  return std::sqrt(s2);
}

// apply_left_factor (include/core/detail/contract.hpp:33-52):
// out = M (rNL x r) @ right_unfold(c) (r x n*rp).
static core apply_left_factor(const mat& M, const core& c)
{
  if (M.cols != c.r) die("apply_left_factor: shape mismatch");
  core out(M.rows, c.n, c.rp);
  gemm(M.rows, c.n * c.rp, c.r, M.v.data(), M.cols, c.data(), c.n * c.rp, out.data(), c.n * c.rp);
  return out;
}

// apply_right_factor (contract.hpp:59-77): out = left_unfold(c) (r*n x rp) @ M (rp x rNR).
static core apply_right_factor(const core& c, const mat& M)
{
  if (M.rows != c.rp) die("apply_right_factor: shape mismatch");
  core out(c.r, c.n, M.cols);
  gemm(c.r * c.n, M.cols, c.rp, c.data(), c.rp, M.v.data(), M.cols, out.data(), M.cols);
  return out;
}

// right_qr_step (include/core/detail/orthogonalize.hpp:63-100): Householder
// QR of the column-major (n*rp x r) view of c_cur; thin Q becomes the new
// (r_eff, n, rp) core, r_eff = min(n*rp, r); L = R^T is absorbed into c_prev.
static void right_qr_step(core& c_cur, core& c_prev)
{
  if (c_cur.r != c_prev.rp) die("right_qr_step: rank mismatch");
  const int r = c_cur.r, n = c_cur.n, rp = c_cur.rp;
  const int m = n * rp;
  // Row-major (m x r) matrix Mt = transpose of right_unfold.
  std::vector<double> Mt((size_t)m * r);
  for (int a = 0; a < r; ++a)
    for (int t = 0; t < m; ++t) Mt[(size_t)t * r + a] = c_cur.v[(size_t)a * m + t];
  mat Q, R;
  qr_thin(Mt.data(), m, r, Q, R);
  const int r_eff = std::min(m, r);
  core nc(r_eff, n, rp);
  for (int a = 0; a < r_eff; ++a)
    for (int t = 0; t < m; ++t) nc.v[(size_t)a * m + t] = Q(t, a);
  mat L(r, r_eff);
  for (int i = 0; i < r_eff; ++i)
    for (int j = 0; j < r; ++j) L(j, i) = R(i, j);
  c_prev = apply_right_factor(c_prev, L);
  c_cur = std::move(nc);
}

// left_qr_step (orthogonalize.hpp:35-60): QR of left_unfold(c_cur); R absorbed into c_next.
static void left_qr_step(core& c_cur, core& c_next)
{
  if (c_cur.rp != c_next.r) die("left_qr_step: rank mismatch");
  const int r = c_cur.r, n = c_cur.n, rp = c_cur.rp;
  const int m = r * n;
  const int k = std::min(m, rp);
  mat Q, R;
  qr_thin(c_cur.data(), m, rp, Q, R);
  core nc(r, n, k);
  std::copy(Q.v.begin(), Q.v.end(), nc.v.begin());
  c_next = apply_left_factor(R, c_next);
  c_cur = std::move(nc);
}

// svd_step (include/core/detail/truncate.hpp:61-99).
static int svd_step(core& c_cur, core& c_next, double delta, int max_rank)
{
  const int r = c_cur.r, n = c_cur.n, rp = c_cur.rp;
  const int rows = r * n, cols = rp;
  mat U, Vt;
  std::vector<double> s;
  if (rows >= 2 * cols && cols > 0)
  {
    mat Q, R, Ur, Vtr;
    qr_thin(c_cur.data(), rows, cols, Q, R);
    svd_thin(R.v.data(), cols, cols, Ur, s, Vtr);
    U = mat(rows, cols);
    gemm(rows, cols, cols, Q.v.data(), Q.cols, Ur.v.data(), Ur.cols, U.v.data(), cols);
    Vt = std::move(Vtr);
  }
  else
  {
    svd_thin(c_cur.data(), rows, cols, U, s, Vt);
  }
  const int r_new = truncate_eps_rank(s.data(), (int)s.size(), delta, max_rank);
  core nc(r, n, r_new);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < r_new; ++j) nc.v[(size_t)i * r_new + j] = U(i, j);
  c_cur = std::move(nc);
  // diag_scale_rows (truncate.hpp:27-38).
  mat F(r_new, Vt.cols);
  for (int i = 0; i < r_new; ++i)
    for (int j = 0; j < Vt.cols; ++j) F(i, j) = s[i] * Vt(i, j);
  c_next = apply_left_factor(F, c_next);
  return r_new;
}

// svd_step_right (truncate.hpp:114-166): truncate the left-rank axis of c_cur
// via the SVD of right_unfold(c_cur); absorb U_head * diag(s_head) into c_prev.
static int svd_step_right(core& c_cur, core& c_prev, double delta, int max_rank)
{
  const int r = c_cur.r, n = c_cur.n, rp = c_cur.rp;
  const int rows = r, cols = n * rp;
  mat U, Vt;
  std::vector<double> s;
  if (cols >= 2 * rows && rows > 0)
  {
    std::vector<double> MT((size_t)cols * rows);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) MT[(size_t)j * rows + i] = c_cur.v[(size_t)i * cols + j];
    mat Q, R;
    qr_thin(MT.data(), cols, rows, Q, R);
    mat Rt(rows, rows);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < rows; ++j) Rt(i, j) = R(j, i);
    mat Ur, Vtr;
    svd_thin(Rt.v.data(), rows, rows, Ur, s, Vtr);
    U = std::move(Ur);
    // Vt = Vt_r * Q^T  (rows x cols).
    Vt = mat(rows, cols);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j)
      {
        double acc = 0.0;
        for (int t = 0; t < rows; ++t) acc += Vtr(i, t) * Q(j, t);
        Vt(i, j) = acc;
      }
  }
  else if (rows >= 2 * cols && cols > 0)
  {
    mat Q, R, Ur, Vtr;
    qr_thin(c_cur.data(), rows, cols, Q, R);
    svd_thin(R.v.data(), cols, cols, Ur, s, Vtr);
    U = mat(rows, cols);
    gemm(rows, cols, cols, Q.v.data(), Q.cols, Ur.v.data(), Ur.cols, U.v.data(), cols);
    Vt = std::move(Vtr);
  }
  else
  {
    svd_thin(c_cur.data(), rows, cols, U, s, Vt);
  }
  const int r_new = truncate_eps_rank(s.data(), (int)s.size(), delta, max_rank);
  core nc(r_new, n, rp);
  for (int i = 0; i < r_new; ++i)
    for (int j = 0; j < cols; ++j) nc.v[(size_t)i * cols + j] = Vt(i, j);
  c_cur = std::move(nc);
  // diag_scale_cols (truncate.hpp:42-53).
  mat F(U.rows, r_new);
  for (int i = 0; i < U.rows; ++i)
    for (int j = 0; j < r_new; ++j) F(i, j) = U(i, j) * s[j];
  c_prev = apply_right_factor(c_prev, F);
  return r_new;
}

// right_orthogonalize (include/core/gauge/tt_orthogonalize.hpp:26-36).
static void right_orthogonalize(tt& a)
{
  const int d = a.d();
  if (d <= 1) return;
  for (int k = d - 1; k >= 1; --k) right_qr_step(a.c[k], a.c[k - 1]);
}

// detail::left_orthogonalize (tt_orthogonalize.hpp:47-57).
static void left_orthogonalize(tt& a)
{
  const int d = a.d();
  if (d <= 1) return;
  for (int k = 0; k < d - 1; ++k) left_qr_step(a.c[k], a.c[k + 1]);
}

// make_rank1_zero_tt (shape_check.hpp:176-188).
static tt rank1_zero(const tt& like)
{
  tt z;
  for (auto& c : like.c) z.c.emplace_back(1, c.n, 1);
  z.m = like.m;
  z.nn = like.nn;
  return z;
}

// detail::round (include/core/round/tt_round.hpp:36-127).
static tt round_tt(tt acc, double eps, int max_rank)
{
  const int d = acc.d();
  if (d <= 1) return acc;
  right_orthogonalize(acc);
  auto& cs = acc.c;
  const double norm = frob_norm_buffer(cs[0].data(), cs[0].size());
  if (norm == 0.0) return rank1_zero(acc);
  const double delta = eps * norm / std::sqrt((double)std::max(d - 1, 1));
  for (int k = 0; k < d - 1; ++k) svd_step(cs[k], cs[k + 1], delta, max_rank);
  // Norm balancing (tt_round.hpp:76-124): column norms of left_unfold for
  // cores 0..d-2, geometric mean over the nonzero ones, scale column c by
  // geo/||col|| and divide row c of the next core by the same factor.
  std::vector<std::vector<double>> col_norms(d);
  double total_log = 0.0;
  int total_cols = 0;
  for (int k = 0; k < d - 1; ++k)
  {
    const core& c = cs[k];
    const int rows = c.r * c.n, ncols = c.rp;
    col_norms[k].assign(ncols, 0.0);
    for (int j = 0; j < ncols; ++j)
    {
      double s2 = 0.0;
      for (int i = 0; i < rows; ++i) s2 += c.v[(size_t)i * ncols + j] * c.v[(size_t)i * ncols + j];
      const double nc = std::sqrt(s2);
      col_norms[k][j] = nc;
      if (nc > 0.0)
      {
        total_log += std::log(nc);
        ++total_cols;
      }
    }
  }
  if (total_cols > 0)
  {
    const double geo = std::exp(total_log / (double)total_cols);
    for (int k = 0; k < d - 1; ++k)
    {
      core& c = cs[k];
      core& cn = cs[k + 1];
      const int rows = c.r * c.n, ncols = c.rp;
      const int ncn = cn.n * cn.rp;
      for (int j = 0; j < ncols; ++j)
      {
        const double nc = col_norms[k][j];
        if (nc > 0.0)
        {
          const double scl = geo / nc;
          for (int i = 0; i < rows; ++i) c.v[(size_t)i * ncols + j] *= scl;
          for (int t = 0; t < ncn; ++t) cn.v[(size_t)j * ncn + t] /= scl;
        }
      }
    }
  }
  return acc;
}

// round_right_canonical (tt_round.hpp:151-191).
static tt round_right_canonical(tt acc, double eps, int max_rank)
{
  const int d = acc.d();
  if (d <= 1) return acc;
  left_orthogonalize(acc);
  auto& cs = acc.c;
  const double norm = frob_norm_buffer(cs[d - 1].data(), cs[d - 1].size());
  if (norm == 0.0) return rank1_zero(acc);
  const double delta = eps * norm / std::sqrt((double)std::max(d - 1, 1));
  for (int k = d - 1; k >= 1; --k) svd_step_right(cs[k], cs[k - 1], delta, max_rank);
  return acc;
}

// random(shape, max_rank, seed) (include/core/factory/tt_factory.hpp:141-207)
// and random(row_shape, col_shape, max_rank, seed)
// (include/core/factory/tt_matrix_factory.hpp:133-191).  Rank rule: forward
// cap r_{k+1} <= r_k * N_k, backward cap r_k <= r_{k+1} * N_k (each also
// <= max_rank when max_rank > 0), then r_0 = r_d = 1 set last.  Entries are
// libstdc++ std::normal_distribution<double>(0,1) over std::mt19937_64(seed),
// drawn core by core in buffer order.
static tt random_impl(const std::vector<int>& N, int max_rank, uint64_t seed)
{
  const int d = (int)N.size();
  for (int k = 0; k < d; ++k)
    if (N[k] <= 0) die("random: mode size is not positive");
  std::vector<int> ranks(d + 1, 1);
  for (int k = 0; k < d; ++k)
  {
    long long cap = (long long)ranks[k] * N[k];
    if (cap > max_rank && max_rank > 0) cap = max_rank;
    if (cap > std::numeric_limits<int>::max()) die("random: rank cap exceeds INT_MAX");
    ranks[k + 1] = (int)cap;
  }
  for (int k = d - 1; k >= 0; --k)
  {
    long long cap = (long long)ranks[k + 1] * N[k];
    if (cap > max_rank && max_rank > 0) cap = max_rank;
    if (cap > std::numeric_limits<int>::max()) die("random: rank cap exceeds INT_MAX");
    if (cap < ranks[k]) ranks[k] = (int)cap;
  }
  ranks[0] = 1;
  ranks[d] = 1;
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  tt out;
  for (int k = 0; k < d; ++k)
  {
    core c(ranks[k], N[k], ranks[k + 1]);
    for (size_t i = 0; i < c.size(); ++i) c.v[i] = nd(rng);
    out.c.push_back(std::move(c));
  }
  return out;
}

// axpby (include/core/algebra/tt_ops.hpp:72-110) via axpby_core
// (include/core/detail/block_concat.hpp:40-156): block-diagonal
// concatenation; alpha/beta folded into the head core.  Works for TT-matrix
// cores too since (r, m, n, rp) == (r, m*n, rp) and the blocks are whole
// rank slabs.
static tt axpby(double alpha, const tt& a, double beta, const tt& b)
{
  if (a.d() != b.d()) die("axpby: core count mismatch");
  const int d = a.d();
  tt out;
  out.m = a.m;
  out.nn = a.nn;
  if (d == 0) return out;
  for (int k = 0; k < d; ++k)
    if (a.c[k].n != b.c[k].n) die("axpby: mode size mismatch");
  if (d == 1)
  {
    const int n = a.c[0].n;
    core c(1, n, 1);
    for (int i = 0; i < n; ++i) c.v[i] = alpha * a.c[0].v[i] + beta * b.c[0].v[i];
    out.c.push_back(std::move(c));
    return out;
  }
  for (int k = 0; k < d; ++k)
  {
    const core& ca = a.c[k];
    const core& cb = b.c[k];
    const double ak = k == 0 ? alpha : 1.0, bk = k == 0 ? beta : 1.0;
    const bool head = k == 0, tail = k == d - 1;
    const int n = ca.n;
    const int rL = head ? 1 : ca.r + cb.r;
    const int rR = tail ? 1 : ca.rp + cb.rp;
    core c(rL, n, rR);
    for (int aa = 0; aa < ca.r; ++aa)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < ca.rp; ++j)
        {
          const double v = ca.v[((size_t)aa * n + i) * ca.rp + j];
          c.v[((size_t)aa * n + i) * rR + j] = ak == 1.0 ? v : ak * v;
        }
    const int boff = head ? 0 : ca.r;
    const int coff = tail ? 0 : ca.rp;
    for (int bb = 0; bb < cb.r; ++bb)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < cb.rp; ++j)
        {
          const double v = cb.v[((size_t)bb * n + i) * cb.rp + j];
          c.v[((size_t)(boff + bb) * n + i) * rR + coff + j] = bk == 1.0 ? v : bk * v;
        }
    out.c.push_back(std::move(c));
  }
  return out;
}

// scale (tt_ops.hpp:35-50): multiply core 0 by alpha.
static tt scale(const tt& a, double alpha)
{
  tt out = a;
  if (!out.c.empty())
    for (auto& v : out.c[0].v) v *= alpha;
  return out;
}

// matvec_core (include/core/apply/tt_matrix_apply.hpp:40-125):
// y(a*s+ax, i, b*sp+bx) = sum_j A(a,i,j,b) * x(ax,j,bx).
static core matvec_core(const core& A, int m, int n, const core& x)
{
  const int r = A.r, rp = A.rp, s = x.r, sp = x.rp;
  if (x.n != n) die("matvec: mode mismatch");
  core out(r * s, m, rp * sp);
  for (int a = 0; a < r; ++a)
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j)
        for (int b = 0; b < rp; ++b)
        {
          const double av = A.v[(((size_t)a * m + i) * n + j) * rp + b];
          if (av == 0.0) continue;
          for (int ax = 0; ax < s; ++ax)
          {
            const double* xr = &x.v[((size_t)ax * n + j) * sp];
            double* yr = &out.v[(((size_t)a * s + ax) * m + i) * ((size_t)rp * sp) + (size_t)b * sp];
            for (int bx = 0; bx < sp; ++bx) yr[bx] += av * xr[bx];
          }
        }
  return out;
}

// matmat_core (tt_matrix_apply.hpp:128-224):
// C(a*s+ax, i, j, b*sp+bx) = sum_p A(a,i,p,b) * B(ax,p,j,bx).
static core matmat_core(const core& A, int m, int p, const core& B, int n)
{
  const int r = A.r, rp = A.rp, s = B.r, sp = B.rp;
  core out(r * s, m * n, rp * sp);
  for (int a = 0; a < r; ++a)
    for (int i = 0; i < m; ++i)
      for (int q = 0; q < p; ++q)
        for (int b = 0; b < rp; ++b)
        {
          const double av = A.v[(((size_t)a * m + i) * p + q) * rp + b];
          if (av == 0.0) continue;
          for (int ax = 0; ax < s; ++ax)
            for (int j = 0; j < n; ++j)
            {
              const double* br = &B.v[(((size_t)ax * p + q) * n + j) * sp];
              double* cr = &out.v[((((size_t)a * s + ax) * m + i) * n + j) * ((size_t)rp * sp) + (size_t)b * sp];
              for (int bx = 0; bx < sp; ++bx) cr[bx] += av * br[bx];
            }
        }
  return out;
}

static tt matvec(const tt& A, const tt& x)
{
  if (A.d() != x.d()) die("matvec: core count mismatch");
  tt y;
  for (int k = 0; k < A.d(); ++k) y.c.push_back(matvec_core(A.c[k], A.m[k], A.nn[k], x.c[k]));
  return y;
}

static tt matmat(const tt& A, const tt& B)
{
  if (A.d() != B.d()) die("matmat: core count mismatch");
  tt C;
  for (int k = 0; k < A.d(); ++k)
  {
    if (A.nn[k] != B.m[k]) die("matmat: mode mismatch");
    C.c.push_back(matmat_core(A.c[k], A.m[k], A.nn[k], B.c[k], B.nn[k]));
    C.m.push_back(A.m[k]);
    C.nn.push_back(B.nn[k]);
  }
  return C;
}

// fused_matvec_core (include/core/detail/fused_apply_core.hpp:33-131):
// out(al, i, b*bx_dim+bx) = sum_{aa,ax,j} F(al, aa*sx+ax) A(aa,i,j,b) x(ax,j,bx).
static core fused_matvec_core(const mat& F, const core& A, int m, int n, const core& x)
{
  const int alpha = F.rows, a = A.r, b = A.rp, sx = x.r, bx = x.rp;
  if (F.cols != a * sx) die("fused_matvec: F shape mismatch");
  // M1(al, ax, i, j, b) = sum_aa F(al, aa*sx+ax) * A(aa, i, j, b)
  const size_t mnb = (size_t)m * n * b;
  std::vector<double> M1((size_t)alpha * sx * mnb, 0.0);
  for (int al = 0; al < alpha; ++al)
    for (int ax = 0; ax < sx; ++ax)
    {
      double* dst = &M1[((size_t)al * sx + ax) * mnb];
      for (int aa = 0; aa < a; ++aa)
      {
        const double f = F(al, aa * sx + ax);
        if (f == 0.0) continue;
        const double* src = &A.v[(size_t)aa * mnb];
        for (size_t t = 0; t < mnb; ++t) dst[t] += f * src[t];
      }
    }
  core out(alpha, m, b * bx);
  for (int al = 0; al < alpha; ++al)
    for (int i = 0; i < m; ++i)
      for (int bb = 0; bb < b; ++bb)
      {
        double* yr = &out.v[(((size_t)al * m + i) * b + bb) * bx];
        for (int ax = 0; ax < sx; ++ax)
          for (int j = 0; j < n; ++j)
          {
            const double w = M1[((size_t)al * sx + ax) * mnb + ((size_t)i * n + j) * b + bb];
            if (w == 0.0) continue;
            const double* xr = &x.v[((size_t)ax * n + j) * bx];
            for (int t = 0; t < bx; ++t) yr[t] += w * xr[t];
          }
      }
  return out;
}

// fused_matmat_core (fused_apply_core.hpp:136-272):
// out(al, i, j, bb*bx_dim+bx) = sum_{aa,ax,q} F(al, aa*sx+ax) A(aa,i,q,bb) B(ax,q,j,bx).
static core fused_matmat_core(const mat& F, const core& A, int m, int p, const core& B, int n)
{
  const int alpha = F.rows, a = A.r, b = A.rp, sx = B.r, bx = B.rp;
  if (F.cols != a * sx) die("fused_matmat: F shape mismatch");
  const size_t mpb = (size_t)m * p * b;
  std::vector<double> M1((size_t)alpha * sx * mpb, 0.0);
  for (int al = 0; al < alpha; ++al)
    for (int ax = 0; ax < sx; ++ax)
    {
      double* dst = &M1[((size_t)al * sx + ax) * mpb];
      for (int aa = 0; aa < a; ++aa)
      {
        const double f = F(al, aa * sx + ax);
        if (f == 0.0) continue;
        const double* src = &A.v[(size_t)aa * mpb];
        for (size_t t = 0; t < mpb; ++t) dst[t] += f * src[t];
      }
    }
  core out(alpha, m * n, b * bx);
  for (int al = 0; al < alpha; ++al)
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j)
        for (int bb = 0; bb < b; ++bb)
        {
          double* yr = &out.v[((((size_t)al * m + i) * n + j) * b + bb) * bx];
          for (int ax = 0; ax < sx; ++ax)
            for (int q = 0; q < p; ++q)
            {
              const double w = M1[((size_t)al * sx + ax) * mpb + ((size_t)i * p + q) * b + bb];
              if (w == 0.0) continue;
              const double* br = &B.v[(((size_t)ax * p + q) * n + j) * bx];
              for (int t = 0; t < bx; ++t) yr[t] += w * br[t];
            }
        }
  return out;
}

// build_streaming_matvec / build_streaming_matmat
// (include/core/detail/streaming_apply.hpp:39-110, :115-202).  `is_matmat`
// selects the fused core.  Right-orthogonalize copies of both operands,
// delta from ||A0|| * ||x0||, then one L->R pass: fused core, thin SVD of its
// left unfold, truncate, F = diag(s) Vt; the last fused core is kept whole.
static tt build_streaming(const tt& A_in, const tt& X_in, bool is_matmat, double eps, int max_rank)
{
  const int d = A_in.d();
  if (d == 0 || X_in.d() != d) die("build_streaming: bad core count");
  tt A = A_in, X = X_in;
  right_orthogonalize(A);
  right_orthogonalize(X);
  const double nA = frob_norm_buffer(A.c[0].data(), A.c[0].size());
  const double nX = frob_norm_buffer(X.c[0].data(), X.c[0].size());
  const double norm = nA * nX;
  tt out;
  if (is_matmat)
  {
    out.m = A_in.m;
    out.nn = X_in.nn;
  }
  if (norm == 0.0)
  {
    for (int k = 0; k < d; ++k)
      out.c.emplace_back(1, is_matmat ? A_in.m[k] * X_in.nn[k] : A_in.m[k], 1);
    return out;
  }
  const double delta = eps * norm / std::sqrt((double)std::max(d - 1, 1));
  mat F(1, 1);
  F(0, 0) = 1.0;
  for (int k = 0; k < d; ++k)
  {
    core C = is_matmat ? fused_matmat_core(F, A.c[k], A.m[k], A.nn[k], X.c[k], X.nn[k])
                       : fused_matvec_core(F, A.c[k], A.m[k], A.nn[k], X.c[k]);
    if (k == d - 1)
    {
      out.c.push_back(std::move(C));
      break;
    }
    const int rows = C.r * C.n, br = C.rp;
    mat U, Vt;
    std::vector<double> s;
    svd_thin(C.data(), rows, br, U, s, Vt);
    const int r_new = truncate_eps_rank(s.data(), (int)s.size(), delta, max_rank);
    core nc(C.r, C.n, r_new);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < r_new; ++j) nc.v[(size_t)i * r_new + j] = U(i, j);
    out.c.push_back(std::move(nc));
    mat Fn(r_new, br);
    for (int i = 0; i < r_new; ++i)
      for (int j = 0; j < br; ++j) Fn(i, j) = s[i] * Vt(i, j);
    F = std::move(Fn);
  }
  return out;
}

// contract_inner_step (include/core/detail/contract.hpp:155-205) and inner
// (include/core/algebra/tt_inner.hpp:33-48).
static double inner(const tt& a, const tt& b)
{
  if (a.d() != b.d()) die("inner: core count mismatch");
  const int d = a.d();
  if (d == 0) return 0.0;
  mat E(1, 1);
  E(0, 0) = 1.0;
  for (int k = 0; k < d; ++k)
  {
    const core& ca = a.c[k];
    const core& cb = b.c[k];
    if (ca.n != cb.n) die("inner: mode mismatch");
    const int n = ca.n, ra = ca.r, rap = ca.rp, rb = cb.r, rbp = cb.rp;
    // tmp (rb, n*rap) = E^T (rb x ra) @ right_unfold(ca) (ra x n*rap)
    std::vector<double> tmp((size_t)rb * n * rap, 0.0);
    for (int aa = 0; aa < ra; ++aa)
      for (int bb = 0; bb < rb; ++bb)
      {
        const double e = E(aa, bb);
        if (e == 0.0) continue;
        const double* src = &ca.v[(size_t)aa * n * rap];
        double* dst = &tmp[(size_t)bb * n * rap];
        for (int t = 0; t < n * rap; ++t) dst[t] += e * src[t];
      }
    // E' (rap, rbp) = tmp_as^T (rap x rb*n) @ left_unfold(cb) (rb*n x rbp)
    mat En(rap, rbp);
    for (int row = 0; row < rb * n; ++row)
      for (int p = 0; p < rap; ++p)
      {
        const double t = tmp[(size_t)row * rap + p];
        if (t == 0.0) continue;
        const double* cbr = &cb.v[(size_t)row * rbp];
        for (int q = 0; q < rbp; ++q) En(p, q) += t * cbr[q];
      }
    E = std::move(En);
  }
  return E(0, 0);
}

// to_dense (include/core/types/tt.hpp:145-198): left-to-right product,
// flat row-major over (i_0, ..., i_{d-1}) (for matrices: over the packed
// (i_k, j_k) pairs in core order).
static std::vector<double> to_dense(const tt& a)
{
  const int d = a.d();
  if (d == 0) return {1.0};
  std::vector<double> T(a.c[0].v);
  size_t rows = a.c[0].n;
  int cols = a.c[0].rp;
  for (int k = 1; k < d; ++k)
  {
    const core& c = a.c[k];
    std::vector<double> Tn(rows * c.n * c.rp, 0.0);
    gemm((int)rows, c.n * c.rp, c.r, T.data(), cols, c.data(), c.n * c.rp, Tn.data(), c.n * c.rp);
    T.swap(Tn);
    rows *= c.n;
    cols = c.rp;
  }
  return T;
}

}  // namespace ora

// ===========================================================================
// C ABI for ctypes.  Handles are heap `ora::tt` objects; outputs are fresh
// handles owned by the caller (free with ora_free).
// ===========================================================================
using ora::tt;

namespace ora
{
// contract_inner_step_right (include/core/detail/contract.hpp:90-139):
// F_out (ca.r x cb.r) = (left_unfold(ca) F as (ca.r x n F.cols))
// right_unfold(cb)^T.
static mat env_right(const core& ca, const core& cb, const mat& F)
{
  const int ral = ca.r, n = ca.n, rar = ca.rp, rbl = cb.r, rbr = cb.rp;
  if (F.rows != rar || F.cols != rbr || ca.n != cb.n) die("env_right: shape mismatch");
  std::vector<double> T1((size_t)ral * n * rbr, 0.0);
  gemm(ral * n, rbr, rar, ca.data(), rar, F.v.data(), rbr, T1.data(), rbr);
  mat out(ral, rbl);
  for (int a = 0; a < ral; ++a)
    for (int b = 0; b < rbl; ++b)
    {
      double s = 0.0;
      for (int t = 0; t < n * rbr; ++t) s += T1[(size_t)a * n * rbr + t] * cb.v[(size_t)b * n * rbr + t];
      out(a, b) = s;
    }
  return out;
}

// contract_inner_step (contract.hpp:141-205): E' (ca.rp x cb.rp) from
// E (ca.r x cb.r): tmp = E^T right_unfold(ca) as (cb.r n x ca.rp), then
// tmp^T left_unfold(cb).
static mat env_left(const mat& E, const core& ca, const core& cb)
{
  const int ra = ca.r, n = ca.n, rap = ca.rp, rb = cb.r, rbp = cb.rp;
  if (E.rows != ra || E.cols != rb || ca.n != cb.n) die("env_left: shape mismatch");
  std::vector<double> tmp((size_t)rb * n * rap, 0.0);
  for (int b = 0; b < rb; ++b)
    for (int a = 0; a < ra; ++a)
    {
      const double e = E(a, b);
      for (int t = 0; t < n * rap; ++t) tmp[(size_t)b * n * rap + t] += e * ca.v[(size_t)a * n * rap + t];
    }
  mat out(rap, rbp);
  for (int bj = 0; bj < rb * n; ++bj)
    for (int p = 0; p < rap; ++p)
    {
      const double x = tmp[(size_t)bj * rap + p];
      for (int q = 0; q < rbp; ++q) out(p, q) += x * cb.v[(size_t)bj * rbp + q];
    }
  return out;
}

// als_local_update (include/core/detail/als_local_solve.hpp:24-52):
// left_unfold(X_k) = (L right_unfold(Y_k) as (L.rows n x Y.rp)) R.
static core als_local(const mat& L, const core& Y, const mat& R)
{
  const int rxl = L.rows, n = Y.n, ryr = Y.rp, rxr = R.cols;
  if (L.cols != Y.r || R.rows != ryr) die("als_local: shape mismatch");
  std::vector<double> T1((size_t)rxl * n * ryr, 0.0);
  gemm(rxl, n * ryr, Y.r, L.v.data(), L.cols, Y.data(), n * ryr, T1.data(), n * ryr);
  core X(rxl, n, rxr);
  gemm(rxl * n, rxr, ryr, T1.data(), ryr, R.v.data(), rxr, X.data(), rxr);
  return X;
}

// detail::als_engine (include/core/round/tt_als.hpp:54-206): X right-
// orthogonalized unless already right-canonical; right environments; per
// iteration one L->R sweep (local update, left QR, left environment) and one
// R->L sweep (local update except at d-1, right QR, right environment); the
// change is the largest relative change of the singular values of every
// left_unfold(X_k), k < d-1 (1.0 with no baseline or a shape change); stop
// when change <= tol.
static tt als_engine(const tt& Y, tt X, int max_iters, double tol, bool right_canonical, int* iters_out,
                     double* change_out, int* conv_out)
{
  const int d = Y.d();
  if (!right_canonical) right_orthogonalize(X);
  std::vector<mat> R(d + 1), L(d + 1);
  R[d] = mat(1, 1);
  R[d](0, 0) = 1.0;
  for (int k = d - 1; k >= 1; --k) R[k] = env_right(Y.c[k], X.c[k], R[k + 1]);
  L[0] = mat(1, 1);
  L[0](0, 0) = 1.0;
  std::vector<std::vector<double>> s_prev(d - 1);
  bool have_prev = false, converged = false;
  double change = 0.0;
  int iters = 0;
  for (int it = 0; it < max_iters; ++it)
  {
    for (int k = 0; k < d; ++k)
    {
      X.c[k] = als_local(L[k], Y.c[k], R[k + 1]);
      if (k < d - 1)
      {
        left_qr_step(X.c[k], X.c[k + 1]);
        L[k + 1] = env_left(L[k], X.c[k], Y.c[k]);
      }
    }
    for (int k = d - 1; k >= 0; --k)
    {
      if (k < d - 1) X.c[k] = als_local(L[k], Y.c[k], R[k + 1]);
      if (k > 0)
      {
        right_qr_step(X.c[k], X.c[k - 1]);
        R[k] = env_right(Y.c[k], X.c[k], R[k + 1]);
      }
    }
    ++iters;
    std::vector<std::vector<double>> s_now(d - 1);
    for (int k = 0; k < d - 1; ++k)
    {
      mat U, Vt;
      svd_thin(X.c[k].data(), X.c[k].r * X.c[k].n, X.c[k].rp, U, s_now[k], Vt);
    }
    if (have_prev)
    {
      double mx = 0.0;
      for (int k = 0; k < d - 1; ++k)
      {
        double nn = 0.0, df = 0.0;
        for (double v : s_now[k]) nn += v * v;
        nn = std::sqrt(nn);
        if (s_prev[k].size() == s_now[k].size())
        {
          for (size_t i = 0; i < s_now[k].size(); ++i) df += (s_now[k][i] - s_prev[k][i]) * (s_now[k][i] - s_prev[k][i]);
          mx = std::max(mx, std::sqrt(df) / (nn + 1.0e-300));
        }
        else
          mx = 1.0;
      }
      change = mx;
    }
    else
      change = 1.0;
    s_prev = std::move(s_now);
    have_prev = true;
    if (change <= tol)
    {
      converged = true;
      break;
    }
  }
  *iters_out = iters;
  *change_out = change;
  *conv_out = converged ? 1 : 0;
  return X;
}

// detail::als_round (tt_als.hpp:208-290): d <= 1 -> SVD round with
// {0, 0, true}; without max_rank the target is the largest rank of the SVD
// round at eps, whose result is also the warm start unless one is given; a
// warm start must match d and the modes (else it is ignored) and stay within
// max_rank; without one the start is round_right_canonical(Y, 0, max_rank).
// cold_random selects the start of the apply-then-ALS paths instead:
// matvec_als_round (tt_matvec_als_round.hpp:30-113) and matmat_als_round
// (tt_matmat_als_round.hpp:39-141) start from random(shape, max_rank, seed)
// and right-orthogonalize every start, the SVD reference included.  For a
// TT-matrix warm start the per-site (m, n) must match
// (warm_start.hpp:44-60); Y and warm carry m in .m when they are matrices.
static tt als_round(const tt& Y, double eps, int max_rank, int max_iters, double tol, const tt* warm, bool cold_random,
                    uint64_t seed, int* iters, double* change, int* conv)
{
  const int d = Y.d();
  if (std::isnan(eps)) die("als_round: eps is NaN");
  if (d <= 1)
  {
    *iters = 0;
    *change = 0.0;
    *conv = 1;
    return round_tt(Y, eps, max_rank);
  }
  tt svd_ref;
  if (max_rank <= 0)
  {
    if (eps <= 0.0) die("als_round: max_rank <= 0 and eps <= 0");
    svd_ref = round_tt(Y, eps, 0);
    for (auto& c : svd_ref.c) max_rank = std::max(max_rank, std::max(c.r, c.rp));
    if (!warm) warm = &svd_ref;
  }
  bool use_warm = warm && warm->d() == d;
  if (use_warm)
    for (int k = 0; k < d; ++k) use_warm = use_warm && warm->c[k].n == Y.c[k].n;
  if (use_warm && !Y.m.empty() && warm != &svd_ref) use_warm = warm->m == Y.m;
  if (use_warm)
    for (auto& c : warm->c)
      if (c.r > max_rank || c.rp > max_rank) die("als_round: warm start rank above max_rank");
  tt X0;
  if (use_warm)
    X0 = *warm;
  else if (cold_random)
  {
    std::vector<int> shape(d);
    for (int k = 0; k < d; ++k) shape[k] = Y.c[k].n;
    X0 = random_impl(shape, max_rank, seed);
  }
  else
    X0 = round_right_canonical(Y, 0.0, max_rank);
  X0.m = Y.m;
  X0.nn = Y.nn;
  return als_engine(Y, std::move(X0), max_iters, tol, !use_warm && !cold_random, iters, change, conv);
}

// merge_y_block (include/core/detail/dmrg_local_solve.hpp:34-75): the
// two-site target Theta (rXl n_k x n_{k+1} rXr) = (L right_unfold(Y_k)) as
// (rXl n_k x rYm) times (left_unfold(Y_{k+1}) R) as (rYm x n_{k+1} rXr).
static mat merge_block(const mat& L, const core& Yk, const core& Yk1, const mat& R)
{
  const int rXl = L.rows, nk = Yk.n, rYm = Yk.rp, nk1 = Yk1.n, rXr = R.cols;
  if (L.cols != Yk.r || R.rows != Yk1.rp || Yk1.r != rYm) die("merge_block: shape mismatch");
  std::vector<double> T1((size_t)rXl * nk * rYm, 0.0), T2((size_t)rYm * nk1 * rXr, 0.0);
  gemm(rXl, nk * rYm, Yk.r, L.v.data(), L.cols, Yk.data(), nk * rYm, T1.data(), nk * rYm);
  gemm(rYm * nk1, rXr, Yk1.rp, Yk1.data(), Yk1.rp, R.v.data(), rXr, T2.data(), rXr);
  mat Th(rXl * nk, nk1 * rXr);
  gemm(rXl * nk, nk1 * rXr, rYm, T1.data(), rYm, T2.data(), nk1 * rXr, Th.v.data(), nk1 * rXr);
  return Th;
}

// split_two_site_lr / _rl (dmrg_local_solve.hpp:83-177): truncated SVD of
// Theta; L->R keeps U (left-orthogonal X_k) and diag(s) Vt (X_{k+1}), R->L
// keeps U diag(s) (X_k) and Vt (right-orthogonal X_{k+1}).
static void split_two_site(const mat& Th, int rXl, int nk, int nk1, int rXr, double delta, int max_rank, bool lr,
                           core& Xk, core& Xk1, std::vector<double>* s_head)
{
  mat U, Vt;
  std::vector<double> s;
  svd_thin(Th.v.data(), Th.rows, Th.cols, U, s, Vt);
  const int r = truncate_eps_rank(s.data(), (int)s.size(), delta, max_rank);
  Xk = core(rXl, nk, r);
  Xk1 = core(r, nk1, rXr);
  for (int i = 0; i < rXl * nk; ++i)
    for (int j = 0; j < r; ++j) Xk.v[(size_t)i * r + j] = U(i, j) * (lr ? 1.0 : s[j]);
  for (int j = 0; j < r; ++j)
    for (int t = 0; t < nk1 * rXr; ++t) Xk1.v[(size_t)j * nk1 * rXr + t] = Vt(j, t) * (lr ? s[j] : 1.0);
  if (s_head) s_head->assign(s.begin(), s.begin() + r);
}

// detail::dmrg_engine (include/core/round/tt_dmrg.hpp:53-230): per-bond
// delta = eps ||Y|| / sqrt(d-1); right environments of the (right-
// orthogonalized unless right-canonical) start; each iteration one L->R and
// one R->L two-site sweep; the change is the largest relative change of the
// kept singular values of the L->R sweep (1.0 with no baseline or a changed
// count), converged when change <= tol.
static tt dmrg_engine(const tt& Y, tt X, double eps, int max_rank, int max_iters, double tol, bool right_canonical,
                      int* iters_out, double* change_out, int* conv_out)
{
  const int d = Y.d();
  if (!right_canonical) right_orthogonalize(X);
  const double y_norm = std::sqrt(std::max(inner(Y, Y), 0.0));
  const double delta = eps * y_norm / std::sqrt((double)std::max(d - 1, 1));
  std::vector<mat> R(d + 1), L(d + 1);
  R[d] = mat(1, 1);
  R[d](0, 0) = 1.0;
  for (int k = d - 1; k >= 1; --k) R[k] = env_right(Y.c[k], X.c[k], R[k + 1]);
  L[0] = mat(1, 1);
  L[0](0, 0) = 1.0;
  std::vector<std::vector<double>> s_prev(d - 1), s_now(d - 1);
  bool have_prev = false, converged = false;
  double change = 0.0;
  int iters = 0;
  for (int it = 0; it < max_iters; ++it)
  {
    for (int k = 0; k < d - 1; ++k)
    {
      const mat Th = merge_block(L[k], Y.c[k], Y.c[k + 1], R[k + 2]);
      split_two_site(Th, L[k].rows, Y.c[k].n, Y.c[k + 1].n, R[k + 2].cols, delta, max_rank, true, X.c[k], X.c[k + 1],
                     &s_now[k]);
      L[k + 1] = env_left(L[k], X.c[k], Y.c[k]);
    }
    for (int k = d - 2; k >= 0; --k)
    {
      const mat Th = merge_block(L[k], Y.c[k], Y.c[k + 1], R[k + 2]);
      split_two_site(Th, L[k].rows, Y.c[k].n, Y.c[k + 1].n, R[k + 2].cols, delta, max_rank, false, X.c[k],
                     X.c[k + 1], nullptr);
      R[k + 1] = env_right(Y.c[k + 1], X.c[k + 1], R[k + 2]);
    }
    ++iters;
    if (have_prev)
    {
      double mx = 0.0;
      for (int k = 0; k < d - 1; ++k)
      {
        double nn = 0.0;
        for (double v : s_now[k]) nn += v * v;
        nn = std::sqrt(nn);
        if (s_prev[k].size() == s_now[k].size())
        {
          double df = 0.0;
          for (size_t i = 0; i < s_now[k].size(); ++i) df += (s_now[k][i] - s_prev[k][i]) * (s_now[k][i] - s_prev[k][i]);
          mx = std::max(mx, std::sqrt(df) / (nn + 1.0e-300));
        }
        else
          mx = 1.0;
      }
      change = mx;
    }
    else
      change = 1.0;
    s_prev = s_now;
    have_prev = true;
    if (change <= tol)
    {
      converged = true;
      break;
    }
  }
  *iters_out = iters;
  *change_out = change;
  *conv_out = converged ? 1 : 0;
  return X;
}

// detail::dmrg_round (tt_dmrg.hpp:235-310) and, with apply_path, the starts
// of matvec_dmrg_round (tt_matvec_dmrg_round.hpp:29-105) and
// matmat_dmrg_round (tt_matmat_dmrg_round.hpp:38-130): eps must be > 0;
// d <= 1 -> SVD round with {0, 0, true}; a compatible warm start (ranks
// checked against max_rank when one is set) is right-orthogonalized; cold
// start round_right_canonical(Y, eps, max_rank), or for the apply paths
// random(shape, max_rank, seed) when max_rank > 0, else
// round_right_canonical(Y, eps, 0).
static tt dmrg_round(const tt& Y, double eps, int max_rank, int max_iters, double tol, const tt* warm,
                     bool apply_path, uint64_t seed, int* iters, double* change, int* conv)
{
  if (std::isnan(eps) || eps <= 0.0) die("dmrg_round: eps must be > 0");
  const int d = Y.d();
  if (d <= 1)
  {
    *iters = 0;
    *change = 0.0;
    *conv = 1;
    return round_tt(Y, eps, max_rank);
  }
  bool use_warm = warm && warm->d() == d;
  if (use_warm)
    for (int k = 0; k < d; ++k) use_warm = use_warm && warm->c[k].n == Y.c[k].n;
  if (use_warm && !Y.m.empty()) use_warm = warm->m == Y.m;
  if (use_warm && max_rank > 0)
    for (auto& c : warm->c)
      if (c.r > max_rank || c.rp > max_rank) die("dmrg_round: warm start rank above max_rank");
  tt X0;
  bool rc = false;
  if (use_warm)
    X0 = *warm;
  else if (apply_path && max_rank > 0)
  {
    std::vector<int> shape(d);
    for (int k = 0; k < d; ++k) shape[k] = Y.c[k].n;
    X0 = random_impl(shape, max_rank, seed);
  }
  else
  {
    X0 = round_right_canonical(Y, eps, apply_path ? 0 : max_rank);
    rc = true;
  }
  X0.m = Y.m;
  X0.nn = Y.nn;
  return dmrg_engine(Y, std::move(X0), eps, max_rank, max_iters, tol, rc, iters, change, conv);
}

// soft_threshold (include/core/round/tt_soft_threshold.hpp:42-176): d <= 1
// or tau <= 0 returns a copy; otherwise right-orthogonalize, then for
// k = d-1..1 the thin SVD of left_unfold(c_{k-1}) right_unfold(c_k), the
// singular values shrunk to max(s - tau, 0), the rank cut after the last
// positive one; c_{k-1} <- U diag(s_shrunk), c_k <- Vt.  When every value is
// zeroed both cores become rank-1 zeros at that bond.
static tt soft_threshold(tt acc, double tau)
{
  const int d = acc.d();
  if (d <= 1 || tau <= 0.0) return acc;
  right_orthogonalize(acc);
  for (int k = d - 1; k >= 1; --k)
  {
    core& cp = acc.c[k - 1];
    core& cc = acc.c[k];
    const int rL = cp.r, np = cp.n, r = cp.rp, nc = cc.n, rR = cc.rp;
    const int rows = rL * np, cols = nc * rR;
    std::vector<double> M((size_t)rows * cols, 0.0);
    gemm(rows, cols, r, cp.data(), r, cc.data(), cols, M.data(), cols);
    mat U, Vt;
    std::vector<double> s;
    svd_thin(M.data(), rows, cols, U, s, Vt);
    const int nsv = (int)s.size();
    int r_new = 0;
    std::vector<double> sh(nsv, 0.0);
    for (int j = 0; j < nsv; ++j)
      if (s[j] > tau)
      {
        sh[j] = s[j] - tau;
        r_new = j + 1;
      }
    if (nsv == 0 || sh[0] <= 0.0)
    {
      cp = core(rL, np, 1);
      cc = core(1, nc, rR);
      continue;
    }
    core np_(rL, np, r_new), nc_(r_new, nc, rR);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < r_new; ++j) np_.v[(size_t)i * r_new + j] = U(i, j) * (sh[j] > 0.0 ? sh[j] : 1.0);
    for (int j = 0; j < r_new; ++j)
      for (int t = 0; t < cols; ++t) nc_.v[(size_t)j * cols + t] = Vt(j, t);
    cp = std::move(np_);
    cc = std::move(nc_);
  }
  return acc;
}

// detail::svd (include/core/factory/tt_svd.hpp:46-160), Oseledets' TT-SVD:
// norm of the dense buffer, delta = eps ||T|| / sqrt(max(d-1, 1)); for
// k = 0..d-2 the thin SVD of C viewed as (r_prev n_k) x rest / n_k, core k =
// U[:, :r_new], C <- diag(s) Vt[:r_new]; the remainder is the last core.
// A zero tensor gives rank-1 zero cores.
static tt tt_svd(const double* dense, const std::vector<int>& shape, double eps, int max_rank)
{
  const int d = (int)shape.size();
  tt out;
  if (d == 0) return out;
  size_t total = 1;
  for (int n : shape)
  {
    if (n <= 0) die("tt_svd: non-positive mode size");
    total *= (size_t)n;
  }
  const double norm_T = frob_norm_buffer(dense, total);
  if (!std::isfinite(norm_T)) die("tt_svd: non-finite input");
  if (norm_T == 0.0)
  {
    for (int k = 0; k < d; ++k) out.c.emplace_back(1, shape[k], 1);
    return out;
  }
  const double delta = eps * norm_T / std::sqrt((double)std::max(d - 1, 1));
  std::vector<double> C(dense, dense + total);
  int r_prev = 1;
  size_t rest = total;
  for (int k = 0; k < d - 1; ++k)
  {
    const int n_k = shape[k];
    const int rows = r_prev * n_k;
    const int cols = (int)(rest / n_k);
    mat U, Vt;
    std::vector<double> s;
    svd_thin(C.data(), rows, cols, U, s, Vt);
    const int r_new = truncate_eps_rank(s.data(), (int)s.size(), delta, max_rank);
    core ck(r_prev, n_k, r_new);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < r_new; ++j) ck.v[(size_t)i * r_new + j] = U(i, j);
    out.c.push_back(std::move(ck));
    C.assign((size_t)r_new * cols, 0.0);
    for (int i = 0; i < r_new; ++i)
      for (int j = 0; j < cols; ++j) C[(size_t)i * cols + j] = s[i] * Vt(i, j);
    r_prev = r_new;
    rest = (size_t)cols;
  }
  core last(r_prev, shape[d - 1], 1);
  std::memcpy(last.data(), C.data(), sizeof(double) * last.size());
  out.c.push_back(std::move(last));
  return out;
}

// detail::qtt_svd (include/core/factory/from_dense.hpp:109-199): each mode
// n_k = base^{b_k} splits into b_k modes of size base, the dense data is
// reordered to the QTT multi-index (lexicographic, last fastest; digits of
// an original index most significant first), then TT-SVD.
static tt qtt_svd(const double* data, const std::vector<int>& shape, int base, double eps, int max_rank)
{
  const int d0 = (int)shape.size();
  std::vector<int> bits(d0, 0), qshape;
  for (int k = 0; k < d0; ++k)
  {
    int n = shape[k], b = 0;
    if (n <= 0) die("qtt_svd: non-positive mode size");
    while (n > 1)
    {
      if (n % base) die("qtt_svd: mode size not a power of the base");
      n /= base;
      ++b;
    }
    bits[k] = b;
    for (int i = 0; i < b; ++i) qshape.push_back(base);
  }
  if (qshape.empty())
  {
    tt out;
    core c(1, 1, 1);
    c.v[0] = data[0];
    out.c.push_back(c);
    return out;
  }
  std::vector<long long> stride(d0, 1);
  for (int k = d0 - 2; k >= 0; --k) stride[k] = stride[k + 1] * shape[k + 1];
  size_t total = 1;
  for (int s : qshape) total *= (size_t)s;
  std::vector<double> q(total);
  std::vector<int> idx(qshape.size(), 0);
  for (size_t pos = 0; pos < total; ++pos)
  {
    int bp = 0;
    long long orig = 0;
    for (int k = 0; k < d0; ++k)
    {
      long long v = 0;
      for (int b = 0; b < bits[k]; ++b, ++bp) v = v * base + idx[bp];
      orig += v * stride[k];
    }
    q[pos] = data[orig];
    for (int i = (int)qshape.size() - 1; i >= 0; --i)
    {
      if (++idx[i] < base) break;
      idx[i] = 0;
    }
  }
  return tt_svd(q.data(), qshape, eps, max_rank);
}

// matrix_from_dense (include/core/factory/tt_matrix_factory.hpp:205-290):
// the dense (prod m) x (prod n) matrix is permuted to per-axis (m_k, n_k)
// pairs, axis-major, then TT-SVD over the modes m_k n_k.
static tt matrix_from_dense(const double* dense, const std::vector<int>& ms, const std::vector<int>& ns, double eps,
                            int max_rank)
{
  const int d = (int)ms.size();
  std::vector<int> mn(d);
  size_t M = 1, N = 1;
  for (int k = 0; k < d; ++k)
  {
    mn[k] = ms[k] * ns[k];
    M *= (size_t)ms[k];
    N *= (size_t)ns[k];
  }
  std::vector<double> perm(M * N);
  for (size_t I = 0; I < M; ++I)
    for (size_t J = 0; J < N; ++J)
    {
      size_t i = I, j = J, flat = 0, kstride = 1;
      for (int k = d - 1; k >= 0; --k)
      {
        const size_t ik = i % ms[k], jk = j % ns[k];
        i /= ms[k];
        j /= ns[k];
        flat += (ik * ns[k] + jk) * kstride;
        kstride *= (size_t)mn[k];
      }
      perm[flat] = dense[I * N + J];
    }
  tt out = tt_svd(perm.data(), mn, eps, max_rank);
  out.m = ms;
  out.nn = ns;
  return out;
}
}  // namespace ora

extern "C"
{
  void* ora_new(int d, const int* n, const int* m, const int* nn, const int* ranks, const double* data)
  {
    tt* t = new tt();
    size_t off = 0;
    for (int k = 0; k < d; ++k)
    {
      ora::core c(ranks[k], n[k], ranks[k + 1]);
      std::memcpy(c.data(), data + off, sizeof(double) * c.size());
      off += c.size();
      t->c.push_back(std::move(c));
    }
    if (m && nn)
    {
      t->m.assign(m, m + d);
      t->nn.assign(nn, nn + d);
    }
    return t;
  }
  void ora_free(void* h) { delete static_cast<tt*>(h); }
  int ora_d(void* h) { return static_cast<tt*>(h)->d(); }
  void ora_ranks(void* h, int* out)
  {
    tt* t = static_cast<tt*>(h);
    if (t->d() == 0)
    {
      out[0] = 1;
      return;
    }
    for (int k = 0; k < t->d(); ++k) out[k] = t->c[k].r;
    out[t->d()] = t->c.back().rp;
  }
  void ora_modes(void* h, int* out)
  {
    tt* t = static_cast<tt*>(h);
    for (int k = 0; k < t->d(); ++k) out[k] = t->c[k].n;
  }
  long long ora_size(void* h)
  {
    long long s = 0;
    for (auto& c : static_cast<tt*>(h)->c) s += (long long)c.size();
    return s;
  }
  void ora_data(void* h, double* out)
  {
    size_t off = 0;
    for (auto& c : static_cast<tt*>(h)->c)
    {
      std::memcpy(out + off, c.data(), sizeof(double) * c.size());
      off += c.size();
    }
  }
  void* ora_random(int d, const int* n, int max_rank, unsigned long long seed)
  {
    return new tt(ora::random_impl(std::vector<int>(n, n + d), max_rank, seed));
  }
  void* ora_random_matrix(int d, const int* m, const int* nn, int max_rank, unsigned long long seed)
  {
    std::vector<int> N(d);
    for (int k = 0; k < d; ++k) N[k] = m[k] * nn[k];
    tt* t = new tt(ora::random_impl(N, max_rank, seed));
    t->m.assign(m, m + d);
    t->nn.assign(nn, nn + d);
    return t;
  }
  void* ora_copy(void* h) { return new tt(*static_cast<tt*>(h)); }
  void* ora_axpby(double alpha, void* a, double beta, void* b)
  {
    return new tt(ora::axpby(alpha, *static_cast<tt*>(a), beta, *static_cast<tt*>(b)));
  }
  void* ora_scale(void* a, double alpha) { return new tt(ora::scale(*static_cast<tt*>(a), alpha)); }
  // gauge: 0 = none, 1 = right_canonical (round.hpp:253-260).
  void* ora_round(void* a, double eps, int max_rank, int gauge)
  {
    const tt& in = *static_cast<tt*>(a);
    return new tt(gauge == 1 ? ora::round_right_canonical(in, eps, max_rank) : ora::round_tt(in, eps, max_rank));
  }
  // The CPU baseline of the batched workloads: outs[i] = round(in[i], eps,
  // max_rank) for count TTs on `threads` std::threads, each taking the next
  // index from an atomic counter.  Every TT is rounded single-threaded, as
  // the reference rounds one TT (it has no threads of its own).
  void ora_round_batch(void* const* in, int count, double eps, int max_rank, int threads, void** outs)
  {
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1))
        outs[i] = new tt(ora::round_tt(*static_cast<tt*>(in[i]), eps, max_rank));
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
  }
  void ora_right_orthogonalize(void* a) { ora::right_orthogonalize(*static_cast<tt*>(a)); }
  void ora_left_orthogonalize(void* a) { ora::left_orthogonalize(*static_cast<tt*>(a)); }
  void* ora_matvec(void* A, void* x)
  {
    return new tt(ora::matvec(*static_cast<tt*>(A), *static_cast<tt*>(x)));
  }
  void* ora_matmat(void* A, void* B)
  {
    return new tt(ora::matmat(*static_cast<tt*>(A), *static_cast<tt*>(B)));
  }
  // method: 0 = svd (naive: apply then round, tt_matrix_apply_round.hpp:32-38),
  //         1 = svd_streaming (streaming_apply.hpp:39-110).
  void* ora_matvec_round(void* A, void* x, double eps, int max_rank, int method)
  {
    const tt& a = *static_cast<tt*>(A);
    const tt& xx = *static_cast<tt*>(x);
    if (method == 1 && a.d() > 1) return new tt(ora::build_streaming(a, xx, false, eps, max_rank));
    return new tt(ora::round_tt(ora::matvec(a, xx), eps, max_rank));
  }
  // method: 0 = svd (streaming, tt_matrix_apply_round.hpp:76-92),
  //         2 = svd_naive (round_matrix(matmat), :66-72).
  void* ora_matmat_round(void* A, void* B, double eps, int max_rank, int method)
  {
    const tt& a = *static_cast<tt*>(A);
    const tt& b = *static_cast<tt*>(B);
    if (method == 0 && a.d() > 1) return new tt(ora::build_streaming(a, b, true, eps, max_rank));
    tt out = ora::round_tt(ora::matmat(a, b), eps, max_rank);
    return new tt(std::move(out));
  }
  double ora_inner(void* a, void* b) { return ora::inner(*static_cast<tt*>(a), *static_cast<tt*>(b)); }
  double ora_norm(void* a)
  {
    const double s = ora::inner(*static_cast<tt*>(a), *static_cast<tt*>(a));
    return std::sqrt(std::max(s, 0.0));
  }
  long long ora_dense_size(void* h)
  {
    long long s = 1;
    for (auto& c : static_cast<tt*>(h)->c) s *= c.n;
    return s;
  }
  void ora_to_dense(void* h, double* out)
  {
    auto v = ora::to_dense(*static_cast<tt*>(h));
    std::memcpy(out, v.data(), sizeof(double) * v.size());
  }
  // Raw kernels for golden-vector checks.
  void ora_qr_thin(const double* M, int rows, int cols, double* Q, double* R)
  {
    ora::mat q, r;
    ora::qr_thin(M, rows, cols, q, r);
    std::memcpy(Q, q.v.data(), sizeof(double) * q.v.size());
    std::memcpy(R, r.v.data(), sizeof(double) * r.v.size());
  }
  void ora_set_svd_extended(int on) { ora::g_svd_extended = on != 0; }
  void ora_svd_thin(const double* M, int rows, int cols, double* U, double* s, double* Vt)
  {
    ora::mat u, vt;
    std::vector<double> sv;
    ora::svd_thin(M, rows, cols, u, sv, vt);
    std::memcpy(U, u.v.data(), sizeof(double) * u.v.size());
    std::memcpy(s, sv.data(), sizeof(double) * sv.size());
    std::memcpy(Vt, vt.v.data(), sizeof(double) * vt.v.size());
  }
  int ora_truncate_eps_rank(const double* s, int n, double delta, int max_rank)
  {
    return ora::truncate_eps_rank(s, n, delta, max_rank);
  }
  void* ora_round_als(void* a, double eps, int max_rank, int max_iters, double tol, void* warm, int cold_random,
                      unsigned long long seed, int* info_iters, double* info_change, int* info_conv)
  {
    return new tt(ora::als_round(*static_cast<tt*>(a), eps, max_rank, max_iters, tol, static_cast<tt*>(warm),
                                 cold_random != 0, seed, info_iters, info_change, info_conv));
  }
  void* ora_round_dmrg(void* a, double eps, int max_rank, int max_iters, double tol, void* warm, int apply_path,
                       unsigned long long seed, int* info_iters, double* info_change, int* info_conv)
  {
    return new tt(ora::dmrg_round(*static_cast<tt*>(a), eps, max_rank, max_iters, tol, static_cast<tt*>(warm),
                                  apply_path != 0, seed, info_iters, info_change, info_conv));
  }
  void* ora_soft_threshold(void* a, double tau) { return new tt(ora::soft_threshold(*static_cast<tt*>(a), tau)); }
  void* ora_from_dense(const double* dense, int d, const int* shape, double eps, int max_rank)
  {
    return new tt(ora::tt_svd(dense, std::vector<int>(shape, shape + d), eps, max_rank));
  }
  void* ora_from_dense_qtt(const double* dense, int d, const int* shape, int base, double eps, int max_rank)
  {
    return new tt(ora::qtt_svd(dense, std::vector<int>(shape, shape + d), base, eps, max_rank));
  }
  void* ora_matrix_from_dense(const double* dense, int d, const int* ms, const int* ns, double eps, int max_rank)
  {
    return new tt(ora::matrix_from_dense(dense, std::vector<int>(ms, ms + d), std::vector<int>(ns, ns + d), eps, max_rank));
  }
}
