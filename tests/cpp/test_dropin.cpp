// Drop-in check of include/tensor_train.hpp: the assertions of the
// reference's hot-path C++ tests (bdaszuta/tensor_train tests/core/
// test_tt_round.cpp, test_tt_invariants.cpp, test_tt_orthogonalize.cpp,
// test_tt_matrix_apply_round.cpp, test_tt_ops.cpp, test_tt_validation.cpp),
// written against the mva::tensor_train API and run on the GPU.  Exit code =
// number of failed checks.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "tensor_train.hpp"

namespace tt = mva::tensor_train;

static int failures = 0;
#define CHECK(cond, msg)                                   \
  do                                                       \
  {                                                        \
    if (!(cond))                                           \
    {                                                      \
      std::printf("FAIL: %s (%s:%d)\n", msg, __FILE__, __LINE__); \
      ++failures;                                          \
    }                                                      \
    else                                                   \
      std::printf("ok:   %s\n", msg);                      \
  } while (0)

static double max_abs_diff(const std::vector<double>& a, const std::vector<double>& b)
{
  double m = 0.0;
  for (size_t i = 0; i < a.size(); ++i) m = std::fmax(m, std::fabs(a[i] - b[i]));
  return m;
}

static double frob(const std::vector<double>& a)
{
  double s = 0.0;
  for (double v : a) s += v * v;
  return std::sqrt(s);
}

static tt::round_options opts_svd(double eps, int mr = 0)
{
  tt::round_options o;
  o.eps = eps;
  o.max_rank = mr;
  return o;
}

static bool left_orthogonal(const tt::tt_core& c, double tol)
{
  const int rows = c.r_left() * c.n_phys(), cols = c.r_right();
  double err = 0.0;
  for (int a = 0; a < cols; ++a)
    for (int b = 0; b < cols; ++b)
    {
      double s = 0.0;
      for (int i = 0; i < rows; ++i) s += c.data()[(size_t)i * cols + a] * c.data()[(size_t)i * cols + b];
      s -= (a == b);
      err += s * s;
    }
  return std::sqrt(err) <= tol * cols;
}

static bool right_orthogonal(const tt::tt_core& c, double tol)
{
  const int rows = c.r_left(), cols = c.n_phys() * c.r_right();
  double err = 0.0;
  for (int a = 0; a < rows; ++a)
    for (int b = 0; b < rows; ++b)
    {
      double s = 0.0;
      for (int i = 0; i < cols; ++i) s += c.data()[(size_t)a * cols + i] * c.data()[(size_t)b * cols + i];
      s -= (a == b);
      err += s * s;
    }
  return std::sqrt(err) <= tol * rows;
}

int main()
{
  // Config 1: round(add(add(a, b), a), 1e-10) recovers 2a + b at ranks <= 16.
  {
    const std::vector<int> shape(10, 4);
    auto a = tt::random(shape, 8, 1);
    auto b = tt::random(shape, 8, 2);
    auto c = tt::add(tt::add(a, b), a);
    auto r = tt::round(c, opts_svd(1e-10));
    const std::vector<int> want = {1, 4, 16, 16, 16, 16, 16, 16, 16, 4, 1};
    CHECK(r.ranks() == want, "config 1: ranks [1,4,16,...,16,4,1]");
    auto da = a.to_dense(), db = b.to_dense(), dr = r.to_dense();
    std::vector<double> ref(da.size());
    for (size_t i = 0; i < da.size(); ++i) ref[i] = 2.0 * da[i] + db[i];
    std::vector<double> diff(ref.size());
    for (size_t i = 0; i < ref.size(); ++i) diff[i] = dr[i] - ref[i];
    CHECK(frob(diff) / frob(ref) <= 1e-10, "config 1: rel. error <= 1e-10");
    tt::round_result info;
    tt::round(c, opts_svd(1e-10), &info);
    CHECK(info.iters_run == 1 && info.converged && info.final_resid == 0.0, "round_result {1, 0, true}");
  }
  // test_tt_invariants.cpp: gauge after round, right_orthogonalize norm, monotone norm.
  {
    auto r = tt::round(tt::random({3, 4, 5, 3}, 6, 42), opts_svd(1e-10));
    bool ok = true;
    for (int k = 0; k < r.d() - 1; ++k) ok = ok && left_orthogonal(r.core(k), 1e-13);
    CHECK(ok, "round: cores[0..d-2] left-orthogonal");
    auto a = tt::random({3, 4, 5, 3}, 6, 99);
    const double n1 = tt::norm(a);
    tt::right_orthogonalize(a);
    ok = true;
    for (int k = 1; k < a.d(); ++k) ok = ok && right_orthogonal(a.core(k), 1e-13);
    CHECK(ok, "right_orthogonalize: cores[1..d-1] right-orthogonal");
    CHECK(std::fabs(tt::norm(a) - n1) <= 1e-13 * (n1 + 1.0), "right_orthogonalize: norm preserved");
    auto b = tt::random({2, 3, 4, 2}, 8, 7);
    CHECK(tt::norm(tt::round(b, opts_svd(1e-2))) <= tt::norm(b) * (1 + 1e-14), "round: norm non-increasing");
    auto r3 = tt::round(b, opts_svd(0.0, 3));
    CHECK(r3.max_rank() <= 3, "round: max_rank cap");
    auto rr = tt::round(tt::random({2, 3, 4, 2, 3}, 6, 123), opts_svd(1e-6));
    auto rr2 = tt::round(rr, opts_svd(1e-6));
    CHECK(max_abs_diff(rr.to_dense(), rr2.to_dense()) <= 1e-13 * (frob(rr.to_dense()) + 1), "round idempotence");
    auto z = tt::round(tt::zeros({3, 4, 2}), opts_svd(1e-10));
    CHECK(z.ranks() == std::vector<int>({1, 1, 1, 1}) && frob(z.to_dense()) == 0.0, "round(zeros): rank-1 zero");
    auto o7 = tt::ones({7});
    CHECK(max_abs_diff(tt::round(o7, opts_svd(1e-10)).to_dense(), o7.to_dense()) == 0.0, "round(d=1) identity");
  }
  // right canonical gauge.
  {
    tt::round_options o = opts_svd(1e-10);
    o.gauge = tt::round_gauge::right_canonical;
    auto r = tt::round(tt::random({3, 4, 5, 3}, 6, 42), o);
    bool ok = true;
    for (int k = 1; k < r.d(); ++k) ok = ok && right_orthogonal(r.core(k), 1e-13);
    CHECK(ok, "round(right_canonical): cores[1..d-1] right-orthogonal");
  }
  // test_tt_matrix_apply_round.cpp: matvec_round ~ matvec; streaming vs naive.
  {
    auto A = tt::random({2, 3, 2, 3}, {3, 2, 2, 2}, 4, 11);
    auto x = tt::random({3, 2, 2, 2}, 5, 12);
    auto y = tt::matvec(A, x);
    auto yr = tt::matvec_round(A, x, opts_svd(1e-12));
    const auto dy = y.to_dense();
    std::vector<double> diff(dy.size());
    auto dyr = yr.to_dense();
    for (size_t i = 0; i < dy.size(); ++i) diff[i] = dyr[i] - dy[i];
    CHECK(frob(diff) <= 1e-9 * frob(dy), "matvec_round ~ matvec");
    tt::round_options so = opts_svd(1e-6);
    so.method = tt::round_method::svd_streaming;
    auto ys = tt::matvec_round(A, x, so);
    auto yn = tt::matvec_round(A, x, opts_svd(1e-6));
    auto ds = ys.to_dense(), dn = yn.to_dense();
    for (size_t i = 0; i < dy.size(); ++i) diff[i] = ds[i] - dn[i];
    CHECK(frob(diff) <= 10 * 1e-6 * frob(dy), "streaming vs naive within 10 eps ||y||");
    auto B = tt::random({3, 2, 2, 2}, {2, 2, 3, 2}, 3, 13);
    auto C = tt::matmat_round(A, B, opts_svd(1e-10));
    auto Cn = tt::matmat(A, B);
    CHECK(std::fabs(tt::frob_norm(C) - tt::frob_norm(Cn)) <= 1e-8 * tt::frob_norm(Cn), "matmat_round ~ matmat");
  }
  // test_tt_ops.cpp: round(add(A, A)) = 2A, ranks <= ranks(A).
  {
    auto a = tt::random({3, 4, 3, 2}, 5, 77);
    auto r = tt::round(tt::add(a, a), opts_svd(1e-12));
    auto ra = a.ranks(), rr = r.ranks();
    bool ok = true;
    for (size_t k = 0; k < ra.size(); ++k) ok = ok && rr[k] <= ra[k];
    auto da = a.to_dense(), dr = r.to_dense();
    for (auto& v : da) v *= 2.0;
    CHECK(ok && max_abs_diff(da, dr) <= 1e-9, "round(add(A, A)) = 2A, ranks <= ranks(A)");
  }
  // test_tt_validation.cpp: d = 0 conventions.
  {
    tt::tt e;
    CHECK(e.ranks() == std::vector<int>({1}) && e.to_dense() == std::vector<double>({1.0}), "d=0 conventions");
  }
  // test_tt_svd.cpp: from_dense recovers a rank-1 outer product with ranks 1.
  {
    const double u[3] = {1.0, 2.0, 3.0}, v[4] = {1.0, -1.0, 0.5, 0.25}, w[5] = {2.0, 0.5, -1.0, 1.0, 3.0};
    std::vector<double> T(60);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 5; ++k) T[(i * 4 + j) * 5 + k] = u[i] * v[j] * w[k];
    auto t = tt::from_dense(T.data(), {3, 4, 5}, 1e-12);
    CHECK(t.ranks() == std::vector<int>({1, 1, 1, 1}), "from_dense: rank-1 outer product, ranks 1");
    CHECK(max_abs_diff(t.to_dense(), T) <= 1e-12 * frob(T), "from_dense: reconstruction");
  }
  // test_tt_round.cpp:117-226: ALS on a zero-padded exact TT (cold, warm,
  // against svd, warm with a perturbation).
  {
    std::mt19937_64 rng(13);
    std::normal_distribution<double> N(0.0, 1.0);
    std::vector<double> T(4 * 5 * 6 * 4);
    for (auto& v : T) v = N(rng);
    auto A = tt::from_dense(T.data(), {4, 5, 6, 4}, 1e-12);
    const int cap = A.max_rank();
    std::vector<tt::tt_core> cs;
    for (int k = 0; k < A.d(); ++k)
    {
      const auto& s = A.core(k);
      const int rl = s.r_left() + (k > 0 ? 5 : 0), rr = s.r_right() + (k < A.d() - 1 ? 5 : 0);
      tt::tt_core d(rl, s.n_phys(), rr);
      for (int a = 0; a < rl; ++a)
        for (int i = 0; i < s.n_phys(); ++i)
          for (int b = 0; b < rr; ++b) d(a, i, b) = (a < s.r_left() && b < s.r_right()) ? s(a, i, b) : 0.0;
      cs.push_back(std::move(d));
    }
    tt::tt B(std::move(cs));
    const auto dA = A.to_dense();
    auto rel_to = [&](const std::vector<double>& x, const std::vector<double>& ref)
    {
      std::vector<double> df(ref.size());
      for (size_t i = 0; i < ref.size(); ++i) df[i] = x[i] - ref[i];
      return frob(df) / frob(ref);
    };
    tt::round_options o;
    o.method = tt::round_method::als;
    o.max_rank = cap;
    o.max_iters = 3;
    o.tol = 1e-12;
    tt::round_result info;
    auto C = tt::round(B, o, &info);
    CHECK(rel_to(C.to_dense(), dA) <= 1e-10 && C.max_rank() <= cap && info.iters_run <= 3,
          "als cold: recovers the padded TT within the cap");
    o.warm_start_tt = &A;
    auto W = tt::round(B, o, &info);
    CHECK(rel_to(W.to_dense(), dA) <= 1e-10, "als warm: recovers the padded TT");
    o.warm_start_tt = nullptr;
    o.tol = 1e-13;
    auto Cs = tt::round(B, opts_svd(1e-10, cap));
    auto Ca = tt::round(B, o);
    CHECK(rel_to(Ca.to_dense(), dA) <= 10.0 * std::fmax(rel_to(Cs.to_dense(), dA), 1e-13), "als vs svd");
    auto Bp = tt::axpby(1.0, A, 1.0e-3, tt::random({4, 5, 6, 4}, 4, 99));
    o.tol = 1e-12;
    o.warm_start_tt = &A;
    auto P = tt::round(Bp, o, &info);
    CHECK(info.iters_run <= 5 && rel_to(P.to_dense(), Bp.to_dense()) <= 1e-2, "als warm-pert");
    // test_tt_round.cpp:227-384: the same cases for dmrg, plus eps-only.
    tt::round_options g;
    g.method = tt::round_method::dmrg;
    g.max_rank = cap;
    g.eps = 1e-12;
    g.max_iters = 3;
    g.tol = 1e-12;
    auto D = tt::round(B, g, &info);
    CHECK(rel_to(D.to_dense(), dA) <= 1e-10 && D.max_rank() <= cap, "dmrg cold");
    g.warm_start_tt = &A;
    CHECK(rel_to(tt::round(B, g).to_dense(), dA) <= 1e-10, "dmrg warm");
    g.warm_start_tt = nullptr;
    g.tol = 1e-13;
    CHECK(rel_to(tt::round(B, g).to_dense(), dA) <= 1.1 * rel_to(Cs.to_dense(), dA) + 1e-12, "dmrg vs svd");
    g.tol = 1e-12;
    g.warm_start_tt = &A;
    auto Pd = tt::round(Bp, g, &info);
    CHECK(info.iters_run <= 5 && rel_to(Pd.to_dense(), Bp.to_dense()) <= 1e-2, "dmrg warm-pert");
    g.warm_start_tt = nullptr;
    g.max_rank = 0;
    g.eps = 1e-10;
    auto De = tt::round(B, g, &info);
    CHECK(rel_to(De.to_dense(), dA) <= 1e-8 && De.max_rank() <= B.max_rank(), "dmrg eps-only");
  }
  // test_tt_matvec_round.cpp / test_tt_matrix_round.cpp: ALS apply-round
  // recovers the product at its exact rank from a random cold start.
  {
    auto A = tt::random({2, 3, 2, 3}, {3, 2, 2, 2}, 2, 7);
    auto x = tt::random({3, 2, 2, 2}, 2, 8);
    auto y = tt::matvec(A, x);
    tt::round_options o;
    o.method = tt::round_method::als;
    o.max_rank = tt::round(y, opts_svd(1e-13)).max_rank();
    o.seed = 0x3030;
    tt::round_result info;
    auto ya = tt::matvec_round(A, x, o, &info);
    const auto dy = y.to_dense(), da = ya.to_dense();
    std::vector<double> df(dy.size());
    for (size_t i = 0; i < dy.size(); ++i) df[i] = da[i] - dy[i];
    CHECK(frob(df) <= 1e-10 * frob(dy) && info.converged, "matvec_round(als) recovers A x");
    auto B = tt::random({3, 2, 2, 2}, {2, 2, 3, 2}, 2, 9);
    auto Cn = tt::matmat(A, B);
    o.max_rank = tt::round(Cn, opts_svd(1e-13)).max_rank();
    auto Ca = tt::matmat_round(A, B, o, &info);
    const auto dc = Cn.to_dense(), dca = Ca.to_dense();
    df.assign(dc.size(), 0.0);
    for (size_t i = 0; i < dc.size(); ++i) df[i] = dca[i] - dc[i];
    CHECK(frob(df) <= 1e-10 * frob(dc) && info.converged, "matmat_round(als) recovers A B");
  }
  // tt_matrix_frob_norm_apply.hpp: ||A x|| and ||A B|| without the product.
  {
    auto A = tt::random({2, 3, 2, 3}, {3, 2, 2, 2}, 3, 21);
    auto x = tt::random({3, 2, 2, 2}, 4, 22);
    auto B = tt::random({3, 2, 2, 2}, {2, 2, 3, 2}, 2, 23);
    const double nx = tt::norm(tt::matvec(A, x)), nb = tt::frob_norm(tt::matmat(A, B));
    CHECK(std::fabs(tt::frob_norm_apply(A, x) - nx) <= 1e-12 * nx, "frob_norm_apply(A, x) = norm(matvec(A, x))");
    CHECK(std::fabs(tt::frob_norm_apply(A, B) - nb) <= 1e-12 * nb, "frob_norm_apply(A, B) = frob_norm(matmat(A, B))");
  }
  std::printf("%d failure(s)\n", failures);
  return failures;
}
