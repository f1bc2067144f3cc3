// tensor_train.hpp -- header-only C++20 drop-in for the TT-rounding hot path
// of bdaszuta/tensor_train (namespace mva::tensor_train), running on an
// MI355X (gfx950) through the C ABI of libtt_amd.so (tt_amd.h).
//
// Kept from the reference (include/tensor_train.hpp:11-55 umbrella):
//   * types tt_core, tt, tt_matrix_core, tt_matrix with row-major core
//     layout (a*n + i)*rp + b and ((a*m + i)*n + j)*rp + b
//     (include/core/detail/storage_backend.hpp:237-253), value semantics;
//   * round_method, round_gauge, round_options, round_result
//     (include/core/round/round.hpp:33-119);
//   * round (round.hpp:230-425), right_orthogonalize
//     (gauge/tt_orthogonalize.hpp:26, tt_matrix_orthogonalize.hpp:34),
//     matvec / matmat (apply/tt_matrix_apply.hpp:237, :260), matvec_round
//     (round/matvec_round.hpp:42), matmat_round (round/matmat_round.hpp:41),
//     inner / norm (algebra/tt_inner.hpp:33, :55), frob_inner / frob_norm
//     (algebra/tt_matrix_inner.hpp:31, :44), scale / axpby / add / sub
//     (algebra/tt_ops.hpp:35-122), random / zeros / ones
//     (factory/tt_factory.hpp:31-207, tt_matrix_factory.hpp:133-191),
//     from_dense and its options (factory/from_dense.hpp:27-278);
//   * errors: fprintf(stderr) + std::abort(), as the reference does.
// Each call copies its host operands to the GPU, runs the device path and
// copies the result back; callers that want to keep tensors resident use the
// C ABI (tt_amd.h) directly.  One device context per thread (device 0, or
// TT_AMD_DEVICE).  round_method::als / dmrg run the ALS and two-site DMRG
// engines on the device (round/tt_als.hpp, round/tt_dmrg.hpp).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <utility>
#include <vector>

#include "tt_amd.h"

namespace mva
{
namespace tensor_train
{

[[noreturn]] inline void tta_die(const char* what, const char* msg)
{
  std::fprintf(stderr, "%s: %s\n", what, msg);
  std::abort();
}

namespace detail
{
// Per-thread device context; created on first use.
inline tta_ctx* device_ctx()
{
  thread_local struct holder
  {
    tta_ctx* c = nullptr;
    holder()
    {
      int dev = 0;
      if (const char* e = std::getenv("TT_AMD_DEVICE")) dev = std::atoi(e);
      if (tta_ctx_create(dev, &c) != TTA_OK) tta_die("tensor_train", "no usable gfx950 device (tta_ctx_create failed)");
    }
    ~holder()
    {
      if (c) tta_ctx_destroy(c);
    }
  } h;
  return h.c;
}

inline void check(int status, const char* what)
{
  if (status != TTA_OK) tta_die(what, tta_last_error(device_ctx()));
}
}  // namespace detail

// ---------------------------------------------------------------------------
// Cores and trains.
// ---------------------------------------------------------------------------
class tt_core
{
 public:
  tt_core() = default;
  tt_core(int r_left, int n_phys, int r_right) { allocate(r_left, n_phys, r_right); }
  void allocate(int r_left, int n_phys, int r_right)
  {
    if (r_left <= 0 || n_phys <= 0 || r_right <= 0) tta_die("tt_core", "dimensions must be positive");
    const int64_t total = (int64_t)r_left * n_phys * r_right;
    if (total > std::numeric_limits<int>::max()) tta_die("tt_core", "dimension product exceeds INT_MAX");
    r_ = r_left;
    n_ = n_phys;
    rp_ = r_right;
    v_.assign((size_t)total, 0.0);
  }
  void zero_clear() { std::fill(v_.begin(), v_.end(), 0.0); }
  double& operator()(int a, int i, int b) { return v_[((size_t)a * n_ + i) * rp_ + b]; }
  double operator()(int a, int i, int b) const { return v_[((size_t)a * n_ + i) * rp_ + b]; }
  int r_left() const { return r_; }
  int n_phys() const { return n_; }
  int r_right() const { return rp_; }
  std::size_t size() const { return v_.size(); }
  std::array<int, 3> dims() const { return {r_, n_, rp_}; }
  double* data() { return v_.data(); }
  const double* data() const { return v_.data(); }

 private:
  int r_ = 0, n_ = 0, rp_ = 0;
  std::vector<double> v_;
};

class tt
{
 public:
  tt() = default;
  explicit tt(std::vector<tt_core> cores) : cores_(std::move(cores))
  {
    const int d = (int)cores_.size();
    if (d == 0) return;
    if (cores_[0].r_left() != 1) tta_die("tt", "r_0 must be 1");
    if (cores_[d - 1].r_right() != 1) tta_die("tt", "r_d must be 1");
    for (int k = 0; k < d - 1; ++k)
      if (cores_[k].r_right() != cores_[k + 1].r_left()) tta_die("tt", "rank mismatch between neighbouring cores");
  }
  int d() const { return (int)cores_.size(); }
  std::vector<int> shape() const
  {
    std::vector<int> s;
    for (auto& c : cores_) s.push_back(c.n_phys());
    return s;
  }
  std::vector<int> ranks() const
  {
    if (cores_.empty()) return {1};
    std::vector<int> r;
    for (auto& c : cores_) r.push_back(c.r_left());
    r.push_back(cores_.back().r_right());
    return r;
  }
  int max_rank() const
  {
    int m = 1;
    for (int r : ranks()) m = r > m ? r : m;
    return m;
  }
  std::size_t num_params() const
  {
    std::size_t s = 0;
    for (auto& c : cores_) s += c.size();
    return s;
  }
  // Dense reconstruction, flat row-major over (i_0, ..., i_{d-1}) (tt.hpp:145-198).
  std::vector<double> to_dense() const
  {
    if (cores_.empty()) return {1.0};
    std::vector<double> T(cores_[0].data(), cores_[0].data() + cores_[0].size());
    size_t rows = (size_t)cores_[0].n_phys();
    int cols = cores_[0].r_right();
    for (size_t k = 1; k < cores_.size(); ++k)
    {
      const tt_core& c = cores_[k];
      const size_t nc = (size_t)c.n_phys() * c.r_right();
      std::vector<double> Tn(rows * nc, 0.0);
      for (size_t i = 0; i < rows; ++i)
        for (int a = 0; a < cols; ++a)
        {
          const double t = T[i * cols + a];
          const double* src = c.data() + (size_t)a * nc;
          double* dst = Tn.data() + i * nc;
          for (size_t j = 0; j < nc; ++j) dst[j] += t * src[j];
        }
      T.swap(Tn);
      rows *= (size_t)c.n_phys();
      cols = c.r_right();
    }
    return T;
  }
  const tt_core& core(int k) const { return cores_.at((size_t)k); }
  tt_core& core(int k) { return cores_.at((size_t)k); }
  const std::vector<tt_core>& cores() const { return cores_; }
  std::vector<tt_core>& cores() { return cores_; }

 private:
  std::vector<tt_core> cores_;
};

class tt_matrix_core
{
 public:
  tt_matrix_core() = default;
  tt_matrix_core(int r_left, int m_phys, int n_phys, int r_right) { allocate(r_left, m_phys, n_phys, r_right); }
  void allocate(int r_left, int m_phys, int n_phys, int r_right)
  {
    if (r_left <= 0 || m_phys <= 0 || n_phys <= 0 || r_right <= 0) tta_die("tt_matrix_core", "dimensions must be positive");
    const int64_t total = (int64_t)r_left * m_phys * n_phys * r_right;
    if (total > std::numeric_limits<int>::max()) tta_die("tt_matrix_core", "dimension product exceeds INT_MAX");
    r_ = r_left;
    m_ = m_phys;
    n_ = n_phys;
    rp_ = r_right;
    v_.assign((size_t)total, 0.0);
  }
  void zero_clear() { std::fill(v_.begin(), v_.end(), 0.0); }
  double& operator()(int a, int i, int j, int b) { return v_[(((size_t)a * m_ + i) * n_ + j) * rp_ + b]; }
  double operator()(int a, int i, int j, int b) const { return v_[(((size_t)a * m_ + i) * n_ + j) * rp_ + b]; }
  int r_left() const { return r_; }
  int m_phys() const { return m_; }
  int n_phys() const { return n_; }
  int r_right() const { return rp_; }
  std::size_t size() const { return v_.size(); }
  double* data() { return v_.data(); }
  const double* data() const { return v_.data(); }

 private:
  int r_ = 0, m_ = 0, n_ = 0, rp_ = 0;
  std::vector<double> v_;
};

class tt_matrix
{
 public:
  tt_matrix() = default;
  explicit tt_matrix(std::vector<tt_matrix_core> cores) : cores_(std::move(cores))
  {
    const int d = (int)cores_.size();
    if (d == 0) return;
    if (cores_[0].r_left() != 1) tta_die("tt_matrix", "r_0 must be 1");
    if (cores_[d - 1].r_right() != 1) tta_die("tt_matrix", "r_d must be 1");
    for (int k = 0; k < d - 1; ++k)
      if (cores_[k].r_right() != cores_[k + 1].r_left()) tta_die("tt_matrix", "rank mismatch between neighbouring cores");
  }
  int d() const { return (int)cores_.size(); }
  std::vector<int> row_shape() const
  {
    std::vector<int> s;
    for (auto& c : cores_) s.push_back(c.m_phys());
    return s;
  }
  std::vector<int> col_shape() const
  {
    std::vector<int> s;
    for (auto& c : cores_) s.push_back(c.n_phys());
    return s;
  }
  std::vector<int> ranks() const
  {
    if (cores_.empty()) return {1};
    std::vector<int> r;
    for (auto& c : cores_) r.push_back(c.r_left());
    r.push_back(cores_.back().r_right());
    return r;
  }
  int max_rank() const
  {
    int m = 1;
    for (int r : ranks()) m = r > m ? r : m;
    return m;
  }
  std::size_t num_params() const
  {
    std::size_t s = 0;
    for (auto& c : cores_) s += c.size();
    return s;
  }
  std::size_t total_rows() const
  {
    std::size_t s = 1;
    for (auto& c : cores_) s *= (std::size_t)c.m_phys();
    return s;
  }
  std::size_t total_cols() const
  {
    std::size_t s = 1;
    for (auto& c : cores_) s *= (std::size_t)c.n_phys();
    return s;
  }
  // Dense row-major (total_rows x total_cols) matrix, row index (i_1..i_d),
  // column index (j_1..j_d), both with the last site fastest
  // (include/core/types/tt_matrix.hpp:415-500); d = 0 is [1].
  std::vector<double> to_dense() const
  {
    const int d = (int)cores_.size();
    if (d == 0) return {1.0};
    // T holds prod_k (m_k n_k) rows over the interleaved (i_1 j_1 .. i_k j_k)
    // index times the current right rank.
    std::vector<double> T(cores_[0].size());
    for (std::size_t t = 0; t < T.size(); ++t) T[t] = cores_[0].data()[t];
    std::size_t rows = (std::size_t)cores_[0].m_phys() * cores_[0].n_phys();
    for (int k = 1; k < d; ++k)
    {
      const tt_matrix_core& c = cores_[k];
      const std::size_t r = c.r_left(), w = (std::size_t)c.m_phys() * c.n_phys() * c.r_right();
      std::vector<double> Tn(rows * w, 0.0);
      for (std::size_t i = 0; i < rows; ++i)
        for (std::size_t a = 0; a < r; ++a)
        {
          const double x = T[i * r + a];
          if (x == 0.0) continue;
          const double* src = c.data() + a * w;
          double* dst = Tn.data() + i * w;
          for (std::size_t t = 0; t < w; ++t) dst[t] += x * src[t];
        }
      T.swap(Tn);
      rows *= (std::size_t)c.m_phys() * c.n_phys();
    }
    const std::size_t M = total_rows(), N = total_cols();
    std::vector<double> out(M * N);
    std::vector<int> ii(d, 0), jj(d, 0);
    for (std::size_t flat = 0; flat < rows; ++flat)
    {
      std::size_t rem = flat;
      for (int k = d - 1; k >= 0; --k)
      {
        const int m = cores_[k].m_phys(), n = cores_[k].n_phys();
        const std::size_t ij = rem % ((std::size_t)m * n);
        rem /= (std::size_t)m * n;
        ii[k] = (int)(ij / n);
        jj[k] = (int)(ij % n);
      }
      std::size_t I = 0, J = 0;
      for (int k = 0; k < d; ++k)
      {
        I = I * cores_[k].m_phys() + ii[k];
        J = J * cores_[k].n_phys() + jj[k];
      }
      out[I * N + J] = T[flat];
    }
    return out;
  }
  const tt_matrix_core& core(int k) const { return cores_.at((size_t)k); }
  tt_matrix_core& core(int k) { return cores_.at((size_t)k); }
  const std::vector<tt_matrix_core>& cores() const { return cores_; }
  std::vector<tt_matrix_core>& cores() { return cores_; }

 private:
  std::vector<tt_matrix_core> cores_;
};

// ---------------------------------------------------------------------------
// Options (include/core/round/round.hpp:33-119).
// ---------------------------------------------------------------------------
enum class round_method
{
  svd,
  svd_streaming,
  svd_naive,
  als,
  dmrg
};

enum class round_gauge
{
  none,
  right_canonical
};

struct round_options
{
  double eps = 0.0;
  int max_rank = 0;
  round_method method = round_method::svd;
  round_gauge gauge = round_gauge::none;
  int max_iters = 20;
  double tol = 1.0e-12;
  std::uint64_t seed = 0;
  const tt* warm_start_tt = nullptr;
  const tt_matrix* warm_start_tt_matrix = nullptr;
};

struct round_result
{
  int iters_run = 0;
  double final_resid = 0.0;
  bool converged = false;
};

// ---------------------------------------------------------------------------
// Host <-> device transfer.
// ---------------------------------------------------------------------------
namespace detail
{
struct dev_handle
{
  tta_tt* t = nullptr;
  dev_handle() = default;
  explicit dev_handle(tta_tt* p) : t(p) {}
  dev_handle(const dev_handle&) = delete;
  dev_handle& operator=(const dev_handle&) = delete;
  ~dev_handle()
  {
    if (t) tta_tt_free(device_ctx(), t);
  }
};

inline tta_tt* upload(const tt& a)
{
  const int d = a.d();
  std::vector<int> modes = a.shape(), ranks = a.ranks();
  tta_tt* t = nullptr;
  check(tta_tt_create(device_ctx(), d, modes.data(), ranks.data(), &t), "tensor_train");
  std::vector<double> flat;
  flat.reserve(a.num_params());
  for (auto& c : a.cores()) flat.insert(flat.end(), c.data(), c.data() + c.size());
  check(tta_tt_upload(device_ctx(), t, flat.data()), "tensor_train");
  return t;
}

inline tta_tt* upload(const tt_matrix& a)
{
  const int d = a.d();
  std::vector<int> rm = a.row_shape(), cm = a.col_shape(), ranks = a.ranks();
  tta_tt* t = nullptr;
  check(tta_ttm_create(device_ctx(), d, rm.data(), cm.data(), ranks.data(), &t), "tensor_train");
  std::vector<double> flat;
  flat.reserve(a.num_params());
  for (auto& c : a.cores()) flat.insert(flat.end(), c.data(), c.data() + c.size());
  check(tta_tt_upload(device_ctx(), t, flat.data()), "tensor_train");
  return t;
}

inline tt download_tt(const tta_tt* t)
{
  int d = 0, is_m = 0;
  tta_tt_dims(t, &d, &is_m);
  std::vector<int> modes((size_t)d + 1), ranks((size_t)d + 1);
  tta_tt_shape(t, modes.data(), nullptr, nullptr, ranks.data());
  std::vector<double> flat((size_t)std::max<long long>(tta_tt_numel(t), 1));
  check(tta_tt_download(device_ctx(), t, flat.data()), "tensor_train");
  std::vector<tt_core> cores;
  size_t off = 0;
  for (int k = 0; k < d; ++k)
  {
    tt_core c(ranks[k], modes[k], ranks[k + 1]);
    std::copy(flat.begin() + (long)off, flat.begin() + (long)(off + c.size()), c.data());
    off += c.size();
    cores.push_back(std::move(c));
  }
  return tt(std::move(cores));
}

inline tt_matrix download_ttm(const tta_tt* t)
{
  int d = 0, is_m = 0;
  tta_tt_dims(t, &d, &is_m);
  std::vector<int> rm((size_t)d + 1), cm((size_t)d + 1), ranks((size_t)d + 1);
  tta_tt_shape(t, nullptr, rm.data(), cm.data(), ranks.data());
  std::vector<double> flat((size_t)std::max<long long>(tta_tt_numel(t), 1));
  check(tta_tt_download(device_ctx(), t, flat.data()), "tensor_train");
  std::vector<tt_matrix_core> cores;
  size_t off = 0;
  for (int k = 0; k < d; ++k)
  {
    tt_matrix_core c(ranks[k], rm[k], cm[k], ranks[k + 1]);
    std::copy(flat.begin() + (long)off, flat.begin() + (long)(off + c.size()), c.data());
    off += c.size();
    cores.push_back(std::move(c));
  }
  return tt_matrix(std::move(cores));
}

inline int method_code(round_method m) { return static_cast<int>(m); }

inline bool iterative(round_method m) { return m == round_method::als || m == round_method::dmrg; }

// The warm start travels only to the iterative engines (the SVD paths ignore it).
template <class T>
inline tta_tt* upload_warm(round_method m, const T* w)
{
  return (w && iterative(m)) ? upload(*w) : nullptr;
}

inline void write_result(round_result* r, int iters, double resid, int conv)
{
  if (!r) return;
  r->iters_run = iters;
  r->final_resid = resid;
  r->converged = conv != 0;
}
}  // namespace detail

// ---------------------------------------------------------------------------
// The hot path.
// ---------------------------------------------------------------------------
inline tt round(const tt& a, const round_options& opts, round_result* info = nullptr)
{
  if (opts.method == round_method::svd_streaming || opts.method == round_method::svd_naive)
    tta_die("round(tt)",
            "round_method::svd_streaming / svd_naive are only valid on matvec_round / matmat_round respectively.");
  detail::dev_handle in(detail::upload(a));
  detail::dev_handle w(detail::upload_warm(opts.method, opts.warm_start_tt));
  tta_tt* out = nullptr;
  int it = 0, cv = 0;
  double ch = 0.0;
  detail::check(tta_round_iterative(detail::device_ctx(), in.t, opts.eps, opts.max_rank,
                                    detail::method_code(opts.method), static_cast<int>(opts.gauge), opts.max_iters,
                                    opts.tol, opts.seed, w.t, &out, &it, &ch, &cv),
                "round(tt)");
  detail::dev_handle o(out);
  detail::write_result(info, it, ch, cv);
  return detail::download_tt(out);
}

inline tt round(tt&& a, const round_options& opts, round_result* info = nullptr)
{
  const tt tmp(std::move(a));
  return round(tmp, opts, info);
}

inline tt_matrix round(const tt_matrix& a, const round_options& opts, round_result* info = nullptr)
{
  if (opts.method == round_method::svd_streaming || opts.method == round_method::svd_naive)
    tta_die("round(tt_matrix)",
            "round_method::svd_streaming / svd_naive are only valid on matvec_round / matmat_round respectively.");
  detail::dev_handle in(detail::upload(a));
  detail::dev_handle w(detail::upload_warm(opts.method, opts.warm_start_tt_matrix));
  tta_tt* out = nullptr;
  int it = 0, cv = 0;
  double ch = 0.0;
  detail::check(tta_round_iterative(detail::device_ctx(), in.t, opts.eps, opts.max_rank,
                                    detail::method_code(opts.method), static_cast<int>(opts.gauge), opts.max_iters,
                                    opts.tol, opts.seed, w.t, &out, &it, &ch, &cv),
                "round(tt_matrix)");
  detail::dev_handle o(out);
  detail::write_result(info, it, ch, cv);
  return detail::download_ttm(out);
}

inline tt_matrix round(tt_matrix&& a, const round_options& opts, round_result* info = nullptr)
{
  const tt_matrix tmp(std::move(a));
  return round(tmp, opts, info);
}

inline void right_orthogonalize(tt& a)
{
  if (a.d() <= 1) return;
  detail::dev_handle h(detail::upload(a));
  detail::check(tta_right_orthogonalize(detail::device_ctx(), h.t), "right_orthogonalize");
  a = detail::download_tt(h.t);
}

inline void right_orthogonalize(tt_matrix& a)
{
  if (a.d() <= 1) return;
  detail::dev_handle h(detail::upload(a));
  detail::check(tta_right_orthogonalize(detail::device_ctx(), h.t), "right_orthogonalize(tt_matrix)");
  a = detail::download_ttm(h.t);
}

namespace detail
{
inline void left_orthogonalize(tt& a)
{
  if (a.d() <= 1) return;
  dev_handle h(upload(a));
  check(tta_left_orthogonalize(device_ctx(), h.t), "left_orthogonalize");
  a = download_tt(h.t);
}
}  // namespace detail

inline tt matvec(const tt_matrix& A, const tt& x)
{
  if (A.d() == 0) return tt();
  detail::dev_handle a(detail::upload(A)), b(detail::upload(x));
  tta_tt* out = nullptr;
  detail::check(tta_matvec(detail::device_ctx(), a.t, b.t, &out), "matvec");
  detail::dev_handle o(out);
  return detail::download_tt(out);
}

inline tt_matrix matmat(const tt_matrix& A, const tt_matrix& B)
{
  if (A.d() == 0) return tt_matrix();
  detail::dev_handle a(detail::upload(A)), b(detail::upload(B));
  tta_tt* out = nullptr;
  detail::check(tta_matmat(detail::device_ctx(), a.t, b.t, &out), "matmat");
  detail::dev_handle o(out);
  return detail::download_ttm(out);
}

inline tt matvec_round(const tt_matrix& A, const tt& x, const round_options& opts, round_result* info = nullptr)
{
  if (opts.method == round_method::svd_naive) tta_die("matvec_round", "round_method::svd_naive is matmat-only");
  detail::dev_handle a(detail::upload(A)), b(detail::upload(x));
  detail::dev_handle w(detail::upload_warm(opts.method, opts.warm_start_tt));
  tta_tt* out = nullptr;
  int it = 0, cv = 0;
  double ch = 0.0;
  detail::check(tta_matvec_round_iterative(detail::device_ctx(), a.t, b.t, opts.eps, opts.max_rank,
                                           detail::method_code(opts.method), opts.max_iters, opts.tol, opts.seed,
                                           w.t, &out, &it, &ch, &cv),
                "matvec_round");
  detail::dev_handle o(out);
  detail::write_result(info, it, ch, cv);
  return detail::download_tt(out);
}

inline tt_matrix matmat_round(const tt_matrix& A, const tt_matrix& B, const round_options& opts,
                              round_result* info = nullptr)
{
  if (opts.method == round_method::svd_streaming)
    tta_die("matmat_round", "round_method::svd_streaming is matvec-only");
  detail::dev_handle a(detail::upload(A)), b(detail::upload(B));
  detail::dev_handle w(detail::upload_warm(opts.method, opts.warm_start_tt_matrix));
  tta_tt* out = nullptr;
  int it = 0, cv = 0;
  double ch = 0.0;
  detail::check(tta_matmat_round_iterative(detail::device_ctx(), a.t, b.t, opts.eps, opts.max_rank,
                                           detail::method_code(opts.method), opts.max_iters, opts.tol, opts.seed,
                                           w.t, &out, &it, &ch, &cv),
                "matmat_round");
  detail::dev_handle o(out);
  detail::write_result(info, it, ch, cv);
  return detail::download_ttm(out);
}

inline double inner(const tt& a, const tt& b)
{
  if (a.d() == 0) return 0.0;
  detail::dev_handle x(detail::upload(a)), y(detail::upload(b));
  double v = 0.0;
  detail::check(tta_inner(detail::device_ctx(), x.t, y.t, &v), "inner");
  return v;
}

inline double norm(const tt& a)
{
  const double s = inner(a, a);
  return std::sqrt(std::max(s, 0.0));
}

inline double frob_inner(const tt_matrix& a, const tt_matrix& b)
{
  if (a.d() == 0) return 0.0;
  detail::dev_handle x(detail::upload(a)), y(detail::upload(b));
  double v = 0.0;
  detail::check(tta_inner(detail::device_ctx(), x.t, y.t, &v), "frob_inner");
  return v;
}

inline double frob_norm(const tt_matrix& a)
{
  const double s = frob_inner(a, a);
  return std::sqrt(std::max(s, 0.0));
}

// ||A x||_F and ||A B||_F without the product tensor
// (algebra/tt_matrix_frob_norm_apply.hpp:36-94); d = 0 gives 0.
inline double frob_norm_apply(const tt_matrix& a, const tt& x)
{
  if (a.d() == 0) return 0.0;
  detail::dev_handle p(detail::upload(a)), q(detail::upload(x));
  double v = 0.0;
  detail::check(tta_frob_norm_apply(detail::device_ctx(), p.t, q.t, &v), "frob_norm_apply(matvec)");
  return v;
}

inline double frob_norm_apply(const tt_matrix& a, const tt_matrix& b)
{
  if (a.d() == 0) return 0.0;
  detail::dev_handle p(detail::upload(a)), q(detail::upload(b));
  double v = 0.0;
  detail::check(tta_frob_norm_apply(detail::device_ctx(), p.t, q.t, &v), "frob_norm_apply(matmat)");
  return v;
}

// ---------------------------------------------------------------------------
// Input algebra and factories (round inputs).
// ---------------------------------------------------------------------------
inline tt axpby(double alpha, const tt& a, double beta, const tt& b)
{
  if (a.d() == 0) return tt();
  detail::dev_handle x(detail::upload(a)), y(detail::upload(b));
  tta_tt* out = nullptr;
  detail::check(tta_axpby(detail::device_ctx(), alpha, x.t, beta, y.t, &out), "axpby");
  detail::dev_handle o(out);
  return detail::download_tt(out);
}

inline tt add(const tt& a, const tt& b) { return axpby(1.0, a, 1.0, b); }
inline tt sub(const tt& a, const tt& b) { return axpby(1.0, a, -1.0, b); }

inline tt scale(const tt& a, double alpha)
{
  std::vector<tt_core> cores(a.cores().begin(), a.cores().end());
  if (!cores.empty())
    for (size_t i = 0; i < cores[0].size(); ++i) cores[0].data()[i] *= alpha;
  return tt(std::move(cores));
}

inline tt random(const std::vector<int>& shape, int max_rank, std::uint64_t seed)
{
  const int d = (int)shape.size();
  std::vector<int> ranks((size_t)d + 1);
  long long n = 0;
  if (tta_random_host(d, shape.data(), max_rank, seed, ranks.data(), nullptr, &n) != TTA_OK)
    tta_die("random", "invalid shape (mode sizes must be positive, ranks must fit in int)");
  std::vector<double> flat((size_t)std::max<long long>(n, 1));
  tta_random_host(d, shape.data(), max_rank, seed, nullptr, flat.data(), nullptr);
  std::vector<tt_core> cores;
  size_t off = 0;
  for (int k = 0; k < d; ++k)
  {
    tt_core c(ranks[k], shape[k], ranks[k + 1]);
    std::copy(flat.begin() + (long)off, flat.begin() + (long)(off + c.size()), c.data());
    off += c.size();
    cores.push_back(std::move(c));
  }
  return tt(std::move(cores));
}

inline tt_matrix random(const std::vector<int>& row_shape, const std::vector<int>& col_shape, int max_rank,
                        std::uint64_t seed)
{
  const int d = (int)row_shape.size();
  if ((int)col_shape.size() != d) tta_die("random", "row/col shape length mismatch");
  std::vector<int> N((size_t)d), ranks((size_t)d + 1);
  for (int k = 0; k < d; ++k) N[k] = row_shape[k] * col_shape[k];
  long long n = 0;
  if (tta_random_host(d, N.data(), max_rank, seed, ranks.data(), nullptr, &n) != TTA_OK)
    tta_die("random", "invalid shape (mode sizes must be positive, ranks must fit in int)");
  std::vector<double> flat((size_t)std::max<long long>(n, 1));
  tta_random_host(d, N.data(), max_rank, seed, nullptr, flat.data(), nullptr);
  std::vector<tt_matrix_core> cores;
  size_t off = 0;
  for (int k = 0; k < d; ++k)
  {
    tt_matrix_core c(ranks[k], row_shape[k], col_shape[k], ranks[k + 1]);
    std::copy(flat.begin() + (long)off, flat.begin() + (long)(off + c.size()), c.data());
    off += c.size();
    cores.push_back(std::move(c));
  }
  return tt_matrix(std::move(cores));
}

inline tt zeros(const std::vector<int>& shape)
{
  std::vector<tt_core> cores;
  for (int n : shape)
  {
    if (n <= 0) tta_die("zeros", "mode size is not positive");
    cores.emplace_back(1, n, 1);
  }
  return tt(std::move(cores));
}

inline tt ones(const std::vector<int>& shape)
{
  std::vector<tt_core> cores;
  for (int n : shape)
  {
    if (n <= 0) tta_die("ones", "mode size is not positive");
    tt_core c(1, n, 1);
    for (int i = 0; i < n; ++i) c(0, i, 0) = 1.0;
    cores.push_back(std::move(c));
  }
  return tt(std::move(cores));
}

// ---------------------------------------------------------------------------
// Dense to TT (include/core/factory/from_dense.hpp:27-40, 215-278).
// ---------------------------------------------------------------------------
enum class from_dense_method
{
  svd,  // TT-SVD (Oseledets' algorithm)
  qtt   // quantized TT: modes base^b split into b modes of size base
};

struct from_dense_options
{
  double eps = 1.0e-10;
  int max_rank = 0;
  from_dense_method method = from_dense_method::svd;
  int qtt_base = 2;
  bool validate = true;  // the QTT shape is always checked here
};

inline tt from_dense(const double* data, const std::vector<int>& shape, const from_dense_options& opts)
{
  tta_ctx* c = detail::device_ctx();
  tta_tt* t = nullptr;
  detail::check(tta_from_dense(c, data, (int)shape.size(), shape.data(), opts.eps, opts.max_rank,
                               opts.method == from_dense_method::qtt ? 1 : 0, opts.qtt_base, &t),
                "from_dense");
  tt out = detail::download_tt(t);
  tta_tt_free(c, t);
  return out;
}

inline tt from_dense(const double* data, const std::vector<int>& shape, double eps)
{
  from_dense_options o;
  o.eps = eps;
  return from_dense(data, shape, o);
}

inline tt_matrix from_dense(const double* data, const std::vector<int>& row_shape, const std::vector<int>& col_shape,
                            const from_dense_options& opts)
{
  if (opts.method == from_dense_method::qtt) tta_die("from_dense(tt_matrix, qtt)", "not implemented");
  if (row_shape.size() != col_shape.size()) tta_die("matrix_from_dense", "row/col shape length mismatch");
  tta_ctx* c = detail::device_ctx();
  tta_tt* t = nullptr;
  detail::check(tta_matrix_from_dense(c, data, (int)row_shape.size(), row_shape.data(), col_shape.data(), opts.eps,
                                      opts.max_rank, &t),
                "matrix_from_dense");
  tt_matrix out = detail::download_ttm(t);
  tta_tt_free(c, t);
  return out;
}

inline tt_matrix from_dense(const double* data, const std::vector<int>& row_shape, const std::vector<int>& col_shape,
                            double eps)
{
  from_dense_options o;
  o.eps = eps;
  return from_dense(data, row_shape, col_shape, o);
}

}  // namespace tensor_train
}  // namespace mva
