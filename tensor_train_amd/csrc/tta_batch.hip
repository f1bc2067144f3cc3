// Batched plain FP64 rounding of many TTs of uneven d, mode sizes and
// ranks: the reference algorithm (detail::round, include/core/round/
// tt_round.hpp:36-127) applied to every TT, with the TTs advanced in
// lockstep by core position so each kernel launch covers the whole batch:
//   * right sweep step j: core d_t-1-j of every TT with d_t-1-j >= 1
//     (Householder QR batch, L extraction, grouped GEMM absorb).  The ranks
//     of this sweep are structural (r_eff = min(n r', r)), so the host
//     knows every shape.
//   * SVD sweep step k: core k of every TT with k < d_t-1 (QR-condense,
//     batched block Jacobi, truncation rank decided on the device, one read
//     of the new ranks per step, then gathers and grouped GEMMs).
//   * norm balancing of every TT (tt_round.hpp:76-124) in three launches.
// Per-problem descriptors are built on the host and uploaded once per
// launch group.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

namespace tta
{
namespace
{

using scratch = arena;  // per-step device buffers

double* new_block(tta_ctx* c, size_t n) { return static_cast<double*>(c->mem.get(std::max<size_t>(n, 1) * 8)); }

void sync(tta_ctx* c) { ctx_sync(c); }

}  // namespace

// Rounds `count` working TTs in place (plain FP64, gauge none).
void round_batched_plain(tta_ctx* c, std::vector<tta_tt*>& ts, double eps, int max_rank)
{
  const int count = (int)ts.size();
  int maxd = 0;
  for (auto* t : ts) maxd = std::max(maxd, t->d());
  if (maxd <= 1) return;

  // ------------------------------------------------------------ right sweep
  for (int j = 0; j <= maxd - 2; ++j)
  {
    scratch S(c);
    std::vector<int> idx;
    std::vector<qrp> qr;
    std::vector<ex_desc> ex;
    std::vector<gemm_desc> gm;
    int max_m = 0, max_n = 0, max_gm = 0, max_gn = 0;
    long max_ex = 0;
    std::vector<double*> newQ, newP;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      const int k = t->d() - 1 - j;
      if (k < 1) continue;
      const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      const int m = n * rp, r_eff = std::min(m, r);
      const int kk = std::min(m, r);
      idx.push_back(i);
      qrp p;
      p.A = t->cores[k];
      p.lda = m;
      p.m = m;
      p.n = r;
      p.tau = S.alloc<double>((size_t)kk);
      p.T = S.alloc<double>((size_t)((kk + 15) / 16) * 256);
      p.Q = new_block(c, (size_t)r_eff * m);
      p.ldq = m;
      newQ.push_back(p.Q);
      qr.push_back(p);
      max_m = std::max(max_m, m);
      max_n = std::max(max_n, r);
      double* L = S.alloc<double>((size_t)r * r_eff);
      ex_desc e;
      e.A = t->cores[k];
      e.lda = m;
      e.kr = r_eff;
      e.nc = r;
      e.dst = L;
      e.drs = 1;
      e.dcs = r_eff;
      ex.push_back(e);
      max_ex = std::max(max_ex, (long)r_eff * r);
      const int rows_p = t->ranks[k - 1] * t->modes[k - 1];
      gemm_desc g;
      g.M = rows_p;
      g.N = r_eff;
      g.K = r;
      g.A = t->cores[k - 1];
      g.sar = r;
      g.sac = 1;
      g.B = L;
      g.sbr = r_eff;
      g.sbc = 1;
      g.C = new_block(c, (size_t)rows_p * r_eff);
      g.scr = r_eff;
      g.scc = 1;
      newP.push_back(g.C);
      gm.push_back(g);
      max_gm = std::max(max_gm, rows_p);
      max_gn = std::max(max_gn, r_eff);
    }
    const int nb = (int)idx.size();
    if (nb == 0) continue;
    qr_batch qb;
    qb.descs = S.upload(qr);
    qb.batch = nb;
    qb.max_m = max_m;
    qb.max_n = max_n;
    qb.form_q = true;
    {
      const credit w = qr_credit(qr.data(), qr.size());
      prof_scope ps(c, TTA_K_QR, w.flops, w.bytes);
      launch_qr_blocked(qb, c->stream);
    }
    {
      prof_scope ps(c, TTA_K_OTHER);
      launch_extract_upper_batched(S.upload(ex), nb, max_ex, c->stream);
    }
    {
      const credit w = gemm_credit(gm.data(), gm.size());
      prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
      launch_gemm_grouped(S.upload(gm), nb, max_gm, max_gn, c->stream);
    }
    for (int a = 0; a < nb; ++a)
    {
      tta_tt* t = ts[idx[a]];
      const int k = t->d() - 1 - j;
      c->mem.put(t->cores[k]);
      c->mem.put(t->cores[k - 1]);
      t->cores[k] = newQ[a];
      t->cores[k - 1] = newP[a];
      t->ranks[k] = std::min(t->modes[k] * t->ranks[k + 1], t->ranks[k]);
    }
    // The step's scratch returns to the stream-ordered pool: no wait needed.
  }

  // ------------------------------------------------------------ norms, delta
  scratch G(c);
  double* norms = G.alloc<double>((size_t)count);
  double* deltas = G.alloc<double>((size_t)count);
  int* rk = G.alloc<int>((size_t)count);
  int* flags = G.alloc<int>((size_t)count * 4);
  std::vector<char> live(count, 0);
  {
    std::vector<sq_desc> sq;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      sq_desc s;
      if (t->d() >= 2)
      {
        s.x = t->cores[0];
        s.n = (long)t->ranks[0] * t->modes[0] * t->ranks[1];
      }
      else
      {
        s.x = nullptr;
        s.n = 0;
      }
      s.out_norm = norms + i;
      s.out_delta = deltas + i;
      s.scale = eps / std::sqrt((double)std::max(t->d() - 1, 1));
      sq.push_back(s);
    }
    launch_frob_batched(G.upload(sq), count, c->stream);
    std::vector<double> h(count);
    hip_check(hipMemcpyAsync(h.data(), norms, count * 8, hipMemcpyDeviceToHost, c->stream), "norms");
    sync(c);
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      if (t->d() < 2) continue;
      if (h[i] == 0.0)
      {
        // make_rank1_zero_tt (include/core/detail/shape_check.hpp:176-188).
        for (int k = 0; k < t->d(); ++k)
        {
          double* p = new_block(c, (size_t)t->modes[k]);
          hip_check(hipMemsetAsync(p, 0, (size_t)t->modes[k] * 8, c->stream), "zero");
          c->mem.put(t->cores[k]);
          t->cores[k] = p;
        }
        std::fill(t->ranks.begin(), t->ranks.end(), 1);
        continue;
      }
      live[i] = 1;
    }
  }

  // ------------------------------------------------------------ SVD sweep
  for (int k = 0; k <= maxd - 2; ++k)
  {
    scratch S(c);
    std::vector<int> idx;
    std::vector<char> tall;
    std::vector<tr_desc> tr;
    std::vector<cp_desc> cp;
    std::vector<qrp> qr;
    std::vector<ex_desc> ex;
    std::vector<jac_prob> jp;
    std::vector<double*> Qs, Xs, Js, Ss;
    std::vector<double*> Fs;
    std::vector<int*> perms;
    int max_m = 0, max_n = 0, max_q = 0, max_tr_r = 0, max_tr_c = 0;
    long max_cp = 0, max_ex = 0;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      if (!live[i] || k >= t->d() - 1) continue;
      const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      const int rows = r * n, cols = rp;
      const bool is_tall = rows >= cols;
      const int q = is_tall ? cols : rows;
      const int m = is_tall ? rows : cols;
      idx.push_back(i);
      tall.push_back(is_tall ? 1 : 0);
      double* A = S.alloc<double>((size_t)rows * cols);
      if (is_tall)
      {
        tr_desc d;
        d.in = t->cores[k];
        d.ldi = cols;
        d.rows = rows;
        d.cols = cols;
        d.out = A;
        d.ldo = rows;
        tr.push_back(d);
        max_tr_r = std::max(max_tr_r, rows);
        max_tr_c = std::max(max_tr_c, cols);
      }
      else
      {
        cp_desc d;
        d.src = t->cores[k];
        d.dst = A;
        d.n = (long)rows * cols;
        cp.push_back(d);
        max_cp = std::max(max_cp, d.n);
      }
      qrp p;
      p.A = A;
      p.lda = m;
      p.m = m;
      p.n = q;
      p.tau = S.alloc<double>((size_t)q);
      p.T = S.alloc<double>((size_t)((q + 15) / 16) * 256);
      p.Q = S.alloc<double>((size_t)m * q);
      p.ldq = m;
      qr.push_back(p);
      Qs.push_back(p.Q);
      max_m = std::max(max_m, m);
      max_n = std::max(max_n, q);
      double* X = S.alloc<double>((size_t)q * q);
      ex_desc e;
      e.A = A;
      e.lda = m;
      e.kr = q;
      e.nc = q;
      e.dst = X;  // X = R^T in both orientations (see svd_left)
      e.drs = q;
      e.dcs = 1;
      ex.push_back(e);
      max_ex = std::max(max_ex, (long)q * q);
      Xs.push_back(X);
      jac_prob j;
      j.X = X;
      j.ldx = q;
      j.q = q;
      j.J = S.alloc<double>((size_t)q * q);
      j.ldj = q;
      j.s = S.alloc<double>((size_t)q);
      Ss.push_back(j.s);
      j.perm = S.alloc<int>((size_t)q);
      j.rank_out = rk + i;
      j.delta_ptr = deltas + i;
      j.max_rank = max_rank;
      j.flag = flags + 4 * i;
      j.sweeps = flags + 4 * i + 2;
      jp.push_back(j);
      Js.push_back(j.J);
      perms.push_back(j.perm);
      max_q = std::max(max_q, q);
    }
    const int nb = (int)idx.size();
    if (nb == 0) continue;
    if (max_q > 640) fail(TTA_ERR_UNSUPPORTED, "round_batched: bond rank %d above 640", max_q);
    {
      prof_scope ps(c, TTA_K_OTHER);
      if (!tr.empty()) launch_transpose_batched(S.upload(tr), (int)tr.size(), max_tr_r, max_tr_c, c->stream);
      if (!cp.empty()) launch_copy_batched(S.upload(cp), (int)cp.size(), max_cp, c->stream);
    }
    qr_batch qb;
    qb.descs = S.upload(qr);
    qb.batch = nb;
    qb.max_m = max_m;
    qb.max_n = max_n;
    qb.form_q = true;
    {
      const credit w = qr_credit(qr.data(), qr.size());
      prof_scope ps(c, TTA_K_QR, w.flops, w.bytes);
      launch_qr_blocked(qb, c->stream);
    }
    {
      prof_scope ps(c, TTA_K_OTHER);
      launch_extract_upper_batched(S.upload(ex), nb, max_ex, c->stream);
    }
    jac_batch jb;
    jb.descs = S.upload(jp);
    jb.batch = nb;
    int* done = S.alloc<int>(1);
    {
      const credit w = jacobi_credit(jp.data(), jp.size());
      prof_scope ps(c, TTA_K_JACOBI, w.flops, w.bytes);
      const int sw = launch_jacobi_svd(jb, max_q, done, c->h_int + 2, c, w.flops, w.bytes);
      if (sw > 0)
      {
        c->jac_calls += 1;
        c->jac_sweeps += sw;
      }
    }
    std::vector<int> hr(count);
    hip_check(hipMemcpyAsync(hr.data(), rk, count * sizeof(int), hipMemcpyDeviceToHost, c->stream), "ranks");
    sync(c);

    // Phase 2: U (new core k), F = diag(s) Vt, next core.
    std::vector<ga_desc> g1, g2;
    std::vector<gemm_desc> m1, m2;
    std::vector<double*> Us, Nn;
    long max_g1 = 0, max_g2 = 0;
    int m1_m = 0, m1_n = 0, m2_m = 0, m2_n = 0;
    for (int a = 0; a < nb; ++a)
    {
      tta_tt* t = ts[idx[a]];
      const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      const int rows = r * n, cols = rp;
      const int q = tall[a] ? cols : rows;
      const int r_new = hr[idx[a]];
      double* U = new_block(c, (size_t)rows * r_new);
      double* F = S.alloc<double>((size_t)r_new * cols);
      Us.push_back(U);
      Fs.push_back(F);
      if (tall[a])
      {
        double* Jp = S.alloc<double>((size_t)q * r_new);
        ga_desc d1{Js[a], q, q, perms[a], r_new, Jp, 1, q};
        g1.push_back(d1);
        ga_desc d2{Xs[a], q, q, perms[a], r_new, F, 1, q};
        g2.push_back(d2);
        gemm_desc g;
        g.M = rows;
        g.N = r_new;
        g.K = q;
        g.A = Qs[a];
        g.sar = 1;
        g.sac = rows;
        g.B = Jp;
        g.sbr = 1;
        g.sbc = q;
        g.C = U;
        g.scr = r_new;
        g.scc = 1;
        m1.push_back(g);
        m1_m = std::max(m1_m, rows);
        m1_n = std::max(m1_n, r_new);
      }
      else
      {
        // M = W (Q J)^T: U = W diag(s)^-1, F = diag(s) (Q J)^T.
        ga_desc d1{Xs[a], q, q, perms[a], r_new, U, r_new, 1, Ss[a], GATHER_DIV};
        g1.push_back(d1);
        double* Wp = S.alloc<double>((size_t)q * r_new);
        ga_desc d2{Js[a], q, q, perms[a], r_new, Wp, 1, q, Ss[a], GATHER_MUL};
        g2.push_back(d2);
        gemm_desc g;
        g.M = r_new;
        g.N = cols;
        g.K = q;
        g.A = Wp;
        g.sar = q;
        g.sac = 1;
        g.B = Qs[a];
        g.sbr = cols;
        g.sbc = 1;
        g.C = F;
        g.scr = cols;
        g.scc = 1;
        m1.push_back(g);
        m1_m = std::max(m1_m, r_new);
        m1_n = std::max(m1_n, cols);
      }
      max_g1 = std::max(max_g1, (long)q * r_new);
      max_g2 = std::max(max_g2, (long)q * r_new);
      const int ncols_n = t->modes[k + 1] * t->ranks[k + 2];
      gemm_desc g;
      g.M = r_new;
      g.N = ncols_n;
      g.K = cols;
      g.A = F;
      g.sar = cols;
      g.sac = 1;
      g.B = t->cores[k + 1];
      g.sbr = ncols_n;
      g.sbc = 1;
      g.C = new_block(c, (size_t)r_new * ncols_n);
      g.scr = ncols_n;
      g.scc = 1;
      Nn.push_back(g.C);
      m2.push_back(g);
      m2_m = std::max(m2_m, r_new);
      m2_n = std::max(m2_n, ncols_n);
    }
    {
      prof_scope ps(c, TTA_K_OTHER);
      launch_gather_cols_batched(S.upload(g1), nb, max_g1, c->stream);
      launch_gather_cols_batched(S.upload(g2), nb, max_g2, c->stream);
    }
    {
      credit w = gemm_credit(m1.data(), m1.size());
      const credit w2 = gemm_credit(m2.data(), m2.size());
      w.flops += w2.flops;
      w.bytes += w2.bytes;
      prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
      launch_gemm_grouped(S.upload(m1), nb, m1_m, m1_n, c->stream);
      launch_gemm_grouped(S.upload(m2), nb, m2_m, m2_n, c->stream);
    }
    for (int a = 0; a < nb; ++a)
    {
      tta_tt* t = ts[idx[a]];
      c->mem.put(t->cores[k]);
      c->mem.put(t->cores[k + 1]);
      t->cores[k] = Us[a];
      t->cores[k + 1] = Nn[a];
      t->ranks[k + 1] = hr[idx[a]];
    }
  }

  // ------------------------------------------------------------ balancing
  {
    scratch S(c);
    std::vector<cn_desc> cn;
    std::vector<gm_desc> gmd;
    std::vector<bal_desc> bal;
    size_t total = 0;
    std::vector<size_t> base(count, 0);
    for (int i = 0; i < count; ++i)
    {
      base[i] = total;
      if (!live[i]) continue;
      for (int k = 0; k < ts[i]->d() - 1; ++k) total += (size_t)ts[i]->ranks[k + 1];
    }
    double* nrm = S.alloc<double>(total + 1);
    double* geo = S.alloc<double>((size_t)count);
    int max_cols = 0;
    long max_el = 0;
    for (int i = 0; i < count; ++i)
    {
      if (!live[i]) continue;
      tta_tt* t = ts[i];
      size_t off = base[i];
      std::vector<size_t> offs(t->d(), 0);
      for (int k = 0; k < t->d() - 1; ++k)
      {
        offs[k] = off;
        cn.push_back(cn_desc{t->cores[k], t->ranks[k] * t->modes[k], t->ranks[k + 1], nrm + off});
        max_cols = std::max(max_cols, t->ranks[k + 1]);
        off += (size_t)t->ranks[k + 1];
      }
      gmd.push_back(gm_desc{nrm + base[i], (int)(off - base[i]), geo + i});
      for (int k = 0; k < t->d(); ++k)
      {
        bal_desc b;
        b.x = t->cores[k];
        b.r = t->ranks[k];
        b.n = t->modes[k];
        b.rp = t->ranks[k + 1];
        b.nl = k > 0 ? nrm + offs[k - 1] : nullptr;
        b.nr = k < t->d() - 1 ? nrm + offs[k] : nullptr;
        b.geo = geo + i;
        bal.push_back(b);
        max_el = std::max(max_el, (long)b.r * b.n * b.rp);
      }
    }
    if (!cn.empty())
    {
      launch_col_norms_batched(S.upload(cn), (int)cn.size(), max_cols, c->stream);
      launch_geo_mean_batched(S.upload(gmd), (int)gmd.size(), c->stream);
      launch_balance_batched(S.upload(bal), (int)bal.size(), max_el, c->stream);
    }
    sync(c);
  }
}

}  // namespace tta

// ----------------------------------------------------------------------------
// C ABI.
// ----------------------------------------------------------------------------
extern "C" int tta_round_batched(tta_ctx* c, int count, tta_tt* const* in, double eps, int max_rank, int precision,
                                 tta_tt** out)
{
  using namespace tta;
  if (!c || count < 0 || (count > 0 && (!in || !out))) return TTA_ERR_ARG;
  try
  {
    if (precision != 0) fail(TTA_ERR_UNSUPPORTED, "round_batched: precision %d not available", precision);
    if (std::isnan(eps)) fail(TTA_ERR_ARG, "round_batched: eps is NaN");
    for (int i = 0; i < count; ++i)
      if (!in[i]) fail(TTA_ERR_ARG, "round_batched: null tensor %d", i);
    // Working copies (one batched copy launch).
    std::vector<tta_tt*> ts(count, nullptr);
    std::vector<cp_desc> cp;
    long max_el = 0;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = new tta_tt();
      t->modes = in[i]->modes;
      t->ranks = in[i]->ranks;
      t->row_modes = in[i]->row_modes;
      t->col_modes = in[i]->col_modes;
      t->matrix = in[i]->matrix;
      for (int k = 0; k < in[i]->d(); ++k)
      {
        const long n = (long)in[i]->ranks[k] * in[i]->modes[k] * in[i]->ranks[k + 1];
        double* p = static_cast<double*>(c->mem.get((size_t)std::max(n, 1L) * 8));
        t->cores.push_back(p);
        cp.push_back(cp_desc{in[i]->cores[k], p, n});
        max_el = std::max(max_el, n);
      }
      ts[i] = t;
    }
    {
      void* d = c->mem.get(std::max<size_t>(cp.size(), 1) * sizeof(cp_desc));
      if (!cp.empty())
      {
        hip_check(hipMemcpyAsync(d, cp.data(), cp.size() * sizeof(cp_desc), hipMemcpyHostToDevice, c->stream),
                  "descs");
        launch_copy_batched(static_cast<cp_desc*>(d), (int)cp.size(), max_el, c->stream);
      }
      hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
      c->mem.put(d);
    }
    try
    {
      round_batched_plain(c, ts, eps, max_rank);
    }
    catch (...)
    {
      for (auto* t : ts)
      {
        for (double* p : t->cores) c->mem.put(p);
        delete t;
      }
      throw;
    }
    for (int i = 0; i < count; ++i) out[i] = ts[i];
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
  catch (const std::exception& e)
  {
    c->last_error = e.what();
    return TTA_ERR_HIP;
  }
  return TTA_OK;
}
