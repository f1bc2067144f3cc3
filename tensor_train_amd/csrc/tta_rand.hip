// Batched randomized TT rounding with given target ranks
// ("Randomize-then-Orthogonalize", Al Daas, Ballard, Cazeaux, Hallman,
// Miedlar, Pasha, Reid, Saibaba, "Randomized algorithms for rounding in the
// tensor-train format", SIAM J. Sci. Comput. 45 (2023)).  No reference
// counterpart: the reference's rounding is the deterministic QR + SVD sweep.
//
// For each TT y (ranks r) and targets l_1..l_{d-1}:
//   1. Gaussian sketch TT R with cores (l_{k-1}, n_k, l_k), l_0 = l_d = 1.
//   2. Right-to-left partial contractions W_k = <y_{k..d}, R_{k..d}>
//      (r_k x l_k) by two GEMMs per core.
//   3. Left-to-right: S = Z_k W_{k+1} ((l' n) x l), Householder QR S = Q R,
//      core k <- Q, M = Q^T Z_k, Z_{k+1} = M Y_{k+1}; the last core is Z_d.
//   The result yhat is the orthogonal projection of y onto the span of the
//   left-orthonormal frame, so ||y - yhat||^2 = ||y||^2 - ||yhat||^2 exactly
//   in exact arithmetic; ||y||^2 comes from a left-to-right Gram environment
//   and ||yhat|| = ||Z_d||.  The caller accepts yhat when that discarded norm
//   is <= eps ||y|| and otherwise falls back to the plain rounding.
// All TTs advance in lockstep by core position, as in the plain batched path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <unordered_set>
#include <vector>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

namespace tta
{
void launch_gauss_batched(void* const* dst, const int* m, const int* nc, const unsigned long long* seed, const int* k,
                          int count, bool f32, tta_ctx* c);
void round_batched_plain(tta_ctx* c, std::vector<tta_tt*>& ts, double eps, int max_rank, bool mixed);

namespace
{
gemm_desc rm_gemm(int M, int N, int K, const double* A, long lda, const double* B, long ldb, double* C, long ldc)
{
  gemm_desc g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sar = lda;
  g.sac = 1;
  g.B = B;
  g.sbr = ldb;
  g.sbc = 1;
  g.C = C;
  g.scr = ldc;
  g.scc = 1;
  return g;
}

void run_xgrouped(tta_ctx* c, arena& O, const std::vector<xgemm_desc>& gs, xgemm_types t)
{
  if (gs.empty()) return;
  credit w;
  for (auto& g : gs)
  {
    w.flops += 2.0 * g.M * g.N * (double)g.K;
    w.bytes += 8.0 * ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N);
  }
  prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
  launch_xgemm_grouped(O, gs, t, c->stream);
}

void run_grouped(tta_ctx* c, arena& O, const std::vector<gemm_desc>& gs)
{
  if (gs.empty()) return;
  const credit w = gemm_credit(gs.data(), gs.size());
  prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
  launch_gemm_grouped(O, gs, c->stream);
}
}  // namespace

// Randomized rounding of the working TTs `ts` (replaced in place by yhat).
// ell[i] holds d_i + 1 target ranks (ell[i][0] = ell[i][d] = 1).  Cores in
// `borrowed` belong to the caller's inputs: they are only read, and every
// core of a TT with d >= 2 is replaced by a new one.  On return
// rel_err[i] = ||y - yhat|| / ||y|| (from the norm identity; 0 for a zero y).
void round_batched_randomized(tta_ctx* c, std::vector<tta_tt*>& ts, const std::vector<std::vector<int>>& ell,
                              unsigned long long seed, const std::unordered_set<const double*>& borrowed,
                              std::vector<double>& rel_err)
{
  auto release_core = [&](double* p) {
    if (!borrowed.count(p)) c->mem.put(p);
  };
  const int count = (int)ts.size();
  rel_err.assign(count, 0.0);
  int maxd = 0;
  for (auto* t : ts) maxd = std::max(maxd, t->d());
  if (count == 0) return;
  // Device buffers by lifetime: sketch cores (phases 1-2), W_k (phases 2-4),
  // Gram environments (two live steps at a time), per-step scratch.
  auto Rar = std::make_unique<arena>(c);
  arena War(c);

  // 1. Sketch cores (l_{k-1} n_k x l_k row-major), FP32: the sketch only
  // selects a subspace, and the projection onto it (phases 3-4) is FP64, so
  // sketch rounding perturbs which near-optimal subspace is found, never the
  // accuracy of the result or of its discarded-norm check.
  std::vector<std::vector<float*>> R(count);
  {
    std::vector<void*> dst;
    std::vector<int> m, nc, kk;
    std::vector<unsigned long long> sd;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      for (int k = 0; k < t->d(); ++k)
      {
        const int rows = ell[i][k] * t->modes[k], cols = ell[i][k + 1];
        float* p = Rar->alloc<float>((size_t)rows * cols);
        R[i].push_back(p);
        dst.push_back(p);
        m.push_back(rows);
        nc.push_back(cols);
        kk.push_back(k);
        sd.push_back(seed + 0x9E3779B97F4A7C15ULL * (unsigned long long)(i + 1));
      }
    }
    prof_scope ps(c, TTA_K_OTHER);
    launch_gauss_batched(dst.data(), m.data(), nc.data(), sd.data(), kk.data(), (int)dst.size(), true, c);
  }

  // 2. W_k (r_k x l_k) from the right, FP32 on the matrix cores (the input
  // cores are read in FP64 and rounded on the way in); W_d = [1] implicit.
  std::vector<std::vector<float*>> W(count);
  for (int i = 0; i < count; ++i) W[i].assign(ts[i]->d() + 1, nullptr);
  for (int j = 0; j <= maxd - 1; ++j)
  {
    arena S(c);
    std::vector<xgemm_desc> g1, g2d, g2s;  // T1 = Y W'; W = T1 R^T (T1 = Y or FP32)
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      const int k = t->d() - 1 - j;
      if (k < 1) continue;
      const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      const int lp = ell[i][k + 1], l = ell[i][k];
      const void* T1 = t->cores[k];  // k = d-1: Y_{d-1} W_d = Y_{d-1}
      const bool t1_f32 = k < t->d() - 1;
      if (t1_f32)
      {
        float* tmp = S.alloc<float>((size_t)r * n * lp);
        g1.push_back(xgemm_desc{r * n, lp, rp, t->cores[k], rp, 1, W[i][k + 1], lp, 1, tmp, lp, 1});
        T1 = tmp;
      }
      float* Wk = War.alloc<float>((size_t)r * l);
      W[i][k] = Wk;
      const xgemm_desc g{r, l, n * lp, T1, (long)n * lp, 1, R[i][k], 1, (long)n * lp, Wk, l, 1};
      (t1_f32 ? g2s : g2d).push_back(g);
    }
    run_xgrouped(c, S, g1, XG_D_S_S_S);
    run_xgrouped(c, S, g2d, XG_D_S_S_S);
    run_xgrouped(c, S, g2s, XG_S_S_S_S);
  }
  Rar.reset();  // sketch cores are consumed (stream-ordered release)

  // 3. ||y||^2 by the left-to-right Gram environment (two GEMMs per core).
  // E starts as [1] for every TT; each step's environments live in their own
  // arena, released once the next step has been queued.
  std::vector<double*> E(count, nullptr);
  auto Ecur = std::make_unique<arena>(c, size_t(16) << 20);
  {
    double* e0 = Ecur->upload(std::vector<double>((size_t)count, 1.0));
    for (int i = 0; i < count; ++i) E[i] = e0 + i;
  }
  for (int k = 0; k < maxd; ++k)
  {
    arena S(c);
    auto Enext = std::make_unique<arena>(c);
    std::vector<gemm_desc> g1, g2;
    std::vector<std::pair<int, double*>> next;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      if (k >= t->d()) continue;
      const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      double* tmp = S.alloc<double>((size_t)r * n * rp);
      gemm_desc a;  // tmp (r x n rp) = E^T (r x r) * right_unfold(Y_k)
      a.M = r;
      a.N = n * rp;
      a.K = r;
      a.A = E[i];
      a.sar = 1;
      a.sac = r;
      a.B = t->cores[k];
      a.sbr = (long)n * rp;
      a.sbc = 1;
      a.C = tmp;
      a.scr = (long)n * rp;
      a.scc = 1;
      g1.push_back(a);
      double* En = Enext->alloc<double>((size_t)rp * rp);
      gemm_desc b;  // E' (rp x rp) = tmp_as^T (rp x r n) * left_unfold(Y_k)
      b.M = rp;
      b.N = rp;
      b.K = r * n;
      b.A = tmp;
      b.sar = 1;
      b.sac = rp;
      b.B = t->cores[k];
      b.sbr = rp;
      b.sbc = 1;
      b.C = En;
      b.scr = rp;
      b.scc = 1;
      g2.push_back(b);
      next.push_back({i, En});
    }
    run_grouped(c, S, g1);
    run_grouped(c, S, g2);
    // TTs shorter than k keep their final environment: carry it over.
    std::vector<cp_desc> carry;
    std::vector<char> moved(count, 0);
    for (auto& pr : next) moved[pr.first] = 1;
    for (int i = 0; i < count; ++i)
      if (!moved[i])
      {
        double* keep = Enext->alloc<double>(1);
        carry.push_back(cp_desc{E[i], keep, 1});
        E[i] = keep;
      }
    if (!carry.empty()) launch_copy_batched(S.upload(carry), (int)carry.size(), 1, c->stream);
    for (auto& pr : next) E[pr.first] = pr.second;
    Ecur = std::move(Enext);
  }
  // ||y||^2 = the final 1 x 1 environment of each TT, gathered in one launch.
  std::vector<double> ny2(count);
  arena G(c, size_t(1) << 20);
  {
    double* n2 = G.alloc<double>((size_t)count);
    std::vector<cp_desc> cp;
    for (int i = 0; i < count; ++i) cp.push_back(cp_desc{E[i], n2 + i, 1});
    launch_copy_batched(G.upload(cp), count, 1, c->stream);
    hip_check(hipMemcpyAsync(ny2.data(), n2, (size_t)count * 8, hipMemcpyDeviceToHost, c->stream), "norms");
  }
  Ecur.reset();

  // 4. Left-to-right QR sweep of the sketched cores.
  std::vector<double*> Z(count, nullptr);
  std::vector<int> zl(count, 1);  // left rank of Z
  for (int i = 0; i < count; ++i) Z[i] = ts[i]->cores[0];
  // Z_{k+1} lives in the next step's arena, except at a TT's last step, where
  // it becomes the TT's last core (a pool block of its own).
  auto Zcur = std::make_unique<arena>(c);
  for (int k = 0; k <= maxd - 2; ++k)
  {
    arena S(c);
    auto Znext = std::make_unique<arena>(c);
    std::vector<int> idx;
    std::vector<xgemm_desc> gS;
    std::vector<gemm_desc> gM, gZ;
    std::vector<qrp> qr;
    std::vector<tr_desc> tr;
    std::vector<double*> newcore, newZ;
    std::vector<int> newl;
    int max_m = 0, max_n = 0, max_tr_r = 0, max_tr_c = 0;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      if (k >= t->d() - 1) continue;
      const int lq = zl[i], n = t->modes[k], rp = t->ranks[k + 1];
      const int m = lq * n, lt = ell[i][k + 1];
      const int l2 = std::min(m, lt);
      idx.push_back(i);
      // S (m x lt) column-major = Z (m x rp, FP64) * W_{k+1} (rp x lt, FP32), in FP64.
      double* Sm = S.alloc<double>((size_t)m * lt);
      gS.push_back(xgemm_desc{m, lt, rp, Z[i], rp, 1, W[i][k + 1], lt, 1, Sm, 1, m});
      qrp p;
      p.A = Sm;
      p.lda = m;
      p.m = m;
      p.n = lt;
      p.tau = S.alloc<double>((size_t)l2);
      p.T = S.alloc<double>((size_t)((l2 + 15) / 16) * 256);
      p.Q = S.alloc<double>((size_t)m * l2);
      p.ldq = m;
      qr.push_back(p);
      max_m = std::max(max_m, m);
      max_n = std::max(max_n, lt);
      // core k (m x l2 row-major) <- Q (m x l2 column-major).
      double* core = static_cast<double*>(c->mem.get((size_t)std::max(m * l2, 1) * 8));
      newcore.push_back(core);
      tr.push_back(tr_desc{p.Q, m, l2, m, core, l2});
      max_tr_r = std::max(max_tr_r, l2);
      max_tr_c = std::max(max_tr_c, m);
      // M (l2 x rp) = Q^T Z.
      double* Mm = S.alloc<double>((size_t)l2 * rp);
      gemm_desc gm;
      gm.M = l2;
      gm.N = rp;
      gm.K = m;
      gm.A = p.Q;
      gm.sar = m;
      gm.sac = 1;
      gm.B = Z[i];
      gm.sbr = rp;
      gm.sbc = 1;
      gm.C = Mm;
      gm.scr = rp;
      gm.scc = 1;
      gM.push_back(gm);
      // Z_{k+1} (l2 x n' r'') = M * right_unfold(Y_{k+1}).
      const int nc2 = t->modes[k + 1] * t->ranks[k + 2];
      double* Zn = k == t->d() - 2 ? static_cast<double*>(c->mem.get((size_t)std::max(l2 * nc2, 1) * 8))
                                   : Znext->alloc<double>((size_t)l2 * nc2);
      newZ.push_back(Zn);
      gZ.push_back(rm_gemm(l2, nc2, rp, Mm, rp, t->cores[k + 1], nc2, Zn, nc2));
      newl.push_back(l2);
    }
    if (idx.empty()) continue;
    run_xgrouped(c, S, gS, XG_D_S_D_D);
    qr_batch qb;
    qb.descs = S.upload(qr);
    qb.batch = (int)qr.size();
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
      launch_transpose_batched(S.upload(tr), (int)tr.size(), max_tr_r, max_tr_c, c->stream);
    }
    run_grouped(c, S, gM);
    run_grouped(c, S, gZ);
    // stream-ordered: the next step may reuse this step's scratch
    for (size_t a = 0; a < idx.size(); ++a)
    {
      const int i = idx[a];
      tta_tt* t = ts[i];
      release_core(t->cores[k]);
      t->cores[k] = newcore[a];
      t->ranks[k + 1] = newl[a];
      Z[i] = newZ[a];
      zl[i] = newl[a];
    }
    Zcur = std::move(Znext);
  }
  Zcur.reset();
  // Last core <- Z; ||yhat|| = ||Z||.
  {
    arena S(c, size_t(1) << 20);
    std::vector<sq_desc> sq;
    double* nz = S.alloc<double>((size_t)count);
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = ts[i];
      const int k = t->d() - 1;
      if (t->d() >= 2)
      {
        release_core(t->cores[k]);
        t->cores[k] = Z[i];
        t->ranks[k] = zl[i];
      }
      sq.push_back(sq_desc{t->cores[k], (long)t->ranks[k] * t->modes[k] * t->ranks[k + 1], nz + i, nullptr, 0.0});
    }
    launch_frob_batched(S.upload(sq), count, c->stream);
    std::vector<double> h(count);
    hip_check(hipMemcpyAsync(h.data(), nz, (size_t)count * 8, hipMemcpyDeviceToHost, c->stream), "norms");
    ctx_sync(c);
    for (int i = 0; i < count; ++i)
    {
      const double y2 = ny2[i], yh2 = h[i] * h[i];
      rel_err[i] = y2 > 0.0 ? std::sqrt(std::max(y2 - yh2, 0.0) / y2) : 0.0;
    }
  }
}

}  // namespace tta

using namespace tta;

// Batched randomized rounding with per-TT target ranks, deterministic
// fallback for the TTs whose discarded norm exceeds eps (fallback =
// TTA_FALLBACK_FP64: the plain rounding, TTA_FALLBACK_MIXED: the mixed one,
// TTA_FALLBACK_NONE: keep the randomized result).  `ranks`: for TT i, d_i + 1
// entries (first and last ignored), concatenated.  `rel_err` and `fell_back`
// (each `count` entries, may be null) report the discarded norm estimate and
// whether the deterministic rounding replaced the randomized result.
extern "C" int tta_round_randomized_batched(tta_ctx* c, int count, tta_tt* const* in, const int* ranks, double eps,
                                            unsigned long long seed, int fallback, tta_tt** out, double* rel_err,
                                            int* fell_back)
{
  if (!c || count < 0 || (count > 0 && (!in || !ranks || !out))) return TTA_ERR_ARG;
  if (fallback != TTA_FALLBACK_NONE && fallback != TTA_FALLBACK_FP64 && fallback != TTA_FALLBACK_MIXED)
    return TTA_ERR_ARG;
  std::vector<tta_tt*> ts;
  // TTs with d >= 2 start from the caller's cores (read only; every core is
  // replaced); shorter ones pass through as copies.
  std::unordered_set<const double*> borrowed;
  try
  {
    std::vector<std::vector<int>> ell(count);
    size_t off = 0;
    std::vector<cp_desc> cp;
    long max_el = 0;
    for (int i = 0; i < count; ++i)
    {
      if (!in[i]) fail(TTA_ERR_ARG, "round_randomized_batched: null tensor %d", i);
      const int d = in[i]->d();
      ell[i].assign(ranks + off, ranks + off + d + 1);
      off += (size_t)d + 1;
      if (d >= 1)
      {
        ell[i][0] = 1;
        ell[i][d] = 1;
      }
      for (int k = 1; k < d; ++k)
        if (ell[i][k] < 1) fail(TTA_ERR_ARG, "round_randomized_batched: target rank < 1 (TT %d bond %d)", i, k);
      tta_tt* t = new tta_tt();
      t->modes = in[i]->modes;
      t->ranks = in[i]->ranks;
      ts.push_back(t);
      if (d >= 2)
      {
        t->cores = in[i]->cores;
        borrowed.insert(in[i]->cores.begin(), in[i]->cores.end());
        continue;
      }
      for (int k = 0; k < d; ++k)
      {
        const long n = (long)in[i]->ranks[k] * in[i]->modes[k] * in[i]->ranks[k + 1];
        double* p = static_cast<double*>(c->mem.get((size_t)std::max(n, 1L) * 8));
        t->cores.push_back(p);
        cp.push_back(cp_desc{in[i]->cores[k], p, n});
        max_el = std::max(max_el, n);
      }
    }
    if (!cp.empty())
    {
      void* d = c->mem.get(cp.size() * sizeof(cp_desc));
      upload_async(c, d, cp.data(), cp.size() * sizeof(cp_desc));
      launch_copy_batched(static_cast<cp_desc*>(d), (int)cp.size(), max_el, c->stream);
      c->mem.put(d);
    }
    std::vector<double> err;
    round_batched_randomized(c, ts, ell, seed, borrowed, err);
    std::vector<int> redo;
    for (int i = 0; i < count; ++i)
    {
      const bool bad = !(err[i] <= eps) && in[i]->d() >= 2;
      if (rel_err) rel_err[i] = err[i];
      if (fell_back) fell_back[i] = (bad && fallback) ? 1 : 0;
      if (bad && fallback) redo.push_back(i);
    }
    if (!redo.empty())
    {
      // Deterministic rounding of the rejected TTs from their inputs.
      std::vector<tta_tt*> again;
      std::vector<cp_desc> cp2;
      long me = 0;
      for (int i : redo)
      {
        for (double* p : ts[i]->cores) c->mem.put(p);
        ts[i]->cores.clear();
        ts[i]->ranks = in[i]->ranks;
        for (int k = 0; k < in[i]->d(); ++k)
        {
          const long n = (long)in[i]->ranks[k] * in[i]->modes[k] * in[i]->ranks[k + 1];
          double* p = static_cast<double*>(c->mem.get((size_t)std::max(n, 1L) * 8));
          ts[i]->cores.push_back(p);
          cp2.push_back(cp_desc{in[i]->cores[k], p, n});
          me = std::max(me, n);
        }
        again.push_back(ts[i]);
      }
      void* d = c->mem.get(cp2.size() * sizeof(cp_desc));
      upload_async(c, d, cp2.data(), cp2.size() * sizeof(cp_desc));
      launch_copy_batched(static_cast<cp_desc*>(d), (int)cp2.size(), me, c->stream);
      c->mem.put(d);
      round_batched_plain(c, again, eps, 0, fallback == TTA_FALLBACK_MIXED);
    }
    for (int i = 0; i < count; ++i) out[i] = ts[i];
  }
  catch (const tta::error& e)
  {
    for (auto* t : ts)
    {
      for (double* p : t->cores)
        if (!borrowed.count(p)) c->mem.put(p);
      delete t;
    }
    c->last_error = e.what();
    return e.code;
  }
  return TTA_OK;
}
