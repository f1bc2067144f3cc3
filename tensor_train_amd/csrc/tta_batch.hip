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

// One SVD-sweep step at core position k of the listed TTs (each with
// k < d - 1), as svd_step (include/core/detail/truncate.hpp:61-99): the
// unfolding QR-condensed by Householder, the batched Jacobi SVD of R^T, the
// rank rule on the device, U and diag(s) V^T applied.  rk, deltas and flags
// are per-TT device arrays indexed like ts.
void svd_step_plain(tta_ctx* c, std::vector<tta_tt*>& ts, const std::vector<int>& sel, int k, int max_rank, int* rk,
                    const double* deltas, int* flags)
{
  const int count = (int)ts.size();
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
  for (int i : sel)
  {
    tta_tt* t = ts[i];
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
  if (nb == 0) return;
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
  if (max_q <= 64)
  {
    // Whole-problem Jacobi kernels of uneven sizes: a persistent launch fed
    // from a largest-first work queue.
    std::vector<int> ord(nb);
    for (int a = 0; a < nb; ++a) ord[a] = a;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return jp[x].q > jp[y].q; });
    jb.order = S.upload(ord);
    jb.queue = S.alloc<int>(1);
  }
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
    launch_gemm_grouped(S, m1, c->stream);
    launch_gemm_grouped(S, m2, c->stream);
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

// One right-sweep step (right_qr_step, include/core/detail/orthogonalize.hpp:
// 63-100) at core d - 1 - j of the listed TTs (each with d - 1 - j >= 1):
// Householder QR of the (n r' x r) view, core k <- Q^T, core k-1 <- core_{k-1} R^T.
void right_step_plain(tta_ctx* c, std::vector<tta_tt*>& ts, const std::vector<int>& sel, int j)
{
  scratch S(c);
  std::vector<int> idx;
  std::vector<qrp> qr;
  std::vector<ex_desc> ex;
  std::vector<gemm_desc> gm;
  int max_m = 0, max_n = 0;
  long max_ex = 0;
  std::vector<double*> newQ, newP;
  for (int i : sel)
  {
    tta_tt* t = ts[i];
    const int k = t->d() - 1 - j;
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
  }
  const int nb = (int)idx.size();
  if (nb == 0) return;
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
    launch_gemm_grouped(S, gm, c->stream);
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

// Rank-revealing CholQR2 of a batch of tall-or-square matrices M (m x q,
// element (i, a) at M[i * smr + a * smc], q <= kPcholMaxQ), all in FP64:
//   1. G = M^T M on the matrix cores and its Cholesky with diagonal pivoting,
//      stopped at the first p whose Schur-complement trace (the squared
//      residual of the p pivot columns) is <= stop_f * (*delta)^2, or where
//      the largest remaining pivot falls to rounding level;
//   2. CholQR2 of the p pivot columns: Q1 = M P R1^-1, then U = Q1 R2^-1 with
//      R2 the Cholesky factor of Q1^T Q1;
//   3. the residual correction R~ = U^T M (p x q, row-major) and
//      E = M - U R~, whose Frobenius norm is taken directly (no squaring).
// U goes to a fresh pool block with element (i, j) at U[i * sur + j * suc]
// (a negative stride stands for p, the pivot count).
// The caller decides acceptance from p, the pass-2 status and ||E||_F (and
// ||M||_F), applies R~ to the neighbouring core and owns U (returned to the
// pool when rejected).
struct rr_prob
{
  const double* M = nullptr;
  long smr = 0, smc = 0;
  int m = 0, q = 0;
  const double* delta = nullptr;  // device; null: stop at rounding level only
  double stop_f = 0.0;
  long sur = 0, suc = 0;
  int max_p = 0;  // > 0: a pivot count above it is not carried to stage 2
  // > 0: not carried to stage 2 when the trace left at the stop exceeds
  // max_trace_f * (*delta)^2 (the step would be rejected anyway).
  double max_trace_f = 0.0;
  double hdelta = 0.0;  // host copy of *delta
};
struct rr_out
{
  int p = 0;          // pivots; 0 when stage 2 did not run
  bool pass2 = false;  // CholQR2's second Cholesky succeeded
  double enorm = 0.0, mnorm = 0.0;
  double* U = nullptr;
  double* Rt = nullptr;  // in the caller's arena
};

std::vector<rr_out> rr_cholqr2(tta_ctx* c, scratch& S, const std::vector<rr_prob>& pr)
{
  const int nb = (int)pr.size();
  std::vector<rr_out> out(nb);
  if (nb == 0) return out;
  std::vector<double*> R1(nb);
  std::vector<int*> pivs(nb);
  int* pout = S.alloc<int>((size_t)nb);
  double* trl = S.alloc<double>((size_t)nb);
  std::vector<gemm_desc> g1;
  std::vector<pchol_desc> pc;
  for (int a = 0; a < nb; ++a)
  {
    const rr_prob& P = pr[a];
    double* G = S.alloc<double>((size_t)P.q * P.q);
    R1[a] = S.alloc<double>((size_t)P.q * P.q);
    pivs[a] = S.alloc<int>((size_t)P.q);
    gemm_desc g;
    g.M = P.q, g.N = P.q, g.K = P.m;
    g.A = P.M, g.sar = P.smc, g.sac = P.smr;
    g.B = P.M, g.sbr = P.smr, g.sbc = P.smc;
    g.C = G, g.scr = P.q, g.scc = 1;
    g1.push_back(g);
    pchol_desc d;
    d.G = G;
    d.ldg = P.q;
    d.q = P.q;
    d.delta = P.delta;
    d.stop_f = P.stop_f;
    d.pivot = 1;
    d.piv = pivs[a];
    d.p_out = pout + a;
    d.L = S.alloc<double>((size_t)P.q * P.q);
    d.Rinv = R1[a];
    d.ldr = P.q;
    d.tr_out = trl + a;
    pc.push_back(d);
  }
  {
    const credit w = gemm_credit(g1.data(), g1.size());
    prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
    launch_gemm_grouped(S, g1, c->stream);
  }
  {
    prof_scope ps(c, TTA_K_CHOL);
    launch_pchol_batch(S, S.upload(pc), pc, c->stream);
  }
  std::vector<int> hp(nb);
  std::vector<double> htr(nb);
  hip_check(hipMemcpyAsync(hp.data(), pout, nb * sizeof(int), hipMemcpyDeviceToHost, c->stream), "pivots");
  hip_check(hipMemcpyAsync(htr.data(), trl, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream), "trace");
  sync(c);

  std::vector<int> al;
  for (int a = 0; a < nb; ++a)
  {
    const rr_prob& P = pr[a];
    if (hp[a] < 1 || (P.max_p > 0 && hp[a] > P.max_p)) continue;
    if (P.max_trace_f > 0.0 && htr[a] > P.max_trace_f * P.hdelta * P.hdelta) continue;
    al.push_back(a);
  }
  const int nc = (int)al.size();
  if (nc == 0) return out;
  std::vector<ga_desc> ga;
  std::vector<gemm_desc> gq1, gg2, gu, grt, ge;
  std::vector<pchol_desc> pc2;
  std::vector<cp_desc> cp;
  std::vector<sq_desc> sq;
  int* p2 = S.alloc<int>((size_t)nc);
  int* skip2 = S.alloc<int>((size_t)nc);
  double* en = S.alloc<double>((size_t)nc * 2);
  long mx_ga = 0, mx_cp = 0;
  for (int b = 0; b < nc; ++b)
  {
    const int a = al[b];
    const rr_prob& P = pr[a];
    const int m = P.m, q = P.q, p = hp[a];
    const long sur = P.sur < 0 ? p : P.sur, suc = P.suc < 0 ? p : P.suc;
    double* Mp = S.alloc<double>((size_t)m * p);
    double* Q1 = S.alloc<double>((size_t)m * p);
    double* G2 = S.alloc<double>((size_t)p * p);
    double* Ri2 = S.alloc<double>((size_t)p * p);
    double* E = S.alloc<double>((size_t)m * q);  // a copy of M, M's layout
    rr_out& o = out[a];
    o.p = p;
    o.Rt = S.alloc<double>((size_t)p * q);
    o.U = new_block(c, (size_t)m * p);
    if (P.smr == 1)
      ga.push_back(ga_desc{P.M, (int)P.smc, m, pivs[a], p, Mp, 1, m, nullptr, GATHER_COPY});
    else
      ga.push_back(ga_desc{P.M, (int)P.smr, m, pivs[a], p, Mp, 1, m, nullptr, GATHER_COPY | GATHER_ROWMAJOR});
    mx_ga = std::max(mx_ga, (long)m * p);
    gemm_desc g;
    // Q1 = M_p R1^-1 (column-major m x p).
    g = gemm_desc();
    g.M = m, g.N = p, g.K = p;
    g.A = Mp, g.sar = 1, g.sac = m;
    g.B = R1[a], g.sbr = q, g.sbc = 1;
    g.C = Q1, g.scr = 1, g.scc = m;
    gq1.push_back(g);
    // G2 = Q1^T Q1 (row-major p x p).
    g = gemm_desc();
    g.M = p, g.N = p, g.K = m;
    g.A = Q1, g.sar = m, g.sac = 1;
    g.B = Q1, g.sbr = 1, g.sbc = m;
    g.C = G2, g.scr = p, g.scc = 1;
    gg2.push_back(g);
    pchol_desc d;
    d.G = G2;
    d.ldg = p;
    d.q = p;
    d.pivot = 0;
    d.p_out = p2 + b;
    d.L = S.alloc<double>((size_t)p * p);
    d.Rinv = Ri2;
    d.ldr = p;
    d.skip = skip2 + b;
    pc2.push_back(d);
    // U = Q1 R2^-1 in the caller's layout.
    g = gemm_desc();
    g.M = m, g.N = p, g.K = p;
    g.A = Q1, g.sar = 1, g.sac = m;
    g.B = Ri2, g.sbr = p, g.sbc = 1;
    g.C = o.U, g.scr = sur, g.scc = suc;
    gu.push_back(g);
    // R~ = U^T M (row-major p x q).
    g = gemm_desc();
    g.M = p, g.N = q, g.K = m;
    g.A = o.U, g.sar = suc, g.sac = sur;
    g.B = P.M, g.sbr = P.smr, g.sbc = P.smc;
    g.C = o.Rt, g.scr = q, g.scc = 1;
    grt.push_back(g);
    // E = M - U R~ on a contiguous copy of M (both callers' M are
    // contiguous, row-major or column-major).
    cp.push_back(cp_desc{P.M, E, (long)m * q});
    mx_cp = std::max(mx_cp, (long)m * q);
    g = gemm_desc();
    g.M = m, g.N = q, g.K = p;
    g.A = o.U, g.sar = sur, g.sac = suc;
    g.B = o.Rt, g.sbr = q, g.sbc = 1;
    g.C = E, g.scr = P.smr, g.scc = P.smc;  // E has M's own layout
    g.alpha = -1.0;
    g.beta = 1.0;
    ge.push_back(g);
    sq_desc sd;
    sd.x = E;
    sd.n = (long)m * q;
    sd.out_norm = en + 2 * b;
    sq.push_back(sd);
    sd.x = P.M;
    sd.out_norm = en + 2 * b + 1;
    sq.push_back(sd);
  }
  {
    prof_scope ps(c, TTA_K_OTHER);
    launch_gather_cols_batched(S.upload(ga), nc, mx_ga, c->stream);
    launch_copy_batched(S.upload(cp), nc, mx_cp, c->stream);
  }
  {
    credit w = gemm_credit(gq1.data(), gq1.size());
    for (const auto* v : {&gg2, &gu, &grt, &ge})
    {
      const credit w2 = gemm_credit(v->data(), v->size());
      w.flops += w2.flops;
      w.bytes += w2.bytes;
    }
    prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
    launch_gemm_grouped(S, gq1, c->stream);
    launch_gemm_grouped(S, gg2, c->stream);
  }
  {
    // Second pass: one Newton-Schulz step where Q1 is already orthonormal to
    // 1e-8 / p (rank-deficient sums: Q1's columns are well conditioned),
    // the Cholesky of G2 for the rest.
    prof_scope ps(c, TTA_K_CHOL);
    const pchol_desc* pc2_dev = S.upload(pc2);
    launch_ns_pass2(pc2_dev, nc, skip2, c->stream);
    launch_pchol_batch(S, pc2_dev, pc2, c->stream);
  }
  {
    credit w;
    for (const auto* v : {&gu, &grt, &ge})
    {
      const credit w2 = gemm_credit(v->data(), v->size());
      w.flops += w2.flops;
      w.bytes += w2.bytes;
    }
    prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
    launch_gemm_grouped(S, gu, c->stream);
    launch_gemm_grouped(S, grt, c->stream);
    launch_gemm_grouped(S, ge, c->stream);
  }
  {
    prof_scope ps(c, TTA_K_OTHER);
    launch_frob_batched(S.upload(sq), 2 * nc, c->stream);
  }
  std::vector<int> hp2(nc);
  std::vector<double> hen(2 * nc);
  hip_check(hipMemcpyAsync(hp2.data(), p2, nc * sizeof(int), hipMemcpyDeviceToHost, c->stream), "pass 2");
  hip_check(hipMemcpyAsync(hen.data(), en, 2 * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream), "residuals");
  sync(c);
  for (int b = 0; b < nc; ++b)
  {
    rr_out& o = out[al[b]];
    o.pass2 = hp2[b] == o.p;
    o.enorm = hen[2 * b];
    o.mnorm = hen[2 * b + 1];
  }
  return out;
}

// Mixed SVD-sweep step at core k: a tall unfolding M (r n x r', row-major,
// r' <= kPcholMaxQ) by rr_cholqr2 with the trace stop at (0.4 delta)^2.
// Accepted when ||E||_F <= delta / 2 (and p <= max_rank when max_rank
// binds): core k becomes U (rank p), core k+1 becomes R~ core_{k+1}.  The
// step then discards at most delta / 2 at this bond (the steps' residuals
// are orthogonal: the cores left of k are orthonormal, those right of k
// right-orthonormal), so the round's bound ||a - round(a)|| <= eps ||a||
// of tt_round.hpp:36-127 holds, and p is never below the rank the rule of
// tt_eigen_bridge.hpp:172-197 picks (every smaller rank leaves more than
// delta).  On exactly rank-deficient unfoldings (sums of TTs) the residual
// is at rounding level and p is the plain rank.  Returns the TTs whose step
// was not accepted; their cores are untouched.
std::vector<int> svd_step_certified(tta_ctx* c, std::vector<tta_tt*>& ts, const std::vector<int>& sel, int k,
                                    int max_rank, const double* deltas, const std::vector<double>& hdelta,
                                    std::vector<char>& off)
{
  std::vector<int> rest, el;
  for (int i : sel)
  {
    const tta_tt* t = ts[i];
    const int rows = t->ranks[k] * t->modes[k], q = t->ranks[k + 1];
    if (!off[i] && rows >= q && q <= kPcholMaxQ && hdelta[i] > 0.0)
      el.push_back(i);
    else
      rest.push_back(i);
  }
  if (el.empty()) return rest;
  c->rr_tried += (long long)el.size();
  scratch S(c);
  std::vector<rr_prob> pr;
  for (int i : el)
  {
    const tta_tt* t = ts[i];
    rr_prob P;
    P.M = t->cores[k];
    P.m = t->ranks[k] * t->modes[k];
    P.q = t->ranks[k + 1];
    P.smr = P.q;
    P.smc = 1;
    P.delta = deltas + i;
    P.stop_f = 0.16;
    P.sur = -1;  // U row-major (m x p): row stride p, known after stage 1
    P.suc = 1;
    P.max_p = max_rank;
    P.hdelta = hdelta[i];
    P.max_trace_f = 0.5;  // a residual near sqrt(0.5) delta would fail the delta / 2 test
    pr.push_back(P);
  }
  std::vector<rr_out> o = rr_cholqr2(c, S, pr);
  std::vector<gemm_desc> gn;
  std::vector<int> acc;
  std::vector<double*> Nn;
  for (size_t a = 0; a < el.size(); ++a)
  {
    const int i = el[a];
    tta_tt* t = ts[i];
    const bool ok = o[a].p >= 1 && o[a].pass2 && o[a].enorm <= 0.5 * hdelta[i];
    if (!ok)
    {
      if (o[a].U) c->mem.put(o[a].U);
      rest.push_back(i);
      off[i] = 1;
      continue;
    }
    const int p = o[a].p, q = t->ranks[k + 1];
    const int ncols_n = t->modes[k + 1] * t->ranks[k + 2];
    gemm_desc g;
    g.M = p, g.N = ncols_n, g.K = q;
    g.A = o[a].Rt, g.sar = q, g.sac = 1;
    g.B = t->cores[k + 1], g.sbr = ncols_n, g.sbc = 1;
    g.C = new_block(c, (size_t)p * ncols_n), g.scr = ncols_n, g.scc = 1;
    Nn.push_back(g.C);
    gn.push_back(g);
    acc.push_back((int)a);
  }
  if (!gn.empty())
  {
    {
      const credit w = gemm_credit(gn.data(), gn.size());
      prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
      launch_gemm_grouped(S, gn, c->stream);
    }
    for (size_t x = 0; x < acc.size(); ++x)
    {
      const int a = acc[x];
      tta_tt* t = ts[el[a]];
      c->mem.put(t->cores[k]);
      c->mem.put(t->cores[k + 1]);
      t->cores[k] = o[a].U;
      t->cores[k + 1] = Nn[x];
      t->ranks[k + 1] = o[a].p;
    }
    c->rr_held += (long long)acc.size();
  }
  std::sort(rest.begin(), rest.end());
  return rest;
}

// Mixed right-sweep step at core k >= 1: B = the (n r' x r) column-major view
// of core k (r <= kPcholMaxQ) by rr_cholqr2 stopped at rounding level.
// Accepted when ||E||_F <= min(1e-13, 1e-3 eps) ||B||_F (never at eps = 0,
// where round drops only exact zeros): only directions that B holds at
// rounding level (a few times the backward error of a Householder QR, which
// is about sqrt(n r') u ||B||) are dropped, so the tensor is that of
// right_qr_step (include/core/detail/orthogonalize.hpp:63-100) to rounding
// level while r_k shrinks from min(n r', r) to the numerical rank p (sums of
// TTs: r_u + r_v instead of 2 r_u + r_v).  Core k becomes U^T (p x n r',
// right-orthonormal), core k-1 becomes core_{k-1} R~^T.  Returns the TTs
// whose step was not accepted.
std::vector<int> right_step_certified(tta_ctx* c, std::vector<tta_tt*>& ts, const std::vector<int>& sel, int j,
                                      double eps, std::vector<char>& off)
{
  const double tol = std::min(1e-13, 1e-3 * eps);
  if (!(tol > 0.0)) return sel;
  std::vector<int> rest, el;
  for (int i : sel)
  {
    const tta_tt* t = ts[i];
    const int k = t->d() - 1 - j;
    if (!off[i] && t->ranks[k] <= kPcholMaxQ)
      el.push_back(i);
    else
      rest.push_back(i);
  }
  if (el.empty()) return rest;
  c->rr_tried += (long long)el.size();
  scratch S(c);
  std::vector<rr_prob> pr;
  for (int i : el)
  {
    const tta_tt* t = ts[i];
    const int k = t->d() - 1 - j;
    rr_prob P;
    P.M = t->cores[k];
    P.m = t->modes[k] * t->ranks[k + 1];
    P.q = t->ranks[k];
    P.smr = 1;
    P.smc = P.m;
    P.delta = nullptr;
    P.stop_f = 0.0;
    P.sur = 1;  // U column-major (m x p) = core k's row-major (p, n, r')
    P.suc = P.m;
    pr.push_back(P);
  }
  std::vector<rr_out> o = rr_cholqr2(c, S, pr);
  std::vector<gemm_desc> gn;
  std::vector<int> acc;
  std::vector<double*> Nn;
  for (size_t a = 0; a < el.size(); ++a)
  {
    const int i = el[a];
    tta_tt* t = ts[i];
    const int k = t->d() - 1 - j;
    const bool ok = o[a].p >= 1 && o[a].pass2 && o[a].enorm <= tol * o[a].mnorm;
    if (!ok)
    {
      if (o[a].U) c->mem.put(o[a].U);
      rest.push_back(i);
      off[i] = 1;
      continue;
    }
    const int p = o[a].p, r = t->ranks[k];
    const int rows_p = t->ranks[k - 1] * t->modes[k - 1];
    // core_{k-1} (rows_p x r) * R~^T (r x p), R~ row-major p x r.
    gemm_desc g;
    g.M = rows_p, g.N = p, g.K = r;
    g.A = t->cores[k - 1], g.sar = r, g.sac = 1;
    g.B = o[a].Rt, g.sbr = 1, g.sbc = r;
    g.C = new_block(c, (size_t)rows_p * p), g.scr = p, g.scc = 1;
    Nn.push_back(g.C);
    gn.push_back(g);
    acc.push_back((int)a);
  }
  if (!gn.empty())
  {
    {
      const credit w = gemm_credit(gn.data(), gn.size());
      prof_scope ps(c, TTA_K_GEMM, w.flops, w.bytes);
      launch_gemm_grouped(S, gn, c->stream);
    }
    for (size_t x = 0; x < acc.size(); ++x)
    {
      const int a = acc[x];
      tta_tt* t = ts[el[a]];
      const int k = t->d() - 1 - j;
      c->mem.put(t->cores[k]);
      c->mem.put(t->cores[k - 1]);
      t->cores[k] = o[a].U;
      t->cores[k - 1] = Nn[x];
      t->ranks[k] = o[a].p;
    }
    c->rr_held += (long long)acc.size();
  }
  std::sort(rest.begin(), rest.end());
  return rest;
}

// Rounds `count` working TTs in place (gauge none): plain FP64, or with
// mixed = true the SVD steps first tried by svd_step_certified.
void round_batched_plain(tta_ctx* c, std::vector<tta_tt*>& ts, double eps, int max_rank, bool mixed)
{
  const int count = (int)ts.size();
  int maxd = 0;
  for (auto* t : ts) maxd = std::max(maxd, t->d());
  if (maxd <= 1) return;

  // A TT whose certified step was rejected takes the plain steps for the
  // rest of that sweep: the rejections come where the spectrum runs on below
  // the Gram's resolution (sqrt(q u) of the largest singular value), which
  // persists along the sweep.
  std::vector<char> rr_off(count, 0);

  // ------------------------------------------------------------ right sweep
  for (int j = 0; j <= maxd - 2; ++j)
  {
    std::vector<int> sel;
    for (int i = 0; i < count; ++i)
      if (ts[i]->d() - 1 - j >= 1) sel.push_back(i);
    if (sel.empty()) continue;
    if (mixed) sel = right_step_certified(c, ts, sel, j, eps, rr_off);
    if (!sel.empty()) right_step_plain(c, ts, sel, j);
  }

  // ------------------------------------------------------------ norms, delta
  scratch G(c);
  double* norms = G.alloc<double>((size_t)count);
  double* deltas = G.alloc<double>((size_t)count);
  int* rk = G.alloc<int>((size_t)count);
  int* flags = G.alloc<int>((size_t)count * 4);
  std::vector<char> live(count, 0);
  std::vector<double> hdelta(count, 0.0);
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
    hip_check(hipMemcpyAsync(hdelta.data(), deltas, count * 8, hipMemcpyDeviceToHost, c->stream), "deltas");
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
  std::fill(rr_off.begin(), rr_off.end(), 0);
  for (int k = 0; k <= maxd - 2; ++k)
  {
    std::vector<int> sel;
    for (int i = 0; i < count; ++i)
      if (live[i] && k < ts[i]->d() - 1) sel.push_back(i);
    if (sel.empty()) continue;
    if (mixed) sel = svd_step_certified(c, ts, sel, k, max_rank, deltas, hdelta, rr_off);
    if (!sel.empty()) svd_step_plain(c, ts, sel, k, max_rank, rk, deltas, flags);
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
    if (precision != TTA_PREC_FP64 && precision != TTA_PREC_MIXED)
      fail(TTA_ERR_ARG, "round_batched: precision %d is not TTA_PREC_FP64 or TTA_PREC_MIXED", precision);
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
      round_batched_plain(c, ts, eps, max_rank, precision == TTA_PREC_MIXED);
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
