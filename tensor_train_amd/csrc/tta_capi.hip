// C ABI of the gfx950 TT-rounding library: device tensors, the plain FP64
// round / orthogonalize / apply / inner paths, orchestrated on one HIP
// stream per context.  Kernels live in tta_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

using namespace tta;

// ----------------------------------------------------------------------------
// Context: stream, size-bucketed device pool, workspace, pinned scalars.
// ----------------------------------------------------------------------------
namespace tta
{

[[noreturn]] void fail(int code, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw error(code, buf);
}

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) fail(TTA_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

double* ctx_ws(tta_ctx* c, int slot, size_t n)
{
  auto& w = c->ws[slot];
  if (w.cap < n)
  {
    if (w.p) c->mem.put(w.p);
    w.p = static_cast<double*>(c->mem.get(n * sizeof(double)));
    w.cap = n;
  }
  return w.p;
}


// Event for the timing of one launch on the context stream (recycled).
hipEvent_t prof_event(tta_ctx* c)
{
  if (!c->prof_free.empty())
  {
    hipEvent_t e = c->prof_free.back();
    c->prof_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  return e;
}

}  // namespace tta

// ----------------------------------------------------------------------------
// Device tensor helpers.
// ----------------------------------------------------------------------------
namespace
{

tta_tt* new_tt(tta_ctx* c, const std::vector<int>& modes, const std::vector<int>& ranks, bool zero)
{
  const int d = (int)modes.size();
  auto* t = new tta_tt();
  t->ranks = ranks;
  t->modes = modes;
  t->cores.assign(d, nullptr);
  try
  {
    for (int k = 0; k < d; ++k)
    {
      const size_t n = (size_t)ranks[k] * modes[k] * ranks[k + 1];
      t->cores[k] = static_cast<double*>(c->mem.get(std::max<size_t>(n, 1) * sizeof(double)));
      if (zero) hip_check(hipMemsetAsync(t->cores[k], 0, n * sizeof(double), c->stream), "hipMemsetAsync");
    }
  }
  catch (...)
  {
    for (double* p : t->cores) c->mem.put(p);
    delete t;
    throw;
  }
  return t;
}

void free_tt(tta_ctx* c, tta_tt* t)
{
  if (!t) return;
  for (double* p : t->cores) c->mem.put(p);
  delete t;
}

size_t core_numel(const tta_tt* t, int k) { return (size_t)t->ranks[k] * t->modes[k] * t->ranks[k + 1]; }

tta_tt* copy_tt(tta_ctx* c, const tta_tt* a)
{
  tta_tt* t = new_tt(c, a->modes, a->ranks, false);
  t->row_modes = a->row_modes;
  t->col_modes = a->col_modes;
  for (int k = 0; k < a->d(); ++k)
    hip_check(hipMemcpyAsync(t->cores[k], a->cores[k], core_numel(a, k) * sizeof(double), hipMemcpyDeviceToDevice,
                             c->stream),
              "hipMemcpyAsync");
  return t;
}

void replace_core(tta_ctx* c, tta_tt* t, int k, double* p)
{
  c->mem.put(t->cores[k]);
  t->cores[k] = p;
}

double* alloc_core(tta_ctx* c, size_t n) { return static_cast<double*>(c->mem.get(std::max<size_t>(n, 1) * 8)); }

void prof_collect(tta_ctx* c)
{
  if (c->prof_pending.empty()) return;
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  for (auto& r : c->prof_pending)
  {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
    c->prof_ms[r.cls] += ms;
    c->prof_n[r.cls] += 1;
    c->prof_flops[r.cls] += r.flops;
    c->prof_bytes[r.cls] += r.bytes;
    c->prof_free.push_back(r.a);
    c->prof_free.push_back(r.b);
  }
  c->prof_pending.clear();
}

void gemm_go(tta_ctx* c, const gemm_desc& g);

void gemm_rm(tta_ctx* c, int M, int N, int K, const double* A, long lda, const double* B, long ldb, double* C,
             long ldc)
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
  gemm_go(c, g);
}

// Scratch slots (see tta_internal.hpp).
enum
{
  WS_A = 0,
  WS_Q,
  WS_X,
  WS_J,
  WS_F,
  WS_JP,
  WS_M1,
  WS_M1P,
  WS_PART,
  WS_SMALL,
  WS_T,
  WS_CERT,
  WS_SPLITK,
  WS_SPEC
};

// Flag of the running speculative sweep (1.0 while every decision held).
double* spec_flag(tta_ctx* c) { return ctx_ws(c, WS_SPEC, 1); }

// One GEMM on the context stream, split along K when its output tiles
// alone would leave the device mostly idle (the partial products go
// through the WS_SPLITK scratch slot).  Timed per kernel: the tiled GEMM
// as TTA_K_GEMM, the split-K partial products as TTA_K_GEMM_SPLITK (both
// credited 2 M N K flops and their operand and result bytes) and the
// reduction of the partials as TTA_K_OTHER.
void gemm_go(tta_ctx* c, const gemm_desc& g)
{
  const double mk = (double)g.M * g.K, kn = (double)g.K * g.N, mn = (double)g.M * g.N;
  const double fl = 2.0 * mn * g.K;
  const int S = gemm_splitk_slices(g.M, g.N, g.K);
  if (S > 1)
  {
    double* part = ctx_ws(c, WS_SPLITK, (size_t)S * g.M * g.N);
    int s_eff = 0;
    {
      prof_scope ps(c, TTA_K_GEMM_SPLITK, fl, 8.0 * (mk + kn + S * mn));
      s_eff = launch_gemm_splitk_parts(g, S, part, c->stream);
    }
    prof_scope ps(c, TTA_K_OTHER, 0.0, 8.0 * (s_eff + (g.beta != 0.0 ? 2.0 : 1.0)) * mn);
    launch_splitk_reduce(g, s_eff, part, c->stream);
  }
  else
  {
    prof_scope ps(c, TTA_K_GEMM, fl, 8.0 * (mk + kn + mn));
    launch_gemm(g, 1, c->stream);
  }
}

// C = A B with element strides, without a profiling scope of its own (for
// use inside a scope of another class).
void gemm_raw(tta_ctx* c, int M, int N, int K, const double* A, long sar, long sac, const double* B, long sbr,
              long sbc, double* C, long scr, long scc)
{
  gemm_desc g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sar = sar;
  g.sac = sac;
  g.B = B;
  g.sbr = sbr;
  g.sbc = sbc;
  g.C = C;
  g.scr = scr;
  g.scc = scc;
  gemm_go(c, g);
}

// A GEMM descriptor for C = alpha A B + beta C with element strides.
gemm_desc gdesc(int M, int N, int K, const double* A, long sar, long sac, const double* B, long sbr, long sbc, double* C,
                long scr, long scc, double alpha = 1.0, double beta = 0.0)
{
  gemm_desc g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sar = sar;
  g.sac = sac;
  g.B = B;
  g.sbr = sbr;
  g.sbc = sbc;
  g.C = C;
  g.scr = scr;
  g.scc = scc;
  g.alpha = alpha;
  g.beta = beta;
  return g;
}

// Independent GEMMs (empty ones, M or N or K = 0, skipped) in one launch
// on the context stream; a single one goes through gemm_go.
// With `decide` (its dinfo .. dflag set), the speculative CholQR2
// acceptance decision rides along in one more workgroup of the launch.
void gemm_multi_go(tta_ctx* c, std::initializer_list<gemm_desc> gs, const gemm_multi* decide = nullptr)
{
  gemm_multi m;
  if (decide)
  {
    m.dinfo = decide->dinfo;
    m.dpart = decide->dpart;
    m.dnp = decide->dnp;
    m.dn = decide->dn;
    m.dthr = decide->dthr;
    m.dflag = decide->dflag;
  }
  double fl = 0.0, by = 0.0;
  for (const gemm_desc& g : gs)
  {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) continue;
    if (m.count == 4) throw tta::error(TTA_ERR_ARG, "gemm_multi_go: more than four products");
    m.g[m.count++] = g;
    fl += 2.0 * g.M * g.N * g.K;
    by += 8.0 * ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N);
  }
  if (m.count == 0 && !decide) return;
  if (m.count == 1 && !decide)
  {
    gemm_go(c, m.g[0]);
    return;
  }
  prof_scope ps(c, TTA_K_GEMM_MULTI, fl, by);
  launch_gemm_multi(m, c->stream);
}

// gemm_go with an epilogue folded into the split-K sum (splitk_epi); only
// for products that split (gemm_splitk_slices > 1), so that the sum runs
// after every partial product has read its operands and may overwrite one.
void gemm_go_epi(tta_ctx* c, int M, int N, int K, const double* A, long sar, long sac, const double* B, long sbr,
                 long sbc, double* C, long scr, long scc, double alpha, const splitk_epi& ep)
{
  gemm_desc g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sar = sar;
  g.sac = sac;
  g.B = B;
  g.sbr = sbr;
  g.sbc = sbc;
  g.C = C;
  g.scr = scr;
  g.scc = scc;
  g.alpha = alpha;
  const double mk = (double)M * K, kn = (double)K * N, mn = (double)M * N;
  const int S = gemm_splitk_slices(M, N, K);
  double* part = ctx_ws(c, WS_SPLITK, (size_t)S * mn);
  int s_eff = 0;
  {
    prof_scope ps(c, TTA_K_GEMM_SPLITK, 2.0 * mn * K, 8.0 * (mk + kn + S * mn));
    s_eff = launch_gemm_splitk_parts(g, S, part, c->stream);
  }
  prof_scope ps(c, TTA_K_OTHER, 0.0, 8.0 * (s_eff + 3.0) * mn);
  launch_splitk_reduce_epi(g, s_eff, part, ep, c->stream);
}

// Squared Frobenius norms of a factorization's triangular factor R and of
// R^-1, when the factorization formed R^-1 anyway (CholQR2 does); the
// full-rank certificate then needs no triangular inverse of its own.
struct rinv_norms
{
  bool have = false;
  double xf2 = 0.0, rf2 = 0.0;
  // In a speculative sweep: the column norms of R and then of R^-1 on the
  // device (2n values), for the certificate decided there.
  const double* dev = nullptr;
};

// Blocked Cholesky G = R^T R of an SPD n x n matrix (row-major, leading
// dimension n) and the inverse X = R^-1, both upper triangular with zero
// strict lower part, in 64-row block steps.  Step J: one wave factors and
// inverts the diagonal block (potrf_inv_kernel; D = R_JJ^-1), then on the
// matrix cores
//   R(J, >J)       = D^T G(J, >J)              (block row of R)
//   G(>J, >J)     -= R(J, >J)^T R(J, >J)       (trailing update)
//   X(<J, J)       = -X(<J, <J) R(<J, J) D     (block column of R^-1).
// G is overwritten by the trailing updates; T is scratch of n * 64 doubles;
// info[0..2] as for launch_potrf_inv.
void chol_inv_blocked(tta_ctx* c, double* G, int n, double* R, double* X, double* info)
{
  const int B = potrf_block();
  {
    const int jb = std::min(B, n);
    prof_scope ps(c, TTA_K_CHOL, 2.0 * jb * jb * jb / 3.0, 24.0 * jb * jb);
    launch_potrf_inv(G, n, R, n, X, n, jb, 0, info, c->stream);
  }
  if (n <= B) return;
  double* S = alloc_core(c, chol_step_scratch(n));
  try
  {
    for (int j0 = 0; j0 < n; j0 += B)
    {
      chol_step st;
      int nx = 0;
      chol_step_shape(n, j0, st.jb, st.nt, st.ntri, nx);
      const int jb = st.jb, w = n - j0 - jb;
      st.G = G;
      st.R = R;
      st.X = X;
      st.S = S;
      st.info = info;
      st.n = n;
      st.j0 = j0;
      // Flops and bytes of the plain block step: block row of R, trailing
      // update, X's block column (T and -T D) and the next diagonal block.
      const int nb = std::min(B, w);
      const double fl = 2.0 * jb * jb * w + 2.0 * w * w * jb + 2.0 * j0 * jb * (j0 + jb) + 2.0 * nb * nb * nb / 3.0;
      const double by = 8.0 * (2.0 * jb * w + 2.0 * w * w + 2.0 * j0 * j0 + 3.0 * j0 * jb + 3.0 * nb * nb);
      prof_scope ps(c, TTA_K_CHOL_STEP, fl, by);
      launch_chol_step(st, c->stream);
    }
  }
  catch (...)
  {
    c->mem.put(S);
    throw;
  }
  c->mem.put(S);
}

// CholQR2 of a column-major (m x n) A, m >= n: A = Q R with Q^T Q = I.
// Pass 1: G = A^T A = R1^T R1 and R1^-1 (blocked Cholesky, chol_inv_blocked),
// Q1 = A R1^-1.  Pass 2 corrects Q1 from G2 = Q1^T Q1 = I + E, where E is of
// the order u kappa(A)^2:
//   * polar form (n max|E| <= 1e-5): P = I - E/2 + 3E^2/8 and
//     Pi = I + E/2 - E^2/8, the second-order series of (I + E)^(-1/2) and
//     (I + E)^(1/2); Q = Q1 P, R = Pi R1, R^-1 = R1^-1 P.  Q^T Q = I and
//     Q R = A hold up to O(||E||^3) (below 1e-15) and rounding.  R is full,
//     not triangular: Q's columns are the symmetric orthogonalization of
//     Q1's.  Four products after G2.
//   * triangular form (n max|E| <= 1e-4, condition numbers up to about 1e6):
//     R2 = I + X with X the fixed point of X = Phi(E - X^T X) (Phi: strict
//     upper part plus half the diagonal), reached from X = Phi(E) in three
//     iterations that each gain a factor ~2||E||; R2^-1 by one Newton step
//     from I - X (error ||X||^4); Q = Q1 R2^-1, R = R2 R1,
//     R^-1 = R1^-1 R2^-1.
// The polar form is formed before the acceptance test is read, so that each
// factorization waits on the stream once, and a speculative sweep (which
// accepts the polar form only, deciding on the device) not at all; the
// triangular form runs only when the test asks for it.  A is read only until
// the factorization is accepted; then R goes to A's leading n x n block (all
// of it) and Q to Q (scratch for the caller either way).  Otherwise nothing
// is written and the caller factors A by Householder.  Every product runs on
// the matrix cores.
// The strided form: A(i, k) = A[i sar + k sac] is only read; Q(i, k) =
// Q[i sqr + k sqc] (Q may be null) and R (n x n row-major, full) go to the
// caller's buffers, so callers holding either layout need no transposes.
bool cholqr2_strided(tta_ctx* c, const double* A, long sar, long sac, int m, int n, double* Q, long sqr, long sqc,
                     double* Rf, rinv_norms* cert = nullptr)
{
  const size_t nn = (size_t)n * n;
  double* G = alloc_core(c, nn);  // Gram, then E
  double* X = alloc_core(c, nn);
  double* W = alloc_core(c, nn);
  double* Y = alloc_core(c, nn);
  double* R1 = alloc_core(c, nn);
  double* Xi = alloc_core(c, nn);  // R1^-1
  double* Q1 = alloc_core(c, (size_t)m * n);
  double* aux = alloc_core(c, 8 + (size_t)kMinusIdentParts + 2 * (size_t)n);
  bool ok = false;
  const double cb = 16.0 * (double)nn;
  const int np = minus_ident_parts(n);
  double* cn2 = aux + 8 + kMinusIdentParts;
  auto gm = [&](int M, int N, int K, const double* Am, long sar, long sac, const double* Bm, long sbr, long sbc,
                double* Cm, long scr, long scc)
  { gemm_raw(c, M, N, K, Am, sar, sac, Bm, sbr, sbc, Cm, scr, scc); };
  auto ew = [&](auto&& launch)
  {
    prof_scope ps(c, TTA_K_OTHER, 0.0, cb);
    launch();
  };
  auto release = [&]()
  {
    for (double* p : {G, X, W, Y, R1, Xi, Q1, aux}) c->mem.put(p);
  };
  // Where the products split along K, the elementwise steps ride on the
  // sums of their partial products.
  const bool split = gemm_splitk_slices(n, n, m) > 1 && gemm_splitk_slices(n, n, n) > 1;
  try
  {
    gm(n, n, m, A, sac, sar, A, sar, sac, G, n, 1);  // G = A^T A
    chol_inv_blocked(c, G, n, R1, Xi, aux);           // R1, Xi = R1^-1
    gm(m, n, n, A, sar, sac, Xi, n, 1, Q1, 1, m);    // Q1 = A R1^-1
    if (split)
    {
      splitk_epi ep;
      ep.mode = SPLITK_EPI_GRAM;
      ep.part = aux + 8;
      gemm_go_epi(c, n, n, m, Q1, m, 1, Q1, 1, m, G, n, 1, 1.0, ep);  // G = E, max |E|
      splitk_epi pp;
      pp.mode = SPLITK_EPI_POLAR;
      pp.E = G;
      pp.out2 = W;
      gemm_go_epi(c, n, n, n, G, n, 1, G, n, 1, X, n, 1, 1.0, pp);    // X = P, W = Pi
    }
    else
    {
      gm(n, n, m, Q1, m, 1, Q1, 1, m, G, n, 1);                       // G = Q1^T Q1
      ew([&] { launch_minus_ident(G, n, aux + 8, c->stream); });     // G = E, max |E|
      gm(n, n, n, G, n, 1, G, n, 1, Y, n, 1);                         // Y = E E
      ew([&] { launch_polar_series(G, Y, X, W, n, c->stream); });    // X = P, W = Pi
    }
    // Q = Q1 P, Rf = Pi R1 and (for the certificate) Y = R^-1 = R1^-1 P.
    // In a speculative sweep without a certificate, the acceptance decision
    // rides on this launch.
    gemm_multi dec;
    dec.dinfo = aux;
    dec.dpart = aux + 8;
    dec.dnp = np;
    dec.dn = n;
    dec.dthr = 1.0e-5;
    dec.dflag = spec_flag(c);
    gemm_multi_go(c,
                  {gdesc(Q ? m : 0, n, n, Q1, 1, m, X, n, 1, Q, sqr, sqc), gdesc(n, n, n, W, n, 1, R1, n, 1, Rf, n, 1),
                   gdesc(cert ? n : 0, n, n, Xi, n, 1, X, n, 1, Y, n, 1)},
                  c->spec && !cert ? &dec : nullptr);
    if (c->spec)
    {
      // Decided on the device; a failed acceptance undoes the whole sweep.
      // With a certificate, the decision and the column norms of R and
      // R^-1 share one launch after the products (which the norms read).
      if (cert)
      {
        // Kept until the sweep's certificates are decided together.
        double* cd = alloc_core(c, 2 * (size_t)n);
        c->spec_cert_bufs.push_back(cd);
        ew([&] { launch_spec_cholqr_norms(aux, aux + 8, np, n, 1.0e-5, spec_flag(c), Rf, Y, cd, c->stream); });
        cert->dev = cd;
      }
      release();
      c->cholqr_tried += 1;
      c->cholqr_used += 1;
      return true;
    }
    if (cert)
    {
      ew([&] { launch_col_norms(Rf, n, n, cn2, c->stream); });
      ew([&] { launch_col_norms(Y, n, n, cn2 + n, c->stream); });
    }
    std::vector<double> h(8 + (size_t)kMinusIdentParts + 2 * (size_t)n);
    const size_t nh = cert ? h.size() : 8 + (size_t)np;
    hip_check(hipMemcpyAsync(h.data(), aux, nh * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    double emax = 0.0;
    for (int b = 0; b < np; ++b) emax = std::max(emax, h[8 + b]);
    const bool pass1 = h[0] == 0.0 && std::isfinite(h[1]) && std::isfinite(h[2]);
    ok = pass1 && emax * n <= 1.0e-4;
    if (ok && emax * n > 1.0e-5)
    {
      // Triangular form (the polar one is not accurate enough here).
      c->cholqr_triangular += 1;
      ew([&] { launch_phi(G, nullptr, X, n, c->stream); });          // X = Phi(E)
      for (int it = 0; it < 3; ++it)
      {
        gm(n, n, n, X, 1, n, X, n, 1, W, n, 1);                       // W = X^T X
        ew([&] { launch_phi(G, W, X, n, c->stream); });              // X = Phi(E - W)
      }
      ew([&] { launch_ident_axpy(X, W, n, 1.0, 1.0, c->stream); });  // W = R2 = I + X
      ew([&] { launch_ident_axpy(X, G, n, 1.0, -1.0, c->stream); }); // G = Y0 = I - X
      gm(n, n, n, W, n, 1, G, n, 1, X, n, 1);                         // X = R2 Y0
      ew([&] { launch_ident_axpy(X, X, n, 2.0, -1.0, c->stream); }); // X = 2 I - R2 Y0
      gm(n, n, n, G, n, 1, X, n, 1, Y, n, 1);                         // Y = R2^-1
      if (Q) gm(m, n, n, Q1, 1, m, Y, n, 1, Q, sqr, sqc);            // Q = Q1 R2^-1
      gm(n, n, n, W, n, 1, R1, n, 1, Rf, n, 1);                       // Rf = R2 R1
      if (cert)
      {
        gm(n, n, n, Xi, n, 1, Y, n, 1, X, n, 1);                      // X = R^-1
        ew([&] { launch_col_norms(Rf, n, n, cn2, c->stream); });
        ew([&] { launch_col_norms(X, n, n, cn2 + n, c->stream); });
        hip_check(hipMemcpyAsync(h.data() + 8 + kMinusIdentParts, cn2, 2 * (size_t)n * 8, hipMemcpyDeviceToHost,
                                 c->stream),
                  "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
      }
    }
    if (ok)
    {
      if (cert)
      {
        const double* hn = h.data() + 8 + kMinusIdentParts;
        cert->rf2 = cert->xf2 = 0.0;
        for (int j = 0; j < n; ++j)
        {
          cert->rf2 += hn[j] * hn[j];
          cert->xf2 += hn[n + j] * hn[n + j];
        }
        cert->have = true;
      }
    }
  }
  catch (...)
  {
    release();
    throw;
  }
  release();
  c->cholqr_tried += 1;
  c->cholqr_used += ok ? 1 : 0;
  return ok;
}

// CholQR2 in place on a column-major A (leading dimension lda): R to A's
// leading n x n block (all of it), Q to Q (leading dimension ldq).
bool cholqr2(tta_ctx* c, double* A, int lda, int m, int n, double* Q, int ldq, rinv_norms* cert = nullptr)
{
  double* Rf = alloc_core(c, (size_t)n * n);
  bool ok = false;
  try
  {
    ok = cholqr2_strided(c, A, 1, lda, m, n, Q, 1, ldq, Rf, cert);
    if (ok)
    {
      prof_scope ps(c, TTA_K_OTHER, 0.0, 16.0 * n * n);
      launch_transpose(Rf, n, n, n, A, lda, c->stream);
    }
  }
  catch (...)
  {
    c->mem.put(Rf);
    throw;
  }
  c->mem.put(Rf);
  return ok;
}

// Whether qr_factor(..., cholqr = true) tries CholQR2 on an m x n matrix.
bool cholqr_applies(const tta_ctx* c, int m, int n) { return c->cholqr && m >= n && n >= 32 && n <= chol_max_n(); }

// QR of a column-major (m x n) matrix in place (R in the leading rows),
// thin Q (m x k) into Q when non-null.  With cholqr, matrices with m >= n
// and 32 <= n <= chol_max_n() go through CholQR2 first; otherwise, and
// whenever CholQR2 declines, blocked compact-WY Householder (any m).
// Returns true when CholQR2 made the factorization: R then fills A's
// leading n x n block and need not be triangular (read it with
// launch_extract_upper(..., full = true)).  Householder leaves R in the upper
// triangle with its reflectors below the diagonal and tau.
// With cert non-null, a CholQR2 factorization also reports the norms of R and
// R^-1 there (cert->have); a Householder one leaves cert->have false.
// Two-level blocked Householder QR of one tall matrix (column-major m x n,
// in place, thin Q into Q when non-null): the 16-column panels of each
// 64-column block update only the rest of that block
// (launch_qr_update_block); the trailing matrix then takes the block's
// reflectors at once, C <- (I - V T^T V^T) C as W = V^T C, W <- T^T W and
// C -= V W on the matrix cores (gemm_go, split along K where the output is
// small), and Q is formed from the identity block by block in reverse order
// with C <- (I - V T V^T) C.  Same reflectors, R and Q as the one-level
// form up to rounding.
void qr_two_level(tta_ctx* c, const qr_batch& b)
{
  const int m = b.one.m, n = b.one.n, lda = b.one.lda, k = std::min(m, n);
  constexpr int NB = 64;
  const int nblk = (k + NB - 1) / NB;
  double* V = alloc_core(c, (size_t)m * NB);
  double* G = alloc_core(c, (size_t)NB * NB);
  double* Tb = alloc_core(c, (size_t)nblk * NB * NB);
  double* W = alloc_core(c, (size_t)NB * std::max(n, k));
  double* W2 = alloc_core(c, (size_t)NB * std::max(n, k));
  auto release = [&]() {
    c->mem.put(V);
    c->mem.put(G);
    c->mem.put(Tb);
    c->mem.put(W);
    c->mem.put(W2);
  };
  try
  {
    // C (len x nc, column-major, ldc) <- (I - V op(T) V^T) C; op(T) = T^T
    // when trans, else T.
    auto apply = [&](int len, int nb, const double* T, bool trans, double* C, int ldc, int nc) {
      gemm_desc g;
      g.M = nb, g.N = nc, g.K = len;  // W = V^T C
      g.A = V, g.sar = len, g.sac = 1;
      g.B = C, g.sbr = 1, g.sbc = ldc;
      g.C = W, g.scr = nc, g.scc = 1;
      gemm_go(c, g);
      g = gemm_desc();
      g.M = nb, g.N = nc, g.K = nb;  // W2 = op(T) W
      g.A = T, g.sar = trans ? 1 : nb, g.sac = trans ? nb : 1;
      g.B = W, g.sbr = nc, g.sbc = 1;
      g.C = W2, g.scr = nc, g.scc = 1;
      gemm_go(c, g);
      g = gemm_desc();
      g.M = len, g.N = nc, g.K = nb;  // C -= V W2
      g.A = V, g.sar = 1, g.sac = len;
      g.B = W2, g.sbr = nc, g.sbc = 1;
      g.C = C, g.scr = 1, g.scc = ldc;
      g.alpha = -1.0;
      g.beta = 1.0;
      gemm_go(c, g);
    };
    // V and T of the block at column j0 (nb columns).
    auto block_vt = [&](int j0, int nb, double* T) {
      const int len = m - j0;
      launch_qr_vblock(b.one.A, lda, j0, len, nb, V, c->stream);
      gemm_desc g;
      g.M = nb, g.N = nb, g.K = len;  // G = V^T V (row-major)
      g.A = V, g.sar = len, g.sac = 1;
      g.B = V, g.sbr = 1, g.sbc = len;
      g.C = G, g.scr = nb, g.scc = 1;
      gemm_go(c, g);
      launch_qr_tblock(G, b.one.tau + j0, nb, T, c->stream);
    };
    for (int J0 = 0; J0 < k; J0 += NB)
    {
      const int nb = std::min(NB, k - J0);
      for (int j0 = J0; j0 < J0 + nb; j0 += 16)
      {
        launch_qr_panel_one(b, j0, c->stream);
        launch_qr_update_block(b, j0, J0 + nb, c->stream);
      }
      double* T = Tb + (size_t)(J0 / NB) * NB * NB;
      block_vt(J0, nb, T);
      if (J0 + nb < n) apply(m - J0, nb, T, true, b.one.A + (size_t)(J0 + nb) * lda + J0, lda, n - J0 - nb);
    }
    if (b.one.Q)
    {
      launch_qr_eye(b, c->stream);
      for (int J0 = ((k - 1) / NB) * NB; J0 >= 0; J0 -= NB)
      {
        const int nb = std::min(NB, k - J0);
        launch_qr_vblock(b.one.A, lda, J0, m - J0, nb, V, c->stream);
        apply(m - J0, nb, Tb + (size_t)(J0 / NB) * NB * NB, false, b.one.Q + (size_t)J0 * b.one.ldq + J0, b.one.ldq,
              k - J0);
      }
    }
    hip_check(hipGetLastError(), "qr_two_level");
  }
  catch (...)
  {
    release();
    throw;
  }
  release();
}

bool qr_factor(tta_ctx* c, double* A, int lda, int m, int n, double* tau, double* Q, int ldq, bool cholqr = false,
               rinv_norms* cert = nullptr)
{
  if (cert)
  {
    cert->have = false;
    cert->dev = nullptr;
  }
  if (cholqr && cholqr_applies(c, m, n) && cholqr2(c, A, lda, m, n, Q, ldq, cert)) return true;
  prof_scope ps(c, TTA_K_QR, qr_flops(m, n), 8.0 * ((double)m * n + (double)m * std::min(m, n)));
  const int k = std::min(m, n);
  qr_batch b;
  b.one.A = A;
  b.one.lda = lda;
  b.one.m = m;
  b.one.n = n;
  b.one.tau = tau;
  b.one.T = ctx_ws(c, WS_T, (size_t)((k + 15) / 16) * 256 + 1);
  b.one.Q = Q;
  b.one.ldq = ldq;
  b.max_m = m;
  b.max_n = n;
  // Tall matrices with many columns: the trailing updates dominate, and the
  // two-level form reads the trailing matrix a quarter as often.
  if (m >= 2048 && std::min(m, n) >= 256)
    qr_two_level(c, b);
  else
    launch_qr_blocked(b, c->stream);
  return false;
}

void run_gemm(tta_ctx* c, const gemm_desc& g) { gemm_go(c, g); }

int read_int(tta_ctx* c, const int* dev)
{
  hip_check(hipMemcpyAsync(c->h_int, dev, sizeof(int), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return *c->h_int;
}

double read_double(tta_ctx* c, const double* dev)
{
  hip_check(hipMemcpyAsync(c->h_dbl, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return *c->h_dbl;
}

// Small device scalars live in WS_SMALL: [0] norm, [1] delta, [2..] ints.
double* small_dbl(tta_ctx* c) { return ctx_ws(c, WS_SMALL, 64); }
int* small_int(tta_ctx* c) { return reinterpret_cast<int*>(small_dbl(c) + 16); }

// Frobenius norm of n doubles at x; also writes delta = eps*norm*inv to small[1].
double frob_norm_dev(tta_ctx* c, const double* x, size_t n, double eps, double inv)
{
  double* part = ctx_ws(c, WS_PART, 1024);
  double* sm = small_dbl(c);
  launch_frob_norm(x, (long)n, part, 1024, sm, eps, inv, sm + 1, c->stream);
  return read_double(c, sm);
}

// ---------------------------------------------------------------------------
// Chained sweep blocks.
//
// A run of consecutive QR steps of a sweep, each of them CholQR2's shape, is
// factored with the Choleskys off the sweep's critical path.  Right sweep
// (dir 0; steps kb, kb-1, .., ka): step k factors the right unfolding of
// core k after core k+1's factor R_{k+1} has been absorbed, whose Gram is
//   G_k = sum_i C_k[:, i, :] G_{k+1} C_k[:, i, :]^T     (G_{k+1} = R^T R),
// so the Grams follow from the cores alone, one after another, by two
// products each, with no factorization between them.  Left sweep (dir 1;
// steps ka .. kb, left unfoldings) is the mirror:
//   H_k = sum_i C_k[:, i, :]^T H_{k-1} C_k[:, i, :].
// Then, all steps at once:
//   pass 1  S_k = chol(G_k) and S_k^-1 (batched blocked Cholesky), and the
//           cores N_k = S_k^-T C_k S_{k+1}^T (dir 0) or S_{k-1} C_k S_k^-1
//           (dir 1), whose unfoldings are orthonormal up to the first pass's
//           error E ~ u kappa^2;
//   pass 2  the same chain of Grams over the N_k (G2_k = I + E_k), then
//           P_k = (I + E_k)^(-1/2) and Pi_k = (I + E_k)^(1/2) by their
//           second-order series (CholQR2's polar form, exact to O(|E|^3)),
//           and the final cores P_k N_k Pi_{k+1} (dir 0) or Pi_{k-1} N_k P_k
//           (dir 1); the factor Rf = Pi S of the block's last step is
//           absorbed into the core past the block.
// In exact arithmetic S_k is the R of step k's QR and E = 0, so the result
// is the sequential sweep's.  Rounding: the chain's G_k and the factor
// S_{k+1} it is used with are consistent to the Cholesky's backward error,
// as in the sequential CholQR's Gram of the absorbed core, so E has the
// sequential first pass's size in the measure u ||C_k||^2 ||G_{k+1}|| /
// lambda_min(G_k), which can exceed the sequential u kappa(A_k)^2 where
// graded spectra of neighbouring cores meet.  So the block is accepted on
// the host (one synchronisation per block): every Cholesky held and every
// n max|E| <= 1e-5 (the polar form's range); otherwise nothing has been
// written and its steps run one by one.  For the SVD sweep, every step's
// full-rank certificate (column norms of Rf = Pi S and Rf^-1 = S^-1 P) is
// recorded for the speculative sweep's end, like a single step's; blocks
// run only inside speculative sweeps.
// ---------------------------------------------------------------------------
constexpr int kChainMin = 2;

// Host copy of the device sum item (launch_chain_sum).
struct chain_sum_item_h
{
  const double* part;
  double* out;
  int s;
  long n;
};
// Orthonormality to which the cores of a sweep that only feeds the next one
// (the QR sweep of round) are left after one pass: the next sweep's singular
// values then carry a relative error of at most about 5e-11, far inside the
// rank rule's and the certificate's margins.
constexpr double kChainSkipTol = 1.0e-10;

bool cert_ok(int max_rank, int q);

// One batched launch of independent GEMMs (grouped kernel, descriptors
// uploaded through the arena).
void gemm_batch_go(tta_ctx* c, arena& ar, const std::vector<gemm_desc>& gs)
{
  if (gs.empty()) return;
  const credit cr = gemm_credit(gs.data(), gs.size());
  prof_scope ps(c, TTA_K_GEMM_GROUPED, cr.flops, cr.bytes);
  launch_gemm_grouped(ar, gs, c->stream);
}

// Whether step k of a sweep in direction dir may join a chained block: a
// speculative sweep, CholQR2's shape, a core to absorb into, and (with
// cert) a certificate that max_rank cannot overrule.
bool chain_step_ok(const tta_ctx* c, const tta_tt* t, int k, int dir, bool cert, int max_rank)
{
  // SVD-sweep blocks record their certificates for a speculative sweep's
  // end; QR-sweep blocks are decided on the host and need no speculation.
  if (!c->chain || (cert && !c->spec)) return false;
  const int d = t->d();
  if (dir == 0 ? k < 1 : k > d - 2) return false;
  const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
  const int m = dir == 0 ? n * rp : r * n, q = dir == 0 ? r : rp;
  if (!cholqr_applies(c, m, q)) return false;
  // A QR step with no more rows than columns takes the identity basis
  // (right_qr_step / left_qr_step); the SVD steps keep square ones, whose
  // certificate needs the factorization.
  if (!cert && m <= q) return false;
  if (cert && (!c->rank_certificate || !cert_ok(max_rank, q) || q > kRinvCertMaxQ)) return false;
  return true;
}

// The chained block of steps ka .. kb (see above).  With cert, each step's
// full-rank certificate for delta is recorded for the sweep's end.
// skip_tol > 0 (for sweeps whose cores only serve the next sweep): when the
// first pass is already orthonormal to skip_tol, the second is left out.
// Its E_k obey E_k = F_k + sum_i N_i E_{k+1} N_i^T with F_k = N_k N_k^T - I
// (the core's own Gram), and ||sum_i N_i E N_i^T|| <= ||E|| (1 + ||F_k||),
// so ||E_k|| <= sum_j n_j max|F_j| (to first order) for every k; the host
// reads that sum (one synchronisation) and skips the second pass when it is
// at most skip_tol.
bool chain_block(tta_ctx* c, tta_tt* t, int ka, int kb, int dir, bool cert, double delta, double skip_tol = 0.0)
{
  const int L = kb - ka + 1;
  std::vector<int> ks(L), qs(L);
  for (int p = 0; p < L; ++p)
  {
    ks[p] = dir == 0 ? kb - p : ka + p;
    qs[p] = dir == 0 ? t->ranks[ks[p]] : t->ranks[ks[p] + 1];
  }
  arena ar(c);
  std::vector<chain_prob> pr(L);
  int max_n = 0;
  size_t max_core = 0;
  for (int p = 0; p < L; ++p)
  {
    const size_t qq = (size_t)qs[p] * qs[p];
    chain_prob& b = pr[p];
    b.n = qs[p];
    for (double** m : {&b.G, &b.R, &b.X, &b.E, &b.V, &b.P, &b.Pi, &b.Rf, &b.Y}) *m = ar.alloc<double>(qq);
    b.S = ar.alloc<double>(chol_step_scratch(qs[p]));
    b.info = ar.alloc<double>(8);
    b.part = ar.alloc<double>(kChainParts);
    if (cert)
    {
      // Read by the sweep's certificate launch at its end.
      b.norms = alloc_core(c, 2 * (size_t)qs[p]);
      c->spec_cert_bufs.push_back(b.norms);
    }
    max_n = std::max(max_n, qs[p]);
    max_core = std::max(max_core, core_numel(t, ks[p]));
  }
  double* T = ar.alloc<double>(max_core);
  std::vector<double*> C(L), N(L), U(L), O(L);
  // N from the pool: with the second pass skipped they become the cores.
  for (int p = 0; p < L; ++p)
  {
    C[p] = t->cores[ks[p]];
    N[p] = alloc_core(c, core_numel(t, ks[p]));
    U[p] = ar.alloc<double>(core_numel(t, ks[p]));
  }
  // The chain of Grams over the cores X[p] into out(p).  Each step is two
  // launches on 32 x 32 tiles: the product with the previous Gram, read as
  // the sum of its split-K partials, and the new Gram's partials; the Grams
  // themselves are summed for all steps at once afterwards.
  auto chain = [&](const std::vector<double*>& X, double* chain_prob::*out)
  {
    std::vector<double*> gp(L);
    std::vector<int> gs(L);
    std::vector<chain_sum_item_h> items(L);
    long max_qq = 0;
    for (int p = 0; p < L; ++p)
    {
      const int k = ks[p], r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      const long w = (long)n * rp;
      const int q = qs[p];
      const double* src = X[p];
      if (p > 0)
      {
        const int qp = qs[p - 1];
        gemm_desc g = dir == 0 ? gdesc(r * n, rp, rp, X[p], rp, 1, gp[p - 1], rp, 1, T, rp, 1)   // X G_{k+1}
                               : gdesc(r, n * rp, r, gp[p - 1], r, 1, X[p], w, 1, T, w, 1);     // H_{k-1} X
        prof_scope ps(c, TTA_K_GEMM_CHAIN, 2.0 * g.M * g.N * (double)g.K,
                      8.0 * ((double)g.M * g.K + (double)gs[p - 1] * qp * qp + (double)g.M * g.N));
        if (dir == 0)
          launch_gemm_chain(g, 1, 0, gs[p - 1], (long)qp * qp, 1, nullptr, c->stream);
        else
          launch_gemm_chain(g, gs[p - 1], (long)qp * qp, 1, 0, 1, nullptr, c->stream);
        src = T;
      }
      gemm_desc g = dir == 0 ? gdesc(q, q, n * rp, src, w, 1, X[p], 1, w, nullptr, q, 1)      // (.) X^T
                             : gdesc(q, q, r * n, src, 1, rp, X[p], rp, 1, nullptr, q, 1);  // (.)^T X, left unfoldings
      const int sl = chain_gemm_slices(g.M, g.N, g.K);
      gp[p] = ar.alloc<double>((size_t)sl * q * q);
      prof_scope ps(c, TTA_K_GEMM_CHAIN, 2.0 * g.M * g.N * (double)g.K,
                    8.0 * (2.0 * (double)g.M * g.K + (double)sl * q * q));
      gs[p] = launch_gemm_chain(g, 1, 0, 1, 0, sl, gp[p], c->stream);
      items[p] = {gp[p], pr[p].*out, gs[p], (long)q * q};
      max_qq = std::max(max_qq, (long)q * q);
    }
    prof_scope ps(c, TTA_K_OTHER, 0.0, 0.0);
    launch_chain_sum(ar.upload(items), L, max_qq, c->stream);
  };
  chain(C, &chain_prob::G);
  chain_prob* dps = ar.upload(pr);
  const int B = potrf_block();
  {
    double fl = 0.0, by = 0.0;
    for (int q : qs)
    {
      const double jb = std::min(B, q);
      fl += 2.0 * jb * jb * jb / 3.0;
      by += 24.0 * jb * jb;
    }
    prof_scope ps(c, TTA_K_CHOL, fl, by);
    launch_chain_potrf(dps, L, c->stream);
  }
  for (int j0 = 0; j0 < max_n; j0 += B)
  {
    // Work of each problem's block step, as chol_inv_blocked credits it.
    double fl = 0.0, by = 0.0;
    for (int q : qs)
    {
      if (j0 >= q) continue;
      const double jb = std::min(B, q - j0), w = q - j0 - jb, nb = std::min<double>(B, w);
      fl += 2.0 * jb * jb * w + 2.0 * w * w * jb + 2.0 * j0 * jb * (j0 + jb) + 2.0 * nb * nb * nb / 3.0;
      by += 8.0 * (2.0 * jb * w + 2.0 * w * w + 2.0 * j0 * j0 + 3.0 * j0 * jb + 3.0 * nb * nb);
    }
    prof_scope ps(c, TTA_K_CHOL_STEP, fl, by);
    launch_chain_chol_step(dps, L, max_n, j0, c->stream);
  }
  // Pass 1 cores.
  {
    std::vector<gemm_desc> g1, g2;
    for (int p = 0; p < L; ++p)
    {
      const int k = ks[p], r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      const long w = (long)n * rp;
      const double* src = C[p];
      if (dir == 0)
      {
        if (p > 0)
        {
          g1.push_back(gdesc(r * n, rp, rp, C[p], rp, 1, pr[p - 1].R, 1, rp, U[p], rp, 1));  // C S_{k+1}^T
          src = U[p];
        }
        g2.push_back(gdesc(r, n * rp, r, pr[p].X, 1, r, src, w, 1, N[p], w, 1));  // S_k^-T (.)
      }
      else
      {
        if (p > 0)
        {
          g1.push_back(gdesc(r, n * rp, r, pr[p - 1].R, r, 1, C[p], w, 1, U[p], w, 1));  // S_{k-1} C
          src = U[p];
        }
        g2.push_back(gdesc(r * n, rp, rp, src, rp, 1, pr[p].X, rp, 1, N[p], rp, 1));  // (.) S_k^-1
      }
    }
    gemm_batch_go(c, ar, g1);
    gemm_batch_go(c, ar, g2);
  }
  // The block's status on the host (one synchronisation): E <- E - I with
  // its partial maxima, then the sum and the largest n max|E| over the
  // steps, infinity on a Cholesky breakdown.
  auto status = [&](double& sum, double& most)
  {
    double* st = small_dbl(c) + 10;
    {
      prof_scope ps(c, TTA_K_OTHER, 0.0, 0.0);
      launch_chain_minus_eye(dps, L, max_n, c->stream);
      launch_chain_bound(dps, L, st, c->stream);
    }
    hip_check(hipMemcpyAsync(c->h_dbl, st, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    sum = c->h_dbl[0];
    most = c->h_dbl[1];
  };
  // A rejected block leaves the TT as it was; its steps then run one by
  // one.
  auto reject = [&]()
  {
    for (int p = 0; p < L; ++p) c->mem.put(N[p]);
    c->chain_rejected += 1;
    return false;
  };
  bool skip2 = false;
  if (skip_tol > 0.0)
  {
    // F_k (into E), then the bound on every E_k.
    std::vector<gemm_desc> gf;
    for (int p = 0; p < L; ++p)
    {
      const int k = ks[p], r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
      const long w = (long)n * rp;
      if (dir == 0)
        gf.push_back(gdesc(r, r, n * rp, N[p], w, 1, N[p], 1, w, pr[p].E, r, 1));  // N N^T (right unfoldings)
      else
        gf.push_back(gdesc(rp, rp, r * n, N[p], 1, rp, N[p], rp, 1, pr[p].E, rp, 1));  // N^T N (left unfoldings)
    }
    gemm_batch_go(c, ar, gf);
    double sum = 0.0, most = 0.0;
    status(sum, most);
    if (!(sum < INFINITY)) return reject();
    skip2 = sum <= skip_tol;
  }
  if (skip2)
  {
    // The first pass's cores are the block's cores; the last step's factor
    // is its Cholesky factor.
    const double* Rs = pr[L - 1].R;
    const int q = qs[L - 1];
    if (dir == 0)
    {
      const int kp = ka - 1, rows_p = t->ranks[kp] * t->modes[kp];
      double* Pn = alloc_core(c, (size_t)rows_p * q);
      gemm_go(c, gdesc(rows_p, q, q, t->cores[kp], q, 1, Rs, 1, q, Pn, q, 1));  // core S^T
      replace_core(c, t, kp, Pn);
    }
    else
    {
      const int kn = kb + 1, cols = t->modes[kn] * t->ranks[kn + 1];
      double* Nn = alloc_core(c, (size_t)q * cols);
      gemm_go(c, gdesc(q, cols, q, Rs, q, 1, t->cores[kn], cols, 1, Nn, cols, 1));  // S core
      replace_core(c, t, kn, Nn);
    }
    for (int p = 0; p < L; ++p) replace_core(c, t, ks[p], N[p]);
    c->cholqr_tried += L;
    c->cholqr_used += L;
    c->chain_blocks += 1;
    c->chain_steps += L;
    c->chain_one_pass += 1;
    return true;
  }
  // Pass 2, accepted when every step's n max|E| is within the polar form's
  // range (as a single CholQR2's).
  chain(N, &chain_prob::E);
  {
    double sum = 0.0, most = 0.0;
    status(sum, most);
    if (!(most <= 1.0e-5)) return reject();
  }
  {
    std::vector<gemm_desc> gv;
    for (int p = 0; p < L; ++p)
    {
      const int q = qs[p];
      gv.push_back(gdesc(q, q, q, pr[p].E, q, 1, pr[p].E, q, 1, pr[p].V, q, 1));  // V = E E
    }
    gemm_batch_go(c, ar, gv);
  }
  {
    prof_scope ps(c, TTA_K_OTHER, 0.0, 0.0);
    launch_chain_polar(dps, L, max_n, c->stream);
  }
  for (int p = 0; p < L; ++p) O[p] = alloc_core(c, core_numel(t, ks[p]));
  {
    std::vector<gemm_desc> g1, g2, gr;
    for (int p = 0; p < L; ++p)
    {
      const int k = ks[p], r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1], q = qs[p];
      const long w = (long)n * rp;
      const double* src = N[p];
      if (dir == 0)
      {
        if (p > 0)
        {
          g1.push_back(gdesc(r * n, rp, rp, N[p], rp, 1, pr[p - 1].Pi, rp, 1, U[p], rp, 1));  // N Pi_{k+1}
          src = U[p];
        }
        g2.push_back(gdesc(r, n * rp, r, pr[p].P, r, 1, src, w, 1, O[p], w, 1));  // P_k (.)
      }
      else
      {
        if (p > 0)
        {
          g1.push_back(gdesc(r, n * rp, r, pr[p - 1].Pi, r, 1, N[p], w, 1, U[p], w, 1));  // Pi_{k-1} N
          src = U[p];
        }
        g2.push_back(gdesc(r * n, rp, rp, src, rp, 1, pr[p].P, rp, 1, O[p], rp, 1));  // (.) P_k
      }
      if (cert || p == L - 1)
      {
        gr.push_back(gdesc(q, q, q, pr[p].Pi, q, 1, pr[p].R, q, 1, pr[p].Rf, q, 1));  // Rf = Pi S
        if (cert) gr.push_back(gdesc(q, q, q, pr[p].X, q, 1, pr[p].P, q, 1, pr[p].Y, q, 1));  // Rf^-1 = S^-1 P
      }
    }
    gemm_batch_go(c, ar, g1);
    gemm_batch_go(c, ar, g2);
    gemm_batch_go(c, ar, gr);
  }
  if (cert)
  {
    prof_scope ps(c, TTA_K_OTHER, 0.0, 0.0);
    launch_chain_norms(dps, L, max_n, c->stream);
  }
  // The last step's factor into the core past the block.
  const double* Rf = pr[L - 1].Rf;
  const int q = qs[L - 1];
  if (dir == 0)
  {
    const int kp = ka - 1, rows_p = t->ranks[kp] * t->modes[kp];
    double* Pn = alloc_core(c, (size_t)rows_p * q);
    gemm_go(c, gdesc(rows_p, q, q, t->cores[kp], q, 1, Rf, 1, q, Pn, q, 1));  // core Rf^T
    replace_core(c, t, kp, Pn);
  }
  else
  {
    const int kn = kb + 1, cols = t->modes[kn] * t->ranks[kn + 1];
    double* Nn = alloc_core(c, (size_t)q * cols);
    gemm_go(c, gdesc(q, cols, q, Rf, q, 1, t->cores[kn], cols, 1, Nn, cols, 1));  // Rf core
    replace_core(c, t, kn, Nn);
  }
  for (int p = 0; p < L; ++p)
  {
    replace_core(c, t, ks[p], O[p]);
    c->mem.put(N[p]);
    if (cert)
    {
      c->spec_certs.push_back({pr[p].norms, qs[p], 0, delta});
      c->cert_tried += 1;
      c->cert_held += 1;
    }
  }
  c->cholqr_tried += L;
  c->cholqr_used += L;
  c->chain_blocks += 1;
  c->chain_steps += L;
  return true;
}

// The end of the run of chainable steps that starts at step k of a sweep in
// direction dir (dir 0 runs downward), or k itself when the run is shorter
// than kChainMin.
int chain_run_end(const tta_ctx* c, const tta_tt* t, int k, int dir, bool cert, int max_rank)
{
  const int d = t->d();
  int e = k;
  while (true)
  {
    const int nx = dir == 0 ? e - 1 : e + 1;
    if (nx < 0 || nx > d - 1 || !chain_step_ok(c, t, nx, dir, cert, max_rank)) break;
    e = nx;
  }
  return std::abs(e - k) + 1 >= kChainMin ? e : k;
}

// ---------------------------------------------------------------------------
// right_qr_step (include/core/detail/orthogonalize.hpp:63-100): QR of the
// column-major (n*rp x r) view of core k (its own bytes, factored in place),
// thin Q -> new core (r_eff, n, rp), L = R^T absorbed into core k-1.  With
// absorb false core k-1 is left as it was (its right rank then no longer
// matches ranks[k]): for callers that overwrite it next.
// ---------------------------------------------------------------------------
void right_qr_step(tta_ctx* c, tta_tt* t, int k, bool absorb = true)
{
  const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
  const int m = n * rp;
  const int r_eff = std::min(m, r);
  if (absorb && m <= r)
  {
    // The unfolding's r columns span at most R^m, and Q = I_m is an
    // orthonormal basis of R^m: R = Q^T M^T = M^T, so core k-1 takes core
    // k whole (one product) and core k becomes the identity, r_eff = m.
    const int rows_p = t->ranks[k - 1] * t->modes[k - 1];
    double* Pn = alloc_core(c, (size_t)rows_p * m);
    gemm_rm(c, rows_p, m, r, t->cores[k - 1], r, t->cores[k], m, Pn, m);
    double* Cn = alloc_core(c, (size_t)m * m);
    {
      prof_scope ps(c, TTA_K_OTHER, 0.0, 8.0 * (double)m * m);
      launch_eye(Cn, m, m, c->stream);
    }
    replace_core(c, t, k - 1, Pn);
    replace_core(c, t, k, Cn);
    t->ranks[k] = m;
    return;
  }
  double* A = t->cores[k];
  double* tau = ctx_ws(c, WS_F, (size_t)std::max(r_eff, 1));
  double* Qn = alloc_core(c, (size_t)r_eff * m);
  const bool try_chol = absorb && cholqr_applies(c, m, r);
  if (try_chol)
  {
    // CholQR2 leaves core k as it is and gives R row-major: core k-1 takes
    // L = R^T by the GEMM's strides, with no extraction.
    double* R = ctx_ws(c, WS_X, (size_t)r * r);
    if (cholqr2_strided(c, A, 1, m, m, r, Qn, 1, m, R))
    {
      const int rows_p = t->ranks[k - 1] * t->modes[k - 1];
      double* Pn = alloc_core(c, (size_t)rows_p * r);
      gemm_go(c, gdesc(rows_p, r, r, t->cores[k - 1], r, 1, R, 1, r, Pn, r, 1));  // core k-1 R^T
      replace_core(c, t, k - 1, Pn);
      replace_core(c, t, k, Qn);
      return;
    }
  }
  const bool full = qr_factor(c, A, m, m, r, tau, Qn, m, !try_chol);
  if (!absorb)
  {
    replace_core(c, t, k, Qn);
    t->ranks[k] = r_eff;
    return;
  }
  double* L = ctx_ws(c, WS_X, (size_t)r * r_eff);
  launch_extract_upper(A, m, r_eff, r, L, 1, r_eff, c->stream, full);
  const int rows_p = t->ranks[k - 1] * t->modes[k - 1];
  double* Pn = alloc_core(c, (size_t)rows_p * r_eff);
  gemm_rm(c, rows_p, r_eff, r, t->cores[k - 1], r, L, r_eff, Pn, r_eff);
  replace_core(c, t, k - 1, Pn);
  replace_core(c, t, k, Qn);
  t->ranks[k] = r_eff;
}

// left_qr_step (orthogonalize.hpp:35-60): QR of left_unfold(core k), R
// absorbed into core k+1 (skipped with absorb false, as above).
void left_qr_step(tta_ctx* c, tta_tt* t, int k, bool absorb = true)
{
  const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
  const int m = r * n;
  const int kk = std::min(m, rp);
  if (absorb && m <= rp)
  {
    // As in right_qr_step: Q = I_m, R = the left unfolding, so core k+1
    // takes core k whole and core k becomes the identity, rank m.
    const int ncols_n = t->modes[k + 1] * t->ranks[k + 2];
    double* Nn = alloc_core(c, (size_t)m * ncols_n);
    gemm_rm(c, m, ncols_n, rp, t->cores[k], rp, t->cores[k + 1], ncols_n, Nn, ncols_n);
    double* Cn = alloc_core(c, (size_t)m * m);
    {
      prof_scope ps(c, TTA_K_OTHER, 0.0, 8.0 * (double)m * m);
      launch_eye(Cn, m, m, c->stream);
    }
    replace_core(c, t, k + 1, Nn);
    replace_core(c, t, k, Cn);
    t->ranks[k + 1] = m;
    return;
  }
  double* A = ctx_ws(c, WS_A, (size_t)m * rp);
  launch_transpose(t->cores[k], rp, m, rp, A, m, c->stream);
  double* tau = ctx_ws(c, WS_F, (size_t)std::max(kk, 1));
  double* Q = ctx_ws(c, WS_Q, (size_t)m * kk);
  const bool full = qr_factor(c, A, m, m, rp, tau, Q, m, true);
  double* Cn = alloc_core(c, (size_t)m * kk);
  launch_transpose(Q, m, kk, m, Cn, kk, c->stream);
  if (!absorb)
  {
    replace_core(c, t, k, Cn);
    t->ranks[k + 1] = kk;
    return;
  }
  double* R = ctx_ws(c, WS_X, (size_t)kk * rp);
  launch_extract_upper(A, m, kk, rp, R, rp, 1, c->stream, full);
  const int ncols_n = t->modes[k + 1] * t->ranks[k + 2];
  double* Nn = alloc_core(c, (size_t)kk * ncols_n);
  gemm_rm(c, kk, ncols_n, rp, R, rp, t->cores[k + 1], ncols_n, Nn, ncols_n);
  replace_core(c, t, k + 1, Nn);
  replace_core(c, t, k, Cn);
  t->ranks[k + 1] = kk;
}

void flush_spec_certs(tta_ctx* c);

// Thrown when a speculative sweep is found to have failed before its end.
struct spec_abort
{
};

// A step of a speculative sweep that needs its SVD (no certificate applies)
// synchronises anyway, to read the truncation rank; there the decisions
// recorded so far (the sweep's earlier certificates and CholQR2
// acceptances, this step's included) are made at once.  If all held, the
// rest of the sweep runs checked and nothing needs undoing; otherwise the
// sweep is abandoned at once (run_sweep undoes it and runs it checked), so
// no further SVDs are spent on a sweep that will be discarded.
void spec_checkpoint(tta_ctx* c)
{
  if (!c->spec) return;
  flush_spec_certs(c);
  if (read_double(c, spec_flag(c)) != 1.0) throw spec_abort{};
  c->spec = false;
}

// Jacobi SVD of the square q x q column-major X (ld q), rotations into J
// (ld q): X J = W.  Leaves sorted norms s and order perm in scratch and
// returns the truncation rank decided on the device (or q with keep_all).
int square_svd(tta_ctx* c, double* X, int q, double* J, double delta, int max_rank, bool keep_all, double** s_out,
               int** perm_out)
{
  spec_checkpoint(c);
  int* ints = small_int(c);
  double* sv = ctx_ws(c, WS_PART, 1024 + 2 * (size_t)q + 8) + 1024;
  int* perm = reinterpret_cast<int*>(sv + q);
  jac_batch b;
  b.one.X = X;
  b.one.ldx = q;
  b.one.q = q;
  b.one.J = J;
  b.one.ldj = q;
  b.one.s = sv;
  b.one.perm = perm;
  b.one.rank_out = ints;
  b.one.delta = delta;
  b.one.max_rank = max_rank;
  b.one.flag = ints + 2;
  b.one.sweeps = ints + 5;
  {
    prof_scope ps(c, TTA_K_JACOBI, svd_flops(q, q), 8.0 * 4.0 * (double)q * q);
    const int sw = launch_jacobi_svd(b, q, ints + 4, c->h_int + 2, c, svd_flops(q, q), 8.0 * 4.0 * (double)q * q);
    if (sw > 0)
    {
      c->jac_calls += 1;
      c->jac_sweeps += sw;
    }
  }
  *s_out = sv;
  *perm_out = perm;
  return keep_all ? q : read_int(c, ints);
}

// C (M x N) = alpha A B + beta C with element strides: A(i,k) =
// A[i sar + k sac], B(k,j) = B[k sbr + j sbc], C(i,j) = C[i scr + j scc].
void gemm_s(tta_ctx* c, int M, int N, int K, const double* A, long sar, long sac, const double* B, long sbr, long sbc,
            double* C, long scr, long scc, double alpha = 1.0, double beta = 0.0)
{
  gemm_desc g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sar = sar;
  g.sac = sac;
  g.B = B;
  g.sbr = sbr;
  g.sbc = sbc;
  g.C = C;
  g.scr = scr;
  g.scc = scc;
  g.alpha = alpha;
  g.beta = beta;
  run_gemm(c, g);
}

// One Newton-Schulz step toward the nearest orthogonal matrix, J (q x q,
// column-major) <- J (3 I - J^T J) / 2.  The accumulated Jacobi rotations
// leave J orthogonal only to ~1e-13 after thousands of them; the step
// squares that departure, down to the two GEMMs' own rounding.
void polar_correct(tta_ctx* c, double* J, int q)
{
  double* G = alloc_core(c, (size_t)q * q);
  double* J0 = alloc_core(c, (size_t)q * q);
  hip_check(hipMemcpyAsync(J0, J, (size_t)q * q * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
  gemm_s(c, q, q, q, J0, q, 1, J0, 1, q, G, q, 1);               // G = J^T J
  gemm_s(c, q, q, q, J0, 1, q, G, q, 1, J, 1, q, -0.5, 1.5);     // J = 1.5 J - 0.5 J G
  c->mem.put(G);
  c->mem.put(J0);
}

// Certificate that the rank rule keeps all q singular values of the
// upper-triangular R (row-major in `R`): with X = R^-1 formed on the device,
// sigma_min(R) >= 1 / ||X||_F, and the rule (truncate_eps_rank,
// tt_eigen_bridge.hpp:172-197) keeps every value once sigma_min > delta.
// Margins: the condition number ||R||_F ||X||_F is capped at 1e8, so the
// inverse's relative rounding error (<= q u kappa) and the difference between
// this R and the one of any backward-stable factorization (<= q u ||M||,
// i.e. <= 1e-5 sigma_min) stay far below the 1e-3 slack taken on each side.
// Non-finite norms (a zero pivot) fail the certificate.  When the QR formed
// R^-1 itself (known holds its norms), they are taken from there: CholQR2's
// R^-1 = R1^-1 R2^-1 carries the same relative error bound as the triangular
// inverse.
bool certify_full_rank(tta_ctx* c, const double* R, int q, double delta, const rinv_norms* known = nullptr)
{
  if (!c->rank_certificate || q < 1 || q > kRinvCertMaxQ) return false;
  c->cert_tried += 1;
  if (c->spec)
  {
    // Decided on the device into the sweep's flag, together with the
    // sweep's other certificates before the flag is read (run_sweep); the
    // step goes on as certified.
    if (known && known->dev)
      c->spec_certs.push_back({known->dev, q, 0, delta});
    else
    {
      const int nb = rinv_cert_blocks(q);
      double* part = alloc_core(c, 2 * (size_t)nb);
      c->spec_cert_bufs.push_back(part);
      {
        prof_scope ps(c, TTA_K_RINV, (double)q * q * q / 3.0, 16.0 * (double)q * q);
        launch_rinv_cert(R, q, part, c->stream);
      }
      c->spec_certs.push_back({part, nb, 1, delta});
    }
    c->cert_held += 1;
    return true;
  }
  double xf2 = 0.0, rf2 = 0.0;
  if (known && known->have)
  {
    xf2 = known->xf2;
    rf2 = known->rf2;
  }
  else
  {
    const int nb = rinv_cert_blocks(q);
    double* part = ctx_ws(c, WS_CERT, 2 * (size_t)nb);
    {
      prof_scope ps(c, TTA_K_RINV, (double)q * q * q / 3.0, 16.0 * (double)q * q);
      launch_rinv_cert(R, q, part, c->stream);
    }
    std::vector<double> h(2 * (size_t)nb);
    hip_check(hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (int b = 0; b < nb; ++b)
    {
      xf2 += h[2 * (size_t)b];
      rf2 += h[2 * (size_t)b + 1];
    }
  }
  if (!std::isfinite(xf2) || !std::isfinite(rf2) || !(xf2 > 0.0) || !(rf2 > 0.0)) return false;
  const double xf = std::sqrt(xf2), rf = std::sqrt(rf2);
  if (rf * xf > 1.0e8) return false;
  const bool held = 1.0 / (xf * (1.0 + 1.0e-3)) > delta * (1.0 + 1.0e-3);
  c->cert_held += held ? 1 : 0;
  return held;
}

// Thin SVD of a row-major (rows x cols) matrix M, truncated on the device:
//   U_head (rows x r_new) row-major with orthonormal columns (products of
//          orthogonal factors, exact up to rounding), in a fresh pool block;
//   F      (r_new x cols) row-major = diag(s_head) Vt_head, in scratch.
// The matrix is QR-condensed to its square triangular factor first, and the
// Jacobi iteration always runs on the columns of R^T (rows of R): on graded
// factors that orientation converges in a handful of sweeps where the
// columns of R can take dozens.
//   rows >= cols: M = Q R,   R^T J = W -> M = (Q J) W^T: U = Q J, F = W^T;
//   rows <  cols: M^T = Q R, R^T J = W -> M = W (Q J)^T: U = W diag(s)^-1
//                 (orthonormal to the Jacobi tolerance, which the last sweep
//                 verifies pair by pair), F = diag(s) (Q J)^T.
// keep_all (every singular triplet, zero ones included) keeps, for
// rows < cols, the exact-rotation form R J = W -> M = J (Q W)^T, whose U is
// orthonormal however many singular values vanish.
// With certify (and no singular values asked for, no binding max_rank), a
// step whose full rank is certified (certify_full_rank) keeps the same
// tensor without the Jacobi SVD: U = Q, F = R for rows >= cols, and
// U = I, F = M for rows < cols.
// F is formed from the exactly orthonormal factor by one GEMM, F = J'^T R
// (rows >= cols, J' the polar-corrected rotations) or F = U^T M
// (rows < cols, U = W diag(s)^-1 orthonormal to the Jacobi tolerance), so
// that U F reproduces M (or its projection on the kept vectors) to one
// GEMM's rounding however many rotations the Jacobi applied.
int svd_left(tta_ctx* c, const double* M, int rows, int cols, double delta, int max_rank, bool keep_all,
             double** U_out, double** F_out, double** s_out = nullptr, bool certify = false)
{
  double* sv = nullptr;
  int* perm = nullptr;
  const bool try_cert = certify && !keep_all && !s_out && (max_rank <= 0 || max_rank >= std::min(rows, cols));
  if (rows >= cols)
  {
    const int q = cols;
    rinv_norms rn;
    rinv_norms* rnp = try_cert && c->rank_certificate ? &rn : nullptr;
    double* X = ctx_ws(c, WS_X, (size_t)q * q);
    // Q with element strides (qrs, qcs).  CholQR2 reads M where it lies and
    // writes Q row-major (the layout of U) and R into X; only Householder
    // (CholQR2 not applicable, or declined) needs the column-major copy.
    double* Q = nullptr;
    long qrs = 0, qcs = 0;
    bool own_q = false;
    if (cholqr_applies(c, rows, cols))
    {
      double* Qd = alloc_core(c, (size_t)rows * q);
      if (cholqr2_strided(c, M, cols, 1, rows, cols, Qd, q, 1, X, rnp))
      {
        Q = Qd;
        qrs = q;
        qcs = 1;
        own_q = true;
      }
      else
        c->mem.put(Qd);
    }
    if (!Q)
    {
      double* A = ctx_ws(c, WS_A, (size_t)rows * cols);
      launch_transpose(M, cols, rows, cols, A, rows, c->stream);
      double* tau = ctx_ws(c, WS_F, (size_t)q);
      Q = ctx_ws(c, WS_Q, (size_t)rows * q);
      qrs = 1;
      qcs = rows;
      qr_factor(c, A, rows, rows, cols, tau, Q, rows, false);
      launch_extract_upper(A, rows, q, q, X, q, 1, c->stream);
    }
    if (try_cert && certify_full_rank(c, X, q, delta, &rn))
    {
      double* U = Q;
      if (!own_q)
      {
        U = alloc_core(c, (size_t)rows * q);
        launch_transpose(Q, rows, q, rows, U, q, c->stream);
      }
      *U_out = U;
      *F_out = X;
      return q;
    }
    // The Jacobi iteration overwrites X; R is needed again for F.
    double* Rb = alloc_core(c, (size_t)q * q);
    hip_check(hipMemcpyAsync(Rb, X, (size_t)q * q * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
    double* J = ctx_ws(c, WS_J, (size_t)q * q);
    const int r_new = square_svd(c, X, q, J, delta, max_rank, keep_all, &sv, &perm);
    polar_correct(c, J, q);
    double* Jp = ctx_ws(c, WS_JP, (size_t)q * r_new);
    launch_gather_cols(J, q, q, perm, r_new, Jp, 1, q, c->stream);
    double* U = alloc_core(c, (size_t)rows * r_new);
    gemm_s(c, rows, r_new, q, Q, qrs, qcs, Jp, 1, q, U, r_new, 1);  // U = Q J'_k
    double* F = ctx_ws(c, WS_M1P, (size_t)r_new * cols);
    gemm_s(c, r_new, q, q, Jp, q, 1, Rb, q, 1, F, cols, 1);          // F = J'_k^T R
    c->mem.put(Rb);
    if (own_q) c->mem.put(Q);
    *U_out = U;
    *F_out = F;
    if (s_out) *s_out = sv;
    return r_new;
  }
  const int q = rows;
  double* A = ctx_ws(c, WS_A, (size_t)rows * cols);
  hip_check(hipMemcpyAsync(A, M, (size_t)rows * cols * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
  double* tau = ctx_ws(c, WS_F, (size_t)q);
  double* Q = ctx_ws(c, WS_Q, (size_t)cols * q);
  rinv_norms rn;
  const bool full =
      qr_factor(c, A, cols, cols, rows, tau, Q, cols, true, try_cert && c->rank_certificate ? &rn : nullptr);
  double* Y = ctx_ws(c, WS_X, (size_t)q * q);
  if (keep_all)
    launch_extract_upper(A, cols, q, q, Y, 1, q, c->stream, full);  // Y = R
  else
    launch_extract_upper(A, cols, q, q, Y, q, 1, c->stream, full);  // Y = R^T
  if (try_cert && certify_full_rank(c, Y, q, delta, &rn))
  {
    double* U = alloc_core(c, (size_t)q * q);
    launch_eye(U, q, q, c->stream);
    double* F = ctx_ws(c, WS_M1P, (size_t)rows * cols);
    hip_check(hipMemcpyAsync(F, M, (size_t)rows * cols * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
    *U_out = U;
    *F_out = F;
    return q;
  }
  double* J = ctx_ws(c, WS_J, (size_t)q * q);
  const int r_new = square_svd(c, Y, q, J, delta, max_rank, keep_all, &sv, &perm);
  double* U = alloc_core(c, (size_t)rows * r_new);
  if (keep_all)
  {
    polar_correct(c, J, q);
    launch_gather_cols(J, q, q, perm, r_new, U, r_new, 1, c->stream);
  }
  else
    launch_gather_cols(Y, q, q, perm, r_new, U, r_new, 1, c->stream, sv, GATHER_DIV);
  double* F = ctx_ws(c, WS_M1P, (size_t)r_new * cols);
  gemm_s(c, r_new, cols, q, U, 1, r_new, M, cols, 1, F, cols, 1);  // F = U^T M
  *U_out = U;
  *F_out = F;
  if (s_out) *s_out = sv;
  return r_new;
}

void make_rank1_zero(tta_ctx* c, tta_tt* t);

// detail::svd (include/core/factory/tt_svd.hpp:46-160) on the device: the
// dense tensor (row-major, uploaded once) is swept left to right, each
// unfolding (r_prev n_k) x rest/n_k through svd_left (QR-condensed Jacobi SVD
// with the truncation rank decided on the device); core k = U, the next
// unfolding = diag(s) Vt.  Zero tensor: rank-1 zero cores.
tta_tt* from_dense_impl(tta_ctx* c, const double* host, const std::vector<int>& shape, double eps, int max_rank)
{
  const int d = (int)shape.size();
  long long total = 1;
  for (int k = 0; k < d; ++k)
  {
    if (shape[k] <= 0) fail(TTA_ERR_SHAPE, "from_dense: mode %d has size %d", k, shape[k]);
    total *= shape[k];
    if (total > (long long)std::numeric_limits<int>::max())
      fail(TTA_ERR_SHAPE, "from_dense: dense tensor above INT_MAX entries");
  }
  std::vector<int> ranks(d + 1, 1);
  tta_tt* t = new_tt(c, shape, ranks, false);
  if (d == 0) return t;
  try
  {
    double* C = alloc_core(c, (size_t)total);
    hip_check(hipMemcpyAsync(C, host, (size_t)total * 8, hipMemcpyHostToDevice, c->stream), "upload");
    const double norm = frob_norm_dev(c, C, (size_t)total, 0.0, 0.0);
    if (!std::isfinite(norm))
    {
      c->mem.put(C);
      fail(TTA_ERR_ARG, "from_dense: input contains non-finite values (norm %g)", norm);
    }
    if (norm == 0.0)
    {
      c->mem.put(C);
      make_rank1_zero(c, t);
      return t;
    }
    const double delta = eps * norm / std::sqrt((double)std::max(d - 1, 1));
    int r_prev = 1;
    long long rest = total;
    for (int k = 0; k < d - 1; ++k)
    {
      const int rows = r_prev * shape[k];
      const int cols = (int)(rest / shape[k]);
      double *U = nullptr, *F = nullptr;
      const int r_new = svd_left(c, C, rows, cols, delta, max_rank, false, &U, &F);
      replace_core(c, t, k, U);
      // F lives in a workspace the next step reuses: copy it out.
      double* Cn = alloc_core(c, (size_t)r_new * cols);
      hip_check(hipMemcpyAsync(Cn, F, (size_t)r_new * cols * 8, hipMemcpyDeviceToDevice, c->stream), "copy");
      c->mem.put(C);
      C = Cn;
      t->ranks[k + 1] = r_new;
      r_prev = r_new;
      rest = cols;
    }
    replace_core(c, t, d - 1, C);  // (r_prev, n_{d-1}, 1)
  }
  catch (...)
  {
    for (double* p : t->cores) c->mem.put(p);
    delete t;
    throw;
  }
  return t;
}

// svd_step (include/core/detail/truncate.hpp:61-99).
void svd_step(tta_ctx* c, tta_tt* t, int k, double delta, int max_rank)
{
  const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
  const int rows = r * n, cols = rp;
  double *U = nullptr, *F = nullptr;
  const int r_new = svd_left(c, t->cores[k], rows, cols, delta, max_rank, false, &U, &F, nullptr, true);
  const int ncols_n = t->modes[k + 1] * t->ranks[k + 2];
  double* Nn = alloc_core(c, (size_t)r_new * ncols_n);
  gemm_rm(c, r_new, ncols_n, cols, F, cols, t->cores[k + 1], ncols_n, Nn, ncols_n);
  replace_core(c, t, k, U);
  replace_core(c, t, k + 1, Nn);
  t->ranks[k + 1] = r_new;
}

// A certificate applies when max_rank cannot force a cut.
bool cert_ok(int max_rank, int q) { return max_rank <= 0 || max_rank >= q; }

// Last part of svd_step_right: core k-1 <- core k-1 F (rows x r_new), core k
// <- Cn.
void absorb_right_step(tta_ctx* c, tta_tt* t, int k, int rows, int r_new, double* Cn, const double* F)
{
  const int rows_p = t->ranks[k - 1] * t->modes[k - 1];
  double* Pn = alloc_core(c, (size_t)rows_p * r_new);
  gemm_rm(c, rows_p, r_new, rows, t->cores[k - 1], rows, F, r_new, Pn, r_new);
  replace_core(c, t, k - 1, Pn);
  replace_core(c, t, k, Cn);
  t->ranks[k] = r_new;
}

// svd_step_right (truncate.hpp:114-166): SVD of right_unfold(core k)
// (r x n*rp) with V orthonormal (the new core's rows) and U diag(s)
// absorbed into core k-1.  Jacobi runs on R^T as in svd_left:
//   cols >= rows: M^T = Q R, R^T J = W -> M = W (Q J)^T: V = Q J, U s = W;
//   cols <  rows: M = Q R,   R^T J = W -> M = (Q J) W^T: V = W diag(s)^-1,
//                 U s = Q J diag(s).
// As in svd_left, the absorbed factor is formed by one GEMM from the exactly
// orthonormal one: U s = R^T J' (J' polar-corrected) or U s = M V.
void svd_step_right(tta_ctx* c, tta_tt* t, int k, double delta, int max_rank)
{
  const int r = t->ranks[k], n = t->modes[k], rp = t->ranks[k + 1];
  const int rows = r, cols = n * rp;
  double* sv = nullptr;
  int* perm = nullptr;
  double* Cn = nullptr;
  double* F = nullptr;
  int r_new = 0;
  if (cols >= rows)
  {
    const int q = rows;
    double* A = t->cores[k];  // M^T (cols x rows) column-major; factored in place
    double* tau = ctx_ws(c, WS_F, (size_t)q);
    double* Q = ctx_ws(c, WS_Q, (size_t)q * cols);
    rinv_norms rn;
    const bool full =
        qr_factor(c, A, cols, cols, rows, tau, Q, cols, true, cert_ok(max_rank, q) && c->rank_certificate ? &rn : nullptr);
    double* X = ctx_ws(c, WS_X, (size_t)q * q);
    launch_extract_upper(A, cols, q, q, X, q, 1, c->stream, full);
    if (cert_ok(max_rank, q) && certify_full_rank(c, X, q, delta, &rn))
    {
      // M = R^T Q^T: the new core is Q^T (the column-major Q read row-major),
      // and R^T is absorbed into core k-1.
      r_new = q;
      Cn = alloc_core(c, (size_t)q * cols);
      hip_check(hipMemcpyAsync(Cn, Q, (size_t)q * cols * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
      F = ctx_ws(c, WS_M1P, (size_t)rows * q);
      launch_transpose(X, q, q, q, F, q, c->stream);
      absorb_right_step(c, t, k, rows, r_new, Cn, F);
      return;
    }
    double* J = ctx_ws(c, WS_J, (size_t)q * q);
    r_new = square_svd(c, X, q, J, delta, max_rank, false, &sv, &perm);
    polar_correct(c, J, q);
    double* Jp = ctx_ws(c, WS_JP, (size_t)q * r_new);
    launch_gather_cols(J, q, q, perm, r_new, Jp, 1, q, c->stream);
    Cn = alloc_core(c, (size_t)r_new * cols);
    gemm_s(c, r_new, cols, q, Jp, q, 1, Q, cols, 1, Cn, cols, 1);  // Cn = J'_k^T Q^T
    double* Rb = alloc_core(c, (size_t)q * q);
    launch_extract_upper(A, cols, q, q, Rb, q, 1, c->stream, full);  // R row-major
    F = ctx_ws(c, WS_M1P, (size_t)rows * r_new);
    gemm_s(c, rows, r_new, q, Rb, 1, q, Jp, 1, q, F, r_new, 1);     // F = R^T J'_k
    c->mem.put(Rb);
  }
  else
  {
    const int q = cols;
    double* A = ctx_ws(c, WS_A, (size_t)rows * cols);
    launch_transpose(t->cores[k], cols, rows, cols, A, rows, c->stream);
    double* tau = ctx_ws(c, WS_F, (size_t)q);
    double* Q = ctx_ws(c, WS_Q, (size_t)rows * q);
    rinv_norms rn;
    const bool full =
        qr_factor(c, A, rows, rows, cols, tau, Q, rows, true, cert_ok(max_rank, q) && c->rank_certificate ? &rn : nullptr);
    double* Y = ctx_ws(c, WS_X, (size_t)q * q);
    launch_extract_upper(A, rows, q, q, Y, q, 1, c->stream, full);  // Y = R^T
    if (cert_ok(max_rank, q) && certify_full_rank(c, Y, q, delta, &rn))
    {
      // M = M I: the new core is the identity and M is absorbed into core k-1.
      r_new = q;
      Cn = alloc_core(c, (size_t)q * q);
      launch_eye(Cn, q, q, c->stream);
      F = ctx_ws(c, WS_M1P, (size_t)rows * cols);
      hip_check(hipMemcpyAsync(F, t->cores[k], (size_t)rows * cols * 8, hipMemcpyDeviceToDevice, c->stream),
                "hipMemcpyAsync");
      absorb_right_step(c, t, k, rows, r_new, Cn, F);
      return;
    }
    double* J = ctx_ws(c, WS_J, (size_t)q * q);
    r_new = square_svd(c, Y, q, J, delta, max_rank, false, &sv, &perm);
    Cn = alloc_core(c, (size_t)r_new * cols);
    launch_gather_cols(Y, q, q, perm, r_new, Cn, 1, q, c->stream, sv, GATHER_DIV);
    F = ctx_ws(c, WS_M1P, (size_t)rows * r_new);
    gemm_s(c, rows, r_new, cols, t->cores[k], cols, 1, Cn, 1, cols, F, r_new, 1);  // F = M Cn^T
  }
  absorb_right_step(c, t, k, rows, r_new, Cn, F);
}

// The QR sweeps.  With in_round (the first sweep of round, whose cores
// only feed its second sweep) runs of chainable steps are factored as
// chained blocks whose second pass may be left out (kChainSkipTol); every
// decision is made on the host, so a declined CholQR2 falls back to
// Householder for its step alone.  Otherwise single steps throughout.
void right_orthogonalize_impl(tta_ctx* c, tta_tt* t, bool in_round = false)
{
  const int d = t->d();
  if (d <= 1) return;
  for (int k = d - 1; k >= 1;)
  {
    const int e = in_round && chain_step_ok(c, t, k, 0, false, 0) ? chain_run_end(c, t, k, 0, false, 0) : k;
    if (e != k && chain_block(c, t, e, k, 0, false, 0.0, kChainSkipTol))
    {
      k = e - 1;
      continue;
    }
    for (; k >= e; --k) right_qr_step(c, t, k);
  }
}

void left_orthogonalize_impl(tta_ctx* c, tta_tt* t, bool in_round = false)
{
  const int d = t->d();
  if (d <= 1) return;
  for (int k = 0; k < d - 1;)
  {
    const int e = in_round && chain_step_ok(c, t, k, 1, false, 0) ? chain_run_end(c, t, k, 1, false, 0) : k;
    if (e != k && chain_block(c, t, k, e, 1, false, 0.0, kChainSkipTol))
    {
      k = e + 1;
      continue;
    }
    for (; k <= e; ++k) left_qr_step(c, t, k);
  }
}

void make_rank1_zero(tta_ctx* c, tta_tt* t)
{
  for (int k = 0; k < t->d(); ++k)
  {
    double* p = alloc_core(c, (size_t)t->modes[k]);
    hip_check(hipMemsetAsync(p, 0, (size_t)t->modes[k] * 8, c->stream), "hipMemsetAsync");
    replace_core(c, t, k, p);
  }
  std::fill(t->ranks.begin(), t->ranks.end(), 1);
}

// Norm balancing (include/core/round/tt_round.hpp:76-124).
void balance(tta_ctx* c, tta_tt* t)
{
  const int d = t->d();
  std::vector<size_t> off(d, 0);
  size_t total = 0;
  for (int k = 0; k < d - 1; ++k)
  {
    off[k] = total;
    total += (size_t)t->ranks[k + 1];
  }
  double* nrm = ctx_ws(c, WS_JP, total + 1);
  // Core k's right column norms at off[k] (none for the last core), its
  // left ones at off[k - 1] (none for the first); kCoreBatch cores a launch.
  auto batches = [&](auto&& launch)
  {
    core_batch b;
    for (int k = 0; k < d; ++k)
    {
      b.it[b.count++] = {t->cores[k], t->ranks[k], t->modes[k], t->ranks[k + 1], k < d - 1 ? (int)off[k] : -1,
                         k > 0 ? (int)off[k - 1] : -1};
      if (b.count == kCoreBatch || k + 1 == d)
      {
        launch(b);
        b.count = 0;
      }
    }
  };
  batches([&](const core_batch& b) { launch_core_norms_batch(b, nrm, c->stream); });
  double* geo = small_dbl(c) + 2;
  launch_geo_mean(nrm, (int)total, geo, c->stream);
  batches([&](const core_batch& b) { launch_balance_batch(b, nrm, geo, c->stream); });
}

// Copies of a TT's cores and its ranks, to undo a speculative sweep.
struct tt_snapshot
{
  std::vector<double*> cores;
  std::vector<int> ranks;
};

tt_snapshot take_snapshot(tta_ctx* c, const tta_tt* t)
{
  tt_snapshot s;
  s.ranks = t->ranks;
  copy_batch b;
  double total = 0.0;
  for (int k = 0; k < t->d(); ++k) total += (double)core_numel(t, k);
  prof_scope ps(c, TTA_K_OTHER, 0.0, 16.0 * total);
  for (int k = 0; k < t->d(); ++k)
  {
    const size_t n = core_numel(t, k);
    double* p = alloc_core(c, n);
    s.cores.push_back(p);
    b.it[b.count++] = {t->cores[k], p, (long)n};
    if (b.count == kCopyBatch || k + 1 == t->d())
    {
      launch_copy_batch(b, c->stream);
      b.count = 0;
    }
  }
  return s;
}

// Runs one sweep of round over t.  Speculative (c->speculate and the last
// sweep of this kind needed no fallback): CholQR2 acceptance and the rank
// certificates are decided on the device into one flag, the host reads it
// once at the end, and if any decision failed the sweep is undone from a
// snapshot and run again with host decisions, so the result is always that
// of the checked sweep.  Otherwise the checked sweep runs directly, and
// speculation resumes once such a sweep has needed no fallback.
// The recorded certificate decisions of a speculative sweep, kCertBatch per
// launch, then their input buffers back to the pool (stream-ordered).
void flush_spec_certs(tta_ctx* c)
{
  cert_batch b;
  {
    prof_scope ps(c, TTA_K_OTHER, 0.0, 16.0 * (double)c->spec_certs.size());
    for (size_t i = 0; i < c->spec_certs.size(); ++i)
    {
      const tta_ctx::spec_cert_rec& r = c->spec_certs[i];
      b.it[b.count++] = {r.a, r.n, r.mode, r.delta};
      if (b.count == kCertBatch || i + 1 == c->spec_certs.size())
      {
        launch_spec_cert_batch(b, spec_flag(c), c->stream);
        b.count = 0;
      }
    }
  }
  c->spec_certs.clear();
  for (void* p : c->spec_cert_bufs) c->mem.put(p);
  c->spec_cert_bufs.clear();
}

void drop_spec_certs(tta_ctx* c)
{
  c->spec_certs.clear();
  for (void* p : c->spec_cert_bufs) c->mem.put(p);
  c->spec_cert_bufs.clear();
}

template <class F>
void run_sweep(tta_ctx* c, tta_tt* t, bool& last_ok, F&& sweep)
{
  if (!c->speculate || !last_ok)
  {
    const long long ct = c->cholqr_tried, cu = c->cholqr_used, et = c->cert_tried, eh = c->cert_held;
    sweep();
    last_ok = c->cholqr_tried - ct == c->cholqr_used - cu && c->cert_tried - et == c->cert_held - eh;
    return;
  }
  tt_snapshot snap = take_snapshot(c, t);
  launch_set_value(spec_flag(c), 1.0, c->stream);
  c->spec = true;
  try
  {
    sweep();
    flush_spec_certs(c);
  }
  catch (const spec_abort&)
  {
    // Failed before its end (spec_checkpoint): undone and run checked.
    c->spec = false;
    drop_spec_certs(c);
    c->spec_sweeps += 1;
    c->spec_failed += 1;
    last_ok = false;
    for (int k = 0; k < t->d(); ++k) c->mem.put(t->cores[k]);
    t->cores = snap.cores;
    t->ranks = snap.ranks;
    sweep();
    return;
  }
  catch (...)
  {
    c->spec = false;
    drop_spec_certs(c);
    for (double* p : snap.cores) c->mem.put(p);
    throw;
  }
  c->spec = false;
  c->spec_sweeps += 1;
  if (read_double(c, spec_flag(c)) == 1.0)
  {
    for (double* p : snap.cores) c->mem.put(p);
    return;
  }
  c->spec_failed += 1;
  last_ok = false;
  for (int k = 0; k < t->d(); ++k) c->mem.put(t->cores[k]);
  t->cores = snap.cores;
  t->ranks = snap.ranks;
  sweep();
}

// detail::round (tt_round.hpp:36-127) / round_right_canonical (:151-191) on
// a working copy `t` (consumed in place).
void round_inplace(tta_ctx* c, tta_tt* t, double eps, int max_rank, int gauge)
{
  const int d = t->d();
  if (d <= 1) return;
  const double inv = 1.0 / std::sqrt((double)std::max(d - 1, 1));
  if (gauge == TTA_GAUGE_RIGHT_CANONICAL)
  {
    left_orthogonalize_impl(c, t, true);
    const double nrm = frob_norm_dev(c, t->cores[d - 1], core_numel(t, d - 1), eps, inv);
    if (nrm == 0.0)
    {
      make_rank1_zero(c, t);
      return;
    }
    const double delta = eps * nrm * inv;
    run_sweep(c, t, c->spec_svd_ok,
              [&]
              {
                for (int k = d - 1; k >= 1;)
                {
                  const int e =
                      chain_step_ok(c, t, k, 0, true, max_rank) ? chain_run_end(c, t, k, 0, true, max_rank) : k;
                  if (e != k && chain_block(c, t, e, k, 0, true, delta))
                  {
                    k = e - 1;
                    continue;
                  }
                  for (; k >= e; --k) svd_step_right(c, t, k, delta, max_rank);
                }
              });
    return;
  }
  right_orthogonalize_impl(c, t, true);
  const double nrm = frob_norm_dev(c, t->cores[0], core_numel(t, 0), eps, inv);
  if (nrm == 0.0)
  {
    make_rank1_zero(c, t);
    return;
  }
  const double delta = eps * nrm * inv;
  run_sweep(c, t, c->spec_svd_ok,
            [&]
            {
              for (int k = 0; k < d - 1;)
              {
                const int e = chain_step_ok(c, t, k, 1, true, max_rank) ? chain_run_end(c, t, k, 1, true, max_rank) : k;
                if (e != k && chain_block(c, t, k, e, 1, true, delta))
                {
                  k = e + 1;
                  continue;
                }
                for (; k <= e; ++k) svd_step(c, t, k, delta, max_rank);
              }
            });
  balance(c, t);
}

// matvec (tt_matrix_apply.hpp:237-248) into a fresh tensor.
tta_tt* matvec_impl(tta_ctx* c, const tta_tt* A, const tta_tt* x)
{
  const int d = A->d();
  std::vector<int> modes(d), ranks(d + 1);
  for (int k = 0; k < d; ++k) modes[k] = A->row_modes[k];
  for (int k = 0; k <= d; ++k) ranks[k] = A->ranks[k] * x->ranks[k];
  tta_tt* y = new_tt(c, modes, ranks, false);
  double fl = 0.0, by = 0.0;
  for (int k = 0; k < d; ++k)
  {
    fl += 2.0 * ranks[k] * modes[k] * ranks[k + 1] * A->col_modes[k];
    by += 8.0 * ((double)core_numel(A, k) + (double)core_numel(x, k) + (double)ranks[k] * modes[k] * ranks[k + 1]);
  }
  prof_scope ps(c, TTA_K_APPLY, fl, by);
  matvec_batch b;
  for (int k = 0; k < d; ++k)
  {
    b.it[b.count++] = {A->cores[k], x->cores[k], y->cores[k], A->ranks[k], A->row_modes[k], A->col_modes[k],
                       A->ranks[k + 1], x->ranks[k], x->ranks[k + 1]};
    if (b.count == kCoreBatch || k + 1 == d)
    {
      launch_matvec_core_batch(b, c->stream);
      b.count = 0;
    }
  }
  return y;
}

// matmat (tt_matrix_apply.hpp:260-271).
tta_tt* matmat_impl(tta_ctx* c, const tta_tt* A, const tta_tt* B)
{
  const int d = A->d();
  std::vector<int> modes(d), ranks(d + 1);
  for (int k = 0; k < d; ++k) modes[k] = A->row_modes[k] * B->col_modes[k];
  for (int k = 0; k <= d; ++k) ranks[k] = A->ranks[k] * B->ranks[k];
  tta_tt* C = new_tt(c, modes, ranks, false);
  C->row_modes = A->row_modes;
  C->col_modes = B->col_modes;
  double fl = 0.0, by = 0.0;
  for (int k = 0; k < d; ++k)
  {
    fl += 2.0 * ranks[k] * modes[k] * ranks[k + 1] * A->col_modes[k];
    by += 8.0 * ((double)core_numel(A, k) + (double)core_numel(B, k) + (double)ranks[k] * modes[k] * ranks[k + 1]);
  }
  prof_scope ps(c, TTA_K_APPLY, fl, by);
  for (int k = 0; k < d; ++k)
    launch_matmat_core(A->cores[k], B->cores[k], C->cores[k], A->ranks[k], A->row_modes[k], A->col_modes[k],
                       A->ranks[k + 1], B->ranks[k], B->col_modes[k], B->ranks[k + 1], c->stream);
  return C;
}

// fused_matvec_core / fused_matmat_core (include/core/detail/fused_apply_core.hpp:33-131,
// :136-272).  F (alpha x a*sx) row-major; A_k (a, m, p, b); X_k (sx, p, nx, bx)
// with nx = 1 for a vector.  Output (alpha, m, nx, b*bx) into a fresh buffer.
double* fused_core(tta_ctx* c, const double* F, int alpha, const double* Ak, int a, int m, int p, int b,
                   const double* Xk, int sx, int nx, int bx)
{
  const size_t mpb = (size_t)m * p * b;
  // M1 (alpha, ax, i, q, bb) = sum_aa F(al, aa*sx + ax) A(aa, i, q, bb): batch over al.
  double* M1 = ctx_ws(c, WS_M1, (size_t)alpha * sx * mpb);
  gemm_desc g;
  g.M = sx;
  g.N = (int)mpb;
  g.K = a;
  g.A = F;
  g.sar = 1;
  g.sac = sx;
  g.B = Ak;
  g.sbr = (long)mpb;
  g.sbc = 1;
  g.C = M1;
  g.scr = (long)mpb;
  g.scc = 1;
  g.bsa = (long)a * sx;
  g.bsb = 0;
  g.bsc = (long)sx * mpb;
  {
    prof_scope ps(c, TTA_K_GEMM, 2.0 * alpha * g.M * g.N * g.K,
                  8.0 * alpha * ((double)g.M * g.K + (double)g.M * g.N) + 8.0 * g.K * g.N);
    launch_gemm(g, alpha, c->stream);
  }
  // M1p (alpha, i, bb, ax, q) <- M1 (alpha, ax, i, q, bb).
  double* M1p = ctx_ws(c, WS_M1P, (size_t)alpha * sx * mpb);
  permute_desc pd;
  pd.src = M1;
  pd.dst = M1p;
  pd.n[0] = 1;
  pd.n[1] = alpha;
  pd.n[2] = m;
  pd.n[3] = b;
  pd.n[4] = sx;
  pd.n[5] = p;
  pd.ss[1] = (long)sx * mpb;
  pd.ss[2] = (long)p * b;
  pd.ss[3] = 1;
  pd.ss[4] = (long)mpb;
  pd.ss[5] = b;
  launch_permute(pd, c->stream);
  // R (alpha*m*b x nx*bx) = M1p (alpha*m*b x sx*p) * X_view (sx*p x nx*bx).
  const long rowsR = (long)alpha * m * b, colsR = (long)nx * bx;
  double* out = alloc_core(c, (size_t)rowsR * colsR);
  if (nx == 1)
  {
    gemm_rm(c, (int)rowsR, (int)colsR, sx * p, M1p, sx * p, Xk, colsR, out, colsR);
    return out;
  }
  double* R = ctx_ws(c, WS_A, (size_t)rowsR * colsR);
  gemm_rm(c, (int)rowsR, (int)colsR, sx * p, M1p, sx * p, Xk, colsR, R, colsR);
  // out (alpha, i, j, bb, bx) <- R (alpha, i, bb, j, bx).
  permute_desc p2;
  p2.src = R;
  p2.dst = out;
  p2.n[1] = alpha;
  p2.n[2] = m;
  p2.n[3] = nx;
  p2.n[4] = b;
  p2.n[5] = bx;
  p2.ss[1] = (long)m * b * nx * bx;
  p2.ss[2] = (long)b * nx * bx;
  p2.ss[3] = bx;
  p2.ss[4] = (long)nx * bx;
  p2.ss[5] = 1;
  launch_permute(p2, c->stream);
  return out;
}

// build_streaming_matvec / build_streaming_matmat
// (include/core/detail/streaming_apply.hpp:39-110, :115-202).
tta_tt* streaming_impl(tta_ctx* c, const tta_tt* A_in, const tta_tt* X_in, bool is_matmat, double eps, int max_rank)
{
  const int d = A_in->d();
  tta_tt* A = copy_tt(c, A_in);
  tta_tt* X = copy_tt(c, X_in);
  tta_tt* out = nullptr;
  try
  {
    right_orthogonalize_impl(c, A);
    right_orthogonalize_impl(c, X);
    const double nA = frob_norm_dev(c, A->cores[0], core_numel(A, 0), 0.0, 0.0);
    const double nX = frob_norm_dev(c, X->cores[0], core_numel(X, 0), 0.0, 0.0);
    const double norm = nA * nX;
    std::vector<int> modes(d), ranks(d + 1, 1);
    for (int k = 0; k < d; ++k)
      modes[k] = is_matmat ? A_in->row_modes[k] * X_in->col_modes[k] : A_in->row_modes[k];
    out = new_tt(c, modes, ranks, true);
    if (is_matmat)
    {
      out->row_modes = A_in->row_modes;
      out->col_modes = X_in->col_modes;
    }
    if (norm == 0.0)
    {
      free_tt(c, A);
      free_tt(c, X);
      return out;
    }
    const double delta = eps * norm / std::sqrt((double)std::max(d - 1, 1));
    // F starts as [[1]].
    double* Fbuf = ctx_ws(c, WS_JP, 1);
    const double one = 1.0;
    hip_check(hipMemcpyAsync(Fbuf, &one, 8, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    // The F environment is carried in its own pool block between steps.
    double* F = alloc_core(c, 1);
    hip_check(hipMemcpyAsync(F, Fbuf, 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
    int alpha = 1;
    for (int k = 0; k < d; ++k)
    {
      const int a = A->ranks[k], b = A->ranks[k + 1];
      const int m = A->row_modes[k], p = A->col_modes[k];
      const int sx = X->ranks[k], bx = X->ranks[k + 1];
      const int nx = is_matmat ? X->col_modes[k] : 1;
      double* C = fused_core(c, F, alpha, A->cores[k], a, m, p, b, X->cores[k], sx, nx, bx);
      const int mk = m * nx, br = b * bx;
      if (k == d - 1)
      {
        replace_core(c, out, k, C);
        out->ranks[k] = alpha;
        out->ranks[k + 1] = br;
        break;
      }
      double *U = nullptr, *Fs = nullptr;
      const int r_new = svd_left(c, C, alpha * mk, br, delta, max_rank, false, &U, &Fs);
      c->mem.put(C);
      replace_core(c, out, k, U);
      out->ranks[k] = alpha;
      out->ranks[k + 1] = r_new;
      double* Fn = alloc_core(c, (size_t)r_new * br);
      hip_check(hipMemcpyAsync(Fn, Fs, (size_t)r_new * br * 8, hipMemcpyDeviceToDevice, c->stream),
                "hipMemcpyAsync");
      c->mem.put(F);
      F = Fn;
      alpha = r_new;
    }
    c->mem.put(F);
  }
  catch (...)
  {
    free_tt(c, A);
    free_tt(c, X);
    free_tt(c, out);
    throw;
  }
  free_tt(c, A);
  free_tt(c, X);
  return out;
}

// One contract_inner_step (contract.hpp:155-205): E1 (rap x rbp) =
// (E0^T right_unfold(ca))_as^T left_unfold(cb), with E0 (ra x rb) and the
// (rb x n rap) intermediate in T.
void inner_step_dev(tta_ctx* c, const double* E0, double* E1, double* T, const double* ca, int ra, int n, int rap,
                    const double* cb, int rb, int rbp)
{
  gemm_desc g;
  g.M = rb;
  g.N = n * rap;
  g.K = ra;
  g.A = E0;
  g.sar = 1;
  g.sac = rb;
  g.B = ca;
  g.sbr = (long)n * rap;
  g.sbc = 1;
  g.C = T;
  g.scr = (long)n * rap;
  g.scc = 1;
  run_gemm(c, g);
  gemm_desc h;
  h.M = rap;
  h.N = rbp;
  h.K = rb * n;
  h.A = T;
  h.sar = 1;
  h.sac = rap;
  h.B = cb;
  h.sbr = rbp;
  h.sbc = 1;
  h.C = E1;
  h.scr = rbp;
  h.scc = 1;
  run_gemm(c, h);
}

// inner (tt_inner.hpp:33-48) via contract_inner_step (contract.hpp:155-205).
double inner_impl(tta_ctx* c, const tta_tt* a, const tta_tt* b)
{
  const int d = a->d();
  if (d == 0) return 0.0;
  size_t emax = 1, tmax = 1;
  for (int k = 0; k <= d; ++k) emax = std::max(emax, (size_t)a->ranks[k] * b->ranks[k]);
  for (int k = 0; k < d; ++k) tmax = std::max(tmax, (size_t)b->ranks[k] * a->modes[k] * a->ranks[k + 1]);
  double* E0 = ctx_ws(c, WS_X, emax);
  double* E1 = ctx_ws(c, WS_J, emax);
  double* T = ctx_ws(c, WS_M1, tmax);
  const double one = 1.0;
  hip_check(hipMemcpyAsync(E0, &one, 8, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  for (int k = 0; k < d; ++k)
  {
    inner_step_dev(c, E0, E1, T, a->cores[k], a->ranks[k], a->modes[k], a->ranks[k + 1], b->cores[k], b->ranks[k],
                   b->ranks[k + 1]);
    std::swap(E0, E1);
  }
  return read_double(c, E0);
}

// frob_norm_apply (include/core/algebra/tt_matrix_frob_norm_apply.hpp:36-94):
// ||A x|| (or ||A B||) by one environment sweep over the product cores, each
// formed in a reused scratch core and consumed at once; the rank-multiplied
// product never exists as a whole.
double frob_norm_apply_impl(tta_ctx* c, const tta_tt* A, const tta_tt* X, bool is_matmat)
{
  const int d = A->d();
  if (d == 0) return 0.0;
  size_t emax = 1, tmax = 1, ymax = 1;
  for (int k = 0; k <= d; ++k)
    emax = std::max(emax, (size_t)A->ranks[k] * X->ranks[k] * A->ranks[k] * X->ranks[k]);
  std::vector<int> nk(d);
  for (int k = 0; k < d; ++k)
  {
    nk[k] = A->row_modes[k] * (is_matmat ? X->col_modes[k] : 1);
    const size_t r = (size_t)A->ranks[k] * X->ranks[k], rp = (size_t)A->ranks[k + 1] * X->ranks[k + 1];
    ymax = std::max(ymax, r * nk[k] * rp);
    tmax = std::max(tmax, r * nk[k] * rp);
  }
  double* E0 = ctx_ws(c, WS_X, emax);
  double* E1 = ctx_ws(c, WS_J, emax);
  double* T = ctx_ws(c, WS_M1, tmax);
  double* Yk = ctx_ws(c, WS_A, ymax);
  const double one = 1.0;
  hip_check(hipMemcpyAsync(E0, &one, 8, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  for (int k = 0; k < d; ++k)
  {
    const int r = A->ranks[k] * X->ranks[k], rp = A->ranks[k + 1] * X->ranks[k + 1];
    {
      prof_scope ps(c, TTA_K_APPLY, 2.0 * r * nk[k] * rp * A->col_modes[k],
                    8.0 * ((double)core_numel(A, k) + (double)core_numel(X, k) + (double)r * nk[k] * rp));
      if (is_matmat)
        launch_matmat_core(A->cores[k], X->cores[k], Yk, A->ranks[k], A->row_modes[k], A->col_modes[k],
                           A->ranks[k + 1], X->ranks[k], X->col_modes[k], X->ranks[k + 1], c->stream);
      else
        launch_matvec_core(A->cores[k], X->cores[k], Yk, A->ranks[k], A->row_modes[k], A->col_modes[k],
                           A->ranks[k + 1], X->ranks[k], X->ranks[k + 1], c->stream);
    }
    inner_step_dev(c, E0, E1, T, Yk, r, nk[k], rp, Yk, r, rp);
    std::swap(E0, E1);
  }
  return std::sqrt(std::max(read_double(c, E0), 0.0));
}

// ---------------------------------------------------------------------------
// ALS rounding (include/core/round/tt_als.hpp:54-290; the apply-then-ALS
// paths tt_matvec_als_round.hpp:30-113 and tt_matmat_als_round.hpp:39-141).
// Environments are small row-major pool blocks: L_k (X.r_k x Y.r_k) and
// R_k (Y.r_k x X.r_k).  The reference absorbs the triangular factor of every
// QR step into the neighbouring core, which the next local update overwrites
// before anything reads it; here the QR steps keep only Q.
// ---------------------------------------------------------------------------
struct iter_info
{
  int iters = 0;
  double change = 0.0;
  int converged = 0;
};

struct env_chain
{
  tta_ctx* c;
  std::vector<double*> L, R;
  double* one = nullptr;  // the 1x1 boundary environment [[1]]
  env_chain(tta_ctx* cc, int d) : c(cc), L(d + 1, nullptr), R(d + 1, nullptr)
  {
    one = alloc_core(c, 1);
    upload_async(c, one, &kOne, sizeof(double));
    L[0] = one;
    R[d] = one;
  }
  ~env_chain()
  {
    for (double* p : L)
      if (p && p != one) c->mem.put(p);
    for (double* p : R)
      if (p && p != one) c->mem.put(p);
    if (one) c->mem.put(one);
  }
  static void put(tta_ctx* c, std::vector<double*>& E, int k, double* p, double* one)
  {
    if (E[k] && E[k] != one) c->mem.put(E[k]);
    E[k] = p;
  }
  void set_L(int k, double* p) { put(c, L, k, p, one); }
  void set_R(int k, double* p) { put(c, R, k, p, one); }
  static constexpr double kOne = 1.0;
};

// als_local_update (include/core/detail/als_local_solve.hpp:24-52):
// left_unfold(X_k) = (L right_unfold(Y_k) as (rx n x Y.r_{k+1})) R.
double* als_local_dev(tta_ctx* c, const double* L, int rx, const tta_tt* Y, int k, const double* R, int rxp)
{
  const int ry = Y->ranks[k], n = Y->modes[k], ryp = Y->ranks[k + 1];
  double* T1 = ctx_ws(c, WS_M1, (size_t)rx * n * ryp);
  gemm_rm(c, rx, n * ryp, ry, L, ry, Y->cores[k], (long)n * ryp, T1, (long)n * ryp);
  double* Xn = alloc_core(c, (size_t)rx * n * rxp);
  gemm_rm(c, rx * n, rxp, ryp, T1, ryp, R, rxp, Xn, rxp);
  return Xn;
}

// contract_inner_step (contract.hpp:141-205) with ca = X_k, cb = Y_k:
// L_{k+1} = (E^T right_unfold(X_k))_as^T left_unfold(Y_k).
double* env_left_dev(tta_ctx* c, const double* E, const tta_tt* X, int k, const tta_tt* Y)
{
  const int rx = X->ranks[k], n = X->modes[k], rxp = X->ranks[k + 1];
  const int ry = Y->ranks[k], ryp = Y->ranks[k + 1];
  double* T = ctx_ws(c, WS_M1, (size_t)ry * n * rxp);
  gemm_desc g;
  g.M = ry;
  g.N = n * rxp;
  g.K = rx;
  g.A = E;
  g.sar = 1;
  g.sac = ry;
  g.B = X->cores[k];
  g.sbr = (long)n * rxp;
  g.sbc = 1;
  g.C = T;
  g.scr = (long)n * rxp;
  g.scc = 1;
  run_gemm(c, g);
  double* out = alloc_core(c, (size_t)rxp * ryp);
  gemm_desc h;
  h.M = rxp;
  h.N = ryp;
  h.K = ry * n;
  h.A = T;
  h.sar = 1;
  h.sac = rxp;
  h.B = Y->cores[k];
  h.sbr = ryp;
  h.sbc = 1;
  h.C = out;
  h.scr = ryp;
  h.scc = 1;
  run_gemm(c, h);
  return out;
}

// contract_inner_step_right (contract.hpp:90-139) with ca = Y_k, cb = X_k:
// R_k = (left_unfold(Y_k) F as (Y.r_k x n X.r_{k+1})) right_unfold(X_k)^T.
double* env_right_dev(tta_ctx* c, const tta_tt* Y, int k, const tta_tt* X, const double* F)
{
  const int ry = Y->ranks[k], n = Y->modes[k], ryp = Y->ranks[k + 1];
  const int rx = X->ranks[k], rxp = X->ranks[k + 1];
  double* T1 = ctx_ws(c, WS_M1, (size_t)ry * n * rxp);
  gemm_rm(c, ry * n, rxp, ryp, Y->cores[k], ryp, F, rxp, T1, rxp);
  double* out = alloc_core(c, (size_t)ry * rx);
  gemm_desc g;
  g.M = ry;
  g.N = rx;
  g.K = n * rxp;
  g.A = T1;
  g.sar = (long)n * rxp;
  g.sac = 1;
  g.B = X->cores[k];
  g.sbr = 1;
  g.sbc = (long)n * rxp;
  g.C = out;
  g.scr = rx;
  g.scc = 1;
  run_gemm(c, g);
  return out;
}

// Singular values (descending) of the row-major rows x cols matrix M into
// dst[0, min(rows, cols)): QR condensation, then one-sided Jacobi on R^T run
// to convergence (delta 0, every value kept).
int singular_values_dev(tta_ctx* c, const double* M, int rows, int cols, double* dst)
{
  const int q = std::min(rows, cols);
  double* A = ctx_ws(c, WS_A, (size_t)rows * cols);
  double* tau = ctx_ws(c, WS_F, (size_t)q);
  double* X = ctx_ws(c, WS_X, (size_t)q * q);
  if (rows >= cols)
  {
    launch_transpose(M, cols, rows, cols, A, rows, c->stream);
    const bool full = qr_factor(c, A, rows, rows, cols, tau, nullptr, 0, true);
    launch_extract_upper(A, rows, q, q, X, q, 1, c->stream, full);
  }
  else
  {
    hip_check(hipMemcpyAsync(A, M, (size_t)rows * cols * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
    const bool full = qr_factor(c, A, cols, cols, rows, tau, nullptr, 0, true);
    launch_extract_upper(A, cols, q, q, X, q, 1, c->stream, full);
  }
  double* J = ctx_ws(c, WS_J, (size_t)q * q);
  double* sv = nullptr;
  int* perm = nullptr;
  square_svd(c, X, q, J, 0.0, 0, true, &sv, &perm);
  hip_check(hipMemcpyAsync(dst, sv, (size_t)q * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
  return q;
}

// Largest relative change of the per-core singular values between two
// sweeps (tt_als.hpp:160-205): ||s_now - s_prev|| / (||s_now|| + 1e-300) per
// core, 1.0 for a core whose count changed.
double sv_change(const std::vector<std::vector<double>>& now, const std::vector<std::vector<double>>& prev)
{
  double mx = 0.0;
  for (size_t k = 0; k < now.size(); ++k)
  {
    double nn = 0.0;
    for (double v : now[k]) nn += v * v;
    nn = std::sqrt(nn);
    if (now[k].size() == prev[k].size())
    {
      double df = 0.0;
      for (size_t i = 0; i < now[k].size(); ++i) df += (now[k][i] - prev[k][i]) * (now[k][i] - prev[k][i]);
      mx = std::max(mx, std::sqrt(df) / (nn + 1.0e-300));
    }
    else
      mx = 1.0;
  }
  return mx;
}

// Singular values of left_unfold(X_k), k = 0 .. d-2, on the host.
std::vector<std::vector<double>> core_singular_values(tta_ctx* c, const tta_tt* X)
{
  const int d = X->d();
  std::vector<int> q(d - 1);
  size_t total = 0;
  for (int k = 0; k < d - 1; ++k)
  {
    q[k] = std::min(X->ranks[k] * X->modes[k], X->ranks[k + 1]);
    total += (size_t)q[k];
  }
  double* sv = alloc_core(c, total);
  std::vector<double> host(total);
  try
  {
    size_t off = 0;
    for (int k = 0; k < d - 1; ++k)
    {
      singular_values_dev(c, X->cores[k], X->ranks[k] * X->modes[k], X->ranks[k + 1], sv + off);
      off += (size_t)q[k];
    }
    hip_check(hipMemcpyAsync(host.data(), sv, total * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  catch (...)
  {
    c->mem.put(sv);
    throw;
  }
  c->mem.put(sv);
  std::vector<std::vector<double>> out(d - 1);
  size_t off = 0;
  for (int k = 0; k < d - 1; ++k)
  {
    out[k].assign(host.begin() + off, host.begin() + off + q[k]);
    off += (size_t)q[k];
  }
  return out;
}

// detail::als_engine (tt_als.hpp:54-224) on X in place.
void als_engine_dev(tta_ctx* c, const tta_tt* Y, tta_tt* X, int max_iters, double tol, bool right_canonical,
                    iter_info* info)
{
  const int d = Y->d();
  if (!right_canonical) right_orthogonalize_impl(c, X);
  env_chain E(c, d);
  for (int k = d - 1; k >= 1; --k) E.set_R(k, env_right_dev(c, Y, k, X, E.R[k + 1]));
  std::vector<std::vector<double>> s_prev;
  bool have_prev = false, converged = false;
  double change = 0.0;
  int iters = 0;
  for (int it = 0; it < max_iters; ++it)
  {
    for (int k = 0; k < d; ++k)
    {
      replace_core(c, X, k, als_local_dev(c, E.L[k], X->ranks[k], Y, k, E.R[k + 1], X->ranks[k + 1]));
      if (k < d - 1)
      {
        left_qr_step(c, X, k, false);
        E.set_L(k + 1, env_left_dev(c, E.L[k], X, k, Y));
      }
    }
    for (int k = d - 1; k >= 0; --k)
    {
      if (k < d - 1)
        replace_core(c, X, k, als_local_dev(c, E.L[k], X->ranks[k], Y, k, E.R[k + 1], X->ranks[k + 1]));
      if (k > 0)
      {
        right_qr_step(c, X, k, false);
        E.set_R(k, env_right_dev(c, Y, k, X, E.R[k + 1]));
      }
    }
    ++iters;
    std::vector<std::vector<double>> s_now = core_singular_values(c, X);
    change = have_prev ? sv_change(s_now, s_prev) : 1.0;
    s_prev = std::move(s_now);
    have_prev = true;
    if (change <= tol)
    {
      converged = true;
      break;
    }
  }
  info->iters = iters;
  info->change = change;
  info->converged = converged ? 1 : 0;
}

// A warm start is used when it has the core count and modes of Y, per-site
// (m, n) for TT-matrices (warm_start.hpp:22-60).
bool warm_compatible(const tta_tt* w, const tta_tt* Y)
{
  if (!w || w->d() != Y->d() || w->is_matrix() != Y->is_matrix()) return false;
  for (int k = 0; k < Y->d(); ++k)
    if (w->modes[k] != Y->modes[k]) return false;
  if (Y->is_matrix())
    for (int k = 0; k < Y->d(); ++k)
      if (w->row_modes[k] != Y->row_modes[k]) return false;
  return true;
}

// random(shape, max_rank, seed) (tt_factory.hpp:141-207) over Y's (packed)
// modes, uploaded.
tta_tt* random_like(tta_ctx* c, const tta_tt* Y, int max_rank, unsigned long long seed)
{
  const int d = Y->d();
  std::vector<int> ranks(d + 1);
  long long n = 0;
  int st = tta_random_host(d, Y->modes.data(), max_rank, seed, ranks.data(), nullptr, &n);
  if (st != TTA_OK) fail(st, "random: invalid shape or rank cap");
  std::vector<double> host((size_t)n);
  tta_random_host(d, Y->modes.data(), max_rank, seed, nullptr, host.data(), nullptr);
  tta_tt* t = new_tt(c, Y->modes, ranks, false);
  size_t off = 0;
  for (int k = 0; k < d; ++k)
  {
    upload_async(c, t->cores[k], host.data() + off, core_numel(t, k) * 8);
    off += core_numel(t, k);
  }
  return t;
}

// detail::als_round (tt_als.hpp:226-300), and with cold_random the start of
// matvec_als_round / matmat_als_round (random(shape, max_rank, seed), every
// start right-orthogonalized by the engine).
tta_tt* als_round_dev(tta_ctx* c, const tta_tt* Y, double eps, int max_rank, int max_iters, double tol,
                      const tta_tt* warm, bool cold_random, unsigned long long seed, iter_info* info, const char* what)
{
  if (std::isnan(eps)) fail(TTA_ERR_ARG, "%s: eps is NaN; must be a valid number", what);
  const int d = Y->d();
  if (d <= 1)
  {
    tta_tt* t = copy_tt(c, Y);
    try
    {
      round_inplace(c, t, eps, max_rank, TTA_GAUGE_NONE);
    }
    catch (...)
    {
      free_tt(c, t);
      throw;
    }
    *info = iter_info{0, 0.0, 1};
    return t;
  }
  tta_tt* svd_ref = nullptr;
  tta_tt* X = nullptr;
  try
  {
    if (max_rank <= 0)
    {
      if (eps <= 0.0)
        fail(TTA_ERR_ARG,
             "%s: max_rank <= 0 and eps <= 0; ALS requires either max_rank > 0 or eps > 0 to determine "
             "compression target",
             what);
      svd_ref = copy_tt(c, Y);
      round_inplace(c, svd_ref, eps, 0, TTA_GAUGE_NONE);
      max_rank = *std::max_element(svd_ref->ranks.begin(), svd_ref->ranks.end());
      if (!warm) warm = svd_ref;
    }
    const bool use_warm = warm_compatible(warm, Y);
    if (use_warm)
      for (int r : warm->ranks)
        if (r > max_rank) fail(TTA_ERR_ARG, "%s: warm_start max rank %d exceeds max_rank %d", what, r, max_rank);
    if (use_warm)
      X = copy_tt(c, warm);
    else if (cold_random)
      X = random_like(c, Y, max_rank, seed);
    else
    {
      X = copy_tt(c, Y);
      round_inplace(c, X, 0.0, max_rank, TTA_GAUGE_RIGHT_CANONICAL);
    }
    X->row_modes = Y->row_modes;
    X->col_modes = Y->col_modes;
    if (svd_ref)
    {
      free_tt(c, svd_ref);
      svd_ref = nullptr;
    }
    als_engine_dev(c, Y, X, max_iters, tol, !use_warm && !cold_random, info);
  }
  catch (...)
  {
    free_tt(c, svd_ref);
    free_tt(c, X);
    throw;
  }
  return X;
}

// ---------------------------------------------------------------------------
// Two-site DMRG rounding (include/core/round/tt_dmrg.hpp:53-310; the apply
// paths tt_matvec_dmrg_round.hpp:29-105, tt_matmat_dmrg_round.hpp:38-130).
// The L->R split keeps U (X_k) and the kept singular values; the R->L split
// factors Theta^T, whose left factor V is the right-orthogonal X_{k+1}.  The
// factors the reference forms and the next bond overwrites unread
// (diag(s) V^T going right, U diag(s) going left except at bond 0) and the
// environments no bond reads (L_{d-1}, R_1) are not formed.
// ---------------------------------------------------------------------------

// merge_y_block (include/core/detail/dmrg_local_solve.hpp:34-75): Theta
// (rXl n_k x n_{k+1} rXr) row-major into Th, or its transpose.
void merge_block_dev(tta_ctx* c, const double* L, int rXl, const tta_tt* Y, int k, const double* R, int rXr,
                     bool transposed, double* Th)
{
  const int rYl = Y->ranks[k], nk = Y->modes[k], rYm = Y->ranks[k + 1], nk1 = Y->modes[k + 1],
            rYr = Y->ranks[k + 2];
  const int rows = rXl * nk, cols = nk1 * rXr;
  double* T1 = alloc_core(c, (size_t)rXl * nk * rYm);
  double* T2 = alloc_core(c, (size_t)rYm * nk1 * rXr);
  gemm_rm(c, rXl, nk * rYm, rYl, L, rYl, Y->cores[k], (long)nk * rYm, T1, (long)nk * rYm);
  gemm_rm(c, rYm * nk1, rXr, rYr, Y->cores[k + 1], rYr, R, rXr, T2, rXr);
  gemm_desc g;
  g.M = rows;
  g.N = cols;
  g.K = rYm;
  g.A = T1;
  g.sar = rYm;
  g.sac = 1;
  g.B = T2;
  g.sbr = cols;
  g.sbc = 1;
  g.C = Th;
  g.scr = transposed ? 1 : cols;
  g.scc = transposed ? rows : 1;
  run_gemm(c, g);
  c->mem.put(T1);
  c->mem.put(T2);
}

// detail::dmrg_engine (tt_dmrg.hpp:53-230) on X in place.
void dmrg_engine_dev(tta_ctx* c, const tta_tt* Y, tta_tt* X, double eps, int max_rank, int max_iters, double tol,
                     bool right_canonical, iter_info* info)
{
  const int d = Y->d();
  if (!right_canonical) right_orthogonalize_impl(c, X);
  const double y_norm = std::sqrt(std::max(inner_impl(c, Y, Y), 0.0));
  const double delta = eps * y_norm / std::sqrt((double)std::max(d - 1, 1));
  if (y_norm == 0.0 && max_iters > 0)
  {
    // Every bond target is zero: each split keeps one zero singular value,
    // so the result is the rank-1 zero tensor and the change is exactly 0
    // from the second sweep on.
    make_rank1_zero(c, X);
    const bool conv = max_iters >= 2 && 0.0 <= tol;
    const int it = conv ? 2 : max_iters;
    *info = iter_info{it, it >= 2 ? 0.0 : 1.0, conv ? 1 : 0};
    return;
  }
  env_chain E(c, d);
  for (int k = d - 1; k >= 2; --k) E.set_R(k, env_right_dev(c, Y, k, X, E.R[k + 1]));
  std::vector<std::vector<double>> s_prev;
  bool have_prev = false, converged = false;
  double change = 0.0;
  int iters = 0;
  std::vector<double*> sblk(d - 1, nullptr);
  std::vector<int> skept(d - 1, 0);
  auto release = [&]()
  {
    for (double*& p : sblk)
      if (p)
      {
        c->mem.put(p);
        p = nullptr;
      }
  };
  try
  {
    for (int it = 0; it < max_iters; ++it)
    {
      for (int k = 0; k < d - 1; ++k)
      {
        const int rXl = X->ranks[k], rXr = X->ranks[k + 2];
        const int rows = rXl * Y->modes[k], cols = Y->modes[k + 1] * rXr;
        double* Th = ctx_ws(c, WS_M1, (size_t)rows * cols);
        merge_block_dev(c, E.L[k], rXl, Y, k, E.R[k + 2], rXr, false, Th);
        double *U = nullptr, *F = nullptr, *s = nullptr;
        const int r = svd_left(c, Th, rows, cols, delta, max_rank, false, &U, &F, &s);
        replace_core(c, X, k, U);
        X->ranks[k + 1] = r;
        if (sblk[k]) c->mem.put(sblk[k]);
        sblk[k] = alloc_core(c, (size_t)r);
        skept[k] = r;
        hip_check(hipMemcpyAsync(sblk[k], s, (size_t)r * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
        if (k < d - 2) E.set_L(k + 1, env_left_dev(c, E.L[k], X, k, Y));
      }
      for (int k = d - 2; k >= 0; --k)
      {
        const int rXl = X->ranks[k], rXr = X->ranks[k + 2];
        const int rows = rXl * Y->modes[k], cols = Y->modes[k + 1] * rXr;
        double* ThT = ctx_ws(c, WS_M1, (size_t)rows * cols);
        merge_block_dev(c, E.L[k], rXl, Y, k, E.R[k + 2], rXr, true, ThT);
        double *V = nullptr, *G = nullptr;
        const int r = svd_left(c, ThT, cols, rows, delta, max_rank, false, &V, &G);
        double* Xn1 = alloc_core(c, (size_t)r * cols);
        launch_transpose(V, r, cols, r, Xn1, cols, c->stream);
        c->mem.put(V);
        replace_core(c, X, k + 1, Xn1);
        X->ranks[k + 1] = r;
        if (k == 0)
        {
          double* Xn0 = alloc_core(c, (size_t)rows * r);
          launch_transpose(G, rows, r, rows, Xn0, r, c->stream);
          replace_core(c, X, 0, Xn0);
        }
        else
          E.set_R(k + 1, env_right_dev(c, Y, k + 1, X, E.R[k + 2]));
      }
      ++iters;
      std::vector<std::vector<double>> s_now(d - 1);
      for (int k = 0; k < d - 1; ++k)
      {
        s_now[k].resize((size_t)skept[k]);
        hip_check(hipMemcpyAsync(s_now[k].data(), sblk[k], (size_t)skept[k] * 8, hipMemcpyDeviceToHost, c->stream),
                  "hipMemcpyAsync");
      }
      hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
      change = have_prev ? sv_change(s_now, s_prev) : 1.0;
      s_prev = std::move(s_now);
      have_prev = true;
      if (change <= tol)
      {
        converged = true;
        break;
      }
    }
  }
  catch (...)
  {
    release();
    throw;
  }
  release();
  info->iters = iters;
  info->change = change;
  info->converged = converged ? 1 : 0;
}

// detail::dmrg_round (tt_dmrg.hpp:235-310), and with apply_path the starts
// of matvec_dmrg_round / matmat_dmrg_round: random(modes, max_rank, seed)
// when max_rank > 0, else round_right_canonical(Y, eps, 0).
tta_tt* dmrg_round_dev(tta_ctx* c, const tta_tt* Y, double eps, int max_rank, int max_iters, double tol,
                       const tta_tt* warm, bool apply_path, unsigned long long seed, iter_info* info, const char* what)
{
  if (std::isnan(eps) || eps <= 0.0)
    fail(TTA_ERR_ARG,
         "%s: eps must be > 0 for DMRG compression (got %.15e); eps=0 means 'no truncation' only for the SVD path",
         what, eps);
  const int d = Y->d();
  if (d <= 1)
  {
    tta_tt* t = copy_tt(c, Y);
    try
    {
      round_inplace(c, t, eps, max_rank, TTA_GAUGE_NONE);
    }
    catch (...)
    {
      free_tt(c, t);
      throw;
    }
    *info = iter_info{0, 0.0, 1};
    return t;
  }
  const bool use_warm = warm_compatible(warm, Y);
  if (use_warm && max_rank > 0)
    for (int r : warm->ranks)
      if (r > max_rank) fail(TTA_ERR_ARG, "%s: warm_start max rank %d exceeds max_rank %d", what, r, max_rank);
  tta_tt* X = nullptr;
  bool rc = false;
  try
  {
    if (use_warm)
      X = copy_tt(c, warm);
    else if (apply_path && max_rank > 0)
      X = random_like(c, Y, max_rank, seed);
    else
    {
      X = copy_tt(c, Y);
      round_inplace(c, X, eps, apply_path ? 0 : max_rank, TTA_GAUGE_RIGHT_CANONICAL);
      rc = true;
    }
    X->row_modes = Y->row_modes;
    X->col_modes = Y->col_modes;
    dmrg_engine_dev(c, Y, X, eps, max_rank, max_iters, tol, rc, info);
  }
  catch (...)
  {
    free_tt(c, X);
    throw;
  }
  return X;
}

void check_tt(const tta_tt* t, const char* what)
{
  if (!t) fail(TTA_ERR_ARG, "%s: null tensor", what);
}

void check_same_shape(const tta_tt* a, const tta_tt* b, const char* what)
{
  check_tt(a, what);
  check_tt(b, what);
  if (a->d() != b->d()) fail(TTA_ERR_SHAPE, "%s: core count mismatch %d vs %d", what, a->d(), b->d());
  for (int k = 0; k < a->d(); ++k)
    if (a->modes[k] != b->modes[k])
      fail(TTA_ERR_SHAPE, "%s: mode size mismatch at core %d (%d vs %d)", what, k, a->modes[k], b->modes[k]);
}

void check_matvec(const tta_tt* A, const tta_tt* x, const char* what)
{
  check_tt(A, what);
  check_tt(x, what);
  if (!A->is_matrix()) fail(TTA_ERR_ARG, "%s: A is not a TT-matrix", what);
  if (x->is_matrix()) fail(TTA_ERR_ARG, "%s: x is a TT-matrix", what);
  if (A->d() != x->d()) fail(TTA_ERR_SHAPE, "%s: core count mismatch %d vs %d", what, A->d(), x->d());
  for (int k = 0; k < A->d(); ++k)
    if (A->col_modes[k] != x->modes[k])
      fail(TTA_ERR_SHAPE, "%s: mode mismatch at core %d (A col %d vs x %d)", what, k, A->col_modes[k], x->modes[k]);
}

void check_matmat(const tta_tt* A, const tta_tt* B, const char* what)
{
  check_tt(A, what);
  check_tt(B, what);
  if (!A->is_matrix() || !B->is_matrix()) fail(TTA_ERR_ARG, "%s: operands must be TT-matrices", what);
  if (A->d() != B->d()) fail(TTA_ERR_SHAPE, "%s: core count mismatch %d vs %d", what, A->d(), B->d());
  for (int k = 0; k < A->d(); ++k)
    if (A->col_modes[k] != B->row_modes[k])
      fail(TTA_ERR_SHAPE, "%s: mode mismatch at core %d (%d vs %d)", what, k, A->col_modes[k], B->row_modes[k]);
}

// round(tt / tt_matrix, round_options, round_result*) (round.hpp:230-420)
// for every method; iteration diagnostics in info (svd_result_default:
// {1, 0, true}).
tta_tt* round_dispatch(tta_ctx* c, const tta_tt* a, double eps, int max_rank, int method, int gauge, int max_iters,
                       double tol, unsigned long long seed, const tta_tt* warm, iter_info* info)
{
  if (method == TTA_METHOD_SVD_STREAMING || method == TTA_METHOD_SVD_NAIVE)
    fail(TTA_ERR_METHOD,
         "round: round_method::svd_streaming / svd_naive are only valid on matvec_round / matmat_round "
         "respectively.");
  if (method != TTA_METHOD_SVD && method != TTA_METHOD_ALS && method != TTA_METHOD_DMRG)
    fail(TTA_ERR_METHOD, "round: unknown method %d", method);
  if (method == TTA_METHOD_DMRG)
  {
    // round_options_to_dmrg (round.hpp:155-176).
    if (std::isnan(eps) || eps <= 0.0)
      fail(TTA_ERR_ARG, "round: eps must be > 0 for method=dmrg (eps=0 means 'no truncation' only for the SVD path)");
    if (max_iters < 0) fail(TTA_ERR_ARG, "round: max_iters is negative");
    tta_tt* t = dmrg_round_dev(c, a, eps, max_rank, max_iters, tol, warm, false, seed, info, "dmrg_round");
    t->matrix = a->matrix;
    return t;
  }
  if (std::isnan(eps)) fail(TTA_ERR_ARG, "round: eps is NaN");
  if (method == TTA_METHOD_ALS)
  {
    if (max_iters < 0) fail(TTA_ERR_ARG, "round: max_iters is negative");
    tta_tt* t = als_round_dev(c, a, eps, max_rank, max_iters, tol, warm, false, seed, info,
                              a->is_matrix() ? "als_round_matrix" : "als_round");
    t->matrix = a->matrix;
    return t;
  }
  tta_tt* t = copy_tt(c, a);
  t->matrix = a->matrix;
  try
  {
    round_inplace(c, t, eps, max_rank, gauge);
  }
  catch (...)
  {
    free_tt(c, t);
    throw;
  }
  *info = iter_info{1, 0.0, 1};
  return t;
}

// matvec_round (matvec_round.hpp:42-91) for every method.
tta_tt* matvec_round_dispatch(tta_ctx* c, const tta_tt* A, const tta_tt* x, double eps, int max_rank, int method,
                              int max_iters, double tol, unsigned long long seed, const tta_tt* warm, iter_info* info)
{
  if (method == TTA_METHOD_SVD_NAIVE) fail(TTA_ERR_METHOD, "matvec_round: round_method::svd_naive is matmat-only");
  if (method != TTA_METHOD_SVD && method != TTA_METHOD_SVD_STREAMING && method != TTA_METHOD_ALS &&
      method != TTA_METHOD_DMRG)
    fail(TTA_ERR_METHOD, "matvec_round: unknown round_method %d", method);
  check_matvec(A, x, "matvec_round");
  if (method == TTA_METHOD_SVD_STREAMING && A->d() > 1)
  {
    *info = iter_info{1, 0.0, 1};
    return streaming_impl(c, A, x, false, eps, max_rank);
  }
  if (method == TTA_METHOD_ALS && std::isnan(eps))
    fail(TTA_ERR_ARG, "matvec_als_round: eps is NaN; must be a valid number");
  if (method == TTA_METHOD_DMRG && (std::isnan(eps) || eps <= 0.0))
    fail(TTA_ERR_ARG,
         "matvec_dmrg_round: eps must be > 0 for DMRG compression (got %.15e); eps=0 means 'no truncation' only "
         "for the SVD path",
         eps);
  if ((method == TTA_METHOD_ALS || method == TTA_METHOD_DMRG) && max_iters < 0)
    fail(TTA_ERR_ARG, "matvec_round: max_iters is negative");
  tta_tt* Y = matvec_impl(c, A, x);
  if (method == TTA_METHOD_ALS || method == TTA_METHOD_DMRG)
  {
    tta_tt* r = nullptr;
    try
    {
      r = method == TTA_METHOD_ALS
            ? als_round_dev(c, Y, eps, max_rank, max_iters, tol, warm, true, seed, info, "matvec_als_round")
            : dmrg_round_dev(c, Y, eps, max_rank, max_iters, tol, warm, true, seed, info, "matvec_dmrg_round");
    }
    catch (...)
    {
      free_tt(c, Y);
      throw;
    }
    free_tt(c, Y);
    return r;
  }
  try
  {
    round_inplace(c, Y, eps, max_rank, TTA_GAUGE_NONE);
  }
  catch (...)
  {
    free_tt(c, Y);
    throw;
  }
  *info = iter_info{1, 0.0, 1};
  return Y;
}

// matmat_round (matmat_round.hpp:41-94) for every method.
tta_tt* matmat_round_dispatch(tta_ctx* c, const tta_tt* A, const tta_tt* B, double eps, int max_rank, int method,
                              int max_iters, double tol, unsigned long long seed, const tta_tt* warm, iter_info* info)
{
  if (method == TTA_METHOD_SVD_STREAMING)
    fail(TTA_ERR_METHOD, "matmat_round: round_method::svd_streaming is matvec-only");
  if (method != TTA_METHOD_SVD && method != TTA_METHOD_SVD_NAIVE && method != TTA_METHOD_ALS &&
      method != TTA_METHOD_DMRG)
    fail(TTA_ERR_METHOD, "matmat_round: unknown round_method %d", method);
  check_matmat(A, B, "matmat_round");
  tta_tt* r = nullptr;
  if (method == TTA_METHOD_SVD && A->d() > 1)
  {
    r = streaming_impl(c, A, B, true, eps, max_rank);
    r->matrix = true;
    *info = iter_info{1, 0.0, 1};
    return r;
  }
  if (method == TTA_METHOD_ALS && std::isnan(eps))
    fail(TTA_ERR_ARG, "matmat_als_round: eps is NaN; must be a valid number");
  if (method == TTA_METHOD_DMRG && (std::isnan(eps) || eps <= 0.0))
    fail(TTA_ERR_ARG,
         "matmat_dmrg_round: eps must be > 0 for DMRG compression (got %.15e); eps=0 means 'no truncation' only "
         "for the SVD path",
         eps);
  if ((method == TTA_METHOD_ALS || method == TTA_METHOD_DMRG) && max_iters < 0)
    fail(TTA_ERR_ARG, "matmat_round: max_iters is negative");
  tta_tt* Y = matmat_impl(c, A, B);
  Y->matrix = true;
  if (method == TTA_METHOD_ALS || method == TTA_METHOD_DMRG)
  {
    try
    {
      r = method == TTA_METHOD_ALS
            ? als_round_dev(c, Y, eps, max_rank, max_iters, tol, warm, true, seed, info, "matmat_als_round")
            : dmrg_round_dev(c, Y, eps, max_rank, max_iters, tol, warm, true, seed, info, "matmat_dmrg_round");
    }
    catch (...)
    {
      free_tt(c, Y);
      throw;
    }
    free_tt(c, Y);
    r->matrix = true;
    return r;
  }
  try
  {
    round_inplace(c, Y, eps, max_rank, TTA_GAUGE_NONE);
  }
  catch (...)
  {
    free_tt(c, Y);
    throw;
  }
  *info = iter_info{1, 0.0, 1};
  return Y;
}

void write_info(const iter_info& in, int* iters_run, double* final_resid, int* converged)
{
  if (iters_run) *iters_run = in.iters;
  if (final_resid) *final_resid = in.change;
  if (converged) *converged = in.converged;
}

void validate_ranks(int d, const int* modes, const int* ranks)
{
  if (d < 0) fail(TTA_ERR_ARG, "negative core count");
  if (d == 0) return;
  if (!modes || !ranks) fail(TTA_ERR_ARG, "null modes/ranks");
  if (ranks[0] != 1 || ranks[d] != 1)
    fail(TTA_ERR_SHAPE, "boundary ranks must be 1 (r_0=%d, r_d=%d)", ranks[0], ranks[d]);
  for (int k = 0; k < d; ++k)
  {
    if (modes[k] <= 0) fail(TTA_ERR_SHAPE, "mode size %d at core %d is not positive", modes[k], k);
    if (ranks[k] <= 0) fail(TTA_ERR_SHAPE, "rank %d at bond %d is not positive", ranks[k], k);
  }
}

}  // namespace

// ----------------------------------------------------------------------------
// extern "C" entry points.
// ----------------------------------------------------------------------------
#define TTA_GUARD_BEGIN \
  try               \
  {
#define TTA_GUARD_END(ctx)                                  \
  }                                                         \
  catch (const tta::error& e)                               \
  {                                                         \
    if (ctx) (ctx)->last_error = e.what();                  \
    return e.code;                                          \
  }                                                         \
  catch (const std::exception& e)                           \
  {                                                         \
    if (ctx) (ctx)->last_error = e.what();                  \
    return TTA_ERR_HIP;                                     \
  }                                                         \
  return TTA_OK;

extern "C"
{

int tta_ctx_create(int device, tta_ctx** out)
{
  if (!out) return TTA_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return TTA_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return TTA_ERR_HIP;
  auto* c = new tta_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
  {
    delete c;
    return TTA_ERR_HIP;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_int), 64, hipHostMallocDefault) != hipSuccess)
  {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return TTA_ERR_HIP;
  }
  c->h_dbl = reinterpret_cast<double*>(c->h_int + 8);
  try
  {
    tta::noalloc_init(c);
  }
  catch (...)
  {
    (void)hipHostFree(c->h_int);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return TTA_ERR_HIP;
  }
  *out = c;
  return TTA_OK;
}

int tta_ctx_destroy(tta_ctx* c)
{
  if (!c) return TTA_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& w : c->ws)
    if (w.p) c->mem.put(w.p);
  free_stage(c);
  (void)hipHostFree(c->h_int);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return TTA_OK;
}

const char* tta_last_error(const tta_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int tta_synchronize(tta_ctx* c)
{
  if (!c) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  ctx_sync(c);
  TTA_GUARD_END(c)
}

void* tta_stream(tta_ctx* c) { return c ? (void*)c->stream : nullptr; }

int tta_tt_create(tta_ctx* c, int d, const int* modes, const int* ranks, tta_tt** out)
{
  if (!c || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  validate_ranks(d, modes, ranks);
  std::vector<int> md(modes, modes + d), rk(ranks, ranks + d + 1);
  if (d == 0) rk.assign(1, 1);
  *out = new_tt(c, md, rk, true);
  TTA_GUARD_END(c)
}

int tta_ttm_create(tta_ctx* c, int d, const int* row_modes, const int* col_modes, const int* ranks, tta_tt** out)
{
  if (!c || !out || (d > 0 && (!row_modes || !col_modes))) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  std::vector<int> md(d);
  for (int k = 0; k < d; ++k)
  {
    if (row_modes[k] <= 0 || col_modes[k] <= 0) fail(TTA_ERR_SHAPE, "mode sizes must be positive");
    md[k] = row_modes[k] * col_modes[k];
  }
  validate_ranks(d, md.data(), ranks);
  std::vector<int> rk(ranks, ranks + d + 1);
  if (d == 0) rk.assign(1, 1);
  tta_tt* t = new_tt(c, md, rk, true);
  t->row_modes.assign(row_modes, row_modes + d);
  t->col_modes.assign(col_modes, col_modes + d);
  t->matrix = true;
  *out = t;
  TTA_GUARD_END(c)
}

int tta_tt_free(tta_ctx* c, tta_tt* t)
{
  if (!c) return TTA_ERR_ARG;
  free_tt(c, t);
  return TTA_OK;
}

int tta_tt_dims(const tta_tt* t, int* d, int* is_matrix)
{
  if (!t) return TTA_ERR_ARG;
  if (d) *d = t->d();
  if (is_matrix) *is_matrix = t->is_matrix() ? 1 : 0;
  return TTA_OK;
}

int tta_tt_shape(const tta_tt* t, int* modes, int* row_modes, int* col_modes, int* ranks)
{
  if (!t) return TTA_ERR_ARG;
  const int d = t->d();
  for (int k = 0; k < d; ++k)
  {
    if (modes) modes[k] = t->modes[k];
    if (row_modes && t->is_matrix()) row_modes[k] = t->row_modes[k];
    if (col_modes && t->is_matrix()) col_modes[k] = t->col_modes[k];
  }
  if (ranks)
    for (int k = 0; k <= d; ++k) ranks[k] = t->ranks[k];
  return TTA_OK;
}

long long tta_tt_numel(const tta_tt* t)
{
  if (!t) return -1;
  long long s = 0;
  for (int k = 0; k < t->d(); ++k) s += (long long)core_numel(t, k);
  return s;
}

int tta_tt_upload(tta_ctx* c, tta_tt* t, const double* host)
{
  if (!c || !t || (!host && t->d() > 0)) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  size_t off = 0;
  for (int k = 0; k < t->d(); ++k)
  {
    const size_t n = core_numel(t, k);
    hip_check(hipMemcpyAsync(t->cores[k], host + off, n * 8, hipMemcpyHostToDevice, c->stream), "upload");
    off += n;
  }
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  TTA_GUARD_END(c)
}

int tta_tt_download(tta_ctx* c, const tta_tt* t, double* host)
{
  if (!c || !t || (!host && t->d() > 0)) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  size_t off = 0;
  for (int k = 0; k < t->d(); ++k)
  {
    const size_t n = core_numel(t, k);
    hip_check(hipMemcpyAsync(host + off, t->cores[k], n * 8, hipMemcpyDeviceToHost, c->stream), "download");
    off += n;
  }
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  TTA_GUARD_END(c)
}

int tta_tt_core_ptr(const tta_tt* t, int k, double** dev)
{
  if (!t || !dev || k < 0 || k >= t->d()) return TTA_ERR_ARG;
  *dev = t->cores[k];
  return TTA_OK;
}

int tta_tt_copy(tta_ctx* c, const tta_tt* t, tta_tt** out)
{
  if (!c || !t || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  tta_tt* r = copy_tt(c, t);
  r->matrix = t->matrix;
  *out = r;
  TTA_GUARD_END(c)
}

int tta_round(tta_ctx* c, const tta_tt* a, double eps, int max_rank, int method, int gauge, tta_tt** out)
{
  return tta_round_iterative(c, a, eps, max_rank, method, gauge, 20, 1.0e-12, 0, nullptr, out, nullptr, nullptr,
                             nullptr);
}

int tta_round_iterative(tta_ctx* c, const tta_tt* a, double eps, int max_rank, int method, int gauge, int max_iters,
                        double tol, unsigned long long seed, const tta_tt* warm, tta_tt** out, int* iters_run,
                        double* final_resid, int* converged)
{
  if (!c || !a || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  iter_info info;
  tta_tt* t = round_dispatch(c, a, eps, max_rank, method, gauge, max_iters, tol, seed, warm, &info);
  try
  {
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  catch (...)
  {
    free_tt(c, t);
    throw;
  }
  write_info(info, iters_run, final_resid, converged);
  *out = t;
  TTA_GUARD_END(c)
}

int tta_right_orthogonalize(tta_ctx* c, tta_tt* a)
{
  if (!c || !a) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  right_orthogonalize_impl(c, a);
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  TTA_GUARD_END(c)
}

int tta_left_orthogonalize(tta_ctx* c, tta_tt* a)
{
  if (!c || !a) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  left_orthogonalize_impl(c, a);
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  TTA_GUARD_END(c)
}

int tta_matvec(tta_ctx* c, const tta_tt* A, const tta_tt* x, tta_tt** y)
{
  if (!c || !y) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  check_matvec(A, x, "matvec");
  *y = matvec_impl(c, A, x);
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  TTA_GUARD_END(c)
}

int tta_matmat(tta_ctx* c, const tta_tt* A, const tta_tt* B, tta_tt** C)
{
  if (!c || !C) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  check_matmat(A, B, "matmat");
  tta_tt* r = matmat_impl(c, A, B);
  r->matrix = true;
  *C = r;
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  TTA_GUARD_END(c)
}

int tta_matvec_round(tta_ctx* c, const tta_tt* A, const tta_tt* x, double eps, int max_rank, int method, tta_tt** y)
{
  return tta_matvec_round_iterative(c, A, x, eps, max_rank, method, 20, 1.0e-12, 0, nullptr, y, nullptr, nullptr,
                                    nullptr);
}

int tta_matvec_round_iterative(tta_ctx* c, const tta_tt* A, const tta_tt* x, double eps, int max_rank, int method,
                               int max_iters, double tol, unsigned long long seed, const tta_tt* warm, tta_tt** y,
                               int* iters_run, double* final_resid, int* converged)
{
  if (!c || !y) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  iter_info info;
  tta_tt* r = matvec_round_dispatch(c, A, x, eps, max_rank, method, max_iters, tol, seed, warm, &info);
  try
  {
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  catch (...)
  {
    free_tt(c, r);
    throw;
  }
  write_info(info, iters_run, final_resid, converged);
  *y = r;
  TTA_GUARD_END(c)
}

int tta_matmat_round(tta_ctx* c, const tta_tt* A, const tta_tt* B, double eps, int max_rank, int method, tta_tt** C)
{
  return tta_matmat_round_iterative(c, A, B, eps, max_rank, method, 20, 1.0e-12, 0, nullptr, C, nullptr, nullptr,
                                    nullptr);
}

int tta_matmat_round_iterative(tta_ctx* c, const tta_tt* A, const tta_tt* B, double eps, int max_rank, int method,
                               int max_iters, double tol, unsigned long long seed, const tta_tt* warm, tta_tt** C,
                               int* iters_run, double* final_resid, int* converged)
{
  if (!c || !C) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  iter_info info;
  tta_tt* r = matmat_round_dispatch(c, A, B, eps, max_rank, method, max_iters, tol, seed, warm, &info);
  r->matrix = true;
  try
  {
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  catch (...)
  {
    free_tt(c, r);
    throw;
  }
  write_info(info, iters_run, final_resid, converged);
  *C = r;
  TTA_GUARD_END(c)
}

int tta_inner(tta_ctx* c, const tta_tt* a, const tta_tt* b, double* out)
{
  if (!c || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  check_same_shape(a, b, "inner");
  *out = inner_impl(c, a, b);
  TTA_GUARD_END(c)
}

int tta_frob_norm_apply(tta_ctx* c, const tta_tt* A, const tta_tt* X, double* out)
{
  if (!c || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  check_tt(X, "frob_norm_apply");
  const bool mm = X->is_matrix();
  if (mm)
    check_matmat(A, X, "frob_norm_apply(matmat)");
  else
    check_matvec(A, X, "frob_norm_apply(matvec)");
  *out = frob_norm_apply_impl(c, A, X, mm);
  TTA_GUARD_END(c)
}

int tta_norm(tta_ctx* c, const tta_tt* a, double* out)
{
  if (!c || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  check_tt(a, "norm");
  const double s = inner_impl(c, a, a);
  *out = std::sqrt(std::max(s, 0.0));
  TTA_GUARD_END(c)
}

int tta_axpby(tta_ctx* c, double alpha, const tta_tt* a, double beta, const tta_tt* b, tta_tt** out)
{
  if (!c || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  check_same_shape(a, b, "axpby");
  const int d = a->d();
  std::vector<int> ranks(d + 1, 1);
  for (int k = 1; k < d; ++k) ranks[k] = a->ranks[k] + b->ranks[k];
  tta_tt* r = new_tt(c, a->modes, ranks, true);
  r->row_modes = a->row_modes;
  r->col_modes = a->col_modes;
  r->matrix = a->matrix;
  // Block-diagonal concatenation (include/core/detail/block_concat.hpp:40-156)
  // as scaled 2-D slab copies: the a-slab at (row 0, col 0), the b-slab at
  // (left rank raL, right rank raR) except on the head (left 0) and the
  // tail (right 0); alpha / beta folded into the head core only.
  if (d == 1)
  {
    // Single core: elementwise alpha*a + beta*b (tt_ops.hpp:85-101).
    const int n = a->modes[0];
    launch_copy2d(a->cores[0], n, r->cores[0], n, 1, n, alpha, 0.0, c->stream);
    launch_copy2d(b->cores[0], n, r->cores[0], n, 1, n, beta, 1.0, c->stream);
  }
  for (int k = 0; k < d && d > 1; ++k)
  {
    const int n = a->modes[k];
    const bool head = k == 0, tail = k == d - 1;
    const int rR = ranks[k + 1];
    for (int which = 0; which < 2; ++which)
    {
      const tta_tt* s = which == 0 ? a : b;
      const int sL = s->ranks[k], sR = s->ranks[k + 1];
      const int row_off = which == 0 ? 0 : (head ? 0 : a->ranks[k]);
      const int col_off = which == 0 ? 0 : (tail ? 0 : a->ranks[k + 1]);
      const double f = head ? (which == 0 ? alpha : beta) : 1.0;
      // Tail core: both slabs share column 0, so the b-slab accumulates.
      const double keep = (tail && which == 1) ? 1.0 : 0.0;
      launch_copy2d(s->cores[k], sR, r->cores[k] + ((size_t)row_off * n) * rR + col_off, rR, sL * n, sR, f, keep,
                    c->stream);
    }
  }
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  *out = r;
  TTA_GUARD_END(c)
}

int tta_scale(tta_ctx* c, const tta_tt* a, double alpha, tta_tt** out)
{
  if (!c || !a || !out) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  tta_tt* r = copy_tt(c, a);
  r->matrix = a->matrix;
  if (r->d() > 0) launch_scale(r->cores[0], (long)core_numel(r, 0), alpha, c->stream);
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  *out = r;
  TTA_GUARD_END(c)
}

int tta_scale_core(tta_ctx* c, tta_tt* t, int k, double alpha)
{
  if (!c || !t || k < 0 || k >= t->d()) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  launch_scale(t->cores[k], (long)core_numel(t, k), alpha, c->stream);
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  TTA_GUARD_END(c)
}

int tta_prof_enable(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  prof_collect(c);
  c->prof = on != 0;
  for (int i = 0; i < TTA_K_COUNT; ++i)
  {
    c->prof_ms[i] = 0.0;
    c->prof_n[i] = 0;
    c->prof_flops[i] = 0.0;
    c->prof_bytes[i] = 0.0;
  }
  c->jac_calls = 0;
  c->jac_sweeps = 0;
  c->jac_unconverged = 0;
  c->rr_tried = 0;
  c->rr_held = 0;
  c->cert_tried = 0;
  c->cert_held = 0;
  c->cholqr_tried = 0;
  c->cholqr_used = 0;
  c->cholqr_triangular = 0;
  c->spec_sweeps = 0;
  c->spec_failed = 0;
  c->chain_blocks = 0;
  c->chain_steps = 0;
  c->chain_one_pass = 0;
  c->chain_rejected = 0;
  TTA_GUARD_END(c)
}

int tta_set_rank_certificate(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  c->rank_certificate = on != 0;
  return TTA_OK;
}

int tta_set_cholqr(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  c->cholqr = on != 0;
  return TTA_OK;
}

int tta_set_speculation(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  c->speculate = on != 0;
  c->spec_svd_ok = true;
  return TTA_OK;
}

int tta_prof_speculation(tta_ctx* c, long long* sweeps, long long* failed)
{
  if (!c) return TTA_ERR_ARG;
  if (sweeps) *sweeps = c->spec_sweeps;
  if (failed) *failed = c->spec_failed;
  return TTA_OK;
}

int tta_set_chain(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  c->chain = on != 0;
  return TTA_OK;
}

int tta_prof_chain(tta_ctx* c, long long* blocks, long long* steps, long long* one_pass, long long* rejected)
{
  if (!c) return TTA_ERR_ARG;
  if (blocks) *blocks = c->chain_blocks;
  if (steps) *steps = c->chain_steps;
  if (one_pass) *one_pass = c->chain_one_pass;
  if (rejected) *rejected = c->chain_rejected;
  return TTA_OK;
}

int tta_prof_cholqr_triangular(tta_ctx* c, long long* triangular)
{
  if (!c) return TTA_ERR_ARG;
  if (triangular) *triangular = c->cholqr_triangular;
  return TTA_OK;
}

int tta_prof_cholqr(tta_ctx* c, long long* tried, long long* used)
{
  if (!c) return TTA_ERR_ARG;
  if (tried) *tried = c->cholqr_tried;
  if (used) *used = c->cholqr_used;
  return TTA_OK;
}

int tta_prof_rank_certificates(tta_ctx* c, long long* tried, long long* held)
{
  if (!c) return TTA_ERR_ARG;
  if (tried) *tried = c->cert_tried;
  if (held) *held = c->cert_held;
  return TTA_OK;
}

int tta_prof_jacobi_sweeps(tta_ctx* c, long long* calls, long long* sweeps)
{
  if (!c) return TTA_ERR_ARG;
  if (calls) *calls = c->jac_calls;
  if (sweeps) *sweeps = c->jac_sweeps;
  return TTA_OK;
}

int tta_prof_jacobi_unconverged(tta_ctx* c, long long* problems)
{
  if (!c || !problems) return TTA_ERR_ARG;
  *problems = c->jac_unconverged;
  return TTA_OK;
}

int tta_set_autotune(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  tta::set_gemm_autotune(on != 0);
  return TTA_OK;
}

int tta_prof_autotune(tta_ctx* c, long long* keys, long long* decided, long long* chose_small)
{
  if (!c) return TTA_ERR_ARG;
  tta::gemm_autotune_stats(keys, decided, chose_small);
  return TTA_OK;
}

int tta_set_graphs(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  c->graphs = on != 0;
  return TTA_OK;
}

int tta_set_jacobi_gram(tta_ctx* c, int on)
{
  if (!c) return TTA_ERR_ARG;
  c->jac_gram = on != 0;
  return TTA_OK;
}

int tta_prof_rank_revealing(tta_ctx* c, long long* tried, long long* held)
{
  if (!c) return TTA_ERR_ARG;
  if (tried) *tried = c->rr_tried;
  if (held) *held = c->rr_held;
  return TTA_OK;
}

int tta_prof_read(tta_ctx* c, int cls, long long* launches, double* total_ms, double* flops, double* bytes)
{
  if (!c || cls < 0 || cls >= TTA_K_COUNT) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  prof_collect(c);
  if (launches) *launches = c->prof_n[cls];
  if (total_ms) *total_ms = c->prof_ms[cls];
  if (flops) *flops = c->prof_flops[cls];
  if (bytes) *bytes = c->prof_bytes[cls];
  TTA_GUARD_END(c)
}

int tta_mem_stats(tta_ctx* c, unsigned long long* live_bytes, unsigned long long* cached_bytes,
                  unsigned long long* device_allocs, unsigned long long* trims)
{
  if (!c) return TTA_ERR_ARG;
  if (live_bytes) *live_bytes = c->mem.live_bytes;
  if (cached_bytes) *cached_bytes = c->mem.cached_bytes;
  if (device_allocs) *device_allocs = c->mem.device_allocs;
  if (trims) *trims = c->mem.trims;
  return TTA_OK;
}

int tta_mem_trim(tta_ctx* c)
{
  if (!c) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  c->mem.trim();
  TTA_GUARD_END(c)
}

// detail::qr_thin (include/core/types/tt_eigen_bridge.hpp:42-60) on host
// buffers: M (rows x cols) row-major -> Q (rows x k), R (k x cols), k = min.
int tta_qr_thin(tta_ctx* c, const double* M, int rows, int cols, double* Q, double* R)
{
  if (!c || rows < 0 || cols < 0) return TTA_ERR_ARG;
  if (rows == 0 || cols == 0) return TTA_OK;
  if (!M || !Q || !R) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  const int k = std::min(rows, cols);
  double* Md = alloc_core(c, (size_t)rows * cols);
  hip_check(hipMemcpyAsync(Md, M, (size_t)rows * cols * 8, hipMemcpyHostToDevice, c->stream), "upload");
  double* A = ctx_ws(c, WS_A, (size_t)rows * cols);
  launch_transpose(Md, cols, rows, cols, A, rows, c->stream);
  double* tau = ctx_ws(c, WS_F, (size_t)k);
  double* Qc = ctx_ws(c, WS_Q, (size_t)rows * k);
  qr_factor(c, A, rows, rows, cols, tau, Qc, rows);
  double* Rd = ctx_ws(c, WS_X, (size_t)k * cols);
  launch_extract_upper(A, rows, k, cols, Rd, cols, 1, c->stream);
  launch_transpose(Qc, rows, k, rows, Md, k, c->stream);
  hip_check(hipMemcpyAsync(Q, Md, (size_t)rows * k * 8, hipMemcpyDeviceToHost, c->stream), "download");
  hip_check(hipMemcpyAsync(R, Rd, (size_t)k * cols * 8, hipMemcpyDeviceToHost, c->stream), "download");
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  c->mem.put(Md);
  TTA_GUARD_END(c)
}

// detail::svd_thin (tt_eigen_bridge.hpp:67-104) on host buffers: U (rows x
// k), s (k, descending), Vt (k x cols).  Vt rows of zero singular values are 0.
int tta_svd_thin(tta_ctx* c, const double* M, int rows, int cols, double* U, double* s, double* Vt)
{
  if (!c || rows < 0 || cols < 0) return TTA_ERR_ARG;
  if (rows == 0 || cols == 0) return TTA_OK;
  if (!M || !U || !s || !Vt) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  const int k = std::min(rows, cols);
  double* Md = alloc_core(c, (size_t)rows * cols);
  hip_check(hipMemcpyAsync(Md, M, (size_t)rows * cols * 8, hipMemcpyHostToDevice, c->stream), "upload");
  double *Ud = nullptr, *Fd = nullptr, *sd = nullptr;
  svd_left(c, Md, rows, cols, 0.0, 0, true, &Ud, &Fd, &sd);
  std::vector<double> F((size_t)k * cols);
  hip_check(hipMemcpyAsync(U, Ud, (size_t)rows * k * 8, hipMemcpyDeviceToHost, c->stream), "download");
  hip_check(hipMemcpyAsync(F.data(), Fd, (size_t)k * cols * 8, hipMemcpyDeviceToHost, c->stream), "download");
  hip_check(hipMemcpyAsync(s, sd, (size_t)k * 8, hipMemcpyDeviceToHost, c->stream), "download");
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < cols; ++j) Vt[(size_t)i * cols + j] = s[i] > 0.0 ? F[(size_t)i * cols + j] / s[i] : 0.0;
  c->mem.put(Md);
  c->mem.put(Ud);
  TTA_GUARD_END(c)
}

// soft_threshold (include/core/round/tt_soft_threshold.hpp:42-176): a copy of
// a when d <= 1 or tau <= 0; otherwise right-orthogonalize, then for
// k = d-1..1 the thin SVD of left_unfold(c_{k-1}) right_unfold(c_k) (device
// GEMM + QR-condensed Jacobi), singular values shrunk to max(s - tau, 0) and
// cut after the last positive one: c_{k-1} <- U diag(s - tau), c_k <- Vt.
// All values zeroed: rank-1 zero cores at that bond.
int tta_soft_threshold(tta_ctx* c, const tta_tt* a, double tau, tta_tt** out)
{
  if (!c || !a || !out) return TTA_ERR_ARG;
  if (a->is_matrix()) return TTA_ERR_SHAPE;
  TTA_GUARD_BEGIN
  tta_tt* t = copy_tt(c, a);
  try
  {
    const int d = t->d();
    if (d > 1 && tau > 0.0)
    {
      right_orthogonalize_impl(c, t);
      for (int k = d - 1; k >= 1; --k)
      {
        const int rL = t->ranks[k - 1], np = t->modes[k - 1], r = t->ranks[k], nc = t->modes[k], rR = t->ranks[k + 1];
        const int rows = rL * np, cols = nc * rR, kk = std::min(rows, cols);
        double* M = alloc_core(c, (size_t)rows * cols);
        gemm_rm(c, rows, cols, r, t->cores[k - 1], r, t->cores[k], cols, M, cols);
        double *U = nullptr, *F = nullptr, *sd = nullptr;
        svd_left(c, M, rows, cols, 0.0, 0, true, &U, &F, &sd);
        c->mem.put(M);
        std::vector<double> s((size_t)kk);
        hip_check(hipMemcpyAsync(s.data(), sd, (size_t)kk * 8, hipMemcpyDeviceToHost, c->stream), "s");
        ctx_sync(c);
        int r_new = 0;
        std::vector<double> sh((size_t)kk, 0.0), inv((size_t)kk, 0.0);
        for (int j = 0; j < kk; ++j)
          if (s[j] > tau)
          {
            sh[j] = s[j] - tau;
            inv[j] = 1.0 / s[j];
            r_new = j + 1;
          }
        if (r_new == 0)
        {
          c->mem.put(U);
          double* zp = alloc_core(c, (size_t)rows);
          double* zc = alloc_core(c, (size_t)cols);
          hip_check(hipMemsetAsync(zp, 0, (size_t)rows * 8, c->stream), "zero");
          hip_check(hipMemsetAsync(zc, 0, (size_t)cols * 8, c->stream), "zero");
          replace_core(c, t, k - 1, zp);
          replace_core(c, t, k, zc);
          t->ranks[k] = 1;
          continue;
        }
        double* f = ctx_ws(c, WS_JP, 2 * (size_t)r_new);
        std::vector<double> fh(sh.begin(), sh.begin() + r_new);
        fh.insert(fh.end(), inv.begin(), inv.begin() + r_new);
        upload_async(c, f, fh.data(), fh.size() * 8);
        double* Pn = alloc_core(c, (size_t)rows * r_new);
        launch_scale_copy(U, rows, r_new, kk, nullptr, f, Pn, c->stream);
        double* Cn = alloc_core(c, (size_t)r_new * cols);
        launch_scale_copy(F, r_new, cols, cols, f + r_new, nullptr, Cn, c->stream);
        c->mem.put(U);
        replace_core(c, t, k - 1, Pn);
        replace_core(c, t, k, Cn);
        t->ranks[k] = r_new;
      }
    }
  }
  catch (...)
  {
    for (double* p : t->cores) c->mem.put(p);
    delete t;
    throw;
  }
  *out = t;
  TTA_GUARD_END(c)
}

// from_dense (include/core/factory/from_dense.hpp:215-225): method 0 = TT-SVD
// (tt_svd.hpp:46-160); method 1 = QTT (from_dense.hpp:109-199: each mode
// n = base^b split into b modes of size base, data reordered on the host,
// then TT-SVD).
int tta_from_dense(tta_ctx* c, const double* dense, int d, const int* shape, double eps, int max_rank, int method,
                   int qtt_base, tta_tt** out)
{
  if (!c || !out || d < 0 || (d > 0 && (!dense || !shape)) || !(eps >= 0.0) || max_rank < 0) return TTA_ERR_ARG;
  if (method != 0 && method != 1) return TTA_ERR_METHOD;
  TTA_GUARD_BEGIN
  std::vector<int> sh(shape, shape + d);
  if (method == 0)
  {
    *out = from_dense_impl(c, dense, sh, eps, max_rank);
  }
  else
  {
    if (qtt_base <= 1 || (qtt_base & (qtt_base - 1)) != 0)
      fail(TTA_ERR_ARG, "from_dense(qtt): base %d must be a power of 2 above 1", qtt_base);
    std::vector<int> bits(d, 0), qshape;
    for (int k = 0; k < d; ++k)
    {
      int n = sh[k], b = 0;
      if (n <= 0) fail(TTA_ERR_SHAPE, "from_dense(qtt): mode %d has size %d", k, n);
      while (n > 1)
      {
        if (n % qtt_base) fail(TTA_ERR_SHAPE, "from_dense(qtt): mode %d size %d is not a power of %d", k, sh[k], qtt_base);
        n /= qtt_base;
        ++b;
      }
      bits[k] = b;
      for (int i = 0; i < b; ++i) qshape.push_back(qtt_base);
    }
    if (qshape.empty())
    {
      // Every mode has size 1: a single scalar, one (1, 1, 1) core.
      *out = from_dense_impl(c, dense, std::vector<int>{1}, 0.0, 0);
    }
    else
    {
      std::vector<long long> stride(d, 1);
      for (int k = d - 2; k >= 0; --k) stride[k] = stride[k + 1] * sh[k + 1];
      size_t total = 1;
      for (int q : qshape) total *= (size_t)q;
      std::vector<double> qd(total);
      std::vector<int> idx(qshape.size(), 0);
      for (size_t pos = 0; pos < total; ++pos)
      {
        int bp = 0;
        long long orig = 0;
        for (int k = 0; k < d; ++k)
        {
          long long v = 0;
          for (int b = 0; b < bits[k]; ++b, ++bp) v = v * qtt_base + idx[bp];
          orig += v * stride[k];
        }
        qd[pos] = dense[orig];
        for (int i = (int)qshape.size() - 1; i >= 0; --i)
        {
          if (++idx[i] < qtt_base) break;
          idx[i] = 0;
        }
      }
      *out = from_dense_impl(c, qd.data(), qshape, eps, max_rank);
    }
  }
  TTA_GUARD_END(c)
}

// matrix_from_dense (include/core/factory/tt_matrix_factory.hpp:205-290): the
// (prod m) x (prod n) row-major matrix is permuted on the host to per-axis
// (m_k, n_k) pairs, then TT-SVD over the modes m_k n_k.
int tta_matrix_from_dense(tta_ctx* c, const double* dense, int d, const int* row_shape, const int* col_shape,
                          double eps, int max_rank, tta_tt** out)
{
  if (!c || !out || d < 1 || !dense || !row_shape || !col_shape || !(eps >= 0.0) || max_rank < 0) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  std::vector<int> ms(row_shape, row_shape + d), ns(col_shape, col_shape + d), mn(d);
  size_t M = 1, N = 1;
  for (int k = 0; k < d; ++k)
  {
    if (ms[k] <= 0 || ns[k] <= 0) fail(TTA_ERR_SHAPE, "matrix_from_dense: axis %d has size %d x %d", k, ms[k], ns[k]);
    mn[k] = ms[k] * ns[k];
    M *= (size_t)ms[k];
    N *= (size_t)ns[k];
  }
  if (M * N > (size_t)std::numeric_limits<int>::max()) fail(TTA_ERR_SHAPE, "matrix_from_dense: above INT_MAX entries");
  std::vector<double> perm(M * N);
  for (size_t I = 0; I < M; ++I)
    for (size_t J = 0; J < N; ++J)
    {
      size_t i = I, j = J, flat = 0, kstride = 1;
      for (int k = d - 1; k >= 0; --k)
      {
        const size_t ik = i % (size_t)ms[k], jk = j % (size_t)ns[k];
        i /= (size_t)ms[k];
        j /= (size_t)ns[k];
        flat += (ik * (size_t)ns[k] + jk) * kstride;
        kstride *= (size_t)mn[k];
      }
      perm[flat] = dense[I * N + J];
    }
  tta_tt* t = from_dense_impl(c, perm.data(), mn, eps, max_rank);
  t->row_modes = ms;
  t->col_modes = ns;
  t->matrix = true;
  *out = t;
  TTA_GUARD_END(c)
}

// Device self-check of the wave reduction used by every kernel (DPP path
// against the ds_bpermute path) on n host values; *mismatches = 0 when good.
int tta_selfcheck_wave_sum(tta_ctx* c, const double* host, int n, int* mismatches)
{
  if (!c || !host || n <= 0 || !mismatches) return TTA_ERR_ARG;
  TTA_GUARD_BEGIN
  double* x = alloc_core(c, (size_t)n);
  int* bad = small_int(c) + 8;
  hip_check(hipMemcpyAsync(x, host, (size_t)n * 8, hipMemcpyHostToDevice, c->stream), "upload");
  hip_check(hipMemsetAsync(bad, 0, sizeof(int), c->stream), "memset");
  launch_wave_sum_check(x, n, bad, c->stream);
  *mismatches = read_int(c, bad);
  c->mem.put(x);
  TTA_GUARD_END(c)
}

}  // extern "C"
