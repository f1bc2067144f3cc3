// Launch wrappers for the FP64 kernels in tta_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

struct tta_ctx;

namespace tta
{
class arena;  // bump allocator of a step's device buffers (tta_internal.hpp)

// C(i,j) = alpha * sum_k A(i,k) B(k,j) + beta * C(i,j), element strides per
// operand, batch strides bs* (elements) over blockIdx.z.
struct gemm_desc
{
  int M = 0, N = 0, K = 0;
  const double* A = nullptr;
  long sar = 0, sac = 0;
  const double* B = nullptr;
  long sbr = 0, sbc = 0;
  double* C = nullptr;
  long scr = 0, scc = 0;
  double alpha = 1.0, beta = 0.0;
  long bsa = 0, bsb = 0, bsc = 0;
};

// GEMM descriptor with untyped operands; the element types of one grouped
// launch are given by an xgemm_types code (operands A, B, C, then the
// compute type; D = FP64, S = FP32).
struct xgemm_desc
{
  int M = 0, N = 0, K = 0;
  const void* A = nullptr;
  long sar = 0, sac = 0;
  const void* B = nullptr;
  long sbr = 0, sbc = 0;
  void* C = nullptr;
  long scr = 0, scc = 0;
};
enum xgemm_types
{
  XG_D_D_D_D = 0,
  XG_D_S_S_S = 1,  // FP64 A, FP32 B -> FP32 C, FP32 matrix cores
  XG_S_S_S_S = 2,
  XG_D_S_D_D = 3   // FP64 A, FP32 B -> FP64 C, FP64 matrix cores
};

// Householder QR of `batch` column-major (m x n) matrices; thin Q optional.
struct qr_desc
{
  double* A = nullptr;
  int lda = 0, m = 0, n = 0;
  double* tau = nullptr;
  double* Q = nullptr;  // null: do not form Q
  int ldq = 0;
  long bsA = 0, bstau = 0, bsQ = 0;
};

// One-sided Jacobi on X (p x q col-major), J (q x q); sorted norms into s,
// order into perm; optional truncation rank (delta from delta_ptr if set).
struct jacobi_desc
{
  double* X = nullptr;
  int ldx = 0, p = 0, q = 0;
  double* J = nullptr;
  int ldj = 0;
  double* s = nullptr;
  int* perm = nullptr;
  int* sweeps = nullptr;
  int* rank_out = nullptr;
  const double* delta_ptr = nullptr;
  double delta = 0.0;
  int max_rank = 0;
  int max_sweeps = 60;
  long bsX = 0, bsJ = 0, bss = 0;
};

// One blocked-QR problem: column-major A (m x n, leading dimension lda),
// factored in place; tau (k), T (ceil(k/16) blocks of 16 x 16), optional
// thin Q (m x k, ldq), k = min(m, n).
struct qrp
{
  double* A = nullptr;
  int lda = 0, m = 0, n = 0;
  double* tau = nullptr;
  double* T = nullptr;
  double* Q = nullptr;
  int ldq = 0;
};

// A batch of problems: `descs` (device array of `batch` entries) or, when
// null, the single problem `one`.  max_m / max_n bound every problem.
struct qr_batch
{
  const qrp* descs = nullptr;
  qrp one;
  int batch = 1;
  int max_m = 0, max_n = 0;
  bool form_q = false;
};

// One square Jacobi SVD problem: X (q x q, column-major, ldx) is replaced by
// W = X J (orthogonal columns), J (q x q, ldj) accumulates the rotations.
// Then s (q, descending column norms of W), perm (their order) and, when
// rank_out is set, the truncation rank for delta (or *delta_ptr).
// flag: 2 device ints (per-sweep rotation flag, converged flag); sweeps:
// optional device counter of sweeps that rotated.
struct jac_prob
{
  double* X = nullptr;
  int ldx = 0, q = 0;
  double* J = nullptr;
  int ldj = 0;
  double* s = nullptr;
  int* perm = nullptr;
  int* rank_out = nullptr;
  const double* delta_ptr = nullptr;
  double delta = 0.0;
  int max_rank = 0;
  int* flag = nullptr;
  int* sweeps = nullptr;
};

struct jac_batch
{
  const jac_prob* descs = nullptr;
  jac_prob one;
  int batch = 1;
  // Optional work queue for the whole-problem kernel (q <= 64): the problems
  // in the order they are taken (device, `batch` ints, largest first) and one
  // device int of scratch; with both set the launch is persistent.
  const int* order = nullptr;
  int* queue = nullptr;
};

// Descriptor-batched helpers (one entry per problem, device arrays).
struct tr_desc  // out[j*ldo + i] = in[i*ldi + j], i < rows, j < cols
{
  const double* in = nullptr;
  int ldi = 0, rows = 0, cols = 0;
  double* out = nullptr;
  int ldo = 0;
};
struct ex_desc  // dst[i*drs + c*dcs] = (i <= c) ? A[c*lda + i] : 0
{
  const double* A = nullptr;
  int lda = 0, kr = 0, nc = 0;
  double* dst = nullptr;
  long drs = 0, dcs = 0;
};
// Column gathers: dst[t*drs + o*dcs] = src[t + perm[o]*lds], optionally
// multiplied by (GATHER_MUL) or divided by (GATHER_DIV) scale[o].
enum
{
  GATHER_COPY = 0,
  GATHER_MUL = 1,
  GATHER_DIV = 2,
  // Flag added to a mode: src is row-major, column perm[o] read as
  // src[t*lds + perm[o]].
  GATHER_ROWMAJOR = 4
};
struct ga_desc
{
  const double* src = nullptr;
  int lds = 0, rows = 0;
  const int* perm = nullptr;
  int ncols = 0;
  double* dst = nullptr;
  long drs = 0, dcs = 0;
  const double* scale = nullptr;
  int mode = GATHER_COPY;
};
struct cp_desc
{
  const double* src = nullptr;
  double* dst = nullptr;
  long n = 0;
};
struct sq_desc  // out_norm = ||x||, out_delta = scale * ||x||
{
  const double* x = nullptr;
  long n = 0;
  double* out_norm = nullptr;
  double* out_delta = nullptr;
  double scale = 0.0;
};
struct cn_desc
{
  const double* M = nullptr;
  int rows = 0, cols = 0;
  double* out = nullptr;
};
struct gm_desc
{
  const double* v = nullptr;
  int n = 0;
  double* out = nullptr;
};
struct bal_desc
{
  double* x = nullptr;
  int r = 0, n = 0, rp = 0;
  const double* nl = nullptr;
  const double* nr = nullptr;
  const double* geo = nullptr;
};

struct permute_desc
{
  const double* src = nullptr;
  double* dst = nullptr;
  long n[6] = {1, 1, 1, 1, 1, 1};
  long ss[6] = {0, 0, 0, 0, 0, 0};
};

// One problem of the batched pivoted Cholesky (tta_rrqr.hip): G (q x q,
// row-major, ldg) factored with diagonal pivoting and a trace stop at
// stop_f * (*delta)^2 (pivot = 1), or in order without pivoting (pivot = 0;
// p_out = -1 on a non-positive pivot).  Writes the pivot order piv (q; the
// identity without pivoting, may be null then), the pivot count p_out, the
// factor's columns L (q x q scratch, column-major, original row index), and
// R^-1 (p x p upper triangular, row-major, ldr) for R = L11^T in pivot order.
struct pchol_desc
{
  const double* G = nullptr;
  int ldg = 0, q = 0;
  const double* delta = nullptr;
  double stop_f = 0.0;
  int pivot = 1;
  int* piv = nullptr;
  int* p_out = nullptr;
  double* L = nullptr;
  double* Rinv = nullptr;
  int ldr = 0;
  double* tr_out = nullptr;  // optional: trace of the Schur complement left at the stop
  // Optional: the problem is skipped when *skip != 0 (its outputs were
  // written by launch_ns_pass2).
  const int* skip = nullptr;
};
constexpr int kPcholMaxQ = 192;  // packed FP64 triangle in LDS
constexpr int kPcholBlock = 4;   // pivots per trailing update (pivoted form)
size_t pchol_lds_bytes(int q);
// Persistent launches fed from work queues (problems taken largest first),
// one per residency class: the problems are grouped by how many of their
// workgroups fit a compute unit's LDS (1, 2, 4 or 8), and each group runs
// with the LDS of its own largest problem, so small problems are not held
// to the residency of the largest.  ds: the descriptors `host` as uploaded.
void launch_pchol_batch(arena& S, const pchol_desc* ds, const std::vector<pchol_desc>& host, hipStream_t st);
// Second pass of CholQR2 by one Newton-Schulz step where it suffices: for
// each problem (descriptor of the unpivoted pass, device array), with
// G2 = Q1^T Q1 = I + E2 in d.G: when p max|E2| <= 1e-8, Rinv = (3 I - G2) / 2
// (row-major p x p, so that Q1 Rinv = Q1 (I - E2 / 2) is orthonormal to
// O(E2^2) <= 1e-16 plus rounding), *p_out = p and *skip = 1; otherwise
// *skip = 0 and the pivoted-Cholesky launch factors it.
void launch_ns_pass2(const pchol_desc* ds, int count, int* skip, hipStream_t st);

void launch_gemm(const gemm_desc& g, int batch, hipStream_t st);
// Split-K form of a single GEMM for shapes with few output tiles:
// gemm_splitk_slices gives the slice count S (1: use launch_gemm);
// launch_gemm_splitk_parts writes the partial products of the slices it
// uses (returned; at most S) into part (S * M * N doubles), and
// launch_splitk_reduce sums them into C with alpha and beta.
int gemm_splitk_slices(int M, int N, int K);
// Up to four independent GEMMs in one launch (descriptors by value;
// blockIdx.z picks the problem, tiles beyond its M x N exit).
struct gemm_multi
{
  gemm_desc g[4];
  int count = 0;
  // Optional: the speculative CholQR2 acceptance decision in one more
  // workgroup of the launch (blockIdx.z == count).
  const double* dinfo = nullptr;
  const double* dpart = nullptr;
  int dnp = 0, dn = 0;
  double dthr = 0.0;
  double* dflag = nullptr;
};
void launch_gemm_multi(const gemm_multi& m, hipStream_t st);
// One block step of the blocked Cholesky with inverse of the n x n
// row-major G (upper triangle and diagonal blocks read): with the diagonal
// block [j0, j0 + jb) already factored into R and X (its R_jj^-1), the block
// row of R right of it, the trailing update of G's upper block rows, the
// factorization of the next diagonal block, and X's block column above the
// diagonal block, in one launch.  S: chol_step_scratch(n) doubles.
struct chol_step
{
  double* G = nullptr;
  double* R = nullptr;
  double* X = nullptr;
  double* S = nullptr;
  double* info = nullptr;
  int n = 0, j0 = 0, jb = 0;
  int nt = 0, ntri = 0;  // trailing block columns and upper-triangle tiles
};
// Block size, trailing block columns, trailing upper-triangle tiles and X
// row blocks of the step at j0.
void chol_step_shape(int n, int j0, int& jb, int& nt, int& ntri, int& nx);
size_t chol_step_scratch(int n);
void launch_chol_step(const chol_step& s, hipStream_t st);

// One factorization of a chained sweep (device arrays of these; every
// matrix n x n row-major):
//   G     SPD Gram, overwritten by the Cholesky's trailing updates;
//   R, X  its Cholesky factor (G = R^T R, upper) and R^-1;
//   S     chol_step_scratch(n) doubles; info 3 doubles (as launch_potrf_inv);
//   E     second-pass Gram on entry to launch_chain_minus_eye, E = G2 - I
//         after it, with kChainParts partial maxima of |E| in part;
//   V     E E; P, Pi the second-order series of (I + E)^(-1/2) and
//         (I + E)^(1/2) (launch_chain_polar);
//   Rf, Y the factor Pi R and its inverse X P, whose column norms go to
//         norms (2n doubles: Rf's, then Y's) for the full-rank certificate.
struct chain_prob
{
  double* G = nullptr;
  double* R = nullptr;
  double* X = nullptr;
  double* S = nullptr;
  double* info = nullptr;
  double* E = nullptr;
  double* V = nullptr;
  double* P = nullptr;
  double* Pi = nullptr;
  double* part = nullptr;
  double* Rf = nullptr;
  double* Y = nullptr;
  double* norms = nullptr;
  int n = 0;
};
constexpr int kChainParts = 64;
// Cross schedule of the staged block-Jacobi kernels (on by default): each
// launch rotates only the pairs across its two blocks, and a block's own
// pairs once per sweep.
extern bool g_jac_cross;
// The same for the grouped kernel of batches.
extern bool g_jac_cross_group;
// XCD-aware tile order of the chain products (on by default), and the
// strip height it picks for a gx x gy x gz tile grid whose A and B operands
// weigh wa and wb in all (0: block order).
extern bool g_chain_xcd;
int xcd_strip_height(int gx, int gy, int gz, double wa, double wb);
// One product of a Gram chain on 32 x 32 tiles.  One operand may be the
// sum of na (or nb) <= 8 buffers psa (psb) elements apart, summed in order
// as it is loaded.  With part null, C = alpha A B + beta C; otherwise the K range is
// cut into at most `slices` slices (chain_gemm_slices) whose partial
// products go to part (row-major M x N each, M N apart), and the number of
// slices used is returned.
int chain_gemm_slices(int M, int N, int K);
int launch_gemm_chain(const gemm_desc& g, int na, long psa, int nb, long psb, int slices, double* part,
                      hipStream_t st);
// For each of `count` items (a device array of {const double* part;
// double* out; int s; long n}): out = P_0 + ... + P_{s-1}, in order.
void launch_chain_sum(const void* items_dev, int count, long max_n, hipStream_t st);
// The blocked Cholesky with inverse (chol_inv_blocked's steps) of `count`
// problems at once: the first diagonal block of each, then block step j0 of
// every problem with n > j0 (max_n bounds the n's).
void launch_chain_potrf(const chain_prob* ps, int count, hipStream_t st);
void launch_chain_chol_step(const chain_prob* ps, int count, int max_n, int j0, hipStream_t st);
// E <- E - I with the partial maxima of |E| (non-finite as infinity).
void launch_chain_minus_eye(const chain_prob* ps, int count, int max_n, hipStream_t st);
// P = I - E/2 + 3/8 V, Pi = I + E/2 - 1/8 V.
void launch_chain_polar(const chain_prob* ps, int count, int max_n, hipStream_t st);
// Column norms of Rf and Y into norms.
void launch_chain_norms(const chain_prob* ps, int count, int max_n, hipStream_t st);
// out[0] = the sum and out[1] = the largest over the problems of n max|E|
// (from the partial maxima); both infinity when any problem's Cholesky
// broke down (info[0] != 0 or non-finite extreme diagonal entries).
void launch_chain_bound(const chain_prob* ps, int count, double* out, hipStream_t st);
// Epilogues folded into the split-K sum of a square row-major n x n product
// v = alpha P + beta C (CholQR2's second pass):
//   SPLITK_EPI_GRAM: C = E = v - I, out2 = Phi(E) (if given), part[b] = max |E| over
//                    block b's entries (non-finite counted as infinity);
//   SPLITK_EPI_PHI:  C = Phi(E - v) for the E given;
//   SPLITK_EPI_PHI_LAST: as PHI, and out2 = I + C, out3 = I - C;
//   SPLITK_EPI_DIAG: C = v + diag I;
//   SPLITK_EPI_POLAR: for v = E E, C = P = I - E/2 + 3/8 v and out2 = Pi =
//                    I + E/2 - 1/8 v, the second-order series of
//                    (I + E)^(-1/2) and (I + E)^(1/2).
// Phi keeps the strict upper part and half the diagonal, zero below.
enum
{
  SPLITK_EPI_GRAM = 1,
  SPLITK_EPI_PHI = 2,
  SPLITK_EPI_PHI_LAST = 3,
  SPLITK_EPI_DIAG = 4,
  SPLITK_EPI_POLAR = 5
};
struct splitk_epi
{
  int mode = 0;
  const double* E = nullptr;
  double* out2 = nullptr;
  double* out3 = nullptr;
  double* part = nullptr;
  double diag = 0.0;
};
// Sum of the partials with an epilogue; with SPLITK_EPI_GRAM the launch has
// minus_ident_parts(n) blocks (one partial maximum each).
void launch_splitk_reduce_epi(const gemm_desc& g, int s_eff, const double* part, const splitk_epi& ep,
                              hipStream_t st);
// The SPLITK_EPI_POLAR epilogue on its own: P and Pi from E and V = E E
// (n x n row-major).
void launch_polar_series(const double* E, const double* V, double* P, double* Pi, int n, hipStream_t st);
int launch_gemm_splitk_parts(const gemm_desc& g, int S, double* part, hipStream_t st);
void launch_splitk_reduce(const gemm_desc& g, int s_eff, const double* part, hipStream_t st);
// The grouped GEMM's tile-size autotuner (per process): on / off (off: the
// 64 x 64 tile always), and the keys seen, decided, and decided for the
// 32 x 32 tile.
void set_gemm_autotune(bool on);
void gemm_autotune_stats(long long* keys, long long* decided, long long* chose_small);
// Grouped GEMM over `count` device descriptors; max_m / max_n bound them.
void launch_gemm_grouped(const gemm_desc* descs_dev, int count, int max_m, int max_n, hipStream_t st);
// The same from host descriptors: uploads them through the arena and, when
// the problems' output shapes differ so much that a grid over the largest
// would be mostly idle workgroups, launches one workgroup per output tile
// that exists.
void launch_gemm_grouped(arena& S, const std::vector<gemm_desc>& v, hipStream_t st);
void launch_xgemm_grouped(const xgemm_desc* descs_dev, int count, int max_m, int max_n, xgemm_types t,
                          hipStream_t st);
// The same from host descriptors, over the output tiles that exist when a
// grid over the largest shape would be mostly idle (as launch_gemm_grouped).
void launch_xgemm_grouped(arena& S, const std::vector<xgemm_desc>& v, xgemm_types t, hipStream_t st);
void launch_qr_blocked(const qr_batch& b, hipStream_t st);
// Pieces of the two-level blocked QR of one matrix (b.one, no descriptors):
// the 16-column panel at j0; its update of the columns before col_end only;
// Q = I; V (len x nb, column-major) of the block of reflectors at column j0;
// T (nb x nb, row-major, nb <= 64) of a block from G = V^T V and tau.
void launch_qr_panel_one(const qr_batch& b, int j0, hipStream_t st);
void launch_qr_update_block(const qr_batch& b, int j0, int col_end, hipStream_t st);
void launch_qr_eye(const qr_batch& b, hipStream_t st);
void launch_qr_vblock(const double* A, int lda, int j0, int len, int nb, double* V, hipStream_t st);
void launch_qr_tblock(const double* G, const double* tau, int nb, double* T, hipStream_t st);
// Largest n of a CholQR2 factorization.
int chol_max_n();
// Diagonal block step of the blocked Cholesky: the jb x jb block (jb <=
// potrf_block()) at (j0, j0) of G factored as R_b^T R_b, R_b into R and
// R_b^-1 into X at the same place, rows j0 .. j0+jb-1 of R and X zeroed left
// of the block; info[0..2] = breakdown flag, min and max diagonal of R
// (started at j0 = 0, folded into afterwards).
int potrf_block();
void launch_potrf_inv(const double* G, int ldg, double* R, int ldr, double* X, int ldx, int jb, int j0, double* info,
                      hipStream_t st);
// done_dev: one device int of scratch; done_host: one pinned host int.
// Returns the number of block sweeps the host ran (q > 64), or -1 when the
// sweeps ran inside one kernel.
// With timing on, each round launch is recorded under TTA_K_JACOBI_ROUND and
// credited credit_flops / credit_bytes divided by the number of rounds run.
int launch_jacobi_svd(const jac_batch& b, int max_q, int* done_dev, int* done_host, tta_ctx* c, double credit_flops,
                      double credit_bytes);
void launch_qr(const qr_desc& q, int batch, hipStream_t st);
void launch_jacobi(const jacobi_desc& j, int batch, hipStream_t st);
void launch_extract_upper(const double* A, int lda, int kr, int nc, double* dst, long drs, long dcs, hipStream_t st,
                          bool full = false);
// dst (rows x cols, packed row-major) = src (row-major, leading dimension ld)
// with row i scaled by rowf[i] and column j by colf[j] (either may be null).
void launch_scale_copy(const double* src, int rows, int cols, long ld, const double* rowf, const double* colf,
                       double* dst, hipStream_t st);
void launch_gather_cols(const double* src, int lds, int rows, const int* perm, int ncols, double* dst, long drs,
                        long dcs, hipStream_t st,
                        const double* scale = nullptr, int mode = GATHER_COPY);
void launch_transpose(const double* in, int ldi, int rows, int cols, double* out, int ldo, hipStream_t st);
// Squared Frobenius norms of R^-1 and R per 16-column block (2 doubles per
// block, rinv_cert_blocks(q) blocks) for an upper-triangular row-major R,
// q <= kRinvCertMaxQ; with X non-null also R^-1 itself (row-major, ldx).
constexpr int kRinvCertMaxQ = 1024;
int rinv_cert_blocks(int q);
void launch_rinv_cert(const double* R, int q, double* out, hipStream_t st, double* X = nullptr, int ldx = 0);
// Decisions of a speculative sweep on the device (flag[0] cleared on the
// first failure): the full-rank certificate from the column norms of R and
// R^-1 (mode 0, a = 2n values) or the n (xf2, rf2) pairs of
// launch_rinv_cert (mode 1).  CholQR2 acceptance (from its info[0..2] and
// the np partial maxima of |E|, accepted when n max|E| <= thr) rides on
// launch_spec_cholqr_norms or on a gemm_multi launch (its dinfo .. dflag).
// Up to kCertBatch certificate decisions (inputs a, n, mode, delta as
// described above) in one launch.
constexpr int kCertBatch = 32;
struct cert_batch
{
  struct item
  {
    const double* a;
    int n, mode;
    double delta;
  };
  item it[kCertBatch];
  int count = 0;
};
void launch_spec_cert_batch(const cert_batch& b, double* flag, hipStream_t st);
// The CholQR2 acceptance decision and the column norms of the n x n
// row-major A and B (into out[0, n) and out[n, 2n)) in one launch.
void launch_spec_cholqr_norms(const double* info, const double* part, int np, int n, double thr, double* flag,
                              const double* A, const double* B, double* out, hipStream_t st);
// p[0] = v.
void launch_set_value(double* p, double v, hipStream_t st);
// Up to kCopyBatch device-to-device copies of doubles in one launch
// (descriptors by value).
constexpr int kCopyBatch = 32;
struct copy_batch
{
  struct item
  {
    const double* src;
    double* dst;
    long n;
  };
  item it[kCopyBatch];
  int count = 0;
};
void launch_copy_batch(const copy_batch& b, hipStream_t st);
// Up to kCoreBatch per-core launches of one TT folded into one (descriptors
// by value): the MPO core contraction of launch_matvec_core, and the column
// norms and balancing of launch_col_norms / launch_balance.
constexpr int kCoreBatch = 32;
struct matvec_batch
{
  struct item
  {
    const double* A;
    const double* x;
    double* y;
    int r, m, n, rp, s, sp;
  };
  item it[kCoreBatch];
  int count = 0;
};
void launch_matvec_core_batch(const matvec_batch& b, hipStream_t st);
// Core x (r, n, rp); off / offl: offsets of its right / left column norms in
// the norm array (negative: none; the norm kernel writes at off).
struct core_batch
{
  struct item
  {
    double* x;
    int r, n, rp;
    int off, offl;
  };
  item it[kCoreBatch];
  int count = 0;
};
void launch_core_norms_batch(const core_batch& b, double* nrm, hipStream_t st);
void launch_balance_batch(const core_batch& b, const double* nrm, const double* geo, hipStream_t st);
// dst (n x n, leading dimension ld) = identity.
void launch_eye(double* dst, int n, long ld, hipStream_t st);
// G <- G - I (row-major n x n); max |G - I| over the entries, as
// minus_ident_parts(n) <= kMinusIdentParts partial maxima in part.
constexpr int kMinusIdentParts = 256;
int minus_ident_parts(int n);
void launch_minus_ident(double* G, int n, double* part, hipStream_t st);
// X = strict upper part of E - W plus half its diagonal (W may be null).
void launch_phi(const double* E, const double* W, double* X, int n, hipStream_t st);
// Y = alpha I + beta X.
void launch_ident_axpy(const double* X, double* Y, int n, double alpha, double beta, hipStream_t st);
void launch_permute(const permute_desc& p, hipStream_t st);
void launch_matvec_core(const double* A, const double* x, double* y, int r, int m, int n, int rp, int s, int sp,
                        hipStream_t st);
void launch_matmat_core(const double* A, const double* B, double* C, int r, int m, int p, int rp, int s, int n,
                        int sp, hipStream_t st);
void launch_frob_norm(const double* x, long n, double* part, int nparts, double* out, double eps, double inv_sqrt_dm1,
                      double* delta_out, hipStream_t st);
void launch_col_norms(const double* M, int rows, int cols, double* out, hipStream_t st);
void launch_geo_mean(const double* v, int n, double* out, hipStream_t st);
void launch_balance(double* x, int r, int n, int rp, const double* nl, const double* nr, const double* geo,
                    hipStream_t st);
void launch_copy2d(const double* src, int lds, double* dst, int ldd, int rows, int cols, double f, double keep,
                   hipStream_t st);
void launch_scale(double* x, long n, double alpha, hipStream_t st);
void launch_wave_sum_check(const double* x, int n, int* bad, hipStream_t st);
void launch_transpose_batched(const tr_desc* ds, int count, int max_rows, int max_cols, hipStream_t st);
void launch_extract_upper_batched(const ex_desc* ds, int count, long max_elems, hipStream_t st);
void launch_gather_cols_batched(const ga_desc* ds, int count, long max_elems, hipStream_t st);
void launch_copy_batched(const cp_desc* ds, int count, long max_elems, hipStream_t st);
void launch_frob_batched(const sq_desc* ds, int count, hipStream_t st);
void launch_col_norms_batched(const cn_desc* ds, int count, int max_cols, hipStream_t st);
void launch_geo_mean_batched(const gm_desc* ds, int count, hipStream_t st);
void launch_balance_batched(const bal_desc* ds, int count, long max_elems, hipStream_t st);

// Flop models (Golub & Van Loan counts).  Householder QR of m x n with
// k = min(m, n) reflectors plus the thin Q: 4 m k^2 - 4/3 k^3 for m >= n,
// 2 n m^2 + 2/3 m^3 for m < n.  Thin SVD with U and V of p x q (p >= q):
// 14 p q^2 + 8 q^3 (22 q^3 for a square matrix).
inline double qr_flops(double m, double n)
{
  return m >= n ? 4.0 * m * n * n - 4.0 / 3.0 * n * n * n : 2.0 * n * m * m + 2.0 / 3.0 * m * m * m;
}
inline double svd_flops(double p, double q)
{
  if (p < q) std::swap(p, q);
  return 14.0 * p * q * q + 8.0 * q * q * q;
}

// Algorithmic work of a batch of launches (credited to timed launches):
// textbook flops at each problem's shape and each operand read / result
// written once.
struct credit
{
  double flops = 0.0, bytes = 0.0;
};
inline credit gemm_credit(const gemm_desc* g, size_t n)
{
  credit c;
  for (size_t i = 0; i < n; ++i)
  {
    c.flops += 2.0 * g[i].M * g[i].N * (double)g[i].K;
    c.bytes += 8.0 * ((double)g[i].M * g[i].K + (double)g[i].K * g[i].N + (double)g[i].M * g[i].N);
  }
  return c;
}
inline credit qr_credit(const qrp* q, size_t n)
{
  credit c;
  for (size_t i = 0; i < n; ++i)
  {
    const double m = q[i].m, nn = q[i].n, k = nn < m ? nn : m;
    c.flops += qr_flops(m, nn);
    c.bytes += 8.0 * (m * nn + m * k);
  }
  return c;
}
inline credit jacobi_credit(const jac_prob* j, size_t n)
{
  credit c;
  for (size_t i = 0; i < n; ++i)
  {
    const double q = j[i].q;
    c.flops += svd_flops(q, q);
    c.bytes += 32.0 * q * q;
  }
  return c;
}

}  // namespace tta
