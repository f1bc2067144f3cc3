/*
 * tt_amd.h -- C ABI of the MI355X (gfx950) TT-rounding library
 * (libtt_amd.so, built from tensor_train_amd/csrc).
 *
 * Every entry point returns an int status (TTA_OK == 0); on failure the
 * message is available from tta_last_error(ctx).  The header-only C++ layer
 * (include/tensor_train.hpp) maps a non-zero status to
 * fprintf(stderr) + abort(), the error behaviour of the reference
 * (round.hpp:124-134, matvec_round.hpp:47-52, shape_check.hpp:29-150 of
 * bdaszuta/tensor_train).
 *
 * Data layout (identical to the reference, storage_backend.hpp:237-253):
 *   TT core k          (r_k, n_k, r_{k+1})       row-major, (a*n + i)*rp + b
 *   TT-matrix core k   (r_k, m_k, n_k, r_{k+1})  row-major, ((a*m + i)*n + j)*rp + b
 * Host buffers passed to upload/download hold the cores concatenated in
 * core order.  Device tensors are owned by the library (tta_tt); outputs are
 * fresh tensors freed with tta_tt_free.
 *
 * Threading: a context owns one HIP stream; calls on one context are
 * serialised by the caller.  All calls are synchronous with respect to the
 * host unless documented otherwise.
 */
#ifndef TT_AMD_H
#define TT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tta_ctx tta_ctx;
typedef struct tta_tt tta_tt;
typedef struct tta_mlp tta_mlp;

enum tta_status
{
  TTA_OK = 0,
  TTA_ERR_ARG = 1,    /* invalid argument (null handle, negative size) */
  TTA_ERR_SHAPE = 2,  /* rank / mode mismatch, boundary ranks != 1 */
  TTA_ERR_METHOD = 3, /* method not valid for this entry point */
  TTA_ERR_HIP = 4,    /* HIP runtime failure */
  TTA_ERR_NOMEM = 5,  /* device allocation failed */
  TTA_ERR_UNSUPPORTED = 6
};

/* round_method, same order as include/core/round/round.hpp:33-44. */
enum tta_method
{
  TTA_METHOD_SVD = 0,
  TTA_METHOD_SVD_STREAMING = 1,
  TTA_METHOD_SVD_NAIVE = 2,
  TTA_METHOD_ALS = 3,
  TTA_METHOD_DMRG = 4
};

/* round_gauge, include/core/round/round.hpp:49-53. */
enum tta_gauge
{
  TTA_GAUGE_NONE = 0,
  TTA_GAUGE_RIGHT_CANONICAL = 1
};

/* ---- context ------------------------------------------------------------ */
int tta_ctx_create(int device, tta_ctx** out);
int tta_ctx_destroy(tta_ctx* ctx);
const char* tta_last_error(const tta_ctx* ctx);
int tta_synchronize(tta_ctx* ctx);
/* The context's hipStream_t, for callers that enqueue their own work. */
void* tta_stream(tta_ctx* ctx);

/* ---- device tensors ----------------------------------------------------- */
/* Allocate a zero-filled TT with d cores, modes[d], ranks[d+1]
 * (ranks[0] == ranks[d] == 1 required, as tt::tt validates, tt.hpp:35-70). */
int tta_tt_create(tta_ctx* ctx, int d, const int* modes, const int* ranks, tta_tt** out);
/* TT-matrix: row_modes[d] (m_k), col_modes[d] (n_k). */
int tta_ttm_create(tta_ctx* ctx, int d, const int* row_modes, const int* col_modes, const int* ranks,
                   tta_tt** out);
int tta_tt_free(tta_ctx* ctx, tta_tt* t);
/* d and whether t is a TT-matrix. */
int tta_tt_dims(const tta_tt* t, int* d, int* is_matrix);
/* Any output pointer may be null: modes[d] (m_k*n_k for matrices),
 * row_modes[d] / col_modes[d] (matrices only), ranks[d+1]. */
int tta_tt_shape(const tta_tt* t, int* modes, int* row_modes, int* col_modes, int* ranks);
long long tta_tt_numel(const tta_tt* t);
int tta_tt_upload(tta_ctx* ctx, tta_tt* t, const double* host);
int tta_tt_download(tta_ctx* ctx, const tta_tt* t, double* host);
/* Device pointer of core k (valid until the tensor is modified or freed). */
int tta_tt_core_ptr(const tta_tt* t, int k, double** dev);
int tta_tt_copy(tta_ctx* ctx, const tta_tt* t, tta_tt** out);

/* ---- the hot path ------------------------------------------------------- */
/* round(const tt&, round_options{eps, max_rank, method, gauge})
 * (include/core/round/round.hpp:230-267) and round(const tt_matrix&, ...)
 * (round.hpp:329-370, via round_matrix, tt_matrix_round.hpp:27-71).
 * method svd, als or dmrg (svd_streaming / svd_naive -> TTA_ERR_METHOD,
 * round.hpp:124-134). */
int tta_round(tta_ctx* ctx, const tta_tt* a, double eps, int max_rank, int method, int gauge, tta_tt** out);
/* round(a, round_options, round_result*) with every option
 * (include/core/round/round.hpp:46-110, 230-420): method svd / als / dmrg;
 * gauge applies to svd; max_iters, tol and warm (warm_start_tt, or
 * warm_start_tt_matrix when a is a TT-matrix) to als and dmrg.  als starts
 * cold from round_right_canonical(a, 0, max_rank) (tt_als.hpp:226-300,
 * tt_matrix_als_round.hpp:40-72); dmrg needs eps > 0 (else TTA_ERR_ARG,
 * round.hpp:155-176) and starts cold from round_right_canonical(a, eps,
 * max_rank) (tt_dmrg.hpp:235-310, tt_matrix_dmrg_round.hpp:30-55).  seed is
 * unused by round (as in the reference).  The round_result fields iters_run,
 * final_resid, converged go to the three outputs (each may be null); the svd
 * method reports {1, 0, 1}.  tta_round is this call with the reference's
 * defaults max_iters = 20, tol = 1e-12, seed = 0, no warm start. */
int tta_round_iterative(tta_ctx* ctx, const tta_tt* a, double eps, int max_rank, int method, int gauge,
                        int max_iters, double tol, unsigned long long seed, const tta_tt* warm, tta_tt** out,
                        int* iters_run, double* final_resid, int* converged);
/* right_orthogonalize(tt&) (include/core/gauge/tt_orthogonalize.hpp:26-36) and
 * right_orthogonalize(tt_matrix&) (tt_matrix_orthogonalize.hpp:34-85): in place;
 * ranks may shrink to r_eff = min(n*r_right, r_left). */
int tta_right_orthogonalize(tta_ctx* ctx, tta_tt* a);
/* detail::left_orthogonalize (tt_orthogonalize.hpp:47-57), in place. */
int tta_left_orthogonalize(tta_ctx* ctx, tta_tt* a);
/* matvec(const tt_matrix&, const tt&) (include/core/apply/tt_matrix_apply.hpp:237-248). */
int tta_matvec(tta_ctx* ctx, const tta_tt* A, const tta_tt* x, tta_tt** y);
/* matmat(const tt_matrix&, const tt_matrix&) (tt_matrix_apply.hpp:260-271). */
int tta_matmat(tta_ctx* ctx, const tta_tt* A, const tta_tt* B, tta_tt** C);
/* matvec_round (include/core/round/matvec_round.hpp:42-91): method svd =
 * round(matvec(A, x)) (tt_matrix_apply_round.hpp:32-38); svd_streaming =
 * fused apply + round (streaming_apply.hpp:39-110); svd_naive -> TTA_ERR_METHOD. */
int tta_matvec_round(tta_ctx* ctx, const tta_tt* A, const tta_tt* x, double eps, int max_rank, int method,
                     tta_tt** y);
/* matvec_round(A, x, round_options, round_result*) with every option
 * (matvec_round.hpp:42-91): method als = matvec_als_round
 * (tt_matvec_als_round.hpp:30-113): the product A x is formed and rounded by
 * ALS from warm, from the SVD round at eps when max_rank <= 0, else from
 * random(modes, max_rank, seed); method dmrg = matvec_dmrg_round
 * (tt_matvec_dmrg_round.hpp:29-105): from warm, else random(modes, max_rank,
 * seed) when max_rank > 0, else round_right_canonical(A x, eps, 0).
 * Outputs as tta_round_iterative. */
int tta_matvec_round_iterative(tta_ctx* ctx, const tta_tt* A, const tta_tt* x, double eps, int max_rank, int method,
                               int max_iters, double tol, unsigned long long seed, const tta_tt* warm, tta_tt** y,
                               int* iters_run, double* final_resid, int* converged);
/* matmat_round (include/core/round/matmat_round.hpp:41-94): method svd =
 * streaming (streaming_apply.hpp:115-202); svd_naive = round_matrix(matmat);
 * svd_streaming -> TTA_ERR_METHOD. */
int tta_matmat_round(tta_ctx* ctx, const tta_tt* A, const tta_tt* B, double eps, int max_rank, int method,
                     tta_tt** C);
/* matmat_round(A, B, round_options, round_result*) with every option
 * (matmat_round.hpp:41-94): method als = matmat_als_round
 * (tt_matmat_als_round.hpp:39-141) and method dmrg = matmat_dmrg_round
 * (tt_matmat_dmrg_round.hpp:38-130) on the packed product, starts as for
 * tta_matvec_round_iterative (warm is a TT-matrix). */
int tta_matmat_round_iterative(tta_ctx* ctx, const tta_tt* A, const tta_tt* B, double eps, int max_rank, int method,
                               int max_iters, double tol, unsigned long long seed, const tta_tt* warm, tta_tt** C,
                               int* iters_run, double* final_resid, int* converged);
/* inner(a, b) (include/core/algebra/tt_inner.hpp:33-48); frob_inner for
 * TT-matrices (tt_matrix_inner.hpp:31-36). */
int tta_inner(tta_ctx* ctx, const tta_tt* a, const tta_tt* b, double* out);
/* inner(a, b) as tta_inner, allocating nothing: no reference counterpart;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The result goes to out, a device double the
 * caller provides; the call only enqueues work on the context's stream (no
 * host read, no synchronization), so the value is ready once the stream
 * reaches it.  The environments live in a fixed library-owned device array of
 * 4 Mi doubles (32 MiB per device, held until the process ends): operands
 * needing more (2 max_k ra_k rb_k + max_k rb_k n_k ra_{k+1} doubles, each
 * part rounded up to 32) are rejected with TTA_ERR_UNSUPPORTED.  Calls that
 * overlap on one device share that array and race. */
int tta_inner_noalloc(tta_ctx* ctx, const tta_tt* a, const tta_tt* b, double* out);
/* norm(a) = sqrt(max(inner(a, a), 0)) (tt_inner.hpp:55-59; tt_matrix_inner.hpp:44-48). */
int tta_norm(tta_ctx* ctx, const tta_tt* a, double* out);
/* frob_norm_apply(A, x) / frob_norm_apply(A, B)
 * (include/core/algebra/tt_matrix_frob_norm_apply.hpp:36-94): ||A x||_F or
 * ||A B||_F by one environment sweep over the product cores, without the
 * product tensor; X is a TT (matvec) or a TT-matrix (matmat). */
int tta_frob_norm_apply(tta_ctx* ctx, const tta_tt* A, const tta_tt* X, double* out);
/* axpby(alpha, a, beta, b) (include/core/algebra/tt_ops.hpp:72-110). */
int tta_axpby(tta_ctx* ctx, double alpha, const tta_tt* a, double beta, const tta_tt* b, tta_tt** out);
/* scale(a, alpha) (tt_ops.hpp:35-50). */
int tta_scale(tta_ctx* ctx, const tta_tt* a, double alpha, tta_tt** out);

/* ---- batched rounding --------------------------------------------------- */
enum tta_precision
{
  TTA_PREC_FP64 = 0,   /* plain: the reference algorithm in FP64 */
  TTA_PREC_MIXED = 1,  /* SVD steps by rank-revealing CholQR2 where certified */
  TTA_PREC_LEARNED = 2 /* learned-rank randomized rounding with plain fallback
                          (tta_round_learned_batched; not a tta_round_batched mode) */
};
/* round(a_i, {eps, max_rank}) for count TTs of any d, modes and ranks
 * (detail::round, include/core/round/tt_round.hpp:36-127, per TT), the
 * TTs advanced in lockstep so each launch covers the batch.  out[i] are
 * fresh tensors.  No reference counterpart for the batching itself.
 * precision TTA_PREC_FP64: the plain algorithm, results as tta_round.
 * precision TTA_PREC_MIXED: each tall SVD-sweep step (unfolding M of at most
 * 192 columns) is first tried as a rank-revealing CholQR2: FP64 Gram M^T M
 * and its Cholesky with diagonal pivoting stopped where the residual of the
 * p pivot columns fits (0.4 delta)^2, CholQR2 of those columns (FP64), then
 * the FP64 residual correction R~ = Q^T M and E = M - Q R~.  The step keeps
 * (Q, R~ core_{k+1}) at rank p when ||E||_F <= delta / 2; otherwise it runs as
 * in the plain mode.  Each accepted step discards at most delta / 2, so the
 * round's error bound holds and ranks are never below the plain ranks; on
 * exactly rank-deficient unfoldings (sums) the ranks are the plain ranks. */
int tta_round_batched(tta_ctx* ctx, int count, tta_tt* const* in, double eps, int max_rank, int precision,
                      tta_tt** out);

/* Deterministic rounding that replaces a randomized result whose discarded
 * norm exceeds eps. */
enum tta_fallback
{
  TTA_FALLBACK_NONE = 0,  /* keep the randomized result */
  TTA_FALLBACK_FP64 = 1,  /* tta_round_batched with TTA_PREC_FP64 */
  TTA_FALLBACK_MIXED = 2  /* tta_round_batched with TTA_PREC_MIXED */
};
/* Randomized rounding with given target ranks (Randomize-then-Orthogonalize:
 * Gaussian sketch TT, right-to-left sketch contractions, one left-to-right
 * Householder QR sweep).  ranks: for TT i, d_i + 1 entries (ends ignored),
 * concatenated.  The discarded norm ||y - yhat|| = sqrt(||y||^2 - ||yhat||^2)
 * is reported in rel_err[i] (relative); every TT whose discarded norm exceeds
 * eps is rounded by the deterministic path that `fallback` names instead
 * (fell_back[i] = 1).  No reference counterpart. */
int tta_round_randomized_batched(tta_ctx* ctx, int count, tta_tt* const* in, const int* ranks, double eps,
                                 unsigned long long seed, int fallback, tta_tt** out, double* rel_err,
                                 int* fell_back);

/* ---- learned ranks (no reference counterpart) ----------------------------- */
/* Rank-prediction MLP n_in -> hidden (x layers, ReLU) -> 1, FP32 parameters
 * (flat layout: per layer W (in x out row-major) then b (out)), He init
 * from seed.  Trained on the device by the library's forward / backward /
 * Adam kernels on mean-squared error. */
int tta_mlp_create(tta_ctx* ctx, int n_in, int hidden, int layers, unsigned long long seed, tta_mlp** out);
int tta_mlp_free(tta_ctx* ctx, tta_mlp* m);
long long tta_mlp_num_params(const tta_mlp* m);
int tta_mlp_get_params(tta_ctx* ctx, const tta_mlp* m, float* host);
int tta_mlp_set_params(tta_ctx* ctx, tta_mlp* m, const float* host);
/* MSE loss and its gradient (flat, parameter layout) on n rows X (n x n_in), y (n). */
int tta_mlp_grad(tta_ctx* ctx, tta_mlp* m, const float* X, const float* y, int n, float* grad, float* loss);
/* `steps` full-batch Adam steps (betas 0.9, 0.999), the step size decayed
 * from lr to lr / 100 on a half cosine over the call; the parameters of the
 * lowest loss seen are kept and *loss is their loss. */
int tta_mlp_train(tta_ctx* ctx, tta_mlp* m, const float* X, const float* y, int n, int steps, float lr,
                  float* loss);
int tta_mlp_predict(tta_ctx* ctx, tta_mlp* m, const float* X, int n, float* y);
/* Inference (tta_mlp_predict and the learned rounding's rank prediction)
 * with weights, biases and activations rounded to BF16 (nearest, ties to
 * even) and products accumulated in FP32 (on = 1), or in FP32 throughout
 * (on = 0, the default).  Training always runs in FP32. */
int tta_mlp_set_bf16_inference(tta_ctx* ctx, tta_mlp* m, int on);
/* Number of per-bond features and the features of every bond k = 1..d_i-1 of
 * every TT (rows concatenated; the feature list is documented with
 * tta_learn.hip). */
int tta_num_bond_features(void);
int tta_bond_features(tta_ctx* ctx, int count, tta_tt* const* in, double eps, float* feats);
/* Learned-rank rounding: target rank of each bond =
 * ceil(prediction * r_k * (1 + slack)) + oversample (clipped to [1, r_k]),
 * randomized rounding with those ranks, deterministic fallback (FP64 or
 * mixed; TTA_FALLBACK_NONE is rejected) when the discarded norm exceeds eps.  The relative slack covers the sketch's error factor over the
 * optimal truncation, which grows as the spectrum decays more slowly.
 * ranks_used (optional) gets the d_i + 1 target ranks per TT. */
int tta_round_learned_batched(tta_ctx* ctx, tta_mlp* m, int count, tta_tt* const* in, double eps, int oversample,
                              double slack, unsigned long long seed, int fallback, tta_tt** out, double* rel_err,
                              int* fell_back, int* ranks_used);

/* ---- dense kernels of the path (host buffers, row-major) --------------- */
/* qr_thin (include/core/types/tt_eigen_bridge.hpp:42-60): Householder QR,
 * Q (rows x k) orthonormal columns, R (k x cols) upper trapezoidal,
 * k = min(rows, cols), dlarfg sign convention. */
int tta_qr_thin(tta_ctx* ctx, const double* M, int rows, int cols, double* Q, double* R);
/* svd_thin (tt_eigen_bridge.hpp:67-104): U (rows x k), s (k, descending),
 * Vt (k x cols). */
int tta_svd_thin(tta_ctx* ctx, const double* M, int rows, int cols, double* U, double* s, double* Vt);

/* soft_threshold(a, tau) (include/core/round/tt_soft_threshold.hpp:42-176):
 * the proximal step of the nuclear norm at every bond, singular values shrunk
 * to max(s - tau, 0); a copy when d <= 1 or tau <= 0. */
int tta_soft_threshold(tta_ctx* ctx, const tta_tt* a, double tau, tta_tt** out);

/* ---- dense to TT (the step before the path) ------------------------------- */
/* from_dense(data, shape, {eps, max_rank, method, qtt_base})
 * (include/core/factory/from_dense.hpp:215-225): method 0 = TT-SVD
 * (tt_svd.hpp:46-160, Oseledets' algorithm, ||T - TT|| <= eps ||T||),
 * method 1 = QTT (from_dense.hpp:109-199: modes base^b split into b modes of
 * size base; base a power of 2).  dense: prod(shape) doubles, row-major (at
 * most INT_MAX).  A zero tensor gives rank-1 zero cores; non-finite input is
 * TTA_ERR_ARG. */
int tta_from_dense(tta_ctx* ctx, const double* dense, int d, const int* shape, double eps, int max_rank, int method,
                   int qtt_base, tta_tt** out);
/* matrix_from_dense (include/core/factory/tt_matrix_factory.hpp:205-290):
 * dense (prod m) x (prod n) row-major -> TT-matrix. */
int tta_matrix_from_dense(tta_ctx* ctx, const double* dense, int d, const int* row_shape, const int* col_shape,
                          double eps, int max_rank, tta_tt** out);

/* ---- seeded generators (inputs of the benchmark configurations) --------- */
/* random(shape, max_rank, seed) (include/core/factory/tt_factory.hpp:141-207),
 * on the host with std::mt19937_64 + std::normal_distribution<double>, drawn
 * core by core in buffer order.  Host-only variant: ranks_out[d+1],
 * numel_out and data_out may each be null (size query).  Needs no GPU. */
int tta_random_host(int d, const int* modes, int max_rank, unsigned long long seed, int* ranks_out,
                    double* data_out, long long* numel_out);
int tta_random(tta_ctx* ctx, int d, const int* modes, int max_rank, unsigned long long seed, tta_tt** out);
/* random(row_shape, col_shape, max_rank, seed) (tt_matrix_factory.hpp:133-191). */
int tta_random_matrix(tta_ctx* ctx, int d, const int* row_modes, const int* col_modes, int max_rank,
                      unsigned long long seed, tta_tt** out);
/* zeros (value 0, tt_factory.hpp:31-52) / ones (value 1, :59-78): rank-1 TT
 * with every core entry equal to value. */
int tta_constant(tta_ctx* ctx, int d, const int* modes, double value, tta_tt** out);
/* Batched benchmark inputs built on the device.  Entries are N(0,1) from
 * the counter-based Philox4x32-10 generator keyed by the given seeds
 * (reproducible, but not the std::mt19937_64 stream of tta_random).
 * Sums: c_i = add(u_i, add(scale(u_i, 0.5), v_i)) with
 * u_i ~ random([n_i]^d_i, ru_i), v_i ~ random([n_i]^d_i, rv_i) (rank rule of
 * tt_factory.hpp:155-190; block concatenation of block_concat.hpp:40-156). */
int tta_build_sum_batch(tta_ctx* ctx, int count, const int* d, const int* n, const int* ru, const int* rv,
                        const unsigned long long* seed_u, const unsigned long long* seed_v, tta_tt** out);
/* Decaying-spectrum TTs: modes [n]^d, ranks min(n^k, n^(d-k), rmax); core
 * k < d-1 = Q_k diag(10^(-j s_i)) with Q_k the thin Q of a Gaussian
 * (r_k n x r_{k+1}) matrix; last core a unit-norm Gaussian. */
int tta_build_decay_batch(tta_ctx* ctx, int count, int d, int n, int rmax, const double* s,
                          const unsigned long long* seeds, tta_tt** out);
/* Multiply core k of t by alpha in place. */
int tta_scale_core(tta_ctx* ctx, tta_tt* t, int k, double alpha);

/* ---- kernel timing (HIP events on the context stream) -------------------- */
enum tta_kernel_class
{
  TTA_K_QR = 0,     /* Householder QR + thin Q (CholQR2 parts: see below) */
  TTA_K_JACOBI = 1, /* one-sided Jacobi SVD + sort + truncation rank */
  TTA_K_GEMM = 2,   /* FP64 contractions in one tiled launch */
  TTA_K_APPLY = 3,  /* MPO core contraction (matvec / matmat cores) */
  TTA_K_OTHER = 4,  /* transposes, gathers, norms, balancing, split-K sums */
  /* Each launch of the block-Jacobi round kernel (q > 64), credited an equal
   * share of its SVD call's textbook flops and bytes.  These launches are
   * also inside TTA_K_JACOBI, which times whole SVD calls. */
  TTA_K_JACOBI_ROUND = 5,
  /* The diagonal-block step of CholQR2's blocked Cholesky (factor and
   * inverse of a b x b block, b <= 64: credited 2 b^3 / 3 flops, 24 b^2
   * bytes) and the triangular inverse of the full-rank certificate where
   * the QR was Householder (n^3/3 flops, 16 n^2 bytes).  The blocked
   * Cholesky's later block steps are TTA_K_CHOL_STEP, CholQR2's elementwise
   * steps TTA_K_OTHER. */
  TTA_K_CHOL = 6,
  TTA_K_RINV = 7,
  /* The partial products of a GEMM split along K (their sum: TTA_K_OTHER). */
  TTA_K_GEMM_SPLITK = 8,
  /* Launches of up to four independent GEMMs together. */
  TTA_K_GEMM_MULTI = 9,
  /* One block step of CholQR2's blocked Cholesky with inverse in one launch:
   * block row of R, trailing update, R^-1's block column and the next
   * diagonal block (credited the flops and bytes of those products and of
   * the block factorization). */
  TTA_K_CHOL_STEP = 10,
  /* The products of the Gram chains of chained sweep blocks (32 x 32 tiles,
   * an operand read as the sum of split-K partials; credited 2 M N K flops
   * and each operand read once, the partials written once). */
  TTA_K_GEMM_CHAIN = 11,
  /* The batched products of chained sweep blocks (grouped launches over the
   * block's steps). */
  TTA_K_GEMM_GROUPED = 12,
  TTA_K_COUNT = 13
};
/* Enable (1) / disable (0) per-launch event timing; enabling resets totals. */
int tta_prof_enable(tta_ctx* ctx, int on);
/* Launch count, summed device milliseconds, and the summed algorithmic
 * flops and minimum HBM bytes credited to those launches (textbook counts at
 * each launch's shape, not hardware counters) for one kernel class. */
int tta_prof_read(tta_ctx* ctx, int kernel_class, long long* launches, double* total_ms, double* flops,
                  double* bytes);
/* Block-Jacobi SVD calls (matrices above 64 columns) since the last
 * tta_prof_enable and the sweeps they ran. */
int tta_prof_jacobi_sweeps(tta_ctx* ctx, long long* calls, long long* sweeps);
/* Block-Jacobi problems still rotating after the sweep limit (30 block
 * sweeps) since the last tta_prof_enable: their singular vectors are
 * orthogonal to working precision, but pairs of columns may remain above
 * the rotation tolerance.  No reference counterpart (Eigen's BDCSVD has no
 * sweep limit to report). */
int tta_prof_jacobi_unconverged(tta_ctx* ctx, long long* problems);
/* SVD-sweep steps of TTA_PREC_MIXED batched rounding tried by the
 * rank-revealing CholQR2 certificate and those accepted, since the last
 * tta_prof_enable. */
int tta_prof_rank_revealing(tta_ctx* ctx, long long* tried, long long* held);

/* ---- full-rank certificate of the SVD steps (no reference counterpart) ---- */
/* Before the Jacobi SVD of an SVD step of round / round_right_canonical, the
 * norms of the step's triangular factor R and of R^-1 are formed on the
 * device (R^-1 from CholQR2's two passes when that QR ran, else by a
 * triangular inverse); when 1/||R^-1||_F (a lower bound on sigma_min)
 * exceeds delta with margin and
 * ||R|| ||R^-1|| <= 1e8, the rank rule provably keeps every singular value,
 * and the step keeps the tensor as Q R without the SVD (same ranks, same
 * tensor, another orthogonal gauge of the cores).  On by default;
 * on = 0 always runs the SVD. */
int tta_set_rank_certificate(tta_ctx* ctx, int on);
/* Steps where the certificate was tried and where it held, since the last
 * tta_prof_enable. */
int tta_prof_rank_certificates(tta_ctx* ctx, long long* tried, long long* held);

/* ---- CholQR2 for the sweeps' QR factorizations (no reference counterpart) */
/* Factorizations with m >= n and 32 <= n <= 512 in the QR and SVD sweeps of
 * single TTs run as CholQR2 (Gram on the matrix cores, blocked Cholesky with
 * inverse, then a second pass from E = Q1^T Q1 - I) when it is accepted: no
 * Cholesky breakdown and n max|E| <= 1e-4, i.e. condition numbers up to
 * about 1e6; otherwise, and with on = 0, blocked Householder.  Both give a
 * backward-stable Q R.  The second pass is the polar form Q = Q1 (I+E)^(-1/2)
 * (to second order; R then not triangular) for n max|E| <= 1e-5, the
 * triangular form (Cholesky factor of I + E) above that. */
int tta_set_cholqr(tta_ctx* ctx, int on);
/* Factorizations where CholQR2 was tried and accepted, since the last
 * tta_prof_enable. */
int tta_prof_cholqr(tta_ctx* ctx, long long* tried, long long* used);
/* Of those, the ones accepted in the triangular form. */
int tta_prof_cholqr_triangular(tta_ctx* ctx, long long* triangular);

/* ---- speculative sweeps (no reference counterpart) ------------------------- */
/* The SVD sweeps of round / round_right_canonical (and of the naive
 * matvec_round / matmat_round built on them) decide CholQR2 acceptance and
 * the rank certificates on the device, into one flag read once per sweep
 * instead of once per factorization.  A sweep with any failed decision is
 * undone from a copy of the TT taken before it and run again with host
 * decisions, so results equal those of on = 0; after such a fallback the
 * next SVD sweep runs checked, and speculation resumes once a checked
 * sweep needed no fallback.  (The QR sweeps decide on the host: a declined
 * CholQR2 falls back to Householder for its own step.)  On by default. */
int tta_set_speculation(tta_ctx* ctx, int on);
/* Speculative sweeps run and those undone, since the last tta_prof_enable. */
int tta_prof_speculation(tta_ctx* ctx, long long* sweeps, long long* failed);

/* ---- chained sweep blocks (no reference counterpart) ----------------------- */
/* In round's sweeps (for the SVD sweeps, speculative ones), a run of two or
 * more consecutive QR steps of CholQR2's shape (and, in the SVD sweeps,
 * steps whose full-rank certificate max_rank cannot overrule) is factored
 * as one block: the steps' Gram
 * matrices follow one another by two products each (G_k = sum_i C_i G_{k+1}
 * C_i^T), and the Choleskys, the orthonormalizing products and a second
 * pass run for all steps of the block at once.  Exact arithmetic gives the
 * sequential sweep's result.  A block is accepted on the host (every
 * Cholesky held, the second pass within the polar form's range); a
 * rejected block leaves the TT untouched and its steps run one by one.
 * The certificates are decided with the sweep's, as above.  On by default. */
int tta_set_chain(tta_ctx* ctx, int on);
/* In the QR sweep of round (whose cores only feed the SVD sweep) the
 * second pass is left out when the first pass's cores are provably
 * orthonormal to 1e-10 already (a bound on every step's Gram error from the
 * cores' own Grams, read with the block's acceptance).
 * Chained blocks accepted, the steps they covered, the accepted blocks that
 * needed only the first pass, and the blocks rejected, since the last
 * tta_prof_enable. */
int tta_prof_chain(tta_ctx* ctx, long long* blocks, long long* steps, long long* one_pass, long long* rejected);

/* ---- HIP graphs (no reference counterpart) ------------------------------ */
/* Block-Jacobi SVD sweeps (more than 64 columns, single problems and
 * batches) captured into a HIP graph at the first sweep of each call and
 * replayed for the later sweeps, instead of enqueuing every round launch.
 * Results are identical either way (the same launches run in the same order).
 * Not used while per-launch event timing (tta_prof_enable) is on.  Off by
 * default. */
int tta_set_graphs(tta_ctx* ctx, int on);

/* ---- Gram-driven block Jacobi (no reference counterpart) ------------------- */
/* Block-Jacobi SVDs of fewer than 32 problems of 65..512 columns (and of
 * one problem of 41..512 columns) decide
 * each round's rotations on the block pair's Gram matrix: the pair's 16
 * columns are staged once, their 16 x 16 Gram formed on the matrix cores,
 * the inner sweep's rotations decided and computed on the Gram alone (the
 * one-sided rule, the same tolerance, two-sided updates of the Gram), and
 * the rows of X and J then rotated pair by pair by the recorded rotations.
 * The Gram is formed afresh from the columns at every launch, so
 * convergence is judged on the columns' own dot products.  The singular
 * values and vectors agree with on = 0 to rounding (not bit-identical).  On
 * by default; on = 0 runs the staged rounds, which compute every rotation
 * from the columns' dot products. */
int tta_set_jacobi_gram(tta_ctx* ctx, int on);

/* ---- autotuner (no reference counterpart) --------------------------------- */
/* The grouped FP64 GEMMs of the batched paths choose their tile size (64 x 64
 * or 32 x 32 outputs per workgroup) per shape class (device, log2 of the
 * largest M, of the largest N, of the problem count) by timing one launch
 * of each on the first two launches of the class and keeping the faster per
 * unit of work.  Both tiles give identical results.  The table is per
 * process; on by default (off: the 64 x 64 tile always). */
int tta_set_autotune(tta_ctx* ctx, int on);
/* Shape classes seen, decided, and decided for the 32 x 32 tile. */
int tta_prof_autotune(tta_ctx* ctx, long long* keys, long long* decided, long long* chose_small);

/* ---- device memory pool --------------------------------------------------- */
/* Statistics of the context's caching allocator: bytes held by live blocks
 * (tensor cores and workspaces), bytes cached for reuse, device allocations
 * made, and cache releases forced by a failed allocation.  No reference
 * counterpart (the reference allocates host memory per core). */
int tta_mem_stats(tta_ctx* ctx, unsigned long long* live_bytes, unsigned long long* cached_bytes,
                  unsigned long long* device_allocs, unsigned long long* trims);
/* Returns every cached (free) block of the context to the device. */
int tta_mem_trim(tta_ctx* ctx);

/* ---- self-checks --------------------------------------------------------- */
/* Compares the device's DPP wave reductions (one value per lane, and 16
 * values per lane at once) with ds_bpermute reductions on n host values
 * (rounded up to whole 64-lane waves); *mismatches = 0 when they agree to
 * rounding.  No reference counterpart: a hardware sanity check. */
int tta_selfcheck_wave_sum(tta_ctx* ctx, const double* host, int n, int* mismatches);

#ifdef __cplusplus
}
#endif

#endif /* TT_AMD_H */
