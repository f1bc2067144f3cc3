// One-sided (Hestenes) Jacobi SVD of batches of square q x q matrices,
// with the rotations accumulated into J:  X_in * J = X_out, X_out having
// mutually orthogonal columns and J orthogonal.
//
// q <= 64: jacobi_small_kernel keeps X and J in LDS and runs whole sweeps in
//   one workgroup (round-robin ordering, one wave per disjoint pair).
// q  > 64: block Jacobi.  The columns are cut into blocks of W; each launch
//   of bj_round_kernel is one round of a round-robin over blocks, one
//   workgroup per block pair: the pair's 2W columns of X and J are staged
//   in LDS, rotated by one inner sweep over pairs of them (by default the
//   cross schedule: the pairs across the two blocks, and each block's own
//   pairs in a sweep's first round), then written back.  A sweep is nb - 1
//   rounds and rotates every column pair once; the host stops after a sweep
//   with no rotation.
// Rotation rule: rotate while |x_i . x_j| > tol ||x_i|| ||x_j|| with
// tol = max(1e-15, q u), the worst-case rounding bound of a length-q dot
// product; below it a "rotation" only chases rounding noise and a sweep
// would never come out clean.  J stays orthogonal to working precision
// whatever tol is, so the singular vectors built from it do too.
// Negligible columns: a column with |x|^2 <= negl2 takes part in no
// rotation, where negl2 = min(delta^2 / (4 q), (q u)^2 ||X||_F^2) with delta
// the truncation threshold.  Leaving such columns alone perturbs X by at
// most the backward error of the QR that produced it (q u ||X||), and all of
// them together hold less than delta^2 / 4, so the truncation rule drops
// every one of them; rotating them only stirs rounding noise between
// columns that are discarded, which costs sweeps.  delta = 0 (no truncation,
// full SVD) disables the rule.
#include <algorithm>
#include <type_traits>

#include "../../include/tt_amd.h"
#include "tta_device.hpp"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

namespace tta
{
using namespace dev;

bool g_jac_cross = true;
bool g_jac_cross_group = true;

namespace
{

// By value: a reference that may point at the kernel argument made the
// compiler copy the argument to scratch memory and read every field through
// generic (flat) loads.
__device__ __forceinline__ jac_prob jpick(const jac_batch& b, int idx)
{
  return b.descs ? b.descs[idx] : b.one;
}

__device__ __forceinline__ double jac_tol(int q) { return fmax(1.0e-15, (double)q * 2.220446049250313e-16); }

// Squared-norm bound of negligible columns from ||X||_F^2 (see above).
__device__ __forceinline__ double jac_negl2(const jac_prob& P, double xnorm2)
{
  const double delta = P.delta_ptr ? *P.delta_ptr : P.delta;
  const double qu = (double)P.q * 2.220446049250313e-16;
  return fmin(delta * delta / (4.0 * P.q), qu * qu * xnorm2);
}

// Sums of three values over the wave, DPP stages interleaved so the three
// chains overlap.
__device__ __forceinline__ void wave_sum3(double& a, double& b, double& g)
{
  a += dpp_mov<0xB1>(a);
  b += dpp_mov<0xB1>(b);
  g += dpp_mov<0xB1>(g);
  a += dpp_mov<0x4E>(a);
  b += dpp_mov<0x4E>(b);
  g += dpp_mov<0x4E>(g);
  a += dpp_mov<0x124>(a);
  b += dpp_mov<0x124>(b);
  g += dpp_mov<0x124>(g);
  a += dpp_mov<0x128>(a);
  b += dpp_mov<0x128>(b);
  g += dpp_mov<0x128>(g);
  a = (read_lane(a, 0) + read_lane(a, 16)) + (read_lane(a, 32) + read_lane(a, 48));
  b = (read_lane(b, 0) + read_lane(b, 16)) + (read_lane(b, 32) + read_lane(b, 48));
  g = (read_lane(g, 0) + read_lane(g, 16)) + (read_lane(g, 32) + read_lane(g, 48));
}

// Orders one wave's LDS accesses: the wave's LDS writes before it have
// completed, and are seen by every lane's reads after it (the memory clobber
// keeps the compiler from moving accesses across).
__device__ __forceinline__ void wave_sync_lds()
{
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

// Rotate columns (xi, xj) of length p and (ji, jj) of length q in LDS by the
// Jacobi rotation that orthogonalises xi and xj.  One wave; returns 1 when a
// rotation was applied.  PL = elements per lane held in registers (p and q
// must be <= 64 * PL).
template <int PL>
__device__ __forceinline__ int rotate_pair(double* xi, double* xj, double* ji, double* jj, int p, int q, double tol,
                                           double negl2, int lane)
{
  double u[PL], v[PL];
  double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll
  for (int e = 0; e < PL; ++e)
  {
    const int t = lane + 64 * e;
    u[e] = t < p ? xi[t] : 0.0;
    v[e] = t < p ? xj[t] : 0.0;
    a = fma(u[e], u[e], a);
    b = fma(v[e], v[e], b);
    g = fma(u[e], v[e], g);
  }
  wave_sum3(a, b, g);
  double cs, sn;
  if (a <= negl2 || b <= negl2 || !jacobi_rotation(a, b, g, tol, cs, sn)) return 0;
#pragma unroll
  for (int e = 0; e < PL; ++e)
  {
    const int t = lane + 64 * e;
    if (t < p)
    {
      xi[t] = cs * u[e] - sn * v[e];
      xj[t] = sn * u[e] + cs * v[e];
    }
  }
#pragma unroll
  for (int e = 0; e < PL; ++e)
  {
    const int t = lane + 64 * e;
    if (t < q)
    {
      const double y1 = ji[t], y2 = jj[t];
      ji[t] = cs * y1 - sn * y2;
      jj[t] = sn * y1 + cs * y2;
    }
  }
  return 1;
}

// Schedules of a block pair's inner sweep (local columns 0..W-1 of block i,
// W..2W-1 of block j; W disjoint pairs per inner round, 2W - 1 rounds or W):
//   kFullSweep: all pairs of the 2W columns by the round-robin over 2W;
//   kCrossIntra / kCross: W - 1 intra rounds pairing each block's own
//     columns by the round-robin over W (pairs 0..W/2-1 in block i,
//     W/2..W-1 in block j; only in a sweep's first round, where every block
//     meets another once), then W cross rounds, round h pairing column a of
//     block i with column (a + h) mod W of block j.
// The cross schedules rotate every column pair of the matrix once per
// sweep, as cyclic Jacobi does; the full inner sweep repeats each block's own
// pairs in every round.  Sweeps to convergence come out the same or a few
// more, for about half the inner rounds per launch.
constexpr int kFullSweep = 0, kCrossIntra = 1, kCross = 2;
template <int W, int SCHED>
constexpr int inner_rounds() { return SCHED == kCross ? W : 2 * W - 1; }

template <int W, int SCHED>
__device__ __forceinline__ void inner_pair(int r2, int pr, int& li, int& lj)
{
  if constexpr (SCHED == kFullSweep)
  {
    int x, y;
    rr_pair(r2, pr, 2 * W, x, y);
    li = min(x, y);
    lj = max(x, y);
  }
  else if (SCHED == kCrossIntra && r2 < W - 1)
  {
    const int blk = pr >= W / 2 ? W : 0;
    int x, y;
    rr_pair(r2, pr - (blk ? W / 2 : 0), W, x, y);
    li = blk + min(x, y);
    lj = blk + max(x, y);
  }
  else
  {
    const int h = SCHED == kCrossIntra ? r2 - (W - 1) : r2;
    li = pr;
    lj = W + ((pr + h) % W);
  }
}

// Deferred J update of one block-pair inner sweep: jr holds one row of the
// pair's 2W columns of J; the sweep's rotations (schedule SCHED, rounds
// applied in sequence) come from rcs / rsn[r2 * W + pair].  Fully unrolled so
// the column indices are constants and jr stays in registers.
template <int W, int SCHED>
__device__ __forceinline__ void apply_rotations(double (&jr)[2 * W], const double* rcs, const double* rsn)
{
#pragma unroll
  for (int r2 = 0; r2 < inner_rounds<W, SCHED>(); ++r2)
#pragma unroll
    for (int pr = 0; pr < W; ++pr)
    {
      int li, lj;
      inner_pair<W, SCHED>(r2, pr, li, lj);
      const double c = rcs[r2 * W + pr], sv = rsn[r2 * W + pr];
      const double y1 = jr[li], y2 = jr[lj];
      jr[li] = c * y1 - sv * y2;
      jr[lj] = sv * y1 + c * y2;
    }
}

// apply_rotations for the cross schedules with runtime loops over the
// inner rounds and compile-time register indices: the row's registers are
// renamed instead of the pairing.  Cross rounds pair column a of block i
// with column (a + h) mod W of block j, so block j's registers shift down by
// one after every round and the pairs stay (a, W + a).  Intra rounds follow
// the circle method (pair 0: W - 1 with s; pair p: (s + p) and (s - p)
// mod (W - 1)), so positions 0..W-2 of each block shift down by one after
// every round and the pairs stay (W - 1, 0) and (p, W - 1 - p), the smaller
// original index taking the first role of the rotation.  W shifts of a
// W-cycle (W - 1 of a (W - 1)-cycle) restore the order.  The fully unrolled
// form kept hundreds of values live (every round's registers and rotations
// scheduled together).
__device__ __forceinline__ void rot2(double& x, double& y, double c, double sv)
{
  const double a = x, b = y;
  x = c * a - sv * b;
  y = sv * a + c * b;
}

template <int W, int SCHED>
__device__ __forceinline__ void apply_rotations_rt(double (&jr)[2 * W], const double* rcs, const double* rsn)
{
  static_assert(SCHED == kCrossIntra || SCHED == kCross, "cross schedules");
  constexpr int M = W - 1;
  int r2 = 0;
  if constexpr (SCHED == kCrossIntra)
  {
#pragma unroll 1
    for (int st = 0; st < M; ++st, ++r2)
    {
#pragma unroll
      for (int blk = 0; blk < 2; ++blk)
      {
        double* v = jr + blk * W;
        const int pr0 = blk * (W / 2);
        // Pair 0: original (M, st), st < M: st first.
        rot2(v[0], v[M], rcs[r2 * W + pr0], rsn[r2 * W + pr0]);
#pragma unroll
        for (int p = 1; p < W / 2; ++p)
        {
          const int x = (st + p) % M, y = (st - p + M) % M;  // originals at p, M - p
          const bool first = x < y;
          double u = first ? v[p] : v[M - p];
          double w = first ? v[M - p] : v[p];
          rot2(u, w, rcs[r2 * W + pr0 + p], rsn[r2 * W + pr0 + p]);
          v[p] = first ? u : w;
          v[M - p] = first ? w : u;
        }
        const double t0 = v[0];
#pragma unroll
        for (int k = 0; k < M - 1; ++k) v[k] = v[k + 1];
        v[M - 1] = t0;
      }
    }
  }
#pragma unroll 1
  for (int h = 0; h < W; ++h, ++r2)
  {
#pragma unroll
    for (int a = 0; a < W; ++a) rot2(jr[a], jr[W + a], rcs[r2 * W + a], rsn[r2 * W + a]);
    const double t0 = jr[W];
#pragma unroll
    for (int k = 0; k < W - 1; ++k) jr[W + k] = jr[W + k + 1];
    jr[2 * W - 1] = t0;
  }
}

// The same for the two cross schedules only.
template <class F>
void launch_cross_sched(int sched, F&& f)
{
  if (sched == kCrossIntra)
    f(std::integral_constant<int, kCrossIntra>{});
  else
    f(std::integral_constant<int, kCross>{});
}

// Calls f with std::integral_constant<int, sched> for the runtime schedule.
template <class F>
void launch_sched(int sched, F&& f)
{
  if (sched == kFullSweep)
    f(std::integral_constant<int, kFullSweep>{});
  else if (sched == kCrossIntra)
    f(std::integral_constant<int, kCrossIntra>{});
  else
    f(std::integral_constant<int, kCross>{});
}

// Whole Jacobi in one workgroup with X, J in LDS (q <= 64).
template <int NT>
__device__ void jacobi_small_one(const jac_prob& P, int max_sweeps)
{
  const int q = P.q;
  extern __shared__ double lds[];
  double* Xs = lds;          // q*q
  double* Js = lds + q * q;  // q*q
  __shared__ int flag;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int e = tid; e < q * q; e += NT)
  {
    const int c = e / q, r = e % q;
    Xs[e] = P.X[(size_t)c * P.ldx + r];
    Js[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  __shared__ double red[NT / 64];
  double part = 0.0;
  for (int e = tid; e < q * q; e += NT) part = fma(Xs[e], Xs[e], part);
  const double negl2 = jac_negl2(P, block_sum<NT>(part, red));
  const double tol = jac_tol(q);
  const int qq = (q + 1) & ~1;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep)
  {
    if (tid == 0) flag = 0;
    __syncthreads();
    int rot = 0;
    for (int s = 0; s < qq - 1; ++s)
    {
      for (int pr = wid; pr < qq / 2; pr += NT / 64)
      {
        int ci, cj;
        rr_pair(s, pr, qq, ci, cj);
        if (ci >= q || cj >= q) continue;
        if (ci > cj)
        {
          const int t = ci;
          ci = cj;
          cj = t;
        }
        rot |= rotate_pair<1>(Xs + ci * q, Xs + cj * q, Js + ci * q, Js + cj * q, q, q, tol, negl2, lane);
      }
      __syncthreads();
    }
    if (rot && lane == 0) atomicOr(&flag, 1);
    __syncthreads();
    const int any = flag;
    __syncthreads();
    if (!any) break;
  }
  for (int e = tid; e < q * q; e += NT)
  {
    const int c = e / q, r = e % q;
    P.X[(size_t)c * P.ldx + r] = Xs[e];
    P.J[(size_t)c * P.ldj + r] = Js[e];
  }
  if (tid == 0 && P.sweeps) *P.sweeps = sweep;
}

// One problem per workgroup, or, with a work queue, a persistent grid whose
// workgroups take problems in bt.order (largest first) until none is left.
template <int NT>
__global__ __launch_bounds__(NT) void jacobi_small_kernel(jac_batch bt, int max_sweeps)
{
  if (!bt.queue)
  {
    jacobi_small_one<NT>(jpick(bt, blockIdx.x), max_sweeps);
    return;
  }
  __shared__ int item;
  for (;;)
  {
    if (threadIdx.x == 0) item = atomicAdd(bt.queue, 1);
    __syncthreads();
    const int it = item;
    __syncthreads();
    if (it >= bt.batch) break;
    jacobi_small_one<NT>(jpick(bt, bt.order[it]), max_sweeps);
    __syncthreads();
  }
}

__global__ void jac_zero_int_kernel(int* p) { *p = 0; }

// Block pair of bj_group_kernel: block pair (bi, bj) of round s
// of the round-robin over nbb (even) blocks of W columns; a block paired with
// a padding block (odd block count) is still orthogonalised internally.
// LDS layout: X columns [2W][q], then the inner sweep's rotations
// rcs / rsn [(2W - 1) * W].  J is not staged: thread t < q keeps row t of the
// pair's 2W columns of J in registers and applies the recorded rotations
// once the inner sweep is done, which halves the LDS traffic of every inner
// round and takes J off its critical path.
template <int W>
struct bj_pair
{
  int ncol0, ncol1, cbeg0, cbeg1;  // scalars: a local array would live in scratch
  __device__ bool setup(const jac_prob& P, int s, int nbb)
  {
    const int q = P.q;
    const int nb = (q + W - 1) / W;
    int bi, bj;
    rr_pair(s, blockIdx.x, nbb, bi, bj);
    if (bi >= nb && bj >= nb) return false;
    cbeg0 = bi * W;
    cbeg1 = bj * W;
    ncol0 = bi < nb ? min(W, q - cbeg0) : 0;
    ncol1 = bj < nb ? min(W, q - cbeg1) : 0;
    return true;
  }
  __device__ bool real(int l) const { return (l % W) < (l < W ? ncol0 : ncol1); }
  __device__ int global_col(int l) const { return (l < W ? cbeg0 : cbeg1) + l % W; }
  // X's columns into LDS ([2W][q]) and row t of J's into jr, for q <= R
  // (a power of two dividing 2 W NT): thread t takes row t % R of
  // 2 W R / NT columns, with every load issued before any LDS store (an
  // element-strided loop with a division by q per element serialized them).
  template <int NT, int R>
  __device__ void load(const jac_prob& P, double* Xs, double (&jr)[2 * W]) const
  {
    static_assert(R <= NT && NT % R == 0 && (2 * W * R) % NT == 0, "row slots");
    constexpr int CPT = 2 * W * R / NT;  // columns per thread
    const int q = P.q;
    const int r = threadIdx.x % R, l0 = (threadIdx.x / R) * CPT;
    double v[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k)
    {
      const bool ok = r < q && real(l0 + k);
      v[k] = ok ? P.X[(size_t)global_col(l0 + k) * P.ldx + r] : 0.0;
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int l = 0; l < 2 * W; ++l) jr[l] = (t < q && real(l)) ? P.J[(size_t)global_col(l) * P.ldj + t] : 0.0;
    if (r < q)
    {
#pragma unroll
      for (int k = 0; k < CPT; ++k) Xs[(l0 + k) * q + r] = v[k];
    }
  }
  template <int NT, int R>
  __device__ void store(const jac_prob& P, const double* Xs, const double (&jr)[2 * W]) const
  {
    constexpr int CPT = 2 * W * R / NT;
    const int q = P.q;
    const int r = threadIdx.x % R, l0 = (threadIdx.x / R) * CPT;
    if (r < q)
    {
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (real(l0 + k)) P.X[(size_t)global_col(l0 + k) * P.ldx + r] = Xs[(l0 + k) * q + r];
    }
    const int t = threadIdx.x;
    if (t < q)
    {
#pragma unroll
      for (int l = 0; l < 2 * W; ++l)
        if (real(l)) P.J[(size_t)global_col(l) * P.ldj + t] = jr[l];
    }
  }
};

// LDS bytes of a block-Jacobi round kernel for problems of up to q columns.
constexpr size_t bj_lds_bytes(int W, int q) { return sizeof(double) * (2 * (size_t)W * q + 2 * (2 * (size_t)W - 1) * W); }

// One round of block Jacobi: workgroup = block pair (bi, bj) of the
// round-robin over nbb (even) blocks of W columns, one wave per column pair,
// X and J of the pair staged in LDS ([2W][q] each) and rotated together.
// With few problems (or q > 128) its 16 waves per workgroup hide the latency
// of each rotation chain better than deferring J would save.  cross: the
// inner sweep follows the cross schedules (inner_pair), else kFullSweep.
template <int NT, int W, int PL>
__global__ __launch_bounds__(NT) void bj_round_kernel(jac_batch bt, int s, int nbb, int inner_sweeps, int cross)
{
  const jac_prob& P = jpick(bt, blockIdx.y);
  const int q = P.q;
  const int nb = (q + W - 1) / W;
  if (P.flag && P.flag[1]) return;  // converged in an earlier sweep
  int bi, bj;
  rr_pair(s, blockIdx.x, nbb, bi, bj);
  // A block paired with a padding block (odd block count) is still
  // orthogonalised internally; the padding half holds no columns.
  if (bi >= nb && bj >= nb) return;
  extern __shared__ double lds[];
  double* Xs = lds;               // [2W][q]
  double* Js = lds + 2 * W * q;   // [2W][q]
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int ncol[2], cbeg[2];
  cbeg[0] = bi * W;
  cbeg[1] = bj * W;
  ncol[0] = bi < nb ? min(W, q - cbeg[0]) : 0;
  ncol[1] = bj < nb ? min(W, q - cbeg[1]) : 0;
  // Local column l (0..2W-1) -> global column.
  for (int e = tid; e < 2 * W * q; e += NT)
  {
    const int l = e / q, r = e % q;
    const int h = l / W, cl = l % W;
    double xv = 0.0, jv = 0.0;
    if (cl < ncol[h])
    {
      const int g = cbeg[h] + cl;
      xv = P.X[(size_t)g * P.ldx + r];
      jv = P.J[(size_t)g * P.ldj + r];
    }
    Xs[e] = xv;
    Js[e] = jv;
  }
  __syncthreads();
  const double tol = jac_tol(q);
  const double negl2 = jac_negl2(P, P.s[0]);  // ||X||_F^2 from jac_init_kernel
  int rot = 0;
  const bool intra = s == 0;
  const int nr = (!cross || intra) ? 2 * W - 1 : W;
  for (int it = 0; it < inner_sweeps; ++it)
  {
    for (int r2 = 0; r2 < nr; ++r2)
    {
      for (int pr = wid; pr < W; pr += NT / 64)
      {
        int li, lj;
        if (!cross)
          inner_pair<W, kFullSweep>(r2, pr, li, lj);
        else if (intra)
          inner_pair<W, kCrossIntra>(r2, pr, li, lj);
        else
          inner_pair<W, kCross>(r2, pr, li, lj);
        if ((li % W) >= ncol[li / W] || (lj % W) >= ncol[lj / W]) continue;
        rot |= rotate_pair<PL>(Xs + li * q, Xs + lj * q, Js + li * q, Js + lj * q, q, q, tol, negl2, lane);
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < 2 * W * q; e += NT)
  {
    const int l = e / q, r = e % q;
    const int h = l / W, cl = l % W;
    if (cl < ncol[h])
    {
      const int g = cbeg[h] + cl;
      P.X[(size_t)g * P.ldx + r] = Xs[e];
      P.J[(size_t)g * P.ldj + r] = Js[e];
    }
  }
  if (rot && lane == 0 && P.flag) atomicOr(P.flag, 1);
}

// One round of block Jacobi for few problems of q <= 64 PL: workgroup =
// block pair, one wave per column pair, X of the pair staged in LDS as in
// bj_round_kernel, but J not staged: thread t < q keeps row t of the pair's
// 2W columns of J in registers, the inner sweep's rotations are recorded in
// LDS (the identity where a pair does not rotate) and applied to those rows
// in order at its end (apply_rotations).  Each inner round then reads and
// writes only X, and J's updates leave the rotation chain.
// The inner sweep follows schedule SCHED (inner_pair).
template <int NT, int W, int PL, int SCHED>
__global__ __launch_bounds__(NT) void bj_defer_kernel(jac_batch bt, int s, int nbb)
{
  static_assert(NT == 64 * W, "one wave per column pair");
  const jac_prob& P = jpick(bt, blockIdx.y);
  const int q = P.q;
  if (P.flag && P.flag[1]) return;  // converged in an earlier sweep
  bj_pair<W> bp;
  if (!bp.setup(P, s, nbb)) return;
  extern __shared__ double lds[];
  double* Xs = lds;
  double* rcs = lds + 2 * W * q;
  double* rsn = rcs + (2 * W - 1) * W;
  double jr[2 * W];
  bp.template load<NT, 64 * PL>(P, Xs, jr);
  __syncthreads();
  const int tid = threadIdx.x, lane = tid & 63, pr = tid >> 6;
  const double tol = jac_tol(q);
  const double negl2 = jac_negl2(P, P.s[0]);  // ||X||_F^2 from jac_init_kernel
  int rot = 0;
  for (int r2 = 0; r2 < inner_rounds<W, SCHED>(); ++r2)
  {
    int li, lj;
    inner_pair<W, SCHED>(r2, pr, li, lj);
    const bool real = bp.real(li) && bp.real(lj);
    double* xi = Xs + li * q;
    double* xj = Xs + lj * q;
    double u[PL], v[PL];
    double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll
    for (int e = 0; e < PL; ++e)
    {
      const int t = lane + 64 * e;
      u[e] = t < q ? xi[t] : 0.0;
      v[e] = t < q ? xj[t] : 0.0;
      a = fma(u[e], u[e], a);
      b = fma(v[e], v[e], b);
      g = fma(u[e], v[e], g);
    }
    wave_sum3(a, b, g);
    double cs = 1.0, sn = 0.0;
    if (real && !(a <= negl2 || b <= negl2) && jacobi_rotation(a, b, g, tol, cs, sn))
    {
      rot = 1;
#pragma unroll
      for (int e = 0; e < PL; ++e)
      {
        const int t = lane + 64 * e;
        if (t < q)
        {
          xi[t] = cs * u[e] - sn * v[e];
          xj[t] = sn * u[e] + cs * v[e];
        }
      }
    }
    else
    {
      cs = 1.0;
      sn = 0.0;
    }
    if (lane == 0)
    {
      rcs[r2 * W + pr] = cs;
      rsn[r2 * W + pr] = sn;
    }
    __syncthreads();
  }
  if (tid < q) apply_rotations<W, SCHED>(jr, rcs, rsn);
  bp.template store<NT, 64 * PL>(P, Xs, jr);
  if (rot && lane == 0 && P.flag) atomicOr(P.flag, 1);
}

// One round of block Jacobi with 16-lane groups: each 16-lane row of a wave
// owns one column pair of the inner round (4 pairs per wave, W / 4 waves),
// holds PL = ceil(q / 16) elements of each column in registers and reduces
// its three dot products within the row (4 DPP stages, no cross-row step).
// The reduction and the rotation formulas are thus shared by the four pairs
// of a wave instead of being paid per pair, which cuts the VALU work per
// rotation about threefold against one pair per wave.  Inner sweep by
// schedule SCHED, rotation rule and convergence flag as bj_round_kernel; J
// deferred (bj_pair).
template <int W, int PL, int SCHED>
__global__ __launch_bounds__(W * 16) void bj_group_kernel(jac_batch bt, int s, int nbb)
{
  constexpr int NT = W * 16;
  const jac_prob& P = jpick(bt, blockIdx.y);
  const int q = P.q;
  if (P.flag && P.flag[1]) return;
  bj_pair<W> bp;
  if (!bp.setup(P, s, nbb)) return;
  extern __shared__ double lds[];
  double* Xs = lds;
  double* rcs = lds + 2 * W * q;
  double* rsn = rcs + (2 * W - 1) * W;
  double jr[2 * W];
  bp.template load<NT, 16 * PL>(P, Xs, jr);
  __syncthreads();
  const int tid = threadIdx.x;
  const double tol = jac_tol(q);
  const double negl2 = jac_negl2(P, P.s[0]);
  const int pr = tid >> 4;  // pair slot of this 16-lane group
  const int li16 = tid & 15;
  int rot = 0;
  for (int r2 = 0; r2 < inner_rounds<W, SCHED>(); ++r2)
  {
    int li, lj;
    inner_pair<W, SCHED>(r2, pr, li, lj);
    const bool real = bp.real(li) && bp.real(lj);
    double* xi = Xs + li * q;
    double* xj = Xs + lj * q;
    double u[PL], v[PL];
    double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll
    for (int e = 0; e < PL; ++e)
    {
      const int t = li16 + 16 * e;
      u[e] = t < q ? xi[t] : 0.0;
      v[e] = t < q ? xj[t] : 0.0;
      a = fma(u[e], u[e], a);
      b = fma(v[e], v[e], b);
      g = fma(u[e], v[e], g);
    }
    // Row sums: every lane of the 16-lane row ends with the pair's totals.
    a += dpp_mov<0xB1>(a);
    b += dpp_mov<0xB1>(b);
    g += dpp_mov<0xB1>(g);
    a += dpp_mov<0x4E>(a);
    b += dpp_mov<0x4E>(b);
    g += dpp_mov<0x4E>(g);
    a += dpp_mov<0x124>(a);
    b += dpp_mov<0x124>(b);
    g += dpp_mov<0x124>(g);
    a += dpp_mov<0x128>(a);
    b += dpp_mov<0x128>(b);
    g += dpp_mov<0x128>(g);
    double cs = 1.0, sn = 0.0;
    if (real && !(a <= negl2 || b <= negl2) && jacobi_rotation(a, b, g, tol, cs, sn))
    {
      rot = 1;
#pragma unroll
      for (int e = 0; e < PL; ++e)
      {
        const int t = li16 + 16 * e;
        if (t < q)
        {
          xi[t] = cs * u[e] - sn * v[e];
          xj[t] = sn * u[e] + cs * v[e];
        }
      }
    }
    else
    {
      cs = 1.0;
      sn = 0.0;
    }
    if (li16 == 0)
    {
      rcs[r2 * W + pr] = cs;
      rsn[r2 * W + pr] = sn;
    }
    __syncthreads();
  }
  if (tid < q) apply_rotations<W, SCHED>(jr, rcs, rsn);
  bp.template store<NT, 16 * PL>(P, Xs, jr);
  if (rot && li16 == 0 && P.flag) atomicOr(P.flag, 1);
}

typedef double dbl4_t __attribute__((ext_vector_type(4)));

// LDS bytes of bj_gram_kernel<*, *, 8, *> for problems of up to q columns.
constexpr size_t bjg_lds_bytes(int q) { return sizeof(double) * (32 * (size_t)q + 5 * 256 + 2 * 15 * 8 + 8); }

// One round of block Jacobi whose rotations are decided on the block pair's
// Gram matrix, for q <= NH = NT RPT / 2: workgroup = block pair (bi, bj) of
// the round-robin over nbb (even) blocks of W = 8 columns.  Row slot
// t + k NT (k < RPT) of thread t loads row t + k NT of the pair's 16 columns
// of X below NH, row t + k NT - NH of J from NH on (every load in flight at
// once), into LDS.
//   1. X_p is staged in LDS and its Gram G = X_p^T X_p formed on the matrix
//      cores by four waves (one 16 x 16 tile, each wave a share of the rows,
//      the partials summed in wave order; G kept exactly symmetric).
//   2. One wave runs the inner sweep (schedule SCHED) on G alone: lane
//      8 P + Q holds the 2 x 2 block of G at the round's row pair P and
//      column pair Q.  The diagonal lanes decide and compute their pair's
//      rotation from their own block by the one-sided rule (the tolerance
//      and negligible-column rule of the staged kernels) and record it;
//      every lane takes the rotations of its two pairs by lane shuffles and
//      updates its block, G <- R^T G R (the rotated pair's entry set to
//      zero, the new diagonal a - t g, b + t g); the blocks are exchanged
//      through LDS for the next round's pairing.  No workgroup barrier per
//      round.
//   3. Every row of X_p and J_p takes the recorded rotations in order
//      (apply_rotations), exactly as the staged kernels rotate the columns,
//      and is written back only when a rotation passed.
// The Gram only decides the rotations: the columns are rotated pair by
// pair, never multiplied by an accumulated 16 x 16 product, so that small
// columns keep the accuracy of the one-sided method (a product's rounding,
// relative to the largest column it mixes in, left graded spectra slower to
// converge and less accurate).  The Gram is formed afresh from the columns at
// every launch, so a sweep with no rotation passing was decided on the
// columns' own dot products, as in the staged kernels.
template <int NT, int RPT, int W, int SCHED>
__global__ __launch_bounds__(NT) void bj_gram_kernel(jac_batch bt, int s, int nbb)
{
  constexpr int C = 2 * W, NH = NT * RPT / 2, NR = inner_rounds<W, SCHED>();
  static_assert(W == 8, "one lane per (row pair, column pair): W * W = 64");
  const jac_prob P = jpick(bt, blockIdx.y);
  const int q = P.q;
  bj_pair<W> bp;
  if (!bp.setup(P, s, nbb)) return;
  extern __shared__ double lds[];
  double* Xs = lds;              // [C][q]
  double* Js = Xs + C * q;       // [C][q]
  double* Gp = Js + C * q;       // [4][C * C]: per-wave partial Grams
  double* G = Gp + 4 * C * C;    // [C * C]
  double* rcs = G + C * C;       // [NR][W]: the round's rotations
  double* rsn = rcs + 15 * W;
  int* rot_lds = (int*)(rsn + 15 * W);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // Slot tid + k NT (k < RPT): row slot of X below NH, of J from NH on;
  // both are staged in LDS ([C][q] each) until the rotations are known.
  {
    double r[RPT][C];
#pragma unroll
    for (int k = 0; k < RPT; ++k)
    {
      // Loads issued before the convergence flag is read, so that their
      // latencies overlap.
      const int slot = tid + k * NT;
      const bool isj = slot >= NH;
      const int row = isj ? slot - NH : slot;
      const double* src = isj ? P.J : P.X;
      const int ld = isj ? P.ldj : P.ldx;
#pragma unroll
      for (int l = 0; l < C; ++l)
      {
        const bool ok = row < q && bp.real(l);
        const size_t gc = ok ? bp.global_col(l) : 0;
        r[k][l] = ok ? src[gc * ld + row] : 0.0;
      }
    }
    if (P.flag && P.flag[1]) return;  // converged in an earlier sweep
#pragma unroll
    for (int k = 0; k < RPT; ++k)
    {
      const int slot = tid + k * NT;
      const bool isj = slot >= NH;
      const int row = isj ? slot - NH : slot;
      if (row < q)
      {
        double* dst = isj ? Js : Xs;
#pragma unroll
        for (int l = 0; l < C; ++l) dst[l * q + row] = r[k][l];
      }
    }
  }
  const double xnorm2 = P.s[0];  // ||X||_F^2 from jac_init_kernel
  __syncthreads();
  if (wid < 4)
  {
    // A(i, k) = X(r + k, i) in lane i + 16 k and B(k, j) = X(r + k, j) in
    // lane j + 16 k: the same element; accumulator r of lane l is
    // G(l / 16 + 4 r, l % 16).  Two accumulators halve the dependent chain.
    dbl4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const double* xc = Xs + (lane & 15) * q;
    const int kk = lane >> 4;
    for (int r0 = 8 * wid; r0 < q; r0 += 32)
    {
      const int ra = r0 + kk, rb = r0 + 4 + kk;
      const double a = ra < q ? xc[ra] : 0.0;
      const double b = rb < q ? xc[rb] : 0.0;
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, acc1, 0, 0, 0);
    }
    double* gw = Gp + wid * C * C;
#pragma unroll
    for (int k = 0; k < 4; ++k) gw[(kk + 4 * k) * C + (lane & 15)] = acc0[k] + acc1[k];
  }
  __syncthreads();
  if (wid == 0)
  {
#pragma unroll
    for (int m = 0; m < 4; ++m)
    {
      const int e = lane + 64 * m;
      const int k = e / C, l = e % C;
      const int o = min(k, l) * C + max(k, l);
      G[e] = ((Gp[o] + Gp[C * C + o]) + Gp[2 * C * C + o]) + Gp[3 * C * C + o];
    }
    wave_sync_lds();
    const double tol = jac_tol(q);
    const double negl2 = jac_negl2(P, xnorm2);
    const int pp = lane >> 3, qp = lane & 7;  // row pair, column pair
    int rot = 0;
#pragma unroll
    for (int r2 = 0; r2 < NR; ++r2)
    {
      int i0, j0, i1, j1;
      inner_pair<W, SCHED>(r2, pp, i0, j0);
      inner_pair<W, SCHED>(r2, qp, i1, j1);
      const double g00 = G[i0 * C + i1], g01 = G[i0 * C + j1];
      const double g10 = G[j0 * C + i1], g11 = G[j0 * C + j1];
      // The diagonal lanes compute their pair's rotation; every lane takes
      // the rotations of its row and column pairs from them.
      double cs = 1.0, sn = 0.0, tP = 0.0;
      if (pp == qp)
      {
        if (!(bp.real(i0) && bp.real(j0) && !(g00 <= negl2 || g11 <= negl2) &&
              jacobi_rotation(g00, g11, g01, tol, cs, sn, tP)))
        {
          cs = 1.0;
          sn = 0.0;
          tP = 0.0;
        }
        rot |= sn != 0.0;
      }
      const double cP = __shfl(cs, 9 * pp), sP = __shfl(sn, 9 * pp);
      const double cQ = __shfl(cs, 9 * qp), sQ = __shfl(sn, 9 * qp);
      // Recorded after the shuffles are issued, so that waiting for their
      // results does not wait for these writes too.
      if (pp == qp)
      {
        rcs[r2 * W + pp] = cs;
        rsn[r2 * W + pp] = sn;
      }
      double n00, n01, n10, n11;
      {
        // G' = R_P^T G R_Q with R = [[c, s], [-s, c]], each entry summed as
        // (t00 + t11) + (t01 + t10) with no contraction, so that the lane
        // holding the transposed block rounds every entry alike.
#pragma clang fp contract(off)
        n00 = ((cP * cQ) * g00 + (sP * sQ) * g11) + ((cP * -sQ) * g01 + (-sP * cQ) * g10);
        n01 = ((cP * sQ) * g00 + (-sP * cQ) * g11) + ((cP * cQ) * g01 + (-sP * sQ) * g10);
        n10 = ((sP * cQ) * g00 + (cP * -sQ) * g11) + ((sP * -sQ) * g01 + (cP * cQ) * g10);
        n11 = ((sP * sQ) * g00 + (cP * cQ) * g11) + ((sP * cQ) * g01 + (cP * sQ) * g10);
      }
      if (pp == qp && sP != 0.0)
      {
        n00 = fma(-tP, g01, g00);
        n11 = fma(tP, g01, g11);
        n01 = 0.0;
        n10 = 0.0;
      }
      G[i0 * C + i1] = n00;
      G[i0 * C + j1] = n01;
      G[j0 * C + i1] = n10;
      G[j0 * C + j1] = n11;
      wave_sync_lds();
    }
    // The vote runs on every lane of the wave (a vote inside a branch
    // taken by lane 0 alone would see only that lane's pair).
    const int any_rot = __any(rot);
    if (lane == 0) *rot_lds = any_rot;
  }
  __syncthreads();
  if (*rot_lds)
  {
#pragma unroll
    for (int k = 0; k < RPT; ++k)
    {
      const int slot = tid + k * NT;
      const bool isj = slot >= NH;
      const int row = isj ? slot - NH : slot;
      if (row < q)
      {
        const double* src = isj ? Js : Xs;
        double r[C];
#pragma unroll
        for (int l = 0; l < C; ++l) r[l] = src[l * q + row];
        apply_rotations_rt<W, SCHED>(r, rcs, rsn);
        double* dst = isj ? P.J : P.X;
        const int ld = isj ? P.ldj : P.ldx;
#pragma unroll
        for (int l = 0; l < C; ++l)
          if (bp.real(l)) dst[(size_t)bp.global_col(l) * ld + row] = r[l];
      }
    }
    if (tid == 0 && P.flag) atomicOr(P.flag, 1);
  }
}

// ---------------------------------------------------------------------------
// Block Jacobi for q above the staged kernels' limit (the X and J columns of a
// block pair no longer fit in LDS).  Blocks of kLW = 16 columns, round-robin
// over block pairs as above; one round is three launches:
//   bjl_gram_kernel  -- partial Grams X_p^T X_p of the pair's 32 columns over
//                       256-row chunks (matrix cores, chunk staged in LDS),
//   bjl_solve_kernel -- the chunks summed in order, then up to two cyclic
//                       Jacobi sweeps on the 32 x 32 Gram in LDS, each rotation
//                       computed by the one-sided rule from the Gram entries
//                       (the same tolerance and negligible-column rule) and
//                       applied as G <- R^T G R, accumulated into V;
//   bjl_apply_kernel -- X_p <- X_p V and J_p <- J_p V by 256-row chunks on the
//                       matrix cores (skipped where V = I).
// The Gram is recomputed from the columns in every round, so the rotations'
// inputs never carry updates from earlier rounds; a pair counts as rotated
// when a rotation of its first inner sweep passed the tolerance.
constexpr int kLW = 16;
constexpr int kLP = 2 * kLW;
constexpr int kLRC = 256;

struct bjl_work
{
  double* gpart = nullptr;  // [problem][pair][chunk][kLP * kLP]
  double* V = nullptr;      // [problem][pair][kLP * kLP]
  int* pflag = nullptr;     // [problem][pair]: V != I
  int pairs = 0, nch = 0;
};

// Global column of local column c (< kLP) of pair `pair` in round s, or -1
// (padding block, or past q).
__device__ __forceinline__ int bjl_col(int s, int pair, int nbb, int c, int q)
{
  int bi, bj;
  rr_pair(s, pair, nbb, bi, bj);
  const int col = (c < kLW ? bi : bj) * kLW + (c % kLW);
  return col < q ? col : -1;
}

__global__ __launch_bounds__(256) void bjl_gram_kernel(jac_batch bt, bjl_work w, int s, int nbb)
{
  const int prob = blockIdx.z;
  const jac_prob& P = jpick(bt, prob);
  if (P.flag && P.flag[1]) return;
  const int q = P.q, pair = blockIdx.y, ch = blockIdx.x;
  const int r0 = ch * kLRC;
  if (r0 >= q) return;
  __shared__ double Xs[kLRC][kLP + 1];
  __shared__ int cols[kLP];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid < kLP) cols[tid] = bjl_col(s, pair, nbb, tid, q);
  __syncthreads();
  for (int e = tid; e < kLRC * kLP; e += 256)
  {
    const int c = e / kLRC, r = e % kLRC;
    Xs[r][c] = (cols[c] >= 0 && r0 + r < q) ? P.X[(size_t)cols[c] * P.ldx + r0 + r] : 0.0;
  }
  __syncthreads();
  const int fl = lane & 15, fk = lane >> 4;
  const int qi = wid >> 1, qj = wid & 1;
  dbl4_t acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < kLRC; k0 += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[k0 + fk][qi * 16 + fl], Xs[k0 + fk][qj * 16 + fl], acc, 0, 0, 0);
  double* G = w.gpart + ((size_t)(prob * w.pairs + pair) * w.nch + ch) * (kLP * kLP);
#pragma unroll
  for (int r = 0; r < 4; ++r) G[(qi * 16 + fk + 4 * r) * kLP + qj * 16 + fl] = acc[r];
}

__global__ __launch_bounds__(256) void bjl_solve_kernel(jac_batch bt, bjl_work w, int s, int nbb)
{
  const int prob = blockIdx.y;
  const jac_prob& P = jpick(bt, prob);
  const int pair = blockIdx.x;
  int* pf = w.pflag + (size_t)prob * w.pairs + pair;
  if (P.flag && P.flag[1])
  {
    if (threadIdx.x == 0) *pf = 0;
    return;
  }
  const int q = P.q;
  __shared__ double G[kLP][kLP + 1];
  __shared__ double V[kLP][kLP + 1];
  __shared__ double rc[kLP / 2], rs[kLP / 2];
  __shared__ int rot_flag;
  const int tid = threadIdx.x;
  const int nch = (q + kLRC - 1) / kLRC;
  const double* Gp = w.gpart + (size_t)(prob * w.pairs + pair) * w.nch * (kLP * kLP);
  for (int e = tid; e < kLP * kLP; e += 256)
  {
    double acc = 0.0;
    for (int ch = 0; ch < nch; ++ch) acc += Gp[(size_t)ch * (kLP * kLP) + e];
    G[e / kLP][e % kLP] = acc;
    V[e / kLP][e % kLP] = (e / kLP == e % kLP) ? 1.0 : 0.0;
  }
  __syncthreads();
  const double negl2 = jac_negl2(P, P.s[0]);
  const double tol = jac_tol(q);
  int any = 0, first = 0;  // V != I; a rotation in the first inner sweep
  for (int inner = 0; inner < 2; ++inner)
  {
    if (tid == 0) rot_flag = 0;
    __syncthreads();
    int rotated = 0;  // thread-local, any pair of this inner sweep
    for (int r2 = 0; r2 < kLP - 1; ++r2)
    {
      if (tid < kLP / 2)
      {
        int li, lj;
        rr_pair(r2, tid, kLP, li, lj);
        if (li > lj)
        {
          const int t = li;
          li = lj;
          lj = t;
        }
        double cs = 1.0, sn = 0.0;
        const double a = G[li][li], b = G[lj][lj], g = G[li][lj];
        if (!(a <= negl2 || b <= negl2) && jacobi_rotation(a, b, g, tol, cs, sn))
          rotated = 1;
        else
        {
          cs = 1.0;
          sn = 0.0;
        }
        rc[tid] = cs;
        rs[tid] = sn;
      }
      __syncthreads();
      // Columns: G <- G R, V <- V R (pair = tid / 16, two rows per thread).
      {
        const int pr = tid >> 4, l16 = tid & 15;
        int li, lj;
        rr_pair(r2, pr, kLP, li, lj);
        if (li > lj)
        {
          const int t = li;
          li = lj;
          lj = t;
        }
        const double cs = rc[pr], sn = rs[pr];
        if (sn != 0.0)
          for (int row = 2 * l16; row < 2 * l16 + 2; ++row)
          {
            const double x = G[row][li], y = G[row][lj];
            G[row][li] = cs * x - sn * y;
            G[row][lj] = sn * x + cs * y;
            const double u = V[row][li], v = V[row][lj];
            V[row][li] = cs * u - sn * v;
            V[row][lj] = sn * u + cs * v;
          }
      }
      __syncthreads();
      // Rows: G <- R^T G.
      {
        const int pr = tid >> 4, l16 = tid & 15;
        int li, lj;
        rr_pair(r2, pr, kLP, li, lj);
        if (li > lj)
        {
          const int t = li;
          li = lj;
          lj = t;
        }
        const double cs = rc[pr], sn = rs[pr];
        if (sn != 0.0)
          for (int col = 2 * l16; col < 2 * l16 + 2; ++col)
          {
            const double x = G[li][col], y = G[lj][col];
            G[li][col] = cs * x - sn * y;
            G[lj][col] = sn * x + cs * y;
          }
      }
      __syncthreads();
    }
    if (rotated) atomicOr(&rot_flag, 1);
    __syncthreads();
    const int r = rot_flag;
    __syncthreads();
    if (!r) break;
    any = 1;
    if (inner == 0) first = 1;
  }
  double* Vg = w.V + (size_t)(prob * w.pairs + pair) * (kLP * kLP);
  for (int e = tid; e < kLP * kLP; e += 256) Vg[e] = V[e / kLP][e % kLP];
  if (tid == 0)
  {
    *pf = any;
    if (first && P.flag) atomicOr(P.flag, 1);
  }
}

__global__ __launch_bounds__(256) void bjl_apply_kernel(jac_batch bt, bjl_work w, int s, int nbb, int nchx)
{
  const int prob = blockIdx.z;
  const jac_prob& P = jpick(bt, prob);
  if (P.flag && P.flag[1]) return;
  const int pair = blockIdx.y;
  if (!w.pflag[(size_t)prob * w.pairs + pair]) return;
  const bool isJ = (int)blockIdx.x >= nchx;
  const int ch = isJ ? (int)blockIdx.x - nchx : (int)blockIdx.x;
  const int q = P.q;
  const int r0 = ch * kLRC;
  if (r0 >= q) return;
  double* M = isJ ? P.J : P.X;
  const int ld = isJ ? P.ldj : P.ldx;
  __shared__ double Xs[kLRC][kLP + 1];
  __shared__ double Vs[kLP][kLP + 1];
  __shared__ int cols[kLP];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double* Vg = w.V + (size_t)(prob * w.pairs + pair) * (kLP * kLP);
  for (int e = tid; e < kLP * kLP; e += 256) Vs[e / kLP][e % kLP] = Vg[e];
  if (tid < kLP) cols[tid] = bjl_col(s, pair, nbb, tid, q);
  __syncthreads();
  for (int e = tid; e < kLRC * kLP; e += 256)
  {
    const int c = e / kLRC, r = e % kLRC;
    Xs[r][c] = (cols[c] >= 0 && r0 + r < q) ? M[(size_t)cols[c] * ld + r0 + r] : 0.0;
  }
  __syncthreads();
  const int fl = lane & 15, fk = lane >> 4;
  // 32 output tiles of 16 x 16 (16 row tiles x 2 column tiles), 8 per wave.
  dbl4_t acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t)
  {
    const int tile = wid + 4 * t;
    const int rb = (tile >> 1) * 16, cb = (tile & 1) * 16;
    acc[t] = dbl4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < kLP; k0 += 4)
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[rb + fl][k0 + fk], Vs[k0 + fk][cb + fl], acc[t], 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 8; ++t)
  {
    const int tile = wid + 4 * t;
    const int rb = (tile >> 1) * 16, cb = (tile & 1) * 16;
#pragma unroll
    for (int r = 0; r < 4; ++r) Xs[rb + fk + 4 * r][cb + fl] = acc[t][r];
  }
  __syncthreads();
  for (int e = tid; e < kLRC * kLP; e += 256)
  {
    const int c = e / kLRC, r = e % kLRC;
    if (cols[c] >= 0 && r0 + r < q) M[(size_t)cols[c] * ld + r0 + r] = Xs[r][c];
  }
}

// J = I for every problem; flags cleared.
__global__ void jac_init_kernel(jac_batch bt)
{
  const jac_prob& P = jpick(bt, blockIdx.y);
  const int q = P.q;
  const long total = (long)q * q;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int c = (int)(e / q), r = (int)(e % q);
    P.J[(size_t)c * P.ldj + r] = (r == c) ? 1.0 : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    if (P.flag)
    {
      P.flag[0] = 0;
      P.flag[1] = 0;
    }
    if (P.sweeps) *P.sweeps = 0;
  }
}

// Squared column norms of X into s[0..q): one wave per column, four
// columns per workgroup, the column's loads in flight eight at a time
// (the loads of one workgroup per problem used to run one dependent column
// after another).  Each lane sums rows lane, lane + 64, ... in order, as
// column_norms_sorted does.
__global__ __launch_bounds__(256) void jac_colnorm2_kernel(jac_batch bt)
{
  const jac_prob P = jpick(bt, blockIdx.y);
  const int q = P.q;
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= q) return;
  const double* w = P.X + (size_t)c * P.ldx;
  double acc = 0.0;
  for (int t0 = 0; t0 < q; t0 += 64 * 8)
  {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
      const int t = t0 + lane + 64 * k;
      v[k] = t < q ? w[t] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += v[k] * v[k];
  }
  acc = wave_sum(acc);
  if (lane == 0) P.s[c] = acc;
}

// s[0] = ||X||_F^2 from the squared column norms in s (one workgroup per
// problem; every read precedes the write through block_sum's barrier); s
// is not otherwise used before jac_finish_kernel writes the singular values.
template <int NT>
__global__ __launch_bounds__(NT) void jac_norm_kernel(jac_batch bt)
{
  const jac_prob P = jpick(bt, blockIdx.x);
  __shared__ double red[NT / 64];
  double part = 0.0;
  for (int c = threadIdx.x; c < P.q; c += NT) part += P.s[c];
  const double t = block_sum<NT>(part, red);
  if (threadIdx.x == 0) P.s[0] = t;
}

// After a sweep: flag[1] = (no rotation in the sweep); flag[0] reset;
// counts converged problems into *done (zeroed before by jac_zero_int_kernel).
__global__ void jac_sweep_end_kernel(jac_batch bt, int batch, int* done)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < batch; i += gridDim.x * blockDim.x)
  {
    const jac_prob& P = jpick(bt, i);
    if (!P.flag) continue;
    if (!P.flag[1] && P.flag[0] == 0) P.flag[1] = 1;
    P.flag[0] = 0;
    if (P.flag[1]) atomicAdd(done, 1);
    if (P.sweeps && !P.flag[1]) *P.sweeps += 1;
  }
}

// The same for batches of at most 256 problems in one workgroup, which
// writes the count itself (no separate zeroing launch).
__global__ __launch_bounds__(256) void jac_sweep_end_one_kernel(jac_batch bt, int batch, int* done)
{
  const int i = threadIdx.x;
  int conv = 0;
  if (i < batch)
  {
    const jac_prob& P = jpick(bt, i);
    if (P.flag)
    {
      if (!P.flag[1] && P.flag[0] == 0) P.flag[1] = 1;
      P.flag[0] = 0;
      conv = P.flag[1] != 0;
      if (P.sweeps && !conv) *P.sweeps += 1;
    }
  }
  const int n = __syncthreads_count(conv);
  if (i == 0) *done = n;
}

// Sorted column norms and truncation rank (one workgroup per problem) from
// the squared column norms in s (jac_colnorm2_kernel): norms descending, a
// NaN norm (from a breakdown upstream) after every number, ties by index, so
// perm is always a permutation of 0..q-1 and every s[o] is written.  The
// rank is decided on the sorted values in LDS.  LDS: 2 q doubles.
template <int NT>
__global__ __launch_bounds__(NT) void jac_finish_kernel(jac_batch bt)
{
  const jac_prob P = jpick(bt, blockIdx.x);
  extern __shared__ double nrm[];
  const int q = P.q;
  double* srt = nrm + q;
  for (int c = threadIdx.x; c < q; c += NT) nrm[c] = sqrt(P.s[c]);
  __syncthreads();
  for (int c = threadIdx.x; c < q; c += NT)
  {
    const double v = nrm[c];
    const double kv = (v == v) ? v : -1.0;
    int rank = 0;
    for (int o = 0; o < q; ++o)
    {
      const double u = nrm[o];
      const double ku = (u == u) ? u : -1.0;
      rank += (ku > kv) || (ku == kv && o < c);
    }
    P.perm[rank] = c;
    P.s[rank] = v;
    srt[rank] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0 && P.rank_out)
  {
    const double delta = P.delta_ptr ? *P.delta_ptr : P.delta;
    *P.rank_out = truncate_rank(srt, q, delta, P.max_rank);
  }
}
// End of a block sweep: flags updated, converged problems counted into done.
void launch_sweep_end(const jac_batch& b, int batch, int* done_dev, hipStream_t st)
{
  if (batch <= 256)
  {
    hipLaunchKernelGGL(jac_sweep_end_one_kernel, dim3(1), dim3(256), 0, st, b, batch, done_dev);
    return;
  }
  hipLaunchKernelGGL(jac_zero_int_kernel, dim3(1), dim3(1), 0, st, done_dev);
  hipLaunchKernelGGL(jac_sweep_end_kernel, dim3((batch + 255) / 256), dim3(256), 0, st, b, batch, done_dev);
}
}  // namespace

int jacobi_block_threshold() { return 64; }

constexpr int kMaxBlockSweeps = 30;

// One block sweep's launches captured into a HIP graph on first use and
// replayed for every later sweep of the same call (the launches and their
// arguments are the same in every sweep).  Off, or with per-launch event
// timing on, the launches are enqueued directly.  The stream is drained
// before the graph is destroyed (every sweep loop ends on a stream wait).
struct sweep_graph
{
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  sweep_graph() = default;
  sweep_graph(const sweep_graph&) = delete;
  sweep_graph& operator=(const sweep_graph&) = delete;
  ~sweep_graph()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  template <class F>
  void run(hipStream_t st, bool on, F&& enqueue)
  {
    if (!on)
    {
      enqueue();
      return;
    }
    if (!exec)
    {
      hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
      enqueue();
      hip_check(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    hip_check(hipGraphLaunch(exec, st), "hipGraphLaunch");
  }
};

int launch_jacobi_svd(const jac_batch& b, int max_q, int* done_dev, int* done_host, tta_ctx* c, double credit_flops,
                      double credit_bytes)
{
  hipStream_t st = c->stream;
  const int batch = b.descs ? b.batch : 1;
  if (batch <= 0 || max_q <= 0) return 0;
  int sweeps_run = -1;
  // A single problem of 40 < q <= 64 takes the Gram-driven block rounds
  // below (a whole-problem workgroup runs each of its 63 rounds per sweep as
  // four dependent pairs per wave: 2 ms at q = 64 against about 0.6 ms).
  // Smaller problems, and batches, stay in the one-launch kernel, whose
  // sweeps need no host read.
  if (max_q <= 64 && !(c->jac_gram && batch == 1 && max_q > 40))
  {
    constexpr int NT = 512;
    const size_t lds = 2 * sizeof(double) * (size_t)max_q * max_q;
    if (b.descs && b.order && b.queue)
    {
      // Persistent: resident workgroups by LDS (at most 4 of 512 threads per
      // compute unit), fed from the work queue.
      static int cus = 0;
      if (cus == 0)
      {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
          cus = 256;
      }
      const int per_cu = std::max(1, std::min(4, (int)((size_t)(160 * 1024) / std::max<size_t>(lds, 1))));
      hipLaunchKernelGGL(jac_zero_int_kernel, dim3(1), dim3(1), 0, st, b.queue);
      hipLaunchKernelGGL(jacobi_small_kernel<NT>, dim3(std::min(batch, cus * per_cu)), dim3(NT), lds, st, b, 60);
    }
    else
      hipLaunchKernelGGL(jacobi_small_kernel<NT>, dim3(batch), dim3(NT), lds, st, b, 60);
  }
  else
  {
    // Block width: the staged 2W columns of X and J (32 W q bytes) must fit
    // in the workgroup's LDS, and q <= 640 bounds the per-lane column slices.
    constexpr int NT = 1024;
    static int max_lds = 0;
    if (max_lds == 0)
    {
      int dev = 0, v = 0;
      (void)hipGetDevice(&dev);
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        v = 65536;
      max_lds = v;
    }
    if ((size_t)32 * 8 * max_q > (size_t)max_lds || max_q > 640)
    {
      // Too tall to stage a block pair: the Gram-based block rounds.
      const int nb = (max_q + kLW - 1) / kLW;
      const int nbb = (nb + 1) & ~1;
      bjl_work w;
      w.pairs = nbb / 2;
      w.nch = (max_q + kLRC - 1) / kLRC;
      const size_t ng = (size_t)batch * w.pairs * w.nch * kLP * kLP;
      const size_t nv = (size_t)batch * w.pairs * kLP * kLP;
      w.gpart = static_cast<double*>(c->mem.get(ng * sizeof(double)));
      w.V = static_cast<double*>(c->mem.get(nv * sizeof(double)));
      w.pflag = static_cast<int*>(c->mem.get((size_t)batch * w.pairs * sizeof(int)));
      {
        long g = ((long)max_q * max_q + 255) / 256;
        if (g > 1024) g = 1024;
        hipLaunchKernelGGL(jac_init_kernel, dim3((int)g, batch), dim3(256), 0, st, b);
        hipLaunchKernelGGL(jac_colnorm2_kernel, dim3((max_q + 3) / 4, batch), dim3(256), 0, st, b);
        hipLaunchKernelGGL(jac_norm_kernel<256>, dim3(batch), dim3(256), 0, st, b);
      }
      const size_t rec0 = c->prof_pending.size();
      sweep_graph sg;
      const bool use_graph = c->graphs && !c->prof;
      for (int sweep = 0; sweep < kMaxBlockSweeps; ++sweep)
      {
        sweeps_run = sweep + 1;
        sg.run(st, use_graph, [&]() {
          for (int r = 0; r < nbb - 1; ++r)
          {
            prof_scope ps(c, TTA_K_JACOBI_ROUND);
            hipLaunchKernelGGL(bjl_gram_kernel, dim3(w.nch, w.pairs, batch), dim3(256), 0, st, b, w, r, nbb);
            hipLaunchKernelGGL(bjl_solve_kernel, dim3(w.pairs, batch), dim3(256), 0, st, b, w, r, nbb);
            hipLaunchKernelGGL(bjl_apply_kernel, dim3(2 * w.nch, w.pairs, batch), dim3(256), 0, st, b, w, r, nbb,
                               w.nch);
          }
          launch_sweep_end(b, batch, done_dev, st);
        });
        if (sweep >= 2 || sweep == kMaxBlockSweeps - 1)
        {
          (void)hipMemcpyAsync(done_host, done_dev, sizeof(int), hipMemcpyDeviceToHost, st);
          (void)hipStreamSynchronize(st);
          if (*done_host >= batch) break;
          if (sweep == kMaxBlockSweeps - 1) c->jac_unconverged += batch - *done_host;
        }
      }
      size_t nrec = 0;
      for (size_t r = rec0; r < c->prof_pending.size(); ++r) nrec += c->prof_pending[r].cls == TTA_K_JACOBI_ROUND;
      for (size_t r = rec0; r < c->prof_pending.size(); ++r)
        if (c->prof_pending[r].cls == TTA_K_JACOBI_ROUND)
        {
          c->prof_pending[r].flops = credit_flops / (double)nrec;
          c->prof_pending[r].bytes = credit_bytes / (double)nrec;
        }
      c->mem.put(w.gpart);
      c->mem.put(w.V);
      c->mem.put(w.pflag);
      constexpr int NF = 256;
      hipLaunchKernelGGL(jac_colnorm2_kernel, dim3((max_q + 3) / 4, batch), dim3(256), 0, st, b);
      hipLaunchKernelGGL(jac_finish_kernel<NF>, dim3(batch), dim3(NF), 2 * sizeof(double) * (size_t)max_q, st, b);
      return sweeps_run;
    }
    // Batches of q <= 256 fill the chip with 256-thread workgroups, where the
    // grouped kernel's lower VALU count and deferred J win; few or larger
    // problems use the 16-wave workgroups of bj_round_kernel.
    const bool grouped = max_q <= 256 && batch >= 32 && bj_lds_bytes(16, max_q) <= (size_t)max_lds;
    // Few problems of q <= 256: blocks of 8 columns, so that a round spreads
    // over twice as many workgroups (CUs); each CU is bound by its LDS and
    // VALU throughput on the rotations, not by the round count.
    const bool narrow = !grouped && batch < 32 && max_q > 128 && max_q <= 256;
    // Few problems of q <= 512: rotations decided on the block pair's Gram
    // matrix (blocks of 8 columns; 2 q row slots of X and J on 256 or 512
    // threads).
    const bool gram = !grouped && c->jac_gram && batch < 32 && max_q <= 512 &&
                      bjg_lds_bytes(max_q) <= (size_t)max_lds;
    const int gnt = max_q <= 128 ? 256 : (max_q <= 256 ? 512 : 1024);
    const int W = grouped ? 16 : ((narrow || gram) ? 8 : ((size_t)32 * 16 * max_q <= (size_t)max_lds ? 16 : 8));
    const bool cross = grouped ? g_jac_cross_group : g_jac_cross;
    const int nb = (max_q + W - 1) / W;
    const int nbb = (nb + 1) & ~1;
    const size_t lds = gram ? bjg_lds_bytes(max_q)
                       : ((grouped || narrow) ? bj_lds_bytes(W, max_q) : 4 * sizeof(double) * (size_t)W * max_q);
    {
      long g = ((long)max_q * max_q + 255) / 256;
      if (g > 1024) g = 1024;
      hipLaunchKernelGGL(jac_init_kernel, dim3((int)g, batch), dim3(256), 0, st, b);
      hipLaunchKernelGGL(jac_colnorm2_kernel, dim3((max_q + 3) / 4, batch), dim3(256), 0, st, b);
      hipLaunchKernelGGL(jac_norm_kernel<256>, dim3(batch), dim3(256), 0, st, b);
    }
    const size_t rec0 = c->prof_pending.size();
    sweep_graph sg;
    const bool use_graph = c->graphs && !c->prof;
    // One problem whose sweep counter sits next to the done count: the flag
    // is read from the hint on, and the counter with it gives the sweeps the
    // problem needed, whatever sweeps ran after its convergence.
    const bool hinted = batch == 1 && !b.descs && b.one.sweeps && b.one.sweeps == done_dev + 1;
    const int check_from = hinted ? std::max(2, c->jac_sweep_hint - 3) : 2;
    for (int sweep = 0; sweep < kMaxBlockSweeps; ++sweep)
    {
      sweeps_run = sweep + 1;
      sg.run(st, use_graph, [&]() {
        for (int s = 0; s < nbb - 1; ++s)
        {
          const dim3 grid(nbb / 2, batch);
          prof_scope ps(c, TTA_K_JACOBI_ROUND);
          const int sched = !cross ? kFullSweep : (s == 0 ? kCrossIntra : kCross);
          // The Gram rounds always take the cross schedules.
          const int gsched = s == 0 ? kCrossIntra : kCross;
          if (gram)
            launch_cross_sched(gsched, [&](auto sc) {
              constexpr int SC = decltype(sc)::value;
              if (gnt == 256)
                hipLaunchKernelGGL((bj_gram_kernel<256, 1, 8, SC>), grid, dim3(256), lds, st, b, s, nbb);
              else if (gnt == 512)
                hipLaunchKernelGGL((bj_gram_kernel<512, 1, 8, SC>), grid, dim3(512), lds, st, b, s, nbb);
              else
                hipLaunchKernelGGL((bj_gram_kernel<512, 2, 8, SC>), grid, dim3(512), lds, st, b, s, nbb);
            });
          else if (grouped && max_q <= 128)
            launch_sched(sched, [&](auto sc) {
              hipLaunchKernelGGL((bj_group_kernel<16, 8, decltype(sc)::value>), grid, dim3(256), lds, st, b, s, nbb);
            });
          else if (grouped)
            launch_sched(sched, [&](auto sc) {
              hipLaunchKernelGGL((bj_group_kernel<16, 16, decltype(sc)::value>), grid, dim3(256), lds, st, b, s, nbb);
            });
          else if (narrow)
            launch_sched(sched, [&](auto sc) {
              hipLaunchKernelGGL((bj_defer_kernel<512, 8, 4, decltype(sc)::value>), grid, dim3(512), lds, st, b, s,
                                 nbb);
            });
          else if (W == 16 && max_q <= 128)
            hipLaunchKernelGGL((bj_round_kernel<NT, 16, 2>), grid, dim3(NT), lds, st, b, s, nbb, 1, cross ? 1 : 0);
          else if (W == 16 && max_q <= 256)
            hipLaunchKernelGGL((bj_round_kernel<NT, 16, 4>), grid, dim3(NT), lds, st, b, s, nbb, 1, cross ? 1 : 0);
          else if (W == 16)
            hipLaunchKernelGGL((bj_round_kernel<NT, 16, 5>), grid, dim3(NT), lds, st, b, s, nbb, 1, cross ? 1 : 0);
          else
            hipLaunchKernelGGL((bj_round_kernel<NT, 8, 10>), grid, dim3(NT), lds, st, b, s, nbb, 1, cross ? 1 : 0);
        }
        launch_sweep_end(b, batch, done_dev, st);
      });
      if (sweep >= check_from || sweep == kMaxBlockSweeps - 1)
      {
        (void)hipMemcpyAsync(done_host, done_dev, (hinted ? 2 : 1) * sizeof(int), hipMemcpyDeviceToHost, st);
        (void)hipStreamSynchronize(st);
        if (*done_host >= batch)
        {
          // The sweeps with a rotation, and the one that found none.
          if (hinted) sweeps_run = done_host[1] + 1;
          break;
        }
        // Problems still rotating after the last sweep: their W = X J is
        // orthogonal only to the last sweep's rotation size, not to the
        // tolerance; counted for tta_prof_jacobi_unconverged.
        if (sweep == kMaxBlockSweeps - 1) c->jac_unconverged += batch - *done_host;
      }
    }
    if (hinted) c->jac_sweep_hint = sweeps_run;
    // Equal share of the call's credit for every round launch recorded.
    size_t nrec = 0;
    for (size_t r = rec0; r < c->prof_pending.size(); ++r) nrec += c->prof_pending[r].cls == TTA_K_JACOBI_ROUND;
    for (size_t r = rec0; r < c->prof_pending.size(); ++r)
      if (c->prof_pending[r].cls == TTA_K_JACOBI_ROUND)
      {
        c->prof_pending[r].flops = credit_flops / (double)nrec;
        c->prof_pending[r].bytes = credit_bytes / (double)nrec;
      }
  }
  constexpr int NF = 256;
  hipLaunchKernelGGL(jac_colnorm2_kernel, dim3((max_q + 3) / 4, batch), dim3(256), 0, st, b);
  hipLaunchKernelGGL(jac_finish_kernel<NF>, dim3(batch), dim3(NF), 2 * sizeof(double) * (size_t)max_q, st, b);
  return sweeps_run;
}

}  // namespace tta
