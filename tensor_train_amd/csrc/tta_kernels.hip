// HIP kernels (gfx950) for the FP64 plain path of TT rounding and MPO
// apply-and-round.  Launch wrappers are declared in tta_kernels.hpp.
#include <algorithm>
#include <array>
#include <map>
#include <mutex>

#include "../../include/tt_amd.h"
#include "tta_device.hpp"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"
#include "tta_potrf.hpp"

namespace tta
{
using namespace dev;

// ---------------------------------------------------------------------------
// FP64 GEMM with arbitrary element strides on the matrix cores:
//   C(i,j) = alpha * sum_k A(i,k) B(k,j) + beta * C(i,j)
// A(i,k) = A[i*sar + k*sac], B(k,j) = B[k*sbr + j*sbc], C(i,j) = C[i*scr + j*scc].
// 64x64 output tile per 256-thread workgroup; K staged through LDS in
// 32-deep slabs, the next slab prefetched into registers while the current
// one is multiplied; each wave owns a 32x32 quarter as 2x2 blocks of
// v_mfma_f64_16x16x4f64.  Operand lanes: A(i,k) in lane i + 16k, B(k,j) in
// lane j + 16k; accumulator element r of lane l is C(l/16 + 4r, l%16).
// blockIdx.z walks a batch with pointer strides bsa/bsb/bsc.
// ---------------------------------------------------------------------------
namespace
{
constexpr int GT = 64, GK = 32;
typedef double dbl4 __attribute__((ext_vector_type(4)));

// Element e (0..7, or 0..3 for 32-wide tiles) of this thread's share of a K
// slab of a TT-wide tile: A(i0 + ii, k0 + kk) and B(k0 + kk, j0 + jj), with
// the thread order matched to the operand's contiguous direction so that
// loads coalesce.
// (Selects rather than branches: with stores to the two outputs in
// branches the compiler kept them in scratch memory.)
// KD: the slab depth (32, or 64 for the Gram chain's products).
template <int TT = GT, int KD = GK>
__device__ __forceinline__ void slab_coords_a(bool kc, int idx, int& ii, int& kk)
{
  constexpr int SH = TT == 64 ? 6 : 5, KS = KD == 64 ? 6 : 5;
  static_assert(TT == 64 || TT == 32, "tile width");
  static_assert(KD == 64 || KD == 32, "slab depth");
  ii = kc ? (idx >> KS) : (idx & (TT - 1));
  kk = kc ? (idx & (KD - 1)) : (idx >> SH);
}
template <int TT = GT, int KD = GK>
__device__ __forceinline__ void slab_coords_b(bool jc, int idx, int& kk, int& jj)
{
  constexpr int SH = TT == 64 ? 6 : 5, KS = KD == 64 ? 6 : 5;
  static_assert(TT == 64 || TT == 32, "tile width");
  static_assert(KD == 64 || KD == 32, "slab depth");
  kk = jc ? (idx >> SH) : (idx & (KD - 1));
  jj = jc ? (idx & (TT - 1)) : (idx >> KS);
}

typedef float flt4 __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core step in the compute type CT.  FP64: accumulator
// element r of lane l is C(l/16 + 4r, l%16); FP32: C(4(l/16) + r, l%16).
template <class CT>
struct mfma_step;
template <>
struct mfma_step<double>
{
  using acc_t = dbl4;
  __device__ static acc_t run(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int fk, int r) { return fk + 4 * r; }
};
template <>
struct mfma_step<float>
{
  using acc_t = flt4;
  __device__ static acc_t run(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int fk, int r) { return 4 * fk + r; }
};

// The TT x TT tile (TT = 64, or 32 for small outputs) at (i0, j0) of
// C = alpha A B + beta C with A of type TA, B of type TB, C of type TC,
// multiplied on the matrix cores in CT (operands converted on the way into
// LDS).  Each wave owns a (TT/2)^2 quarter as blocks of 16 x 16.  Every
// output element accumulates the same sequence of 4-deep matrix-core steps
// over K whatever TT is, so both tile sizes give identical results.  Every
// use of LDS is followed by a barrier, so a workgroup may run several in a
// row.
// NSA / NSB > 1: the operand is the sum of na <= NSA (nb <= NSB) equally
// laid out buffers psa (psb) elements apart, all loaded at once and summed
// in order (split-K partial products read without a separate sum).
template <class TA, class TB, class TC, class CT, int TT = GT, int NSA = 1, int NSB = 1, int KD = GK>
__device__ __forceinline__ void gemm_tile_t(int M, int N, int K, const TA* A, long sar, long sac, const TB* B, long sbr,
                                            long sbc, TC* C, long scr, long scc, double alpha, double beta, int i0,
                                            int j0, int na = 1, long psa = 0, int nb = 1, long psb = 0)
{
  using S = mfma_step<CT>;
  constexpr int WS = TT / 2, NB = WS / 16, NL = KD * TT / 256;
  __shared__ CT As[KD][TT + 1];
  __shared__ CT Bs[KD][TT + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int wr = wid >> 1, wc = wid & 1;
  const bool a_kc = (sac == 1);  // A contiguous along k
  const bool b_jc = (sbc == 1);  // B contiguous along j
  typename S::acc_t acc[NB][NB];
#pragma unroll
  for (int u = 0; u < NB; ++u)
#pragma unroll
    for (int v = 0; v < NB; ++v) acc[u][v] = typename S::acc_t{0, 0, 0, 0};
  const int fl = lane & 15, fk = lane >> 4;
  // Registers holding the next slab while the current one is multiplied.
  CT pa[NL], pb[NL];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int l = 0; l < NL; ++l)
    {
      const int idx = tid + 256 * l;
      int ii, kk, jj, kb;
      slab_coords_a<TT, KD>(a_kc, idx, ii, kk);
      const int gi = i0 + ii, gk = k0 + kk;
      if constexpr (NSA > 1)
      {
        const bool in = gi < M && gk < K;
        const size_t o = (size_t)gi * sar + (size_t)gk * sac;
        CT vs[NSA];
#pragma unroll
        for (int t = 0; t < NSA; ++t) vs[t] = (in && t < na) ? (CT)A[o + (size_t)t * psa] : (CT)0;
        CT v = vs[0];
#pragma unroll
        for (int t = 1; t < NSA; ++t) v += vs[t];
        pa[l] = v;
      }
      else
        pa[l] = (gi < M && gk < K) ? (CT)A[(size_t)gi * sar + (size_t)gk * sac] : (CT)0;
      slab_coords_b<TT, KD>(b_jc, idx, kb, jj);
      const int gj = j0 + jj, gk2 = k0 + kb;
      if constexpr (NSB > 1)
      {
        const bool in = gj < N && gk2 < K;
        const size_t o = (size_t)gk2 * sbr + (size_t)gj * sbc;
        CT vs[NSB];
#pragma unroll
        for (int t = 0; t < NSB; ++t) vs[t] = (in && t < nb) ? (CT)B[o + (size_t)t * psb] : (CT)0;
        CT v = vs[0];
#pragma unroll
        for (int t = 1; t < NSB; ++t) v += vs[t];
        pb[l] = v;
      }
      else
        pb[l] = (gj < N && gk2 < K) ? (CT)B[(size_t)gk2 * sbr + (size_t)gj * sbc] : (CT)0;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += KD)
  {
#pragma unroll
    for (int l = 0; l < NL; ++l)
    {
      const int idx = tid + 256 * l;
      int ii, kk, jj, kb;
      slab_coords_a<TT, KD>(a_kc, idx, ii, kk);
      As[kk][ii] = pa[l];
      slab_coords_b<TT, KD>(b_jc, idx, kb, jj);
      Bs[kb][jj] = pb[l];
    }
    __syncthreads();
    if (k0 + KD < K) fetch(k0 + KD);
#pragma unroll
    for (int ks = 0; ks < KD; ks += 4)
    {
      CT a[NB], b[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) a[u] = As[ks + fk][wr * WS + u * 16 + fl];
#pragma unroll
      for (int v = 0; v < NB; ++v) b[v] = Bs[ks + fk][wc * WS + v * 16 + fl];
#pragma unroll
      for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int v = 0; v < NB; ++v) acc[u][v] = S::run(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < NB; ++u)
#pragma unroll
    for (int v = 0; v < NB; ++v)
    {
      const int gj = j0 + wc * WS + v * 16 + fl;
      if (gj >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        const int gi = i0 + wr * WS + u * 16 + S::row(fk, r);
        if (gi >= M) continue;
        TC* c = C + (size_t)gi * scr + (size_t)gj * scc;
        const double x = (double)acc[u][v][r];
        *c = (TC)((beta == 0.0) ? alpha * x : alpha * x + beta * (double)(*c));
      }
    }
}

__device__ __forceinline__ void gemm_tile(const gemm_desc& g, const double* A, const double* B, double* C)
{
  gemm_tile_t<double, double, double, double>(g.M, g.N, g.K, A, g.sar, g.sac, B, g.sbr, g.sbc, C, g.scr, g.scc, g.alpha,
                                              g.beta, blockIdx.y * GT, blockIdx.x * GT);
}

__global__ __launch_bounds__(256) void gemm_f64_kernel(gemm_desc g)
{
  gemm_tile(g, g.A + (size_t)blockIdx.z * g.bsa, g.B + (size_t)blockIdx.z * g.bsb, g.C + (size_t)blockIdx.z * g.bsc);
}

// Split-K: slice blockIdx.z multiplies the K range [z kper, (z+1) kper) into
// its own partial product P_z (M x N, in C's orientation: row-major when C's
// column stride is 1, column-major otherwise); splitk_reduce_kernel then
// sums P_0 .. P_{S-1} in that order (a fixed order, so the result does not
// depend on scheduling) and applies alpha and beta.
__global__ __launch_bounds__(256) void gemm_f64_splitk_kernel(gemm_desc g, int kper, double* part)
{
  const int k0 = blockIdx.z * kper;
  const int kn = min(g.K - k0, kper);
  const bool rm = g.scc == 1;
  double* P = part + (size_t)blockIdx.z * g.M * g.N;
  gemm_tile_t<double, double, double, double>(g.M, g.N, kn, g.A + (size_t)k0 * g.sac, g.sar, g.sac,
                                              g.B + (size_t)k0 * g.sbr, g.sbr, g.sbc, P, rm ? g.N : 1,
                                              rm ? 1 : g.M, 1.0, 0.0, blockIdx.y * GT, blockIdx.x * GT);
}

__global__ __launch_bounds__(256) void splitk_reduce_kernel(gemm_desc g, int S, const double* part)
{
  const long mn = (long)g.M * g.N;
  const bool rm = g.scc == 1;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < mn; e += (long)gridDim.x * 256)
  {
    double sum = 0.0;
    for (int z = 0; z < S; ++z) sum += part[(size_t)z * mn + e];
    const long i = rm ? e / g.N : e % g.M, j = rm ? e % g.N : e / g.M;
    double* c = g.C + i * g.scr + j * g.scc;
    *c = g.beta == 0.0 ? g.alpha * sum : g.alpha * sum + g.beta * *c;
  }
}

// Split-K sum with one of the epilogues of splitk_epi (square row-major
// n x n output).
__global__ __launch_bounds__(256) void splitk_reduce_epi_kernel(gemm_desc g, int S, const double* part, splitk_epi ep)
{
  __shared__ double red[4];
  const int n = g.N;
  const long mn = (long)n * n;
  double emax = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < mn; e += (long)gridDim.x * 256)
  {
    double sum = 0.0;
    for (int z = 0; z < S; ++z) sum += part[(size_t)z * mn + e];
    const int i = (int)(e / n), j = (int)(e % n);
    double* c = g.C + (size_t)i * g.scr + j;
    const double v = g.beta == 0.0 ? g.alpha * sum : g.alpha * sum + g.beta * *c;
    const double id = i == j ? 1.0 : 0.0;
    switch (ep.mode)
    {
      case SPLITK_EPI_GRAM:
      {
        const double x = v - id;
        *c = x;
        if (ep.out2) ep.out2[e] = i < j ? x : (i == j ? 0.5 * x : 0.0);
        emax = fmax(emax, fabs(x) <= DBL_MAX ? fabs(x) : INFINITY);
        break;
      }
      case SPLITK_EPI_PHI:
      case SPLITK_EPI_PHI_LAST:
      {
        const double d = ep.E[e] - v;
        const double x = i < j ? d : (i == j ? 0.5 * d : 0.0);
        *c = x;
        if (ep.mode == SPLITK_EPI_PHI_LAST)
        {
          ep.out2[e] = id + x;
          ep.out3[e] = id - x;
        }
        break;
      }
      case SPLITK_EPI_POLAR:
      {
        const double x = ep.E[e];
        *c = id - 0.5 * x + 0.375 * v;
        ep.out2[e] = id + 0.5 * x - 0.125 * v;
        break;
      }
      default:
        *c = v + ep.diag * id;
    }
  }
  if (ep.mode == SPLITK_EPI_GRAM)
  {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, 64));
    if (lane == 0) red[wid] = emax;
    __syncthreads();
    if (threadIdx.x == 0) ep.part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  }
}

// Independent GEMMs of one launch, descriptors passed by value.
__global__ __launch_bounds__(256) void gemm_f64_multi_kernel(gemm_multi m)
{
  if ((int)blockIdx.z == m.count)
  {
    // The speculative CholQR2 acceptance decision riding along.
    if (blockIdx.x != 0 || blockIdx.y != 0 || threadIdx.x >= 64) return;
    double emax = 0.0;
    for (int b = threadIdx.x; b < m.dnp; b += 64) emax = fmax(emax, m.dpart[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, 64));
    if (threadIdx.x != 0) return;
    const bool ok = m.dinfo[0] == 0.0 && isfinite(m.dinfo[1]) && isfinite(m.dinfo[2]) && emax * m.dn <= m.dthr;
    if (!ok) m.dflag[0] = 0.0;
    return;
  }
  const gemm_desc& g = m.g[blockIdx.z];
  if ((int)blockIdx.y * GT >= g.M || (int)blockIdx.x * GT >= g.N) return;
  gemm_tile(g, g.A, g.B, g.C);
}

// One block step of the blocked Cholesky with inverse (chol_step in the
// header), one workgroup per output tile:
//   * trailing tile (a, b), a <= b, of the upper triangle (tile (0, 0)
//     first): R_a = D^T G(jb rows from j0, columns of block a) and likewise
//     R_b (the block row of R; tile (a, a) writes its R_a into R, others
//     into their scratch), then G(block a, block b) -= R_a^T R_b.  Tile
//     (0, 0) then factors the next diagonal block (potrf_inv_block).
//     Lower trailing tiles are left stale: only the upper block rows and the
//     diagonal blocks of G are read afterwards.
//   * X tile i (rows 64 i .. of the first j0): T = X(rows, 0:j0) R(0:j0,
//     j0:j1) into scratch, then X(rows, j0:j1) = -T D.
// D = R_jj^-1, the step's diagonal block of X, factored by the launch before.
__device__ __forceinline__ void chol_step_run(const chol_step& s, int t)
{
  const int n = s.n, j0 = s.j0, jb = s.jb, j1 = j0 + jb, w = n - j1;
  const double* D = s.X + (size_t)j0 * n + j0;
  double* S = s.S + (size_t)t * 2 * GT * GT;
  if (t < s.ntri)
  {
    int a = 0, rem = t;
    while (rem >= s.nt - a)
    {
      rem -= s.nt - a;
      ++a;
    }
    const int b = a + rem;
    const int ca = j1 + GT * a, cb = j1 + GT * b;
    const int wa = min(GT, n - ca), wb = min(GT, n - cb);
    double* Ra = a == b ? s.R + (size_t)j0 * n + ca : S;
    const long la = a == b ? n : GT;
    gemm_tile_t<double, double, double, double>(jb, wa, jb, D, 1, n, s.G + (size_t)j0 * n + ca, n, 1, Ra, la, 1, 1.0,
                                                0.0, 0, 0);
    const double* Rb = Ra;
    long lb = la;
    if (a != b)
    {
      double* Rw = S + GT * GT;
      gemm_tile_t<double, double, double, double>(jb, wb, jb, D, 1, n, s.G + (size_t)j0 * n + cb, n, 1, Rw, GT, 1,
                                                  1.0, 0.0, 0, 0);
      Rb = Rw;
      lb = GT;
    }
    __syncthreads();  // R_a and R_b, stored by the whole workgroup, are read by the whole workgroup
    gemm_tile_t<double, double, double, double>(wa, wb, jb, Ra, 1, la, Rb, lb, 1, s.G + (size_t)ca * n + cb, n, 1, -1.0,
                                                1.0, 0, 0);
    if (t == 0 && w > 0)
    {
      __syncthreads();  // the updated diagonal block is visible to the whole workgroup
      potrf_inv_block(s.G, n, s.R, n, s.X, n, min(GT, w), j1, s.info);
    }
  }
  else
  {
    const int ri = GT * (t - s.ntri), hi = min(GT, j0 - ri);
    gemm_tile_t<double, double, double, double>(hi, jb, j0, s.X + (size_t)ri * n, n, 1, s.R + j0, n, 1, S, GT, 1, 1.0,
                                                0.0, 0, 0);
    __syncthreads();  // T is read by the whole workgroup
    gemm_tile_t<double, double, double, double>(hi, jb, jb, S, GT, 1, D, n, 1, s.X + (size_t)ri * n + j0, n, 1, -1.0,
                                                0.0, 0, 0);
  }
}

__global__ __launch_bounds__(256) void chol_step_kernel(chol_step s) { chol_step_run(s, blockIdx.x); }

// The tile (bx, by, bz) this workgroup computes.  With bh = 0 it is its
// block index.  Otherwise the grid (whose size is a multiple of 8) is
// renumbered so that the workgroups the dispatcher deals to one XCD (every
// eighth, in launch order) form one contiguous run of tiles, and that run
// is walked in column strips bh tiles tall: each XCD then covers a compact
// block of bh x (run / bh) tiles of one slice and fetches only the operand
// strips of that block into its own L2.  A bijection whenever gridDim.y is
// a multiple of bh.
__device__ __forceinline__ void xcd_tile(int bh, int& bx, int& by, int& bz)
{
  bx = blockIdx.x;
  by = blockIdx.y;
  bz = blockIdx.z;
  if (bh <= 0) return;
  const int gx = gridDim.x, gy = gridDim.y;
  const int per = gx * gy;
  const int L = bx + gx * (by + gy * bz);
  const int run = per * gridDim.z / 8;
  const int l = (L % 8) * run + L / 8;
  bz = l / per;
  int r = l - bz * per;
  const int grp = r / (bh * gx);
  r -= grp * bh * gx;
  bx = r / bh;
  by = grp * bh + r % bh;
}

// One product of a Gram chain on 32 x 32 tiles: an operand given as the
// sum of split-K partials (na / nb buffers) is summed while loaded; with
// part set, slice z of the K range [z kper, (z+1) kper) writes its partial
// product P_z (row-major M x N, M N apart) instead of C.  Tiles are ordered
// per XCD as xcd_tile describes.
template <int SA, int SB>
__global__ __launch_bounds__(256) void gemm_chain_kernel(gemm_desc g, int na, long psa, int nb, long psb, int kper,
                                                         double* part, int bh)
{
  constexpr int TT = 32;
  int bx, by, bz;
  xcd_tile(bh, bx, by, bz);
  if (!part)
  {
    gemm_tile_t<double, double, double, double, TT, SA, SB, 64>(g.M, g.N, g.K, g.A, g.sar, g.sac, g.B, g.sbr, g.sbc, g.C,
                                                            g.scr, g.scc, g.alpha, g.beta, by * TT, bx * TT, na, psa,
                                                            nb, psb);
    return;
  }
  const int k0 = bz * kper;
  const int kn = min(g.K - k0, kper);
  double* P = part + (size_t)bz * g.M * g.N;
  gemm_tile_t<double, double, double, double, TT, SA, SB, 64>(g.M, g.N, kn, g.A + (size_t)k0 * g.sac, g.sar, g.sac,
                                                          g.B + (size_t)k0 * g.sbr, g.sbr, g.sbc, P, g.N, 1, 1.0, 0.0,
                                                          by * TT, bx * TT, na, psa, nb, psb);
}

// out = the sum of the s partials P_0 .. P_{s-1} (n elements each), in
// order, for each item of a batch (blockIdx.y).
struct chain_sum_item
{
  const double* part;
  double* out;
  int s;
  long n;
};
__global__ __launch_bounds__(256) void chain_sum_kernel(const chain_sum_item* items)
{
  const chain_sum_item it = items[blockIdx.y];
  for (long e = blockIdx.x * 256L + threadIdx.x; e < it.n; e += (long)gridDim.x * 256)
  {
    double v = it.part[e];
    for (int t = 1; t < it.s; ++t) v += it.part[(size_t)t * it.n + e];
    it.out[e] = v;
  }
}

// ---------------------------------------------------------------------------
// Batched pieces of a chained sweep (chain_prob in the header): problem
// blockIdx.y (blockIdx.x for the first diagonal blocks) of a device array.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void chain_potrf_kernel(const chain_prob* ps)
{
  const int p = blockIdx.x;
  double* G = ps[p].G;
  double* R = ps[p].R;
  double* X = ps[p].X;
  double* info = ps[p].info;
  const int n = ps[p].n;
  potrf_inv_block(G, n, R, n, X, n, min(kPotrfB, n), 0, info);
}

__global__ __launch_bounds__(256) void chain_chol_step_kernel(const chain_prob* ps, int j0)
{
  const int p = blockIdx.y;
  const int n = ps[p].n;
  if (j0 >= n) return;
  chol_step s;
  s.n = n;
  s.j0 = j0;
  s.jb = min(kPotrfB, n - j0);
  const int w = n - j0 - s.jb;
  s.nt = (w + GT - 1) / GT;
  s.ntri = s.nt * (s.nt + 1) / 2;
  const int nx = (j0 + GT - 1) / GT;
  if ((int)blockIdx.x >= s.ntri + nx) return;
  s.G = ps[p].G;
  s.R = ps[p].R;
  s.X = ps[p].X;
  s.S = ps[p].S;
  s.info = ps[p].info;
  chol_step_run(s, blockIdx.x);
}

// E <- E - I; block b of problem p writes the largest |E| of rows b, b +
// kChainParts, ... to part[b].
__global__ __launch_bounds__(256) void chain_minus_eye_kernel(const chain_prob* ps)
{
  __shared__ double red[4];
  const int p = blockIdx.y, n = ps[p].n;
  double* E = ps[p].E;
  double m = 0.0;
  for (int i = blockIdx.x; i < n; i += kChainParts)
    for (int j = threadIdx.x; j < n; j += 256)
    {
      const size_t e = (size_t)i * n + j;
      const double v = E[e] - (i == j ? 1.0 : 0.0);
      E[e] = v;
      m = fmax(m, fabs(v) <= DBL_MAX ? fabs(v) : INFINITY);
    }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) ps[p].part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ __launch_bounds__(256) void chain_polar_kernel(const chain_prob* ps)
{
  const int p = blockIdx.y, n = ps[p].n;
  const double* E = ps[p].E;
  const double* V = ps[p].V;
  double* P = ps[p].P;
  double* Pi = ps[p].Pi;
  const int total = n * n;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < total; e += gridDim.x * 256)
  {
    const double id = (e / n == e % n) ? 1.0 : 0.0;
    const double x = E[e], v = V[e];
    P[e] = id - 0.5 * x + 0.375 * v;
    Pi[e] = id + 0.5 * x - 0.125 * v;
  }
}

// Column c < 2n of problem p (four per block, one wave each): Rf's columns,
// then Y's.
__global__ __launch_bounds__(256) void chain_norms_kernel(const chain_prob* ps)
{
  const int p = blockIdx.y, n = ps[p].n;
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (col >= 2 * n) return;
  const double* M = col < n ? ps[p].Rf : ps[p].Y;
  const int c = col < n ? col : col - n;
  double acc = 0.0;
  for (int i = lane; i < n; i += 64) acc = fma(M[(size_t)i * n + c], M[(size_t)i * n + c], acc);
  acc = wave_sum(acc);
  if (lane == 0) ps[p].norms[col] = sqrt(acc);
}

// out[0] = the sum and out[1] = the largest over the problems of n max|E|
// (from the partial maxima, non-finite as infinity), both infinity when
// any problem's Cholesky broke down (info[0] != 0 or a non-finite extreme
// diagonal entry).  The sum bounds the spectral norm of every chained E.
__global__ __launch_bounds__(64) void chain_bound_kernel(const chain_prob* ps, int count, double* out)
{
  const int lane = threadIdx.x;
  double total = 0.0, most = 0.0;
  bool ok = true;
  for (int p = 0; p < count; ++p)
  {
    double emax = ps[p].part[lane];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, 64));
    const double* info = ps[p].info;
    ok = ok && info[0] == 0.0 && isfinite(info[1]) && isfinite(info[2]);
    total += emax * ps[p].n;
    most = fmax(most, emax * ps[p].n);
  }
  if (lane == 0)
  {
    out[0] = ok && total <= DBL_MAX ? total : INFINITY;
    out[1] = ok && most <= DBL_MAX ? most : INFINITY;
  }
}
static_assert(kChainParts == 64, "chain_bound_kernel reads one partial maximum per lane");

// Grouped GEMM: problem blockIdx.z has its own descriptor; tiles beyond its
// M x N exit.  TT: the tile size (64, or 32 for small outputs).
template <int TT>
__global__ __launch_bounds__(256) void gemm_f64_grouped_kernel(const gemm_desc* gs)
{
  const gemm_desc& g = gs[blockIdx.z];
  if (g.M <= 0 || g.N <= 0 || (int)blockIdx.y * TT >= g.M || (int)blockIdx.x * TT >= g.N) return;
  gemm_tile_t<double, double, double, double, TT>(g.M, g.N, g.K, g.A, g.sar, g.sac, g.B, g.sbr, g.sbc, g.C, g.scr,
                                                  g.scc, g.alpha, g.beta, blockIdx.y * TT, blockIdx.x * TT);
}

// Grouped GEMM over a list of the output tiles that exist: workgroup b
// computes tile tiles[b].y (row-major over the problem's tile grid) of
// problem tiles[b].x.
template <int TT>
__global__ __launch_bounds__(256) void gemm_f64_tiles_kernel(const gemm_desc* gs, const int2* tiles)
{
  const int2 t = tiles[blockIdx.x];
  const gemm_desc& g = gs[t.x];
  const int tn = (g.N + TT - 1) / TT;
  const int ty = t.y / tn, tx = t.y - ty * tn;
  gemm_tile_t<double, double, double, double, TT>(g.M, g.N, g.K, g.A, g.sar, g.sac, g.B, g.sbr, g.sbc, g.C, g.scr,
                                                  g.scc, g.alpha, g.beta, ty * TT, tx * TT);
}

// Grouped GEMM with per-operand precision (xgemm_desc.types): A, B and C
// each FP64 or FP32, multiplied in FP32 on the matrix cores when
// types.compute is FP32, in FP64 otherwise.
template <class TA, class TB, class TC, class CT>
__global__ __launch_bounds__(256) void xgemm_grouped_kernel(const xgemm_desc* gs)
{
  const xgemm_desc& g = gs[blockIdx.z];
  if (g.M <= 0 || g.N <= 0 || (int)blockIdx.y * GT >= g.M || (int)blockIdx.x * GT >= g.N) return;
  gemm_tile_t<TA, TB, TC, CT>(g.M, g.N, g.K, static_cast<const TA*>(g.A), g.sar, g.sac, static_cast<const TB*>(g.B),
                              g.sbr, g.sbc, static_cast<TC*>(g.C), g.scr, g.scc, 1.0, 0.0, blockIdx.y * GT,
                              blockIdx.x * GT);
}

// The same over a list of the output tiles that exist (as
// gemm_f64_tiles_kernel).
template <class TA, class TB, class TC, class CT>
__global__ __launch_bounds__(256) void xgemm_tiles_kernel(const xgemm_desc* gs, const int2* tiles)
{
  const int2 w = tiles[blockIdx.x];
  const xgemm_desc& g = gs[w.x];
  const int tn = (g.N + GT - 1) / GT;
  const int ty = w.y / tn, tx = w.y - ty * tn;
  gemm_tile_t<TA, TB, TC, CT>(g.M, g.N, g.K, static_cast<const TA*>(g.A), g.sar, g.sac, static_cast<const TB*>(g.B),
                              g.sbr, g.sbc, static_cast<TC*>(g.C), g.scr, g.scc, 1.0, 0.0, ty * GT, tx * GT);
}

// Householder QR (+ optional thin Q) of one column-major matrix per block.
template <int NT>
__global__ __launch_bounds__(NT) void qr_kernel(qr_desc q)
{
  extern __shared__ double lds[];
  double* red = lds;
  double* vbuf = lds + NT / kWave;
  const size_t b = blockIdx.x;
  double* A = q.A + b * q.bsA;
  double* tau = q.tau + b * q.bstau;
  geqr2<NT>(A, q.lda, q.m, q.n, tau, vbuf, red);
  if (q.Q)
  {
    const int k = q.m < q.n ? q.m : q.n;
    org2r<NT>(A, q.lda, q.m, k, tau, q.Q + b * q.bsQ, q.ldq, vbuf);
  }
}

// One-sided Jacobi + sorted column norms (+ optional truncation rank).
template <int NT>
__global__ __launch_bounds__(NT) void jacobi_kernel(jacobi_desc jd)
{
  __shared__ int flag;
  extern __shared__ double nrm[];
  const size_t b = blockIdx.x;
  double* X = jd.X + b * jd.bsX;
  double* J = jd.J + b * jd.bsJ;
  const int sweeps = jacobi_sweeps<NT>(X, jd.ldx, jd.p, jd.q, J, jd.ldj, jd.max_sweeps, &flag);
  __syncthreads();
  column_norms_sorted<NT>(X, jd.ldx, jd.p, jd.q, nrm, jd.s + b * jd.bss, jd.perm + b * jd.bss);
  if (threadIdx.x == 0)
  {
    if (jd.sweeps) jd.sweeps[b] = sweeps;
    if (jd.rank_out)
    {
      const double delta = jd.delta_ptr ? *jd.delta_ptr : jd.delta;
      jd.rank_out[b] = truncate_rank(jd.s + b * jd.bss, jd.q, delta, jd.max_rank);
    }
  }
}

// dst[i*drs + c*dcs] = (i <= c) ? A[c*lda + i] : 0   for i < kr, c < nc.
__global__ void extract_upper_kernel(const double* A, int lda, int kr, int nc, double* dst, long drs, long dcs,
                                     int full)
{
  const long total = (long)kr * nc;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
  {
    const int c = (int)(t / kr), i = (int)(t % kr);
    dst[i * drs + c * dcs] = (full || i <= c) ? A[(size_t)c * lda + i] : 0.0;
  }
}

// P = I - E/2 + 3/8 V, Pi = I + E/2 - 1/8 V (n x n row-major).
__global__ void polar_series_kernel(const double* E, const double* V, double* P, double* Pi, int n)
{
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const double id = e / n == e % n ? 1.0 : 0.0;
    P[e] = id - 0.5 * E[e] + 0.375 * V[e];
    Pi[e] = id + 0.5 * E[e] - 0.125 * V[e];
  }
}

// dst[t*drs + o*dcs] = src[t + perm[o]*lds] for t < rows, o < ncols.
// Column o of the result is column perm[o] of src, times scale[o]
// (mode GATHER_MUL) or divided by it (GATHER_DIV: unit vector e_o when
// scale[o] == 0, the normalisation of a zero column).
__device__ __forceinline__ double gathered(const double* src, int lds_, int t, int o, const int* perm,
                                           const double* scale, int mode)
{
  const double v = (mode & GATHER_ROWMAJOR) ? src[(size_t)t * lds_ + perm[o]] : src[t + (size_t)perm[o] * lds_];
  mode &= ~GATHER_ROWMAJOR;
  if (mode == GATHER_MUL) return v * scale[o];
  if (mode == GATHER_DIV) return scale[o] != 0.0 ? v / scale[o] : (t == o ? 1.0 : 0.0);
  return v;
}

__global__ void gather_cols_kernel(const double* src, int lds_, int rows, const int* perm, int ncols, double* dst,
                                   long drs, long dcs, const double* scale, int mode)
{
  const long total = (long)rows * ncols;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total; x += (long)gridDim.x * blockDim.x)
  {
    const int o = (int)(x / rows), t = (int)(x % rows);
    dst[t * drs + o * dcs] = gathered(src, lds_, t, o, perm, scale, mode);
  }
}

// 2-D transpose through a 32x33 LDS tile: out[j*ldo + i] = in[i*ldi + j]
// for i < rows, j < cols.
__global__ __launch_bounds__(256) void transpose_kernel(const double* in, int ldi, int rows, int cols, double* out,
                                                        int ldo)
{
  __shared__ double tile[32][33];
  const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8)
  {
    const int i = bi + r, j = bj + tx;
    if (i < rows && j < cols) tile[r][tx] = in[(size_t)i * ldi + j];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8)
  {
    const int j = bj + r, i = bi + tx;
    if (i < rows && j < cols) out[(size_t)j * ldo + i] = tile[tx][r];
  }
}

// Generic strided gather: dst is contiguous over up to 6 dims (sizes in
// `n`, row-major); src offset = sum idx_d * ss[d].
__global__ void permute_kernel(permute_desc p)
{
  long total = 1;
  for (int d = 0; d < 6; ++d) total *= p.n[d];
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total; x += (long)gridDim.x * blockDim.x)
  {
    long rem = x, off = 0;
    for (int d = 5; d >= 0; --d)
    {
      const long i = rem % p.n[d];
      rem /= p.n[d];
      off += i * p.ss[d];
    }
    p.dst[x] = p.src[off];
  }
}

// y(a*s+ax, i, b*sp+bx) = sum_j A(a,i,j,b) x(ax,j,bx), the MPO core
// contraction, one thread per column of outputs: thread t = ((a m + i) s +
// ax) sp + bx keeps x(ax, :, bx) in registers (up to kMvN of them; larger
// mode sizes read x again per output) and writes the rp outputs
// y(a, ax, i, :, bx); lanes with consecutive bx write adjacent doubles.
// When the 64 lanes of a wave share (a, i) (s sp a multiple of 64 in the
// interior of a core) and A(a, i, :, :) has at most 64 entries, the wave
// loads it once, one entry per lane, and each product takes its factor by
// readlane, so the vector memory pipeline carries only x and the stores.
// 32-bit indices: the launcher checks that the core has fewer than 2^31
// entries.  Launch: one thread per column, cols = r s m sp.
constexpr int kMvN = 8;
__device__ __forceinline__ double read_lane(double v, int k)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, k), hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void matvec_core_cols(const double* __restrict__ A, const double* __restrict__ x,
                                                 double* __restrict__ y, unsigned r, unsigned m, unsigned n,
                                                 unsigned rp, unsigned s, unsigned sp, unsigned t)
{
  const unsigned cols = r * s * m * sp, gs = s * sp;
  const unsigned lane = threadIdx.x & 63, t0 = t - lane;
  if (t0 >= cols) return;  // the whole wave is past the core
  const unsigned tl = min(t0 + 63, cols - 1), tt = min(t, cols - 1);
  const unsigned g = tt / gs, rem = tt - g * gs, ax = rem / sp, bx = rem - ax * sp;
  const unsigned a = g / m, i = g - a * m;
  const double* xr = x + (size_t)ax * n * sp + bx;
  double* yr = y + ((size_t)(a * s + ax) * m + i) * rp * sp + bx;
  const bool live = t < cols;
  if (n > (unsigned)kMvN)
  {
    const double* Ar = A + (size_t)g * n * rp;
    for (unsigned bb = 0; bb < rp; ++bb)
    {
      double acc = 0.0;
      for (unsigned j = 0; j < n; ++j) acc = fma(Ar[(size_t)j * rp + bb], xr[(size_t)j * sp], acc);
      if (live) yr[(size_t)bb * sp] = acc;
    }
    return;
  }
  double xv[kMvN];
#pragma unroll
  for (int j = 0; j < kMvN; ++j) xv[j] = j < (int)n ? xr[(size_t)j * sp] : 0.0;
  if (t0 / gs == tl / gs && n * rp <= 64)
  {
    const double* Ag = A + (size_t)(t0 / gs) * n * rp;
    const double mine = lane < n * rp ? Ag[lane] : 0.0;
    for (unsigned bb = 0; bb < rp; ++bb)
    {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < kMvN; ++j)
        if (j < (int)n) acc = fma(read_lane(mine, j * rp + bb), xv[j], acc);
      if (live) yr[(size_t)bb * sp] = acc;
    }
  }
  else
  {
    const double* Ar = A + (size_t)g * n * rp;
    for (unsigned bb = 0; bb < rp; ++bb)
    {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < kMvN; ++j)
        if (j < (int)n) acc = fma(Ar[(size_t)j * rp + bb], xv[j], acc);
      if (live) yr[(size_t)bb * sp] = acc;
    }
  }
}

__global__ __launch_bounds__(256) void matvec_core_kernel(const double* __restrict__ A, const double* __restrict__ x,
                                                          double* __restrict__ y, int r, int m, int n, int rp, int s,
                                                          int sp)
{
  matvec_core_cols(A, x, y, r, m, n, rp, s, sp, blockIdx.x * 256u + threadIdx.x);
}

// C(a*s+ax, i, j, b*sp+bx) = sum_q A(a,i,q,b) B(ax,q,j,bx).
__global__ void matmat_core_kernel(const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C,
                                   int r, int m, int p, int rp, int s, int n, int sp)
{
  const long total = (long)r * s * m * n * rp * sp;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x)
  {
    long t = o;
    const int bx = (int)(t % sp);
    t /= sp;
    const int b = (int)(t % rp);
    t /= rp;
    const int j = (int)(t % n);
    t /= n;
    const int i = (int)(t % m);
    t /= m;
    const int ax = (int)(t % s);
    const int a = (int)(t / s);
    const double* Ar = A + (((size_t)a * m + i) * p) * rp + b;
    const double* Br = B + (((size_t)ax * p) * n + j) * sp + bx;
    double acc = 0.0;
    for (int q = 0; q < p; ++q) acc = fma(Ar[(size_t)q * rp], Br[(size_t)q * n * sp], acc);
    C[o] = acc;
  }
}

// Sum of squares, two-level: per-block partials into `part`.
__global__ __launch_bounds__(256) void sumsq_kernel(const double* x, long n, double* part)
{
  __shared__ double red[4];
  double acc = 0.0;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256L) acc = fma(x[t], x[t], acc);
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void finish_sumsq_kernel(const double* part, int np, double* out, double eps,
                                                           double inv_sqrt_dm1, double* delta_out)
{
  __shared__ double red[4];
  double acc = 0.0;
  for (int t = threadIdx.x; t < np; t += 256) acc += part[t];
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0)
  {
    const double nrm = sqrt(acc);
    out[0] = nrm;
    if (delta_out) delta_out[0] = eps * nrm * inv_sqrt_dm1;
  }
}

// Norms of columns [c0, c0 + kCnCols) of a row-major (rows x cols) matrix by
// a workgroup of 256 threads: cw = min(kCnCols, cols rounded up to a power of
// two) adjacent columns (one 128-byte line) of 256 / cw rows at a time, eight
// such loads in flight per thread, so every entry is read once and the
// workgroup keeps enough requests outstanding to cover memory latency; each
// column's partial sums (one per thread that read it) are then added in a
// fixed order.
constexpr int kCnCols = 16;
__device__ __forceinline__ void col_norms_tile(const double* __restrict__ M, unsigned rows, int cols,
                                               double* __restrict__ out, int c0)
{
  __shared__ double red[256];
  int cw = 1;
  while (cw < kCnCols && cw < cols) cw <<= 1;
  const unsigned per = 256 / cw;
  const int tid = threadIdx.x;
  const int c = c0 + (tid % cw);
  double acc = 0.0;
  if (c < cols)
  {
    unsigned i = tid / cw;
    for (; i + 7 * per < rows; i += 8 * per)
    {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = M[(size_t)(i + k * per) * cols + c];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc = fma(v[k], v[k], acc);
    }
    for (; i < rows; i += per)
    {
      const double v = M[(size_t)i * cols + c];
      acc = fma(v, v, acc);
    }
  }
  red[tid] = acc;
  __syncthreads();
  if (tid < cw && c < cols)
  {
    double sum = 0.0;
    for (unsigned k = 0; k < per; ++k) sum += red[k * cw + tid];
    out[c] = sqrt(sum);
  }
}

// Balancing of one core x (r, n, rp) by a grid-stride loop over groups of 4
// consecutive entries: x(a, i, b) *= (g / nr(b)) / (g / nl(a)), a factor 1
// where a norm is zero or its array null.  One 32-bit decomposition of each
// group's first index (cores have fewer than 2^31 entries), carries for the
// rest, one vector load and store for whole groups of a 32-byte aligned core.
__device__ __forceinline__ void balance_core(double* __restrict__ x, unsigned r, unsigned n, unsigned rp,
                                             const double* __restrict__ nl, const double* __restrict__ nr, double g,
                                             unsigned first, unsigned stride)
{
  const unsigned total = r * n * rp;
  const bool aligned = (reinterpret_cast<uintptr_t>(x) & 31) == 0;
  for (unsigned o = first * 4; o < total; o += stride * 4)
  {
    unsigned row = o / rp, bb = o - row * rp, a = row / n;
    double v[4];
    const bool whole = aligned && o + 4 <= total;
    if (whole)
    {
      const dbl4 w = *reinterpret_cast<const dbl4*>(x + o);
      v[0] = w[0], v[1] = w[1], v[2] = w[2], v[3] = w[3];
    }
    else
    {
      for (int k = 0; k < 4; ++k) v[k] = o + k < total ? x[o + k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
      double f = 1.0;
      if (nr && nr[bb] > 0.0) f *= g / nr[bb];
      if (nl && nl[a] > 0.0) f /= g / nl[a];
      v[k] *= f;
      if (++bb == rp)
      {
        bb = 0;
        a = ++row / n;
      }
    }
    if (whole)
    {
      const dbl4 w = {v[0], v[1], v[2], v[3]};
      *reinterpret_cast<dbl4*>(x + o) = w;
    }
    else
    {
      for (int k = 0; k < 4 && o + k < total; ++k) x[o + k] = v[k];
    }
  }
}

// Column norms of a row-major (rows x cols) matrix (col_norms_tile).
__global__ __launch_bounds__(256) void col_norms_kernel(const double* M, int rows, int cols, double* out)
{
  col_norms_tile(M, rows, cols, out, blockIdx.x * kCnCols);
}

// Geometric mean of the nonzero norms (single block): out = exp(mean(log)).
__global__ __launch_bounds__(256) void geo_mean_kernel(const double* v, int n, double* out)
{
  __shared__ double red[4];
  double s = 0.0, c = 0.0;
  for (int t = threadIdx.x; t < n; t += 256)
    if (v[t] > 0.0)
    {
      s += log(v[t]);
      c += 1.0;
    }
  s = block_sum<256>(s, red);
  c = block_sum<256>(c, red);
  if (threadIdx.x == 0) out[0] = c > 0.0 ? exp(s / c) : 0.0;
}

// Balancing of one core (r, n, rp): x(a,i,b) *= f_right(b) / f_left(a), with
// f(c) = geo / norm(c) for norm(c) > 0 and 1 otherwise; a null norm array
// means all factors are 1.
__global__ __launch_bounds__(256) void balance_kernel(double* x, int r, int n, int rp, const double* nl,
                                                      const double* nr, const double* geo)
{
  const double g = *geo;
  if (g == 0.0) return;
  balance_core(x, r, n, rp, nl, nr, g, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

// Per-core launches folded into one (descriptors by value, blockIdx.y picks
// the core); each core's work is that of the single-core kernel above.
__global__ __launch_bounds__(256) void matvec_core_batch_kernel(matvec_batch b)
{
  const matvec_batch::item& it = b.it[blockIdx.y];
  matvec_core_cols(it.A, it.x, it.y, it.r, it.m, it.n, it.rp, it.s, it.sp, blockIdx.x * 256u + threadIdx.x);
}

// Column norms of each core's left unfolding (r n x rp, row-major) into
// nrm + off (col_norms_tile; workgroup (x, core) takes columns kCnCols x ..).
__global__ __launch_bounds__(256) void core_norms_batch_kernel(core_batch b, double* nrm)
{
  const core_batch::item& it = b.it[blockIdx.y];
  if (it.off < 0 || (int)blockIdx.x * kCnCols >= it.rp) return;
  col_norms_tile(it.x, (unsigned)it.r * it.n, it.rp, nrm + it.off, blockIdx.x * kCnCols);
}

// balance_kernel for each core: left norms at nrm + offl, right ones at
// nrm + off (negative: none).
__global__ __launch_bounds__(256) void balance_batch_kernel(core_batch b, const double* nrm, const double* geo)
{
  const core_batch::item& it = b.it[blockIdx.y];
  const double g = *geo;
  if (g == 0.0) return;
  balance_core(it.x, it.r, it.n, it.rp, it.offl >= 0 ? nrm + it.offl : nullptr, it.off >= 0 ? nrm + it.off : nullptr,
               g, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

// dst[i*ldd + j] = f * src[i*lds + j] + keep * dst[i*ldd + j].
__global__ void copy2d_kernel(const double* src, int lds_, double* dst, int ldd, int rows, int cols, double f,
                              double keep)
{
  const long total = (long)rows * cols;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
  {
    const int i = (int)(t / cols), j = (int)(t % cols);
    double* o = dst + (size_t)i * ldd + j;
    const double v = f * src[(size_t)i * lds_ + j];
    *o = keep == 0.0 ? v : v + keep * (*o);
  }
}

__global__ void scale_kernel(double* x, long n, double alpha)
{
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) x[t] *= alpha;
}

// Self-check of the DPP wave reductions against the ds_bpermute one: each
// wave sums 64 values of x (wave_sum), and 16 vectors of 64 values taken
// from x at strides of 64 (wave_vsum16); mismatches (beyond 8 ulp of the
// sum of magnitudes) are counted.
__global__ __launch_bounds__(256) void wave_sum_check_kernel(const double* x, int n, int* bad)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const double v = i < n ? x[i] : 0.0;
  const double a = wave_sum(v), b = wave_sum_shfl(v);
  double mag = wave_sum(fabs(v));
  if ((threadIdx.x & 63) == 0 && fabs(a - b) > 8.0 * 2.220446049250313e-16 * mag) atomicAdd(bad, 1);
  if (fabs(a - __shfl(a, 0)) != 0.0) atomicAdd(bad, 1000);
  double v16[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) v16[c] = i < n ? x[(i + 64 * c) % n] : 0.0;
  const double s16 = wave_vsum16(v16);
  double ref = 0.0, m16 = 0.0;
#pragma unroll
  for (int c = 0; c < 16; ++c)
  {
    const double rc = wave_sum_shfl(v16[c]), mc = wave_sum_shfl(fabs(v16[c]));
    if ((threadIdx.x & 15) == c)
    {
      ref = rc;
      m16 = mc;
    }
  }
  if (fabs(s16 - ref) > 8.0 * 2.220446049250313e-16 * m16) atomicAdd(bad, 1);
}

// ---- descriptor-batched helpers (problem = blockIdx.y / blockIdx.z) ----
__global__ __launch_bounds__(256) void transpose_batched_kernel(const tr_desc* ds)
{
  const tr_desc& d = ds[blockIdx.z];
  __shared__ double tile[32][33];
  const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;
  if (bi >= d.rows || bj >= d.cols) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8)
  {
    const int i = bi + r, j = bj + tx;
    if (i < d.rows && j < d.cols) tile[r][tx] = d.in[(size_t)i * d.ldi + j];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8)
  {
    const int j = bj + r, i = bi + tx;
    if (i < d.rows && j < d.cols) d.out[(size_t)j * d.ldo + i] = tile[tx][r];
  }
}

__global__ void extract_upper_batched_kernel(const ex_desc* ds)
{
  const ex_desc& d = ds[blockIdx.y];
  const long total = (long)d.kr * d.nc;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
  {
    const int c = (int)(t / d.kr), i = (int)(t % d.kr);
    d.dst[i * d.drs + c * d.dcs] = (i <= c) ? d.A[(size_t)c * d.lda + i] : 0.0;
  }
}

__global__ void gather_cols_batched_kernel(const ga_desc* ds)
{
  const ga_desc& d = ds[blockIdx.y];
  const long total = (long)d.rows * d.ncols;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total; x += (long)gridDim.x * blockDim.x)
  {
    const int o = (int)(x / d.rows), t = (int)(x % d.rows);
    d.dst[t * d.drs + o * d.dcs] = gathered(d.src, d.lds, t, o, d.perm, d.scale, d.mode);
  }
}

__global__ void copy_batched_kernel(const cp_desc* ds)
{
  const cp_desc& d = ds[blockIdx.y];
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < d.n; x += (long)gridDim.x * blockDim.x)
    d.dst[x] = d.src[x];
}

// One workgroup per problem: out_norm = ||x||_F, out_delta = scale * norm.
__global__ __launch_bounds__(256) void frob_batched_kernel(const sq_desc* ds)
{
  const sq_desc& d = ds[blockIdx.x];
  __shared__ double red[4];
  double acc = 0.0;
  for (long t = threadIdx.x; t < d.n; t += 256) acc = fma(d.x[t], d.x[t], acc);
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0)
  {
    const double nrm = sqrt(acc);
    d.out_norm[0] = nrm;
    if (d.out_delta) d.out_delta[0] = d.scale * nrm;
  }
}

// Column norms of a row-major (rows x cols) core per problem (col_norms_tile).
__global__ __launch_bounds__(256) void col_norms_batched_kernel(const cn_desc* ds)
{
  const cn_desc& d = ds[blockIdx.y];
  if ((int)blockIdx.x * kCnCols >= d.cols) return;
  col_norms_tile(d.M, d.rows, d.cols, d.out, blockIdx.x * kCnCols);
}

// Geometric mean of the nonzero entries of v[0..n) per problem.
__global__ __launch_bounds__(256) void geo_mean_batched_kernel(const gm_desc* ds)
{
  const gm_desc& d = ds[blockIdx.x];
  __shared__ double red[4];
  double s = 0.0, c = 0.0;
  for (int t = threadIdx.x; t < d.n; t += 256)
    if (d.v[t] > 0.0)
    {
      s += log(d.v[t]);
      c += 1.0;
    }
  s = block_sum<256>(s, red);
  c = block_sum<256>(c, red);
  if (threadIdx.x == 0) d.out[0] = c > 0.0 ? exp(s / c) : 0.0;
}

__global__ __launch_bounds__(256) void balance_batched_kernel(const bal_desc* ds)
{
  const bal_desc& d = ds[blockIdx.y];
  const double g = *d.geo;
  if (g == 0.0) return;
  balance_core(d.x, d.r, d.n, d.rp, d.nl, d.nr, g, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

int grid_for(long n, int bs = 256)
{
  long g = (n + bs - 1) / bs;
  if (g > 65535L * 4) g = 65535L * 4;
  return (int)(g < 1 ? 1 : g);
}

// Certificate for the rank rule: workgroup J forms columns [16J, 16J + 16)
// of X = R^-1 for an upper-triangular R (row-major, q x q) by blocked back
// substitution, row blocks I = J down to 0.  The off-diagonal sums
// B = E - R(I, I+1:) X(I+1:, J) run on all 256 threads from LDS (R staged in
// 128-column chunks, X(:, J) resident); the 16 x 16 diagonal solve runs on
// one thread per column.  out[2J] = ||X(:, J)||_F^2, out[2J+1] =
// ||R(:, J)||_F^2; a zero or non-finite pivot propagates as inf / NaN.
// With X non-null the columns of R^-1 are also written (row-major, leading
// dimension ldx, zeros below the diagonal).
constexpr int kCertB = 16, kCertChunk = 128;

__global__ __launch_bounds__(256) void rinv_cert_kernel(const double* __restrict__ R, int q, double* __restrict__ out,
                                                        double* __restrict__ X, int ldx)
{
  extern __shared__ double lds[];
  const int J = blockIdx.x;
  const int j0 = J * kCertB;
  const int jn = min(kCertB, q - j0);
  const int rows = j0 + jn;
  double* Xs = lds;                              // rows x 16
  double* Rs = Xs + (size_t)rows * kCertB;       // 16 x kCertChunk
  double* Bs = Rs + kCertB * kCertChunk;         // 16 x 16
  double* Rii = Bs + kCertB * kCertB;            // 16 x 16 diagonal block
  double* red = Rii + kCertB * kCertB;           // 4 waves
  const int t = threadIdx.x, r = t / kCertB, cc = t % kCertB;
  for (int e = t; e < rows * kCertB; e += 256) Xs[e] = 0.0;
  __syncthreads();
  for (int I = J; I >= 0; --I)
  {
    const int i0 = I * kCertB;
    const int in = min(kCertB, q - i0);
    const int i1 = i0 + in;
    double acc = (r < in && cc < jn && i0 + r == j0 + cc) ? 1.0 : 0.0;
    for (int k0 = i1; k0 < rows; k0 += kCertChunk)
    {
      const int kw = min(kCertChunk, rows - k0);
      for (int e = t; e < kCertB * kCertChunk; e += 256)
      {
        const int rr = e / kCertChunk, kk = e % kCertChunk;
        Rs[e] = (rr < in && kk < kw) ? R[(size_t)(i0 + rr) * q + k0 + kk] : 0.0;
      }
      __syncthreads();
      if (r < in && cc < jn)
        for (int kk = 0; kk < kw; ++kk) acc -= Rs[r * kCertChunk + kk] * Xs[(size_t)(k0 + kk) * kCertB + cc];
      __syncthreads();
    }
    Bs[r * kCertB + cc] = acc;
    {
      const double rv = (r < in && cc < in) ? R[(size_t)(i0 + r) * q + i0 + cc] : 0.0;
      // The diagonal holds reciprocals: the substitution below multiplies.
      Rii[r * kCertB + cc] = (r == cc && r < in) ? 1.0 / rv : rv;
    }
    __syncthreads();
    if (t < jn)
    {
      // Back substitution for column t with the column in registers.
      double xc[kCertB];
#pragma unroll
      for (int rr = kCertB - 1; rr >= 0; --rr)
      {
        double v = 0.0;
        if (rr < in)
        {
          v = Bs[rr * kCertB + t];
#pragma unroll
          for (int k = rr + 1; k < kCertB; ++k)
            if (k < in) v = fma(-Rii[rr * kCertB + k], xc[k], v);
          v *= Rii[rr * kCertB + rr];
        }
        xc[rr] = v;
      }
#pragma unroll
      for (int rr = 0; rr < kCertB; ++rr)
        if (rr < in) Xs[(size_t)(i0 + rr) * kCertB + t] = xc[rr];
    }
    __syncthreads();
  }
  double sx = 0.0, sr = 0.0;
  for (int e = t; e < rows * kCertB; e += 256)
  {
    const int row = e / kCertB, col = e % kCertB;
    if (col < jn)
    {
      const double x = Xs[e];
      sx += x * x;
      if (row <= j0 + col)
      {
        const double v = R[(size_t)row * q + j0 + col];
        sr += v * v;
      }
    }
  }
  if (X)
    for (int e = t; e < q * kCertB; e += 256)
    {
      const int row = e / kCertB, col = e % kCertB;
      if (col < jn) X[(size_t)row * ldx + j0 + col] = row < rows ? Xs[e] : 0.0;
    }
  sx = block_sum<256>(sx, red);
  sr = block_sum<256>(sr, red);
  if (t == 0)
  {
    out[2 * J] = sx;
    out[2 * J + 1] = sr;
  }
}

// G <- G - I (n x n, row-major); block b writes the largest |G - I| of its
// share to part[b] (non-finite entries counted as infinity).
__global__ __launch_bounds__(256) void minus_ident_kernel(double* G, int n, double* part)
{
  __shared__ double red[4];
  double m = 0.0;
  for (int i = blockIdx.x; i < n; i += gridDim.x)
    for (int j = threadIdx.x; j < n; j += 256)
    {
      const size_t e = (size_t)i * n + j;
      const double v = G[e] - (i == j ? 1.0 : 0.0);
      G[e] = v;
      m = fmax(m, fabs(v) <= DBL_MAX ? fabs(v) : INFINITY);
    }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Device-side decisions of a speculative sweep: flag[0] starts at 1 and is
// cleared by the first factorization or certificate that does not hold.
// CholQR2 acceptance (here in block 0's first wave, or in one extra
// workgroup of gemm_f64_multi_kernel): no breakdown in pass 1 (info[0] == 0,
// finite extreme diagonal entries) and n max|E| <= thr over the np partial
// maxima.  Blocks 1 .. of this kernel compute the column norms of the n x n
// row-major A and B into out[0, n) and out[n, 2n) (col_norms_kernel's work,
// four columns per block).
__global__ __launch_bounds__(256) void spec_cholqr_norms_kernel(const double* info, const double* part, int np, int n,
                                                               double thr, double* flag, const double* A,
                                                               const double* B, double* out)
{
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0)
  {
    if (threadIdx.x >= 64) return;
    double emax = 0.0;
    for (int b = lane; b < np; b += 64) emax = fmax(emax, part[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, 64));
    if (lane != 0) return;
    const bool ok = info[0] == 0.0 && isfinite(info[1]) && isfinite(info[2]) && emax * n <= thr;
    if (!ok) flag[0] = 0.0;
    return;
  }
  const int col = (blockIdx.x - 1) * 4 + (threadIdx.x >> 6);
  if (col >= 2 * n) return;
  const double* M = col < n ? A : B;
  const int c = col < n ? col : col - n;
  double acc = 0.0;
  for (int i = lane; i < n; i += 64) acc = fma(M[(size_t)i * n + c], M[(size_t)i * n + c], acc);
  acc = wave_sum(acc);
  if (lane == 0) out[col] = sqrt(acc);
}

// Full-rank certificate from ||R||_F^2 and ||R^-1||_F^2: mode 0 reads the
// column norms of R (a[0..n)) and of R^-1 (a[n..2n)), mode 1 the n (xf2,
// rf2) pairs of rinv_cert_kernel.  Same test as the host's: both sums finite
// and positive, ||R|| ||R^-1|| <= 1e8, 1/(||R^-1|| (1+1e-3)) > delta (1+1e-3).
__device__ __forceinline__ void spec_cert_decide(const double* a, int n, int mode, double delta, double* flag)
{
  const int lane = threadIdx.x;
  double x = 0.0, r = 0.0;
  for (int i = lane; i < n; i += 64)
  {
    if (mode == 0)
    {
      r = fma(a[i], a[i], r);
      x = fma(a[n + i], a[n + i], x);
    }
    else
    {
      x += a[2 * i];
      r += a[2 * i + 1];
    }
  }
  x = wave_sum(x);
  r = wave_sum(r);
  if (lane != 0) return;
  bool held = isfinite(x) && isfinite(r) && x > 0.0 && r > 0.0;
  if (held)
  {
    const double xf = sqrt(x), rf = sqrt(r);
    held = rf * xf <= 1.0e8 && 1.0 / (xf * (1.0 + 1.0e-3)) > delta * (1.0 + 1.0e-3);
  }
  if (!held) flag[0] = 0.0;
}

// The certificates of a speculative sweep, decided together at its end: block
// z decides item z (any failure clears the flag).
__global__ __launch_bounds__(64) void spec_cert_batch_kernel(cert_batch b, double* flag)
{
  const cert_batch::item& it = b.it[blockIdx.x];
  spec_cert_decide(it.a, it.n, it.mode, it.delta, flag);
}

__global__ void set_value_kernel(double* p, double v)
{
  if (threadIdx.x == 0) p[0] = v;
}

// X = Phi(E - W) (n x n, row-major; W may be null): the strict upper part of
// E - W, half its diagonal, zero below.
__global__ void phi_kernel(const double* E, const double* W, double* X, int n)
{
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int i = (int)(e / n), j = (int)(e % n);
    const double v = E[e] - (W ? W[e] : 0.0);
    X[e] = i < j ? v : (i == j ? 0.5 * v : 0.0);
  }
}

// Y = alpha I + beta X (n x n, row-major).
__global__ void ident_axpy_kernel(const double* X, double* Y, int n, double alpha, double beta)
{
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int i = (int)(e / n), j = (int)(e % n);
    Y[e] = (i == j ? alpha : 0.0) + beta * X[e];
  }
}

// dst (n x n, leading dimension ld) = identity.
__global__ void eye_kernel(double* dst, int n, long ld)
{
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int i = (int)(e / n), j = (int)(e % n);
    dst[(size_t)i * ld + j] = i == j ? 1.0 : 0.0;
  }
}

}  // namespace

// ------------------------------- launchers --------------------------------
void launch_gemm(const gemm_desc& g, int batch, hipStream_t st)
{
  if (g.M <= 0 || g.N <= 0 || batch <= 0) return;
  dim3 grid((g.N + GT - 1) / GT, (g.M + GT - 1) / GT, batch);
  hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, st, g);
}

// Slices for a single GEMM whose 64 x 64 tiles alone would leave most of
// the 256 compute units idle: enough to reach about 256 workgroups, each
// slice at least 64 deep in K, and at most 8 (beyond that the partials'
// extra traffic costs as much as the shorter slices save); 1 when the tiles
// already fill the device.
int gemm_splitk_slices(int M, int N, int K)
{
  const long tiles = (long)((M + GT - 1) / GT) * ((N + GT - 1) / GT);
  if (tiles >= 128 || K < 128) return 1;
  long s = std::min<long>(256 / tiles, K / 64);
  s = std::min<long>(s, 8);
  return s < 2 ? 1 : (int)s;
}

void chol_step_shape(int n, int j0, int& jb, int& nt, int& ntri, int& nx)
{
  jb = std::min(kPotrfB, n - j0);
  const int w = n - j0 - jb;
  nt = (w + GT - 1) / GT;
  ntri = nt * (nt + 1) / 2;
  nx = (j0 + GT - 1) / GT;
}

size_t chol_step_scratch(int n)
{
  size_t most = 0;
  for (int j0 = 0; j0 < n; j0 += kPotrfB)
  {
    int jb, nt, ntri, nx;
    chol_step_shape(n, j0, jb, nt, ntri, nx);
    most = std::max(most, (size_t)(ntri + nx) * 2 * GT * GT);
  }
  return most;
}

void launch_chol_step(const chol_step& s, hipStream_t st)
{
  const int wgs = s.ntri + (s.j0 + GT - 1) / GT;
  if (wgs <= 0) return;
  hipLaunchKernelGGL(chol_step_kernel, dim3(wgs), dim3(256), 0, st, s);
}

bool g_chain_xcd = true;

int xcd_strip_height(int gx, int gy, int gz, double wa, double wb)
{
  const long per = (long)gx * gy, total = per * gz;
  if (total < 16 || total % 8 != 0) return 0;
  const long run = total / 8;
  if (run >= per) return gy;  // each XCD holds whole slices
  // A run of bh x bw tiles: the A strips (weight wa per tile row) are
  // fetched by gx / bw XCDs each, the B strips (wb per tile column) by
  // gy / bh; the cheapest shape that tiles the grid exactly wins.
  int best = 0;
  double best_cost = 0.0;
  for (int bh = 1; bh <= gy; ++bh)
  {
    if (gy % bh != 0 || run % bh != 0) continue;
    const long bw = run / bh;
    if (bw > gx || gx % bw != 0) continue;
    const double cost = wa * (double)(gx / bw) + wb * (double)(gy / bh);
    if (!best || cost < best_cost)
    {
      best = bh;
      best_cost = cost;
    }
  }
  return best;
}

int chain_gemm_slices(int M, int N, int K)
{
  const long tiles = (long)((M + 31) / 32) * ((N + 31) / 32);
  if (tiles >= 128 || K < 128) return 1;
  long s = std::min<long>(256 / tiles, K / 64);
  s = std::min<long>(s, 8);
  return s < 2 ? 1 : (int)s;
}

int launch_gemm_chain(const gemm_desc& g, int na, long psa, int nb, long psb, int slices, double* part,
                      hipStream_t st)
{
  if (g.M <= 0 || g.N <= 0) return 0;
  int kper = g.K, s_eff = 1;
  if (part)
  {
    kper = ((g.K + slices - 1) / slices + 63) / 64 * 64;
    s_eff = (g.K + kper - 1) / kper;
  }
  dim3 grid((g.N + 31) / 32, (g.M + 31) / 32, part ? s_eff : 1);
  if (na > 8 || nb > 8 || (na > 1 && nb > 1)) throw error(TTA_ERR_ARG, "launch_gemm_chain: summed operands");
  const int bh = g_chain_xcd ? xcd_strip_height((int)grid.x, (int)grid.y, (int)grid.z, (double)g.M * na,
                                                (double)g.N * nb)
                             : 0;
  // Summand capacity rounded up to 2, 4 or 8 (unused ones load nothing).
  auto cap = [](int x) { return x <= 1 ? 1 : (x <= 2 ? 2 : (x <= 4 ? 4 : 8)); };
  const int ca = cap(na), cb = cap(nb);
#define TTA_CHAIN_LAUNCH(A_, B_) \
  hipLaunchKernelGGL((gemm_chain_kernel<A_, B_>), grid, dim3(256), 0, st, g, na, psa, nb, psb, kper, part, bh)
  if (ca == 2)
    TTA_CHAIN_LAUNCH(2, 1);
  else if (ca == 4)
    TTA_CHAIN_LAUNCH(4, 1);
  else if (ca == 8)
    TTA_CHAIN_LAUNCH(8, 1);
  else if (cb == 2)
    TTA_CHAIN_LAUNCH(1, 2);
  else if (cb == 4)
    TTA_CHAIN_LAUNCH(1, 4);
  else if (cb == 8)
    TTA_CHAIN_LAUNCH(1, 8);
  else
    TTA_CHAIN_LAUNCH(1, 1);
#undef TTA_CHAIN_LAUNCH
  return part ? s_eff : 0;
}

void launch_chain_sum(const void* items_dev, int count, long max_n, hipStream_t st)
{
  if (count <= 0 || max_n <= 0) return;
  const int blocks = (int)std::min<long>((max_n + 255) / 256, 512);
  hipLaunchKernelGGL(chain_sum_kernel, dim3(blocks, count), dim3(256), 0, st,
                     static_cast<const chain_sum_item*>(items_dev));
}

void launch_chain_potrf(const chain_prob* ps, int count, hipStream_t st)
{
  if (count <= 0) return;
  hipLaunchKernelGGL(chain_potrf_kernel, dim3(count), dim3(256), 0, st, ps);
}

void launch_chain_chol_step(const chain_prob* ps, int count, int max_n, int j0, hipStream_t st)
{
  if (count <= 0 || j0 >= max_n) return;
  // The most workgroups any problem's step at j0 uses: the largest n has the
  // most trailing tiles, and every problem with n > j0 has the same X tiles.
  int jb, nt, ntri, nx;
  chol_step_shape(max_n, j0, jb, nt, ntri, nx);
  const int wgs = ntri + nx;
  if (wgs <= 0) return;
  hipLaunchKernelGGL(chain_chol_step_kernel, dim3(wgs, count), dim3(256), 0, st, ps, j0);
}

void launch_chain_minus_eye(const chain_prob* ps, int count, int max_n, hipStream_t st)
{
  if (count <= 0 || max_n <= 0) return;
  hipLaunchKernelGGL(chain_minus_eye_kernel, dim3(kChainParts, count), dim3(256), 0, st, ps);
}

void launch_chain_polar(const chain_prob* ps, int count, int max_n, hipStream_t st)
{
  if (count <= 0 || max_n <= 0) return;
  const int blocks = std::min((max_n * max_n + 255) / 256, 256);
  hipLaunchKernelGGL(chain_polar_kernel, dim3(blocks, count), dim3(256), 0, st, ps);
}

void launch_chain_norms(const chain_prob* ps, int count, int max_n, hipStream_t st)
{
  if (count <= 0 || max_n <= 0) return;
  hipLaunchKernelGGL(chain_norms_kernel, dim3((2 * max_n + 3) / 4, count), dim3(256), 0, st, ps);
}

void launch_chain_bound(const chain_prob* ps, int count, double* out, hipStream_t st)
{
  if (count <= 0) return;
  hipLaunchKernelGGL(chain_bound_kernel, dim3(1), dim3(64), 0, st, ps, count, out);
}


void launch_gemm_multi(const gemm_multi& m, hipStream_t st)
{
  int mt = 0, nt = 0;
  for (int i = 0; i < m.count; ++i)
  {
    mt = std::max(mt, (m.g[i].M + GT - 1) / GT);
    nt = std::max(nt, (m.g[i].N + GT - 1) / GT);
  }
  if (!m.dflag && (m.count <= 0 || mt <= 0 || nt <= 0)) return;
  hipLaunchKernelGGL(gemm_f64_multi_kernel, dim3(std::max(nt, 1), std::max(mt, 1), m.count + (m.dflag ? 1 : 0)),
                     dim3(256), 0, st, m);
}

int launch_gemm_splitk_parts(const gemm_desc& g, int S, double* part, hipStream_t st)
{
  if (g.M <= 0 || g.N <= 0) return 0;
  // K per slice: a multiple of the 32-deep slab; trailing empty slices are
  // dropped.
  const int kper = ((g.K + S - 1) / S + GK - 1) / GK * GK;
  const int s_eff = (g.K + kper - 1) / kper;
  dim3 grid((g.N + GT - 1) / GT, (g.M + GT - 1) / GT, s_eff);
  hipLaunchKernelGGL(gemm_f64_splitk_kernel, grid, dim3(256), 0, st, g, kper, part);
  return s_eff;
}

void launch_splitk_reduce(const gemm_desc& g, int s_eff, const double* part, hipStream_t st)
{
  if (g.M <= 0 || g.N <= 0 || s_eff <= 0) return;
  const long mn = (long)g.M * g.N;
  const int rb = (int)std::min<long>((mn + 255) / 256, 2048);
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(rb), dim3(256), 0, st, g, s_eff, part);
}

void launch_splitk_reduce_epi(const gemm_desc& g, int s_eff, const double* part, const splitk_epi& ep,
                              hipStream_t st)
{
  if (g.M <= 0 || g.N <= 0 || s_eff <= 0) return;
  const long mn = (long)g.M * g.N;
  const int rb = ep.mode == SPLITK_EPI_GRAM ? minus_ident_parts(g.N) : (int)std::min<long>((mn + 255) / 256, 2048);
  hipLaunchKernelGGL(splitk_reduce_epi_kernel, dim3(rb), dim3(256), 0, st, g, s_eff, part, ep);
}

void launch_qr(const qr_desc& q, int batch, hipStream_t st)
{
  constexpr int NT = 256;
  const size_t lds = sizeof(double) * (NT / kWave + (size_t)q.m);
  hipLaunchKernelGGL(qr_kernel<NT>, dim3(batch), dim3(NT), lds, st, q);
}

void launch_jacobi(const jacobi_desc& j, int batch, hipStream_t st)
{
  constexpr int NT = 1024;
  const size_t lds = sizeof(double) * (size_t)j.q;
  hipLaunchKernelGGL(jacobi_kernel<NT>, dim3(batch), dim3(NT), lds, st, j);
}

void launch_extract_upper(const double* A, int lda, int kr, int nc, double* dst, long drs, long dcs, hipStream_t st,
                          bool full)
{
  hipLaunchKernelGGL(extract_upper_kernel, dim3(grid_for((long)kr * nc)), dim3(256), 0, st, A, lda, kr, nc, dst, drs,
                     dcs, full ? 1 : 0);
}

void launch_polar_series(const double* E, const double* V, double* P, double* Pi, int n, hipStream_t st)
{
  hipLaunchKernelGGL(polar_series_kernel, dim3(grid_for((long)n * n)), dim3(256), 0, st, E, V, P, Pi, n);
}

__global__ void scale_copy_kernel(const double* src, int rows, int cols, long ld, const double* rowf, const double* colf,
                                  double* dst)
{
  const long total = (long)rows * cols;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total; x += (long)gridDim.x * blockDim.x)
  {
    const int i = (int)(x / cols), j = (int)(x % cols);
    double v = src[(size_t)i * ld + j];
    if (rowf) v *= rowf[i];
    if (colf) v *= colf[j];
    dst[x] = v;
  }
}

void launch_scale_copy(const double* src, int rows, int cols, long ld, const double* rowf, const double* colf,
                       double* dst, hipStream_t st)
{
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(scale_copy_kernel, dim3(grid_for((long)rows * cols)), dim3(256), 0, st, src, rows, cols, ld, rowf,
                     colf, dst);
}

void launch_gather_cols(const double* src, int lds_, int rows, const int* perm, int ncols, double* dst, long drs,
                        long dcs, hipStream_t st, const double* scale, int mode)
{
  hipLaunchKernelGGL(gather_cols_kernel, dim3(grid_for((long)rows * ncols)), dim3(256), 0, st, src, lds_, rows, perm,
                     ncols, dst, drs, dcs, scale, mode);
}

int rinv_cert_blocks(int q) { return (q + kCertB - 1) / kCertB; }

void launch_rinv_cert(const double* R, int q, double* out, hipStream_t st, double* X, int ldx)
{
  const size_t lds = sizeof(double) * ((size_t)q * kCertB + kCertB * kCertChunk + 2 * kCertB * kCertB + 4);
  hipLaunchKernelGGL(rinv_cert_kernel, dim3(rinv_cert_blocks(q)), dim3(256), lds, st, R, q, out, X, ldx);
}

int minus_ident_parts(int n) { return std::max(1, std::min(n, kMinusIdentParts)); }

void launch_minus_ident(double* G, int n, double* part, hipStream_t st)
{
  hipLaunchKernelGGL(minus_ident_kernel, dim3(minus_ident_parts(n)), dim3(256), 0, st, G, n, part);
}

void launch_phi(const double* E, const double* W, double* X, int n, hipStream_t st)
{
  long g = ((long)n * n + 255) / 256;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(phi_kernel, dim3((int)(g < 1 ? 1 : g)), dim3(256), 0, st, E, W, X, n);
}

void launch_ident_axpy(const double* X, double* Y, int n, double alpha, double beta, hipStream_t st)
{
  long g = ((long)n * n + 255) / 256;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(ident_axpy_kernel, dim3((int)(g < 1 ? 1 : g)), dim3(256), 0, st, X, Y, n, alpha, beta);
}

void launch_spec_cholqr_norms(const double* info, const double* part, int np, int n, double thr, double* flag,
                              const double* A, const double* B, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(spec_cholqr_norms_kernel, dim3(1 + (2 * n + 3) / 4), dim3(256), 0, st, info, part, np, n, thr,
                     flag, A, B, out);
}

void launch_spec_cert_batch(const cert_batch& b, double* flag, hipStream_t st)
{
  if (b.count <= 0) return;
  hipLaunchKernelGGL(spec_cert_batch_kernel, dim3(b.count), dim3(64), 0, st, b, flag);
}

// Copy z of the batch: workgroups of blockIdx.y = z stride over its n
// doubles, two per thread per step.
__global__ __launch_bounds__(256) void copy_batch_kernel(copy_batch b)
{
  const copy_batch::item& it = b.it[blockIdx.y];
  const long n = it.n;
  for (long e = 2 * (blockIdx.x * 256L + threadIdx.x); e < n; e += 2L * gridDim.x * 256)
  {
    it.dst[e] = it.src[e];
    if (e + 1 < n) it.dst[e + 1] = it.src[e + 1];
  }
}

void launch_copy_batch(const copy_batch& b, hipStream_t st)
{
  long most = 0;
  for (int z = 0; z < b.count; ++z) most = std::max(most, b.it[z].n);
  if (b.count <= 0 || most <= 0) return;
  const int gx = (int)std::min<long>((most + 511) / 512, 256);
  hipLaunchKernelGGL(copy_batch_kernel, dim3(gx, b.count), dim3(256), 0, st, b);
}

void launch_set_value(double* p, double v, hipStream_t st)
{
  hipLaunchKernelGGL(set_value_kernel, dim3(1), dim3(64), 0, st, p, v);
}

void launch_eye(double* dst, int n, long ld, hipStream_t st)
{
  const long total = (long)n * n;
  long g = (total + 255) / 256;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(eye_kernel, dim3((int)(g < 1 ? 1 : g)), dim3(256), 0, st, dst, n, ld);
}

void launch_transpose(const double* in, int ldi, int rows, int cols, double* out, int ldo, hipStream_t st)
{
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, st, in, ldi, rows, cols, out, ldo);
}

void launch_permute(const permute_desc& p, hipStream_t st)
{
  long total = 1;
  for (int d = 0; d < 6; ++d) total *= p.n[d];
  hipLaunchKernelGGL(permute_kernel, dim3(grid_for(total)), dim3(256), 0, st, p);
}

void launch_matvec_core(const double* A, const double* x, double* y, int r, int m, int n, int rp, int s, int sp,
                        hipStream_t st)
{
  const long total = (long)r * s * m * rp * sp;
  if (total <= 0) return;
  if (total >= (1L << 31)) fail(TTA_ERR_UNSUPPORTED, "matvec core of %ld entries (at most 2^31 - 1)", total);
  const long cols = total / rp;
  hipLaunchKernelGGL(matvec_core_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, st, A, x, y, r, m, n, rp,
                     s, sp);
}

void launch_matmat_core(const double* A, const double* B, double* C, int r, int m, int p, int rp, int s, int n,
                        int sp, hipStream_t st)
{
  const long total = (long)r * s * m * n * rp * sp;
  hipLaunchKernelGGL(matmat_core_kernel, dim3(grid_for(total)), dim3(256), 0, st, A, B, C, r, m, p, rp, s, n, sp);
}

void launch_frob_norm(const double* x, long n, double* part, int nparts, double* out, double eps, double inv_sqrt_dm1,
                      double* delta_out, hipStream_t st)
{
  int g = (int)((n + 255) / 256);
  if (g > nparts) g = nparts;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(sumsq_kernel, dim3(g), dim3(256), 0, st, x, n, part);
  hipLaunchKernelGGL(finish_sumsq_kernel, dim3(1), dim3(256), 0, st, part, g, out, eps, inv_sqrt_dm1, delta_out);
}

void launch_col_norms(const double* M, int rows, int cols, double* out, hipStream_t st)
{
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(col_norms_kernel, dim3((cols + kCnCols - 1) / kCnCols), dim3(256), 0, st, M, rows, cols, out);
}

void launch_geo_mean(const double* v, int n, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(geo_mean_kernel, dim3(1), dim3(256), 0, st, v, n, out);
}

void launch_matvec_core_batch(const matvec_batch& b, hipStream_t st)
{
  long most = 0, most_cols = 0;
  for (int z = 0; z < b.count; ++z)
  {
    const matvec_batch::item& it = b.it[z];
    most = std::max(most, (long)it.r * it.s * it.m * it.rp * it.sp);
    most_cols = std::max(most_cols, (long)it.r * it.s * it.m * it.sp);
  }
  if (b.count <= 0 || most <= 0) return;
  if (most >= (1L << 31)) fail(TTA_ERR_UNSUPPORTED, "matvec core of %ld entries (at most 2^31 - 1)", most);
  hipLaunchKernelGGL(matvec_core_batch_kernel, dim3((unsigned)((most_cols + 255) / 256), b.count), dim3(256), 0, st,
                     b);
}

void launch_core_norms_batch(const core_batch& b, double* nrm, hipStream_t st)
{
  int most = 0;
  for (int z = 0; z < b.count; ++z) most = std::max(most, b.it[z].rp);
  if (b.count <= 0 || most <= 0) return;
  hipLaunchKernelGGL(core_norms_batch_kernel, dim3((most + kCnCols - 1) / kCnCols, b.count), dim3(256), 0, st, b,
                     nrm);
}

void launch_balance_batch(const core_batch& b, const double* nrm, const double* geo, hipStream_t st)
{
  long most = 0;
  for (int z = 0; z < b.count; ++z) most = std::max(most, (long)b.it[z].r * b.it[z].n * b.it[z].rp);
  if (b.count <= 0 || most <= 0) return;
  if (most >= (1L << 31)) fail(TTA_ERR_UNSUPPORTED, "core of %ld entries (at most 2^31 - 1)", most);
  hipLaunchKernelGGL(balance_batch_kernel, dim3(grid_for((most + 3) / 4), b.count), dim3(256), 0, st, b, nrm, geo);
}

void launch_balance(double* x, int r, int n, int rp, const double* nl, const double* nr, const double* geo,
                    hipStream_t st)
{
  const long total = (long)r * n * rp;
  if (total <= 0) return;
  if (total >= (1L << 31)) fail(TTA_ERR_UNSUPPORTED, "core of %ld entries (at most 2^31 - 1)", total);
  hipLaunchKernelGGL(balance_kernel, dim3(grid_for((total + 3) / 4)), dim3(256), 0, st, x, r, n, rp, nl, nr, geo);
}

void launch_copy2d(const double* src, int lds_, double* dst, int ldd, int rows, int cols, double f, double keep,
                   hipStream_t st)
{
  hipLaunchKernelGGL(copy2d_kernel, dim3(grid_for((long)rows * cols)), dim3(256), 0, st, src, lds_, dst, ldd, rows,
                     cols, f, keep);
}

void launch_wave_sum_check(const double* x, int n, int* bad, hipStream_t st)
{
  hipLaunchKernelGGL(wave_sum_check_kernel, dim3((n + 255) / 256), dim3(256), 0, st, x, n, bad);
}

void launch_scale(double* x, long n, double alpha, hipStream_t st)
{
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, n, alpha);
}


void launch_xgemm_grouped(const xgemm_desc* descs_dev, int count, int max_m, int max_n, xgemm_types t,
                          hipStream_t st)
{
  if (count <= 0 || max_m <= 0 || max_n <= 0) return;
  dim3 grid((max_n + GT - 1) / GT, (max_m + GT - 1) / GT, count);
  if (t == XG_D_S_S_S)
    hipLaunchKernelGGL((xgemm_grouped_kernel<double, float, float, float>), grid, dim3(256), 0, st, descs_dev);
  else if (t == XG_S_S_S_S)
    hipLaunchKernelGGL((xgemm_grouped_kernel<float, float, float, float>), grid, dim3(256), 0, st, descs_dev);
  else if (t == XG_D_S_D_D)
    hipLaunchKernelGGL((xgemm_grouped_kernel<double, float, double, double>), grid, dim3(256), 0, st, descs_dev);
  else
    hipLaunchKernelGGL((xgemm_grouped_kernel<double, double, double, double>), grid, dim3(256), 0, st, descs_dev);
}

namespace
{
// Output tiles (problem, tile index) of the problems of v that have any, for
// a T x T tile, and their count.
template <class D>
long count_tiles(const std::vector<D>& v, int T)
{
  long used = 0;
  for (const D& g : v)
    if (g.M > 0 && g.N > 0) used += (long)((g.M + T - 1) / T) * ((g.N + T - 1) / T);
  return used;
}
template <class D>
std::vector<int2> tile_list(const std::vector<D>& v, int T, long used)
{
  std::vector<int2> tiles;
  tiles.reserve(used);
  for (int z = 0; z < (int)v.size(); ++z)
  {
    const D& g = v[z];
    if (g.M <= 0 || g.N <= 0) continue;
    const int nt = ((g.M + T - 1) / T) * ((g.N + T - 1) / T);
    for (int k = 0; k < nt; ++k) tiles.push_back(make_int2(z, k));
  }
  return tiles;
}
// Whether a grid over the largest shape (per tiles per problem) would hold
// few idle workgroups, or cannot be replaced by a one-dimensional grid.
bool full_grid_ok(long used, long per, int count)
{
  return (4 * used > 3 * per * count && count <= 65535) || used > 0x7fffffffL;
}
}  // namespace

void launch_xgemm_grouped(arena& S, const std::vector<xgemm_desc>& v, xgemm_types t, hipStream_t st)
{
  const int count = (int)v.size();
  int max_m = 0, max_n = 0;
  for (const xgemm_desc& g : v)
  {
    max_m = std::max(max_m, g.M);
    max_n = std::max(max_n, g.N);
  }
  if (count <= 0 || max_m <= 0 || max_n <= 0) return;
  const xgemm_desc* dv = S.upload(v);
  const long used = count_tiles(v, GT);
  if (used == 0) return;
  if (full_grid_ok(used, (long)((max_m + GT - 1) / GT) * ((max_n + GT - 1) / GT), count))
  {
    launch_xgemm_grouped(dv, count, max_m, max_n, t, st);
    return;
  }
  const int2* td = S.upload(tile_list(v, GT, used));
  const dim3 grid((unsigned)used);
  if (t == XG_D_S_S_S)
    hipLaunchKernelGGL((xgemm_tiles_kernel<double, float, float, float>), grid, dim3(256), 0, st, dv, td);
  else if (t == XG_S_S_S_S)
    hipLaunchKernelGGL((xgemm_tiles_kernel<float, float, float, float>), grid, dim3(256), 0, st, dv, td);
  else if (t == XG_D_S_D_D)
    hipLaunchKernelGGL((xgemm_tiles_kernel<double, float, double, double>), grid, dim3(256), 0, st, dv, td);
  else
    hipLaunchKernelGGL((xgemm_tiles_kernel<double, double, double, double>), grid, dim3(256), 0, st, dv, td);
}

// ---------------------------------------------------------------------------
// Autotuner of the grouped GEMM's tile size.  The 64 x 64 and 32 x 32 tiles
// give identical results (see gemm_tile_t), so the choice only moves time:
// small outputs waste most of a 64 x 64 tile, large ones lose operand reuse
// with 32 x 32.  Launches are keyed by (device, log2 of max M, of max N, of
// the problem count); the first launches of a key alternate the two tiles,
// each between two events, until each has two timings (read without
// waiting, at later launches of the key); the key then keeps the 32 tile
// when its best time per unit of tile work (count x max M x max N) is at
// least 10 % below the 64 tile's.
// The table is per process and guarded by a mutex.
// ---------------------------------------------------------------------------
namespace
{
struct tune_entry
{
  int trials[2] = {0, 0};
  double ms_per_work[2] = {0.0, 0.0};
  bool pending[2] = {false, false};
  double pending_work[2] = {0.0, 0.0};
  hipEvent_t ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int choice = -1;
};
std::mutex g_tune_mu;
std::map<std::array<int, 4>, tune_entry> g_tune;
bool g_tune_on = true;
constexpr int kTuneTrials = 2;  // timed launches per tile; the faster minimum wins

int log2_bucket(long v)
{
  int b = 0;
  while ((1L << b) < v) ++b;
  return b;
}
}  // namespace

void set_gemm_autotune(bool on)
{
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_on = on;
}

void gemm_autotune_stats(long long* keys, long long* decided, long long* chose_small)
{
  std::lock_guard<std::mutex> lk(g_tune_mu);
  long long k = 0, d = 0, sm = 0;
  for (const auto& kv : g_tune)
  {
    ++k;
    if (kv.second.choice >= 0) ++d;
    if (kv.second.choice == 1) ++sm;
  }
  if (keys) *keys = k;
  if (decided) *decided = d;
  if (chose_small) *chose_small = sm;
}

namespace
{
// Tile choice of a grouped FP64 GEMM (64, or 32 for small outputs), timed
// per shape bucket until each tile has kTuneTrials timings.
struct gemm_pick
{
  int variant = 0;
  bool timed = false;
  tune_entry* e = nullptr;
};

gemm_pick pick_gemm_tile(int count, int max_m, int max_n, hipStream_t st)
{
  int variant = 0;
  bool timed = false;
  tune_entry* e = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune_on)
    {
      int dev = 0;
      (void)hipGetDevice(&dev);
      e = &g_tune[{dev, log2_bucket(max_m), log2_bucket(max_n), log2_bucket(count)}];
      for (int v = 0; v < 2; ++v)
        if (e->pending[v] && hipEventQuery(e->ev[v][1]) == hipSuccess)
        {
          float ms = 0.0f;
          if (hipEventElapsedTime(&ms, e->ev[v][0], e->ev[v][1]) == hipSuccess)
          {
            const double t = ms / e->pending_work[v];
            e->ms_per_work[v] = e->trials[v] == 0 ? t : std::min(e->ms_per_work[v], t);
            e->trials[v] += 1;
          }
          e->pending[v] = false;
        }
      if (e->choice < 0 && e->trials[0] >= kTuneTrials && e->trials[1] >= kTuneTrials)
        e->choice = e->ms_per_work[1] < 0.9 * e->ms_per_work[0] ? 1 : 0;
      if (e->choice >= 0)
        variant = e->choice;
      else
      {
        // Alternate the two tiles until each has kTuneTrials timings.
        const int n0 = e->trials[0] + (e->pending[0] ? 1 : 0), n1 = e->trials[1] + (e->pending[1] ? 1 : 0);
        const int want = n0 <= n1 ? 0 : 1;
        variant = want;
        timed = !e->pending[want] && (want == 0 ? n0 : n1) < kTuneTrials;
      }
      if (timed)
      {
        for (int t = 0; t < 2; ++t)
          if (!e->ev[variant][t] && hipEventCreate(&e->ev[variant][t]) != hipSuccess) timed = false;
        if (timed)
        {
          e->pending[variant] = true;
          e->pending_work[variant] = (double)count * max_m * max_n;
          (void)hipEventRecord(e->ev[variant][0], st);
        }
      }
    }
  }
  return gemm_pick{variant, timed, e};
}

void end_gemm_pick(const gemm_pick& t, hipStream_t st)
{
  if (!t.timed) return;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  (void)hipEventRecord(t.e->ev[t.variant][1], st);
}
}  // namespace

void launch_gemm_grouped(const gemm_desc* descs_dev, int count, int max_m, int max_n, hipStream_t st)
{
  if (count <= 0 || max_m <= 0 || max_n <= 0) return;
  const gemm_pick t = pick_gemm_tile(count, max_m, max_n, st);
  if (t.variant == 1)
  {
    dim3 grid((max_n + 31) / 32, (max_m + 31) / 32, count);
    hipLaunchKernelGGL(gemm_f64_grouped_kernel<32>, grid, dim3(256), 0, st, descs_dev);
  }
  else
  {
    dim3 grid((max_n + GT - 1) / GT, (max_m + GT - 1) / GT, count);
    hipLaunchKernelGGL(gemm_f64_grouped_kernel<GT>, grid, dim3(256), 0, st, descs_dev);
  }
  end_gemm_pick(t, st);
}

void launch_gemm_grouped(arena& S, const std::vector<gemm_desc>& v, hipStream_t st)
{
  const int count = (int)v.size();
  int max_m = 0, max_n = 0;
  for (const gemm_desc& g : v)
  {
    max_m = std::max(max_m, g.M);
    max_n = std::max(max_n, g.N);
  }
  if (count <= 0 || max_m <= 0 || max_n <= 0) return;
  const gemm_desc* dv = S.upload(v);
  const gemm_pick t = pick_gemm_tile(count, max_m, max_n, st);
  const int T = t.variant == 1 ? 32 : GT;
  const long used = count_tiles(v, T);
  if (used == 0)
  {
    end_gemm_pick(t, st);
    return;
  }
  if (full_grid_ok(used, (long)((max_m + T - 1) / T) * ((max_n + T - 1) / T), count))
  {
    // Shapes close to the largest: the grid over it wastes little.
    dim3 grid((max_n + T - 1) / T, (max_m + T - 1) / T, count);
    if (T == 32)
      hipLaunchKernelGGL(gemm_f64_grouped_kernel<32>, grid, dim3(256), 0, st, dv);
    else
      hipLaunchKernelGGL(gemm_f64_grouped_kernel<GT>, grid, dim3(256), 0, st, dv);
  }
  else
  {
    // One workgroup per output tile that exists.
    const int2* td = S.upload(tile_list(v, T, used));
    if (T == 32)
      hipLaunchKernelGGL(gemm_f64_tiles_kernel<32>, dim3((unsigned)used), dim3(256), 0, st, dv, td);
    else
      hipLaunchKernelGGL(gemm_f64_tiles_kernel<GT>, dim3((unsigned)used), dim3(256), 0, st, dv, td);
  }
  end_gemm_pick(t, st);
}

void launch_transpose_batched(const tr_desc* ds, int count, int max_rows, int max_cols, hipStream_t st)
{
  if (count <= 0) return;
  dim3 grid((max_cols + 31) / 32, (max_rows + 31) / 32, count);
  hipLaunchKernelGGL(transpose_batched_kernel, grid, dim3(256), 0, st, ds);
}

void launch_extract_upper_batched(const ex_desc* ds, int count, long max_elems, hipStream_t st)
{
  if (count <= 0) return;
  int g = (int)std::min<long>((max_elems + 255) / 256, 256);
  hipLaunchKernelGGL(extract_upper_batched_kernel, dim3(g < 1 ? 1 : g, count), dim3(256), 0, st, ds);
}

void launch_gather_cols_batched(const ga_desc* ds, int count, long max_elems, hipStream_t st)
{
  if (count <= 0) return;
  int g = (int)std::min<long>((max_elems + 255) / 256, 256);
  hipLaunchKernelGGL(gather_cols_batched_kernel, dim3(g < 1 ? 1 : g, count), dim3(256), 0, st, ds);
}

void launch_copy_batched(const cp_desc* ds, int count, long max_elems, hipStream_t st)
{
  if (count <= 0) return;
  int g = (int)std::min<long>((max_elems + 255) / 256, 512);
  hipLaunchKernelGGL(copy_batched_kernel, dim3(g < 1 ? 1 : g, count), dim3(256), 0, st, ds);
}

void launch_frob_batched(const sq_desc* ds, int count, hipStream_t st)
{
  if (count <= 0) return;
  hipLaunchKernelGGL(frob_batched_kernel, dim3(count), dim3(256), 0, st, ds);
}

void launch_col_norms_batched(const cn_desc* ds, int count, int max_cols, hipStream_t st)
{
  if (count <= 0 || max_cols <= 0) return;
  hipLaunchKernelGGL(col_norms_batched_kernel, dim3((max_cols + kCnCols - 1) / kCnCols, count), dim3(256), 0, st,
                     ds);
}

void launch_geo_mean_batched(const gm_desc* ds, int count, hipStream_t st)
{
  if (count <= 0) return;
  hipLaunchKernelGGL(geo_mean_batched_kernel, dim3(count), dim3(256), 0, st, ds);
}

void launch_balance_batched(const bal_desc* ds, int count, long max_elems, hipStream_t st)
{
  if (count <= 0) return;
  if (max_elems >= (1L << 31)) fail(TTA_ERR_UNSUPPORTED, "core of %ld entries (at most 2^31 - 1)", max_elems);
  int g = (int)std::min<long>((max_elems + 1023) / 1024, 256);
  hipLaunchKernelGGL(balance_batched_kernel, dim3(g < 1 ? 1 : g, count), dim3(256), 0, st, ds);
}

}  // namespace tta
