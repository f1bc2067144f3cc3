// Diagonal-block step of the blocked Cholesky with inverse, as a device
// function for the workgroups that run it.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

namespace tta
{
namespace dev
{

// ---------------------------------------------------------------------------
// Diagonal block of the blocked Cholesky (chol_inv_blocked in the C ABI
// layer): the jb x jb block (jb <= 64) at (j0, j0) of the row-major G is
// factored, G_b = R_b^T R_b, R_b is written into R and D_b = R_b^-1 into X
// at the same place, and rows j0 .. j0+jb-1 of R and X are zeroed left of
// the block.
// Gaussian elimination with symmetric pivots on [G_b | I] (padded with the
// identity where jb < 64): pivot row j at its turn, scaled by 1/sqrt(pivot),
// is row j of [R_b | R_b^-T], and the rows below take M(k, :) -= M(j, k)
// M(j, :) / pivot.  Column c of G_b is needed only before pivot c and
// column c of the identity only from pivot c on (before it, it is e_c), so
// both share one storage column: at pivot c it turns from the one into the
// other (in-place Gauss-Jordan), and the 64 x 64 array holds, in row j at
// pivot j, R_b's row right of the diagonal and R_b^-T's row left of it.
// 256 threads hold it in 4 x 4 sub-blocks (rows tr + 16a, columns tc + 16b).
// At pivot j the owning threads store the unscaled row into the LDS history
// H[j]; after one barrier every thread reads its 4 row and 4 column factors
// from there (all reads issued at once; rows at or above the pivot take a
// zero factor rather than a branch) and updates.  1/pivot comes from the
// hardware reciprocal and two Newton steps; the square roots, the scaling of
// the rows, the diagonal statistics and all global stores happen after the
// last pivot.
// info[0] = 1 on a non-positive or non-finite pivot, info[1] / info[2] =
// smallest / largest diagonal entry of R; the launch with j0 = 0 starts
// them, later ones fold into them.
// ---------------------------------------------------------------------------
constexpr int kPotrfB = 64;

// The block step by the calling workgroup of 256 threads (a kernel of its
// own, or the workgroup of a GEMM launch that has just written the block).
__device__ __forceinline__ void potrf_inv_block(const double* G, int ldg, double* R, int ldr, double* X, int ldx,
                                                int jb, int j0, double* info)
{
  __shared__ double H[kPotrfB][kPotrfB + 1];  // unscaled pivot rows
  __shared__ double sqv[kPotrfB], rsv[kPotrfB];
  __shared__ int s_bad;
  __shared__ double s_min[4], s_max[4];
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const double* Gb = G + (size_t)j0 * ldg + j0;
  double g[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
    {
      const int row = tr + 16 * a, col = tc + 16 * b;
      g[a][b] = (row < jb && col < jb) ? Gb[(size_t)row * ldg + col] : (row == col ? 1.0 : 0.0);
    }
#pragma unroll
  for (int a = 0; a < 4; ++a)
  {
    for (int t = 0; t < 16; ++t)
    {
      const int j = 16 * a + t;
      if (j >= jb) break;
      double* h = H[j];
      if (tr == t)
      {
#pragma unroll
        for (int b = 0; b < 4; ++b) h[tc + 16 * b] = g[a][b];
      }
      __syncthreads();
      const double piv = h[j];
      double cf[4], rfv[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) cf[b] = h[tc + 16 * b];
#pragma unroll
      for (int a2 = a; a2 < 4; ++a2) rfv[a2] = h[tr + 16 * a2];
      double inv = __builtin_amdgcn_rcp(piv);
      inv = fma(inv, fma(-piv, inv, 1.0), inv);
      inv = fma(inv, fma(-piv, inv, 1.0), inv);
      // Column j turns into column j of the identity part: value 1 in the
      // pivot row, 0 below it before this pivot.
      bool turn[4];
#pragma unroll
      for (int b = 0; b < 4; ++b)
      {
        turn[b] = tc + 16 * b == j;
        if (turn[b]) cf[b] = 1.0;
      }
#pragma unroll
      for (int a2 = a; a2 < 4; ++a2)
      {
        double rf = rfv[a2] * inv;
        if (a2 == a) rf = tr > t ? rf : 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) g[a2][b] = fma(-rf, cf[b], turn[b] ? 0.0 : g[a2][b]);
      }
    }
  }
  __syncthreads();
  if (tid < jb)
  {
    const double piv = H[tid][tid];
    const double sq = sqrt(piv);
    sqv[tid] = sq;
    rsv[tid] = 1.0 / sq;
  }
  if (tid < 64)
  {
    // Pivot statistics over the 64 lanes of wave 0.
    const bool ok = tid >= jb || (H[tid][tid] > 0.0 && H[tid][tid] < INFINITY);
    double mn = INFINITY, mx = 0.0;
    if (tid < jb)
    {
      mn = sqrt(H[tid][tid]);
      mx = mn;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
      mn = fmin(mn, __shfl_xor(mn, o, 64));
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    const bool all_ok = __all(ok);
    if (tid == 0)
    {
      s_bad = all_ok ? 0 : 1;
      s_min[0] = mn;
      s_max[0] = mx;
    }
  }
  __syncthreads();
  if (tid == 0)
  {
    if (j0 == 0)
    {
      info[0] = s_bad ? 1.0 : 0.0;
      info[1] = s_min[0];
      info[2] = s_max[0];
    }
    else
    {
      info[0] = (s_bad || info[0] != 0.0) ? 1.0 : 0.0;
      info[1] = fmin(info[1], s_min[0]);
      info[2] = fmax(info[2], s_max[0]);
    }
  }
  // R_b(r, c) = H[r][c] / sqrt(pivot r) right of the diagonal; X_b(r, c) =
  // (R_b^-T)(c, r) = H[c][r] / sqrt(pivot c) left of the diagonal of row c,
  // 1 / sqrt(pivot c) on it.
  double* Rb = R + (size_t)j0 * ldr + j0;
  double* Xb = X + (size_t)j0 * ldx + j0;
  for (int e = tid; e < jb * kPotrfB; e += 256)
  {
    const int r = e / kPotrfB, c = e % kPotrfB;
    if (c < jb)
    {
      Rb[(size_t)r * ldr + c] = c > r ? H[r][c] * rsv[r] : (c == r ? sqv[r] : 0.0);
      Xb[(size_t)r * ldx + c] = r < c ? H[c][r] * rsv[c] : (r == c ? rsv[c] : 0.0);
    }
  }
  for (int e = tid; e < jb * j0; e += 256)
  {
    const int t = e / j0, col = e % j0;
    R[(size_t)(j0 + t) * ldr + col] = 0.0;
    X[(size_t)(j0 + t) * ldx + col] = 0.0;
  }
}

}  // namespace dev
}  // namespace tta
