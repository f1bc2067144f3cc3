// Batched Cholesky with diagonal pivoting and a trace stop, for the
// rank-revealing CholQR2 step of the mixed batched rounding.
//
// One workgroup per problem.  The symmetric q x q matrix G (row-major) is
// held as a packed lower triangle in LDS (q <= kPcholMaxQ, about 147 KiB at
// q = 192) and factored right-looking, one pivot per step:
//   * pivoting on: the pivot is the largest remaining diagonal entry of the
//     Schur complement S (ties to the smaller index), and the factorization
//     stops before step j when trace(S) <= stop_f * delta^2 or the largest
//     entry is not above 64 u times the largest diagonal entry of G.  trace(S)
//     is the squared Frobenius norm of the part of M that the p = j pivot
//     columns leave unexplained when G = M^T M, so the stop picks the
//     smallest p whose residual fits the budget;
//   * pivoting off: pivots in order, all q of them; a non-positive pivot
//     stops with p_out = -1.
// Column j of the factor (by original row index) goes to L (q x q,
// column-major); the p x p leading block R = L11^T in pivot order is then
// inverted in place in LDS and R^-1 written upper triangular, row-major.
// The permutation and R^-1 are all the CholQR2 pass needs: Q1 = M P R^-1.
#include "../../include/tt_amd.h"
#include "tta_device.hpp"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

namespace tta
{
using namespace dev;

namespace
{
__device__ __forceinline__ int pk(int i, int l) { return i * (i + 1) / 2 + l; }  // i >= l

// Sum and arg-max (largest value, smallest index on ties) over the block;
// every thread gets the results.  `red`: LDS of 3 * NT / 64 doubles.
template <int NT>
__device__ __forceinline__ void block_sum_argmax(double& sum, double& mv, int& mi, double* red)
{
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1)
  {
    sum += __shfl_xor(sum, o, 64);
    const double ov = __shfl_xor(mv, o, 64);
    const int oi = __shfl_xor(mi, o, 64);
    if (ov > mv || (ov == mv && oi < mi))
    {
      mv = ov;
      mi = oi;
    }
  }
  __syncthreads();
  if (lane == 0)
  {
    red[wid] = sum;
    red[NT / 64 + wid] = mv;
    red[2 * (NT / 64) + wid] = (double)mi;
  }
  __syncthreads();
  sum = 0.0;
  mv = -1.0;
  mi = 0x7fffffff;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w)
  {
    sum += red[w];
    const double ov = red[NT / 64 + w];
    const int oi = (int)red[2 * (NT / 64) + w];
    if (ov > mv || (ov == mv && oi < mi))
    {
      mv = ov;
      mi = oi;
    }
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void pchol_batch_kernel(const pchol_desc* ds)
{
  const pchol_desc& d = ds[blockIdx.x];
  const int q = d.q;
  const int tid = threadIdx.x;
  extern __shared__ double sm[];
  const int npk = q * (q + 1) / 2;
  double* S = sm;
  double* lv = S + npk;
  double* red = lv + q;
  int* done = reinterpret_cast<int*>(red + 3 * (NT / 64));
  if (q <= 0)
  {
    if (tid == 0) *d.p_out = 0;
    return;
  }
  for (int e = tid; e < npk; e += NT)
  {
    int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while (pk(i + 1, 0) <= e) ++i;
    while (pk(i, 0) > e) --i;
    const int l = e - pk(i, 0);
    S[e] = d.G[(size_t)i * d.ldg + l];
  }
  for (int i = tid; i < q; i += NT) done[i] = 0;
  __syncthreads();
  double dmax = 0.0;
  for (int i = 0; i < q; ++i) dmax = fmax(dmax, S[pk(i, i)]);
  const double floor_piv = 64.0 * 2.220446049250313e-16 * dmax;
  double stop2 = -1.0;
  if (d.pivot)
  {
    const double dl = d.delta ? *d.delta : 0.0;
    stop2 = d.stop_f * dl * dl;
  }
  int p = 0;
  bool broke = false;
  for (int j = 0; j < q; ++j)
  {
    double tr = 0.0, mv = -1.0;
    int mi = 0x7fffffff;
    for (int i = tid; i < q; i += NT)
      if (!done[i])
      {
        const double v = S[pk(i, i)];
        tr += v > 0.0 ? v : 0.0;
        if (v > mv || (v == mv && i < mi))
        {
          mv = v;
          mi = i;
        }
      }
    block_sum_argmax<NT>(tr, mv, mi, red);
    int pj;
    if (d.pivot)
    {
      if (tr <= stop2 || !(mv > floor_piv)) break;
      pj = mi;
    }
    else
    {
      pj = j;
      mv = S[pk(j, j)];
      if (!(mv > 0.0))
      {
        broke = true;
        break;
      }
    }
    const double dj = sqrt(mv);
    const double inv = 1.0 / dj;
    for (int i = tid; i < q; i += NT)
    {
      double li = 0.0;
      if (i == pj)
        li = dj;
      else if (!done[i])
        li = (i > pj ? S[pk(i, pj)] : S[pk(pj, i)]) * inv;
      lv[i] = li;
      d.L[(size_t)j * q + i] = li;
    }
    __syncthreads();
    // Rank-1 update of the remaining Schur complement (rows and columns of
    // finished pivots hold lv = 0; the pivot's own row becomes dead).
    for (int i = tid; i < q; i += NT)
    {
      const double li = lv[i];
      if (li == 0.0 || i == pj) continue;
      double* row = S + pk(i, 0);
      for (int l = 0; l <= i; ++l)
        if (l != pj) row[l] = fma(-li, lv[l], row[l]);
    }
    __syncthreads();
    if (tid == 0)
    {
      done[pj] = 1;
      if (d.piv) d.piv[j] = pj;
    }
    __syncthreads();
    p = j + 1;
  }
  if (broke)
  {
    if (tid == 0) *d.p_out = -1;
    return;
  }
  if (tid == 0) *d.p_out = p;
  if (d.piv && !d.pivot)
    for (int i = tid; i < q; i += NT) d.piv[i] = i;
  __syncthreads();
  // Leading block in pivot order, packed: T(a, b) = L(piv_a, b), a >= b.
  for (int a = tid; a < p; a += NT)
  {
    const int ra = d.pivot ? d.piv[a] : a;
    for (int b = 0; b <= a; ++b) S[pk(a, b)] = d.L[(size_t)b * q + ra];
  }
  __syncthreads();
  // In-place inverse of the lower triangle, last column first:
  //   X(j, j) = 1 / T(j, j),  X(i, j) = -X(j, j) sum_{k=j+1..i} X(i, k) T(k, j).
  for (int j = p - 1; j >= 0; --j)
  {
    for (int i = j + 1 + tid; i < p; i += NT)
    {
      double s = 0.0;
      for (int k = j + 1; k <= i; ++k) s = fma(S[pk(i, k)], S[pk(k, j)], s);
      lv[i] = s;
    }
    __syncthreads();
    const double xjj = 1.0 / S[pk(j, j)];
    for (int i = j + 1 + tid; i < p; i += NT) S[pk(i, j)] = -lv[i] * xjj;
    __syncthreads();
    if (tid == 0) S[pk(j, j)] = xjj;
    __syncthreads();
  }
  // R^-1 = X^T, upper triangular, row-major.
  for (int e = tid; e < p * p; e += NT)
  {
    const int a = e / p, b = e % p;
    d.Rinv[(size_t)a * d.ldr + b] = b >= a ? S[pk(b, a)] : 0.0;
  }
}
}  // namespace

size_t pchol_lds_bytes(int q)
{
  constexpr int NT = 256;
  return sizeof(double) * ((size_t)q * (q + 1) / 2 + q + 3 * (NT / 64)) + sizeof(int) * (size_t)q;
}

void launch_pchol_batch(const pchol_desc* ds, int count, int max_q, hipStream_t st)
{
  if (count <= 0) return;
  if (max_q > kPcholMaxQ) fail(TTA_ERR_UNSUPPORTED, "pivoted Cholesky: %d columns above %d", max_q, kPcholMaxQ);
  hipLaunchKernelGGL(pchol_batch_kernel<256>, dim3(count), dim3(256), pchol_lds_bytes(max_q), st, ds);
}

}  // namespace tta
