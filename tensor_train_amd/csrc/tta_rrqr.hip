// Batched Cholesky with diagonal pivoting and a trace stop, for the
// rank-revealing CholQR2 step of the mixed batched rounding.
//
// One workgroup per problem.  The symmetric q x q matrix G (row-major) is
// held as a packed lower triangle in LDS (q <= kPcholMaxQ, about 152 KiB at
// q = 192) and factored right-looking, each pivot moved to the front by a
// symmetric swap (as LAPACK's dpstf2) so the trailing Schur complement stays
// a contiguous triangle; with pivoting the trailing update is applied once
// per block of kPcholBlock pivots, spread evenly over the workgroup:
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

#include <algorithm>

namespace tta
{
using namespace dev;

namespace
{
__device__ __forceinline__ int pk(int i, int l) { return i * (i + 1) / 2 + l; }  // i >= l

// Sum and arg-max (largest value, smallest position on ties) over the block;
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

// Row and column of entry t of a packed lower triangle (t = i (i + 1) / 2 + l).
__device__ __forceinline__ void unpack(int t, int& i, int& l)
{
  i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (pk(i + 1, 0) <= t) ++i;
  while (pk(i, 0) > t) --i;
  l = t - pk(i, 0);
}

template <int NT>
__device__ void pchol_one(const pchol_desc& d)
{
  if (d.skip && *d.skip) return;  // second pass taken by the Newton-Schulz step
  const int q = d.q;
  const int tid = threadIdx.x;
  extern __shared__ double sm[];
  const int npk = q * (q + 1) / 2;
  double* S = sm;            // packed lower triangle, by position
  double* lv = S + npk;      // columns of the factor, by position (kPcholBlock of q)
  double* red = lv + kPcholBlock * q;
  int* ord = reinterpret_cast<int*>(red + 3 * (NT / 64));  // position -> original index
  if (q <= 0)
  {
    if (tid == 0) *d.p_out = 0;
    return;
  }
  for (int e = tid; e < npk; e += NT)
  {
    int i, l;
    unpack(e, i, l);
    S[e] = d.G[(size_t)i * d.ldg + l];
  }
  for (int i = tid; i < q; i += NT) ord[i] = i;
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
  double tr_left = 0.0;
  // Symmetric swap of positions j and pj (pj > j) in the trailing part of S
  // (entries of positions above j and the diagonal), the permutation and
  // the first ncol factor columns held in lv.
  auto swap_pos = [&](int j, int pj, int ncol) {
    for (int l = j + 1 + tid; l < q; l += NT)
    {
      if (l < pj)
      {
        const double x = S[pk(l, j)];
        S[pk(l, j)] = S[pk(pj, l)];
        S[pk(pj, l)] = x;
      }
      else if (l > pj)
      {
        const double x = S[pk(l, j)];
        S[pk(l, j)] = S[pk(l, pj)];
        S[pk(l, pj)] = x;
      }
    }
    if (tid == 0)
    {
      const double x = S[pk(j, j)];
      S[pk(j, j)] = S[pk(pj, pj)];
      S[pk(pj, pj)] = x;
      const int o = ord[j];
      ord[j] = ord[pj];
      ord[pj] = o;
      for (int c = 0; c < ncol; ++c)
      {
        const double y = lv[c * q + j];
        lv[c * q + j] = lv[c * q + pj];
        lv[c * q + pj] = y;
      }
    }
  };
  // Trace and largest diagonal entry of the trailing Schur complement from
  // position j on (every thread returns them).
  auto diag_scan = [&](int j, double& tr, double& mv, int& mi) {
    tr = 0.0;
    mv = -1.0;
    mi = 0x7fffffff;
    for (int i = j + tid; i < q; i += NT)
    {
      const double v = S[pk(i, i)];
      tr += v > 0.0 ? v : 0.0;
      if (v > mv)
      {
        mv = v;
        mi = i;
      }
    }
    block_sum_argmax<NT>(tr, mv, mi, red);
  };
  if (d.pivot)
  {
    // Pivots in blocks of kPcholBlock: each pivot of a block is chosen on the
    // diagonal updated by the block's earlier columns, its column taken from
    // the raw entries minus those columns' terms; then one pass over the
    // trailing triangle applies the whole block's rank-1 updates (a quarter
    // of the passes, and of their barriers, of one pass per pivot).  The stop
    // and the pivot choices are those of the one-at-a-time form.
    for (int j = 0; j < q; j += kPcholBlock)
    {
      bool stop = false;
      for (int m = 0; m < kPcholBlock && j + m < q; ++m)
      {
        const int jm = j + m;
        double tr, mv;
        int mi;
        diag_scan(jm, tr, mv, mi);
        tr_left = tr;
        if (tr <= stop2 || !(mv > floor_piv))
        {
          stop = true;
          break;
        }
        if (mi != jm)
        {
          swap_pos(jm, mi, m);
          __syncthreads();
        }
        const double dj = sqrt(mv), inv = 1.0 / dj;
        for (int i = jm + tid; i < q; i += NT)
        {
          double li = dj;
          if (i > jm)
          {
            double x = S[pk(i, jm)];
            for (int c = 0; c < m; ++c) x = fma(-lv[c * q + i], lv[c * q + jm], x);
            li = x * inv;
            S[pk(i, i)] = fma(-li, li, S[pk(i, i)]);
          }
          lv[m * q + i] = li;
          d.L[(size_t)jm * q + ord[i]] = li;
        }
        __syncthreads();
        p = jm + 1;
      }
      if (stop || p == q) break;
      // Off-diagonal trailing entries S(i, l), p <= l < i < q, spread evenly
      // over the block by the packed index of the strictly lower triangle.
      const int nt = q - p;
      const int cnt = nt * (nt - 1) / 2;
      if (tid < cnt)
      {
        int ii, ll;
        unpack(tid, ii, ll);
        for (int t = tid; t < cnt; t += NT)
        {
          const int i = p + 1 + ii, l = p + ll;
          double* e = S + pk(i, l);
          double x = *e;
#pragma unroll
          for (int c = 0; c < kPcholBlock; ++c) x = fma(-lv[c * q + i], lv[c * q + l], x);
          *e = x;
          ll += NT;
          while (ll > ii)
          {
            ll -= ii + 1;
            ++ii;
          }
        }
      }
      __syncthreads();
    }
  }
  else
  {
    for (int j = 0; j < q; ++j)
    {
      // Trace and largest diagonal entry of the trailing Schur complement.
      double tr = 0.0, mv = -1.0;
      int mi = 0x7fffffff;
      for (int i = j + tid; i < q; i += NT)
      {
        const double v = S[pk(i, i)];
        tr += v > 0.0 ? v : 0.0;
        if (v > mv)
        {
          mv = v;
          mi = i;
        }
      }
      block_sum_argmax<NT>(tr, mv, mi, red);
      tr_left = tr;
      int pj = j;
      if (d.pivot)
      {
        if (tr <= stop2 || !(mv > floor_piv)) break;
        pj = mi;
      }
      else
      {
        mv = S[pk(j, j)];
        if (!(mv > 0.0))
        {
          broke = true;
          break;
        }
      }
      // Symmetric swap of positions j and pj (pj > j) in the trailing part.
      if (pj != j)
      {
        for (int l = j + 1 + tid; l < q; l += NT)
        {
          if (l < pj)
          {
            const double x = S[pk(l, j)];
            S[pk(l, j)] = S[pk(pj, l)];
            S[pk(pj, l)] = x;
          }
          else if (l > pj)
          {
            const double x = S[pk(l, j)];
            S[pk(l, j)] = S[pk(l, pj)];
            S[pk(l, pj)] = x;
          }
        }
        if (tid == 0)
        {
          const double x = S[pk(j, j)];
          S[pk(j, j)] = S[pk(pj, pj)];
          S[pk(pj, pj)] = x;
          const int o = ord[j];
          ord[j] = ord[pj];
          ord[pj] = o;
        }
        __syncthreads();
      }
      const double dj = sqrt(mv);
      const double inv = 1.0 / dj;
      for (int i = j + tid; i < q; i += NT)
      {
        const double li = i == j ? dj : S[pk(i, j)] * inv;
        lv[i] = li;
        d.L[(size_t)j * q + ord[i]] = li;
      }
      __syncthreads();
      // Trailing update S(i, l) -= lv_i lv_l, j < l <= i < q, spread evenly
      // over the block by the packed index of the trailing triangle.
      const int nt = q - j - 1;
      const int cnt = nt * (nt + 1) / 2;
      if (tid < cnt)
      {
        int ii, ll;
        unpack(tid, ii, ll);
        for (int t = tid; t < cnt; t += NT)
        {
          const int i = j + 1 + ii, l = j + 1 + ll;
          double* e = S + pk(i, l);
          *e = fma(-lv[i], lv[l], *e);
          ll += NT;
          while (ll > ii)
          {
            ll -= ii + 1;
            ++ii;
          }
        }
      }
      __syncthreads();
      p = j + 1;
    }
  }
  if (broke)
  {
    if (tid == 0) *d.p_out = -1;
    return;
  }
  if (tid == 0)
  {
    *d.p_out = p;
    if (d.tr_out) *d.tr_out = p == q ? 0.0 : tr_left;
  }
  if (d.piv)
    for (int i = tid; i < q; i += NT) d.piv[i] = ord[i];
  // Leading block of the factor in pivot order, packed: T(a, b) = L(ord_a, b).
  // The L columns were written by this workgroup in this launch; the barrier
  // orders those stores before the loads.
  __syncthreads();
  for (int e = tid; e < p * (p + 1) / 2; e += NT)
  {
    int a, b;
    unpack(e, a, b);
    S[e] = d.L[(size_t)b * q + ord[a]];
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
// Persistent form: a grid sized to the device's resident capacity; each
// workgroup takes the next problem of `order` (the problems by decreasing
// size, so the largest start first) from the work-queue counter *queue,
// which the launch wrapper zeroes beforehand, until the queue is empty.
template <int NT>
__global__ __launch_bounds__(NT) void pchol_batch_kernel(const pchol_desc* ds, int count, const int* order, int* queue)
{
  __shared__ int item;
  for (;;)
  {
    if (threadIdx.x == 0) item = atomicAdd(queue, 1);
    __syncthreads();
    const int it = item;
    __syncthreads();
    if (it >= count) break;
    pchol_one<NT>(ds[order[it]]);
    __syncthreads();
  }
}

__global__ void zero_int_kernel(int* p, int n)
{
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__device__ __forceinline__ double wave_max_all(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// One workgroup per problem: max |G2 - I|, then either Rinv = (3 I - G2) / 2
// with the problem marked done, or the problem left to the Cholesky.
template <int NT>
__global__ __launch_bounds__(NT) void ns_pass2_kernel(const pchol_desc* ds, int* skip)
{
  const pchol_desc d = ds[blockIdx.x];
  const int p = d.q;
  __shared__ double red[NT / 64];
  double mx = 0.0;
  for (int e = threadIdx.x; e < p * p; e += NT)
  {
    const int i = e / p, j = e % p;
    const double v = d.G[(size_t)i * d.ldg + j] - (i == j ? 1.0 : 0.0);
    mx = fmax(mx, fabs(v));  // a NaN leaves mx, and fails the test below
    if (v != v) mx = 1.0e300;
  }
  mx = wave_max_all(mx);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  double m = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) m = fmax(m, red[w]);
  const bool ns = p > 0 && (double)p * m <= 1.0e-8;
  if (ns)
    for (int e = threadIdx.x; e < p * p; e += NT)
    {
      const int i = e / p, j = e % p;
      d.Rinv[(size_t)i * d.ldr + j] = 0.5 * ((i == j ? 3.0 : 0.0) - d.G[(size_t)i * d.ldg + j]);
    }
  if (threadIdx.x == 0)
  {
    skip[blockIdx.x] = ns ? 1 : 0;
    if (ns) *d.p_out = p;
  }
}
}  // namespace

void launch_ns_pass2(const pchol_desc* ds, int count, int* skip, hipStream_t st)
{
  if (count <= 0) return;
  hipLaunchKernelGGL(ns_pass2_kernel<256>, dim3(count), dim3(256), 0, st, ds, skip);
}

size_t pchol_lds_bytes(int q)
{
  constexpr int NT = 256;
  return sizeof(double) * ((size_t)q * (q + 1) / 2 + kPcholBlock * (size_t)q + 3 * (NT / 64)) + sizeof(int) * (size_t)q;
}

void launch_pchol_batch(arena& S, const pchol_desc* ds, const std::vector<pchol_desc>& host, hipStream_t st)
{
  const int count = (int)host.size();
  if (count <= 0) return;
  static int cus = 0, lds_cu = 0;
  if (cus == 0)
  {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        lds_cu <= 0)
      lds_cu = 160 * 1024;
  }
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return host[a].q > host[b].q; });
  const int max_q = host[order[0]].q;
  if (max_q > kPcholMaxQ) fail(TTA_ERR_UNSUPPORTED, "pivoted Cholesky: %d columns above %d", max_q, kPcholMaxQ);
  // Resident workgroups of 256 threads per compute unit that a problem's LDS
  // allows, rounded down to 1, 2, 4 or 8.  It never decreases along the
  // largest-first order, so its runs are the classes.
  auto resident = [&](int q) {
    const int r = std::max(1, std::min(8, (int)((size_t)lds_cu / std::max<size_t>(pchol_lds_bytes(q), 1))));
    int b = 1;
    while (2 * b <= r) b *= 2;
    return b;
  };
  std::vector<int> start, per_cu;
  for (int i = 0; i < count; ++i)
  {
    const int r = resident(host[order[i]].q);
    if (per_cu.empty() || r != per_cu.back())
    {
      start.push_back(i);
      per_cu.push_back(r);
    }
  }
  start.push_back(count);
  const int nclass = (int)per_cu.size();
  const int* od = S.upload(order);
  int* queues = S.alloc<int>(nclass);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, st, queues, nclass);
  for (int k = 0; k < nclass; ++k)
  {
    const int n = start[k + 1] - start[k];
    const size_t lds = pchol_lds_bytes(host[order[start[k]]].q);
    const int grid = std::min(n, cus * per_cu[k]);
    hipLaunchKernelGGL(pchol_batch_kernel<256>, dim3(grid), dim3(256), lds, st, ds, n, od + start[k], queues + k);
  }
}

}  // namespace tta
