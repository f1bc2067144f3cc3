// Blocked Householder QR (compact WY) of batches of column-major matrices.
//
// Per panel of QB = 16 columns:
//   * qr_panel_kernel  -- one workgroup per matrix; the panel lives in
//     registers (thread t owns rows t, t + NT, ...), unblocked Householder
//     with the dlarfg sign convention, then the triangular T of the block
//     reflector H = H_0 ... H_{QB-1} = I - V T V^T (LAPACK dlarft, forward,
//     columnwise) from the Gram V^T V.  Column j's block reduction sums both
//     v_j^T x_c (c > j, for the update) and v_c^T v_j (c < j, the Gram), so
//     the Gram costs no pass of its own.
//   * qr_update_kernel -- many workgroups per matrix (CT = 16 columns each):
//     C <- (I - V T^T V^T) C on the trailing columns (factorisation) or
//     C <- (I - V T V^T) C on the thin Q (formation, panels in reverse).
//     V is staged through LDS in 128-row chunks.
// The thin Q (m x k, k = min(m, n)) is formed from Q = I by the same update
// kernel (LAPACK dorgqr order).
#include <algorithm>

#include "tta_device.hpp"
#include "tta_kernels.hpp"
#include "tta_potrf.hpp"

namespace tta
{
using namespace dev;

namespace
{
constexpr int QB = 16;   // panel width
constexpr int QCT = 8;   // columns per update workgroup (2 per wave)
constexpr int QCPW = QCT / 4;
constexpr int QRC = 128; // rows per V chunk in the update kernel

// By value: a reference that may point at the kernel argument made the
// compiler copy the argument to scratch memory and read every field through
// generic (flat) loads.
__device__ __forceinline__ qrp pick(const qr_batch& b)
{
  return b.descs ? b.descs[blockIdx.y] : b.one;
}

// Sum v[c] for c in [lo, hi) across the block; results to out[c] (LDS),
// visible to all threads on return.  `red`: LDS of (NT/64) * NV doubles.
// The loop over c is unrolled so the register array is indexed statically.
template <int NT, int NV>
__device__ __forceinline__ void block_vsum(const double (&v)[NV], int lo, int hi, double* red, double* out)
{
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if constexpr (NV == 16)
  {
    // All 16 at once; lane c holds the wave's sum of v[c].
    const double s = wave_vsum16(v);
    if (lane < 16 && lane >= lo && lane < hi) red[wid * NV + lane] = s;
  }
  else
  {
#pragma unroll
    for (int c = 0; c < NV; ++c)
    {
      if (c >= lo && c < hi)
      {
        const double s = wave_sum(v[c]);
        if (lane == 0) red[wid * NV + c] = s;
      }
    }
  }
  __syncthreads();
  for (int c = lo + (int)threadIdx.x; c < hi; c += NT)
  {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += red[w * NV + c];
    out[c] = s;
  }
  __syncthreads();
}

template <int NT, int RPT>
__global__ __launch_bounds__(NT) void qr_panel_kernel(qr_batch bt, int j0)
{
  const qrp P = bt.descs ? bt.descs[blockIdx.x] : bt.one;
  const int m = P.m, n = P.n;
  const int k = m < n ? m : n;
  if (j0 >= k) return;
  const int bw = (k - j0) < QB ? (k - j0) : QB;
  const int len = m - j0;
  const int lda = P.lda;
  double* A = P.A;
  __shared__ double red[(NT / 64) * QB];
  __shared__ double bc[QB];
  __shared__ double Gm[QB][QB + 1];
  __shared__ double Tl[QB][QB];
  __shared__ double taus[QB];
  __shared__ double sc[2];
  const int tid = threadIdx.x;

  double x[RPT][QB];
#pragma unroll
  for (int i = 0; i < RPT; ++i)
  {
    const int r = tid + NT * i;
#pragma unroll
    for (int c = 0; c < QB; ++c) x[i][c] = (r < len && c < bw) ? A[(size_t)(j0 + c) * lda + j0 + r] : 0.0;
  }
  if (tid < QB) taus[tid] = 0.0;
  __syncthreads();
  for (int jj = 0; jj < bw; ++jj)
  {
    // Column jj of the rows this thread owns (static indices + selects).
    double xj[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i)
    {
      xj[i] = 0.0;
#pragma unroll
      for (int c = 0; c < QB; ++c)
        if (c == jj) xj[i] = x[i][c];
    }
    double acc[1] = {0.0};
#pragma unroll
    for (int i = 0; i < RPT; ++i)
    {
      const int r = tid + NT * i;
      if (r > jj && r < len) acc[0] += xj[i] * xj[i];
      if (r == jj) sc[1] = xj[i];
    }
    block_vsum<NT, 1>(acc, 0, 1, red, sc);
    const double tail = sc[0];
    const double alpha = sc[1];
    double tj = 0.0, scal = 0.0, beta = alpha;
    if (tail > DBL_MIN)
    {
      beta = sqrt(alpha * alpha + tail);
      if (alpha >= 0.0) beta = -beta;
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (tid == 0) taus[jj] = tj;
    double vv[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i)
    {
      const int r = tid + NT * i;
      vv[i] = (r == jj) ? 1.0 : ((r > jj && r < len) ? xj[i] * scal : 0.0);
    }
    // part[c] = v_jj^T x_c for c > jj (update), v_c^T v_jj for c < jj (Gram;
    // x holds v_c below the diagonal, the implicit 1 on it).
    double part[QB];
#pragma unroll
    for (int c = 0; c < QB; ++c)
    {
      part[c] = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i)
      {
        const int r = tid + NT * i;
        if (c > jj)
          part[c] += vv[i] * x[i][c];
        else if (c < jj)
          part[c] += vv[i] * ((r == c) ? 1.0 : (r > c ? x[i][c] : 0.0));
      }
    }
    if (bw > 1) block_vsum<NT, QB>(part, 0, bw, red, bc);
    if (tid < jj) Gm[tid][jj] = bc[tid];
#pragma unroll
    for (int c = 0; c < QB; ++c)
    {
      if (c > jj && c < bw)
      {
        const double w = tj * bc[c];
#pragma unroll
        for (int i = 0; i < RPT; ++i) x[i][c] -= w * vv[i];
      }
      else if (c == jj)
      {
#pragma unroll
        for (int i = 0; i < RPT; ++i)
        {
          const int r = tid + NT * i;
          if (r == jj) x[i][c] = beta;
          else if (r > jj) x[i][c] = vv[i];
        }
      }
    }
    __syncthreads();
  }
  // Store the panel (R on/above the diagonal, V below) and tau.
#pragma unroll
  for (int i = 0; i < RPT; ++i)
  {
    const int r = tid + NT * i;
    if (r < len)
    {
#pragma unroll
      for (int c = 0; c < QB; ++c)
        if (c < bw) A[(size_t)(j0 + c) * lda + j0 + r] = x[i][c];
    }
  }
  if (tid < bw) P.tau[j0 + tid] = taus[tid];
  __syncthreads();
  // T of the block reflector (LAPACK dlarft, forward, columnwise):
  //   T[a][a] = tau_a,  T[a][j] = -tau_j * sum_{b=a}^{j-1} T[a][b] (V^T V)[b][j].
  // Row a only needs row a, so thread a builds its row in registers.
  if (tid < QB)
  {
    const int a = tid;
    double row[QB];
#pragma unroll
    for (int j = 0; j < QB; ++j) row[j] = 0.0;
    if (a < bw)
    {
#pragma unroll
      for (int j = 0; j < QB; ++j)
      {
        if (j == a) row[j] = taus[a];
        if (j > a && j < bw)
        {
          double sacc = 0.0;
#pragma unroll
          for (int bb = 0; bb < QB; ++bb)
            if (bb >= a && bb < j) sacc += row[bb] * Gm[bb][j];
          row[j] = -taus[j] * sacc;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < QB; ++j) Tl[a][j] = row[j];
  }
  __syncthreads();
  double* Tout = P.T + (size_t)(j0 / QB) * QB * QB;
  for (int e = tid; e < QB * QB; e += NT) Tout[e] = Tl[e / QB][e % QB];
}

// The same panel factorization for panels of any height: the panel stays in
// place in global memory (in L2 for the heights this serves) and each
// column costs two passes over the rows a thread owns (tid, tid + NT, ...):
//   pass A  -- v_jj^T x_c for c > jj (update) and v_c^T v_jj for c < jj (Gram),
//   pass B  -- x_c -= tau v_jj (v_jj^T x_c) for c > jj, column jj <- (beta, v_jj),
//              and the tail norm and diagonal entry of column jj + 1 from the
//              updated values (so the next column needs no pass of its own).
// The reflector formulas are those of qr_panel_kernel; only the summation
// order differs.
template <int NT>
__global__ __launch_bounds__(NT) void qr_panel_tall_kernel(qr_batch bt, int j0)
{
  const qrp P = bt.descs ? bt.descs[blockIdx.x] : bt.one;
  const int m = P.m, n = P.n;
  const int k = m < n ? m : n;
  if (j0 >= k) return;
  const int bw = (k - j0) < QB ? (k - j0) : QB;
  const int len = m - j0;
  const int lda = P.lda;
  double* A = P.A + (size_t)j0 * lda + j0;  // panel origin: A(j0, j0)
  __shared__ double red[(NT / 64) * QB];
  __shared__ double bc[QB];
  __shared__ double Gm[QB][QB + 1];
  __shared__ double Tl[QB][QB];
  __shared__ double taus[QB];
  __shared__ double sc[2];
  const int tid = threadIdx.x;
  if (tid < QB) taus[tid] = 0.0;
  // Tail norm and diagonal entry of column 0.
  {
    double acc[1] = {0.0};
    for (int r = tid; r < len; r += NT)
    {
      const double v = A[r];
      if (r > 0) acc[0] += v * v;
      else sc[1] = v;
    }
    block_vsum<NT, 1>(acc, 0, 1, red, sc);
  }
  for (int jj = 0; jj < bw; ++jj)
  {
    const double tail = sc[0];
    const double alpha = sc[1];
    double tj = 0.0, scal = 0.0, beta = alpha;
    if (tail > DBL_MIN)
    {
      beta = sqrt(alpha * alpha + tail);
      if (alpha >= 0.0) beta = -beta;
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (tid == 0) taus[jj] = tj;
    const double* xj = A + (size_t)jj * lda;
    // Pass A.
    double part[QB];
#pragma unroll
    for (int c = 0; c < QB; ++c) part[c] = 0.0;
    for (int r = jj + tid; r < len; r += NT)
    {
      const double v = (r == jj) ? 1.0 : xj[r] * scal;
#pragma unroll
      for (int c = 0; c < QB; ++c)
      {
        if (c >= bw || c == jj) continue;
        const double* xc = A + (size_t)c * lda;
        // Column c < jj holds v_c below its diagonal (implicit 1 on it); rows
        // r >= jj > c are all below c's diagonal.
        part[c] += v * xc[r];
      }
    }
    if (bw > 1) block_vsum<NT, QB>(part, 0, bw, red, bc);
    if (tid < jj) Gm[tid][jj] = bc[tid];
    // Pass B.
    double acc[1] = {0.0};
    const int nx = jj + 1;
    for (int r = jj + tid; r < len; r += NT)
    {
      const double xv = xj[r];
      const double v = (r == jj) ? 1.0 : xv * scal;
#pragma unroll
      for (int c = 0; c < QB; ++c)
      {
        if (c <= jj || c >= bw) continue;
        double* xc = A + (size_t)c * lda;
        const double u = xc[r] - tj * bc[c] * v;
        xc[r] = u;
        if (c == nx)
        {
          if (r > nx) acc[0] += u * u;
          else if (r == nx) sc[1] = u;
        }
      }
      A[(size_t)jj * lda + r] = (r == jj) ? beta : v;
    }
    __syncthreads();  // sc[1] written above is read after the sum
    if (nx < bw) block_vsum<NT, 1>(acc, 0, 1, red, sc);
    else __syncthreads();
  }
  if (tid < bw) P.tau[j0 + tid] = taus[tid];
  __syncthreads();
  if (tid < QB)
  {
    const int a = tid;
    double row[QB];
#pragma unroll
    for (int j = 0; j < QB; ++j) row[j] = 0.0;
    if (a < bw)
    {
#pragma unroll
      for (int j = 0; j < QB; ++j)
      {
        if (j == a) row[j] = taus[a];
        if (j > a && j < bw)
        {
          double sacc = 0.0;
#pragma unroll
          for (int bb = 0; bb < QB; ++bb)
            if (bb >= a && bb < j) sacc += row[bb] * Gm[bb][j];
          row[j] = -taus[j] * sacc;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < QB; ++j) Tl[a][j] = row[j];
  }
  __syncthreads();
  double* Tout = P.T + (size_t)(j0 / QB) * QB * QB;
  for (int e = tid; e < QB * QB; e += NT) Tout[e] = Tl[e / QB][e % QB];
}

// C <- (I - V op(T) V^T) C for columns [cbeg, cend) of C, rows j0 .. m-1.
// mode 0: factorisation (C = A, op(T) = T^T, cbeg = j0 + bw, cend = n).
// mode 1: Q formation (C = Q, op(T) = T, cbeg = j0, cend = k).
// mode 2: as mode 1 with C taken to be the identity (Q is only written; the
//         sums are the same as mode 1's after an identity fill).
__global__ __launch_bounds__(256) void qr_update_kernel(qr_batch bt, int j0, int mode)
{
  const qrp& P = pick(bt);
  const int m = P.m, n = P.n;
  const int k = m < n ? m : n;
  if (j0 >= k) return;
  const int bw = (k - j0) < QB ? (k - j0) : QB;
  const int len = m - j0;
  const int cbeg = mode == 0 ? j0 + bw : j0;
  const int cend = mode == 0 ? n : k;
  const int c0 = cbeg + blockIdx.x * QCT;
  if (c0 >= cend) return;
  const int lda = P.lda;
  const double* A = P.A;
  double* C = mode == 0 ? P.A : P.Q;
  const int ldc = mode == 0 ? lda : P.ldq;
  __shared__ double Vs[QRC][QB + 1];
  __shared__ double Ts[QB][QB];
  __shared__ double Ws[QB][QCT];
  __shared__ double W2[QB][QCT];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double* Tg = P.T + (size_t)(j0 / QB) * QB * QB;
  Ts[tid / QB][tid % QB] = Tg[tid];
  // Each wave owns QCPW columns.
  double acc[QCPW][QB];
#pragma unroll
  for (int u = 0; u < QCPW; ++u)
#pragma unroll
    for (int jj = 0; jj < QB; ++jj) acc[u][jj] = 0.0;
  for (int r0 = 0; r0 < len; r0 += QRC)
  {
    __syncthreads();
    for (int e = tid; e < QRC * QB; e += 256)
    {
      const int rr = e % QRC, jj = e / QRC;
      const int r = r0 + rr;
      double v = 0.0;
      if (jj < bw && r < len) v = (r == jj) ? 1.0 : (r > jj ? A[(size_t)(j0 + jj) * lda + j0 + r] : 0.0);
      Vs[rr][jj] = v;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < QCPW; ++u)
    {
      const int c = c0 + wid * QCPW + u;
      if (c >= cend) continue;
      const double* col = C + (size_t)c * ldc + j0;
      for (int rr = lane; rr < QRC; rr += 64)
      {
        const int r = r0 + rr;
        if (r >= len) break;
        const double a = mode == 2 ? (j0 + r == c ? 1.0 : 0.0) : col[r];
#pragma unroll
        for (int jj = 0; jj < QB; ++jj) acc[u][jj] = fma(Vs[rr][jj], a, acc[u][jj]);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < QCPW; ++u)
#pragma unroll
    for (int jj = 0; jj < QB; ++jj)
    {
      const double s = wave_sum(acc[u][jj]);
      if (lane == 0) Ws[jj][wid * QCPW + u] = s;
    }
  __syncthreads();
  if (tid < QB * QCT)
  {
    const int jj = tid / QCT, cc = tid % QCT;
    double s = 0.0;
    if (mode == 0)
    {
#pragma unroll
      for (int a = 0; a < QB; ++a) s += Ts[a][jj] * Ws[a][cc];
    }
    else
    {
#pragma unroll
      for (int a = 0; a < QB; ++a) s += Ts[jj][a] * Ws[a][cc];
    }
    W2[jj][cc] = s;
  }
  for (int r0 = 0; r0 < len; r0 += QRC)
  {
    __syncthreads();
    for (int e = tid; e < QRC * QB; e += 256)
    {
      const int rr = e % QRC, jj = e / QRC;
      const int r = r0 + rr;
      double v = 0.0;
      if (jj < bw && r < len) v = (r == jj) ? 1.0 : (r > jj ? A[(size_t)(j0 + jj) * lda + j0 + r] : 0.0);
      Vs[rr][jj] = v;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < QCPW; ++u)
    {
      const int c = c0 + wid * QCPW + u;
      if (c >= cend) continue;
      double* col = C + (size_t)c * ldc + j0;
      const int cc = wid * QCPW + u;
      for (int rr = lane; rr < QRC; rr += 64)
      {
        const int r = r0 + rr;
        if (r >= len) break;
        double s = 0.0;
#pragma unroll
        for (int jj = 0; jj < QB; ++jj) s = fma(Vs[rr][jj], W2[jj][cc], s);
        if (mode == 2)
          col[r] = (j0 + r == c ? 1.0 : 0.0) - s;
        else
          col[r] -= s;
      }
    }
  }
}

// Matrix-core form of qr_update_kernel for 16 columns per workgroup:
//   W  = V^T C        (16 x 16; each wave accumulates a quarter of the rows
//                      with v_mfma_f64_16x16x4f64, summed through LDS),
//   W2 = op(T) W      (16 x 16, one element per thread),
//   C -= V W2         (computed as (W2^T V^T) so that the accumulator of a
//                      lane runs down a column of C: coalesced updates).
// V and C pass through LDS in 64-row chunks.  Same modes as qr_update_kernel.
typedef double qdbl4 __attribute__((ext_vector_type(4)));
constexpr int UCT = 16;   // columns per workgroup
constexpr int URC = 64;   // rows per LDS chunk (16 per wave)

__global__ __launch_bounds__(256) void qr_update_mfma_kernel(qr_batch bt, int j0, int mode, int col_end)
{
  const qrp& P = pick(bt);
  const int m = P.m, n = P.n;
  const int k = m < n ? m : n;
  if (j0 >= k) return;
  const int bw = (k - j0) < QB ? (k - j0) : QB;
  const int len = m - j0;
  const int cbeg = mode == 0 ? j0 + bw : j0;
  const int cend = mode == 0 ? (col_end > 0 && col_end < n ? col_end : n) : k;
  const int c0 = cbeg + blockIdx.x * UCT;
  if (c0 >= cend) return;
  const int nc = (cend - c0) < UCT ? (cend - c0) : UCT;
  const int lda = P.lda;
  const double* A = P.A;
  double* C = mode == 0 ? P.A : P.Q;
  const int ldc = mode == 0 ? lda : P.ldq;
  __shared__ double Vs[URC][QB + 1];
  __shared__ double Cs[URC][UCT + 1];
  __shared__ double Wp[4][QB][UCT + 1];
  __shared__ double W2[QB][UCT + 1];
  __shared__ double Ts[QB][QB];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int fl = lane & 15, fk = lane >> 4;
  const double* Tg = P.T + (size_t)(j0 / QB) * QB * QB;
  Ts[tid / QB][tid % QB] = Tg[tid];
  // V(r, jj): unit diagonal, zero above, stored below the diagonal of A.
  auto vload = [&](int r, int jj) -> double {
    if (jj >= bw || r >= len) return 0.0;
    return r == jj ? 1.0 : (r > jj ? A[(size_t)(j0 + jj) * lda + j0 + r] : 0.0);
  };
  qdbl4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = 0; r0 < len; r0 += URC)
  {
    __syncthreads();
    for (int e = tid; e < URC * QB; e += 256)
    {
      const int rr = e % URC, cc = e / URC;
      Vs[rr][cc] = vload(r0 + rr, cc);
      Cs[rr][cc] = (cc < nc && r0 + rr < len) ? C[(size_t)(c0 + cc) * ldc + j0 + r0 + rr] : 0.0;
    }
    __syncthreads();
    // Rows 16 wid .. 16 wid + 15 of the chunk, 4 per step: A = V^T, B = C.
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4)
    {
      const int rr = wid * 16 + ks + fk;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[rr][fl], Cs[rr][fl], acc, 0, 0, 0);
    }
  }
  // acc[r] = W(fk + 4r, fl) over this wave's rows.
#pragma unroll
  for (int r = 0; r < 4; ++r) Wp[wid][fk + 4 * r][fl] = acc[r];
  __syncthreads();
  {
    const int jj = tid / UCT, cc = tid % UCT;
    double s = 0.0;
    if (mode == 0)
    {
#pragma unroll
      for (int a = 0; a < QB; ++a) s += Ts[a][jj] * (Wp[0][a][cc] + Wp[1][a][cc] + Wp[2][a][cc] + Wp[3][a][cc]);
    }
    else
    {
#pragma unroll
      for (int a = 0; a < QB; ++a) s += Ts[jj][a] * (Wp[0][a][cc] + Wp[1][a][cc] + Wp[2][a][cc] + Wp[3][a][cc]);
    }
    W2[jj][cc] = s;
  }
  for (int r0 = 0; r0 < len; r0 += URC)
  {
    __syncthreads();
    for (int e = tid; e < URC * QB; e += 256)
    {
      const int rr = e % URC, cc = e / URC;
      Vs[rr][cc] = vload(r0 + rr, cc);
    }
    __syncthreads();
    // Z(col, row) = sum_jj W2(jj, col) V(row, jj) for the wave's 16 rows:
    // A(i = col, k = jj) = W2(jj, col), B(k = jj, j = row) = V(row, jj).
    qdbl4 z = {0.0, 0.0, 0.0, 0.0};
    const int rb = wid * 16;
#pragma unroll
    for (int ks = 0; ks < QB; ks += 4)
      z = __builtin_amdgcn_mfma_f64_16x16x4f64(W2[ks + fk][fl], Vs[rb + fl][ks + fk], z, 0, 0, 0);
    // z[r] = Z(col = fk + 4r, row = rb + fl).
    const int row = r0 + rb + fl;
    if (row < len)
    {
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        const int cc = fk + 4 * r;
        if (cc < nc) C[(size_t)(c0 + cc) * ldc + j0 + row] -= z[r];
      }
    }
  }
}

// Q = I[:, :k] (m x k, leading dimension ldq).
__global__ void qr_eye_kernel(qr_batch bt)
{
  const qrp& P = pick(bt);
  const int m = P.m, k = P.m < P.n ? P.m : P.n;
  const long total = (long)m * k;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int c = (int)(e / m), r = (int)(e % m);
    P.Q[(size_t)c * P.ldq + r] = (r == c) ? 1.0 : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Two-level blocking for single tall matrices: the 16-column panels of a
// 64-column block update only the rest of that block; the trailing matrix
// then takes one update with the block's compact-WY form, V (len x nb, unit
// lower trapezoidal) and T (nb x nb upper triangular, H_1 .. H_nb = I - V T
// V^T), as three GEMMs on the matrix cores.  The trailing matrix is then read
// and written once per 64 columns instead of once per 16.
// ---------------------------------------------------------------------------
// V (column-major, len x nb) of the block at column j0 from the reflectors
// stored below A's diagonal: 1 on the diagonal, 0 above.
__global__ void qr_vblock_kernel(const double* A, int lda, int j0, int len, int nb, double* V)
{
  const long total = (long)len * nb;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int c = (int)(e / len), r = (int)(e % len);
    V[e] = r == c ? 1.0 : (r > c ? A[(size_t)(j0 + c) * lda + j0 + r] : 0.0);
  }
}

// T of the block (row-major nb x nb, nb <= 64) from G = V^T V (row-major)
// and tau (LAPACK dlarft, forward, columnwise): T(i, i) = tau_i and
// T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).  One workgroup of 64 threads;
// thread r keeps row r of T in registers.
__global__ __launch_bounds__(64) void qr_tblock_kernel(const double* G, const double* tau, int nb, double* T)
{
  __shared__ double Gs[64][65];
  __shared__ double taus[64];
  const int r = threadIdx.x;
  for (int e = r; e < nb * nb; e += 64) Gs[e / nb][e % nb] = G[e];
  if (r < nb) taus[r] = tau[r];
  __syncthreads();
  double t[64];
#pragma unroll
  for (int c = 0; c < 64; ++c) t[c] = 0.0;
  for (int i = 0; i < nb; ++i)
  {
    double z = 0.0;
#pragma unroll
    for (int c = 0; c < 64; ++c)
      if (c >= r && c < i) z = fma(t[c], Gs[c][i], z);
    const double v = r < i ? -taus[i] * z : (r == i ? taus[i] : 0.0);
#pragma unroll
    for (int c = 0; c < 64; ++c)
      if (c == i) t[c] = v;
  }
  if (r < nb)
  {
#pragma unroll
    for (int c = 0; c < 64; ++c)
      if (c < nb) T[(size_t)r * nb + c] = t[c];
  }
}

// Largest n that CholQR2 takes (the blocked Cholesky itself has no limit;
// beyond this the Grams cost more than Householder saves).
constexpr int kCholMaxN = 512;

__global__ __launch_bounds__(256) void potrf_inv_kernel(const double* G, int ldg, double* R, int ldr, double* X,
                                                        int ldx, int jb, int j0, double* info)
{
  potrf_inv_block(G, ldg, R, ldr, X, ldx, jb, j0, info);
}
}  // namespace

int potrf_block() { return kPotrfB; }

void launch_potrf_inv(const double* G, int ldg, double* R, int ldr, double* X, int ldx, int jb, int j0, double* info,
                      hipStream_t st)
{
  hipLaunchKernelGGL(potrf_inv_kernel, dim3(1), dim3(256), 0, st, G, ldg, R, ldr, X, ldx, jb, j0, info);
}

int chol_max_n() { return kCholMaxN; }


void launch_qr_panel_one(const qr_batch& b, int j0, hipStream_t st)
{
  constexpr int NT = 512;
  if (b.max_m > 4 * NT)
    hipLaunchKernelGGL((qr_panel_tall_kernel<NT>), dim3(1), dim3(NT), 0, st, b, j0);
  else if (b.max_m <= 64)
    hipLaunchKernelGGL((qr_panel_kernel<64, 1>), dim3(1), dim3(64), 0, st, b, j0);
  else if (b.max_m <= 256)
    hipLaunchKernelGGL((qr_panel_kernel<256, 1>), dim3(1), dim3(256), 0, st, b, j0);
  else if (b.max_m <= NT)
    hipLaunchKernelGGL((qr_panel_kernel<NT, 1>), dim3(1), dim3(NT), 0, st, b, j0);
  else if (b.max_m <= 2 * NT)
    hipLaunchKernelGGL((qr_panel_kernel<NT, 2>), dim3(1), dim3(NT), 0, st, b, j0);
  else
    hipLaunchKernelGGL((qr_panel_kernel<NT, 4>), dim3(1), dim3(NT), 0, st, b, j0);
}

void launch_qr_update_block(const qr_batch& b, int j0, int col_end, hipStream_t st)
{
  const int k = std::min(b.one.m, b.one.n);
  const int bw = std::min(k - j0, QB);
  const int ncols = std::min(col_end, b.one.n) - (j0 + bw);
  if (ncols <= 0) return;
  hipLaunchKernelGGL(qr_update_mfma_kernel, dim3((ncols + UCT - 1) / UCT, 1), dim3(256), 0, st, b, j0, 0, col_end);
}

void launch_qr_eye(const qr_batch& b, hipStream_t st)
{
  const long total = (long)b.max_m * std::min(b.max_m, b.max_n);
  long g = (total + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(qr_eye_kernel, dim3((int)g, 1), dim3(256), 0, st, b);
}

void launch_qr_vblock(const double* A, int lda, int j0, int len, int nb, double* V, hipStream_t st)
{
  long g = ((long)len * nb + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(qr_vblock_kernel, dim3((int)g), dim3(256), 0, st, A, lda, j0, len, nb, V);
}

void launch_qr_tblock(const double* G, const double* tau, int nb, double* T, hipStream_t st)
{
  hipLaunchKernelGGL(qr_tblock_kernel, dim3(1), dim3(64), 0, st, G, tau, nb, T);
}

void launch_qr_blocked(const qr_batch& b, hipStream_t st)
{
  const int batch = b.descs ? b.batch : 1;
  if (batch <= 0) return;
  const int kmax = b.max_m < b.max_n ? b.max_m : b.max_n;
  constexpr int NT = 512;
  // Batches and wide matrices: 16-column matrix-core updates (enough
  // workgroups anyway); a single narrower matrix keeps the 8-column VALU
  // update, whose twice as many workgroups matter more there than the
  // per-column cost.
  const bool mfma_update = batch >= 16 || b.max_n >= 512;
  for (int j0 = 0; j0 < kmax; j0 += QB)
  {
    // A single small matrix: up to 64 rows on one wave (the per-column
    // reductions need no cross-wave barrier), up to 256 one row per thread
    // (measured: 66 against 110 us per 16-column panel with four rows per
    // lane of one wave).
    if (b.max_m > 4 * NT)
      hipLaunchKernelGGL((qr_panel_tall_kernel<NT>), dim3(batch), dim3(NT), 0, st, b, j0);
    else if (batch == 1 && b.max_m <= 64)
      hipLaunchKernelGGL((qr_panel_kernel<64, 1>), dim3(1), dim3(64), 0, st, b, j0);
    else if (batch == 1 && b.max_m <= 256)
      hipLaunchKernelGGL((qr_panel_kernel<256, 1>), dim3(1), dim3(256), 0, st, b, j0);
    else if (b.max_m <= NT)
      hipLaunchKernelGGL((qr_panel_kernel<NT, 1>), dim3(batch), dim3(NT), 0, st, b, j0);
    else if (b.max_m <= 2 * NT)
      hipLaunchKernelGGL((qr_panel_kernel<NT, 2>), dim3(batch), dim3(NT), 0, st, b, j0);
    else
      hipLaunchKernelGGL((qr_panel_kernel<NT, 4>), dim3(batch), dim3(NT), 0, st, b, j0);
    // Columns right of the panel: for the single problem `one`, exactly
    // those beyond its panel width (min(m, n) - j0 capped at QB); otherwise
    // the update runs whenever any column may remain.
    int ncols = b.max_n - j0 - 1;
    if (!b.descs) ncols = b.one.n - j0 - std::min(std::min(b.one.m, b.one.n) - j0, QB);
    if (ncols > 0)
    {
      if (mfma_update)
        hipLaunchKernelGGL(qr_update_mfma_kernel, dim3((ncols + UCT - 1) / UCT, batch), dim3(256), 0, st, b, j0, 0, 0);
      else
        hipLaunchKernelGGL(qr_update_kernel, dim3((ncols + QCT - 1) / QCT, batch), dim3(256), 0, st, b, j0, 0);
    }
  }
  if (b.descs ? b.form_q : b.one.Q != nullptr)
  {
    if (!mfma_update && kmax <= QB)
    {
      // One panel: its update reads C as the identity (mode 2), so Q needs
      // no separate fill.
      hipLaunchKernelGGL(qr_update_kernel, dim3((kmax + QCT - 1) / QCT, batch), dim3(256), 0, st, b, 0, 2);
      return;
    }
    const long total = (long)b.max_m * kmax;
    long g = (total + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(qr_eye_kernel, dim3((int)g, batch), dim3(256), 0, st, b);
    const int last = ((kmax - 1) / QB) * QB;
    for (int j0 = last; j0 >= 0; j0 -= QB)
    {
      if (mfma_update)
        hipLaunchKernelGGL(qr_update_mfma_kernel, dim3((kmax - j0 + UCT - 1) / UCT, batch), dim3(256), 0, st, b, j0, 1,
                           0);
      else
        hipLaunchKernelGGL(qr_update_kernel, dim3((kmax - j0 + QCT - 1) / QCT, batch), dim3(256), 0, st, b, j0, 1);
    }
  }
}

}  // namespace tta
