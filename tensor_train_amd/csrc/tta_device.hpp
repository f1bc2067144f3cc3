// Device-side building blocks for TT rounding on gfx950 (CDNA4, wave64).
//
// Matrices handed to these routines are column-major with a leading
// dimension unless stated otherwise.  A TT core (r, n, rp) stored row-major
// is, byte for byte:
//   * right_unfold  (r x n*rp) row-major  ==  (n*rp x r) column-major;
//   * left_unfold   (r*n x rp) row-major  ==  (rp x r*n) column-major.
// so the right-to-left QR sweep works on the core bytes directly.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace tta
{
namespace dev
{

constexpr int kWave = 64;

// One DPP move of a double (two 32-bit halves); `ctrl` is a DPP control
// word (quad_perm, row_ror, ...), all rows and banks enabled.
template <int ctrl>
__device__ __forceinline__ double dpp_mov(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wave, returned (wave-uniform) to every lane.
// Within each 16-lane row: quad_perm [1,0,3,2] and [2,3,0,1] sum each quad,
// row_ror:4 and row_ror:8 sum the four quads; the four row sums are then
// read from lanes 0, 16, 32, 48.  All on the ALU/DPP path (no LDS).
__device__ __forceinline__ double wave_sum(double v)
{
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x128>(v);  // row_ror:8
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Sums of 16 values over the 64 lanes of a wave at once: lane l returns the
// sum over all lanes of v[l & 15].  Within each 16-lane row a butterfly
// halves the values at each step, with partners l ^ 15, l ^ 7, l ^ 3 and
// l ^ 1 (DPP row_mirror, row_half_mirror, quad_perm [3,2,1,0] and
// [1,0,3,2]): a lane keeps the half picked by its own bit 3, 2, 1, 0 and
// adds the partner's copy of that half, which the partner gives away because
// its bit is the opposite.  The four rows are then added by two lane
// exchanges (xor 16, xor 32).  15 DPP moves instead of 16 separate
// reductions.
__device__ __forceinline__ double wave_vsum16(const double (&v)[16])
{
  const int lane = threadIdx.x & (kWave - 1);
  double a[8], b[4], c[2];
  const bool h3 = lane & 8, h2 = lane & 4, h1 = lane & 2, h0 = lane & 1;
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (h3 ? v[8 + i] : v[i]) + dpp_mov<0x140>(h3 ? v[i] : v[8 + i]);
#pragma unroll
  for (int i = 0; i < 4; ++i) b[i] = (h2 ? a[4 + i] : a[i]) + dpp_mov<0x141>(h2 ? a[i] : a[4 + i]);
#pragma unroll
  for (int i = 0; i < 2; ++i) c[i] = (h1 ? b[2 + i] : b[i]) + dpp_mov<0x1B>(h1 ? b[i] : b[2 + i]);
  double r = (h0 ? c[1] : c[0]) + dpp_mov<0xB1>(h0 ? c[0] : c[1]);
  r += __shfl_xor(r, 16, kWave);
  r += __shfl_xor(r, 32, kWave);
  return r;
}

// Reference reduction through ds_bpermute, kept for cross-checking the DPP
// path in the tests.
__device__ __forceinline__ double wave_sum_shfl(double v)
{
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Block-wide sum; `red` must hold at least NT / 64 doubles of LDS.  Every
// thread returns the total.  Contains two barriers.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red)
{
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w) t += red[w];
  return t;
}

// ---------------------------------------------------------------------------
// Householder QR (unblocked, LAPACK dgeqr2 semantics) of a column-major
// (m x nc) matrix A with leading dimension lda, by one workgroup of NT
// threads.  On return the upper trapezoid of A holds R, the part below the
// diagonal holds the essential Householder vectors (v[0] = 1 implicit) and
// tau[j] the scalar factors, j < k = min(m, nc).  Sign convention of dlarfg
// (and Eigen's makeHouseholder): beta = -sign(alpha) * ||x||.
// `vbuf` is LDS of at least m doubles; `red` LDS of NT/64 doubles.
// ---------------------------------------------------------------------------
template <int NT>
__device__ void geqr2(double* __restrict__ A, int lda, int m, int nc, double* __restrict__ tau,
                      double* vbuf, double* red)
{
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  constexpr int NW = NT / kWave;
  const int k = m < nc ? m : nc;
  for (int j = 0; j < k; ++j)
  {
    double* colj = A + (size_t)j * lda + j;
    const int len = m - j;
    double acc = 0.0;
    for (int i = tid; i < len; i += NT)
    {
      const double x = colj[i];
      vbuf[i] = x;
      if (i > 0) acc += x * x;
    }
    const double tail = block_sum<NT>(acc, red);  // barrier also publishes vbuf
    const double alpha = vbuf[0];
    double tj, scal, beta;
    if (tail <= DBL_MIN)
    {
      tj = 0.0;
      scal = 0.0;
      beta = alpha;
    }
    else
    {
      beta = sqrt(alpha * alpha + tail);
      if (alpha >= 0.0) beta = -beta;
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();
    for (int i = tid; i < len; i += NT)
    {
      const double v = (i == 0) ? 1.0 : vbuf[i] * scal;
      vbuf[i] = v;
      colj[i] = (i == 0) ? beta : v;
    }
    if (tid == 0) tau[j] = tj;
    __syncthreads();
    if (tj != 0.0)
    {
      // Trailing update: one wave per column.
      for (int c = j + 1 + wid; c < nc; c += NW)
      {
        double* y = A + (size_t)c * lda + j;
        double dot = 0.0;
        for (int i = lane; i < len; i += kWave) dot += vbuf[i] * y[i];
        dot = wave_sum(dot) * tj;
        for (int i = lane; i < len; i += kWave) y[i] -= dot * vbuf[i];
      }
    }
    __syncthreads();
  }
}

// Form the thin Q (m x k, column-major, leading dimension ldq) of a QR
// factorisation stored by geqr2 (LAPACK dorg2r: backward accumulation of the
// reflectors on the identity).  `vbuf` LDS of m doubles.
template <int NT>
__device__ void org2r(const double* __restrict__ A, int lda, int m, int k, const double* __restrict__ tau,
                      double* __restrict__ Q, int ldq, double* vbuf)
{
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  constexpr int NW = NT / kWave;
  for (int c = 0; c < k; ++c)
    for (int i = tid; i < m; i += NT) Q[(size_t)c * ldq + i] = (i == c) ? 1.0 : 0.0;
  __syncthreads();
  for (int j = k - 1; j >= 0; --j)
  {
    const double tj = tau[j];
    if (tj == 0.0) continue;
    const int len = m - j;
    for (int i = tid; i < len; i += NT) vbuf[i] = (i == 0) ? 1.0 : A[(size_t)j * lda + j + i];
    __syncthreads();
    for (int c = j + wid; c < k; c += NW)
    {
      double* y = Q + (size_t)c * ldq + j;
      double dot = 0.0;
      for (int i = lane; i < len; i += kWave) dot += vbuf[i] * y[i];
      dot = wave_sum(dot) * tj;
      for (int i = lane; i < len; i += kWave) y[i] -= dot * vbuf[i];
    }
    __syncthreads();
  }
}

// Hardware reciprocal and reciprocal square root, each refined by two
// Newton steps (within a few ulps).
__device__ __forceinline__ double rcp_newton(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}

// 1 / sqrt(y) for 1e-300 <= y <= 1e300 (y r and y r^2 stay in range).
__device__ __forceinline__ double rsqrt_newton(double y)
{
  double r = __builtin_amdgcn_rsq(y);
  r = r * fma(-0.5 * y * r, r, 1.5);
  return r * fma(-0.5 * y * r, r, 1.5);
}

// Jacobi rotation that orthogonalises columns x_i, x_j from a = |x_i|^2,
// b = |x_j|^2, g = x_i . x_j: the rotation x_i' = cs x_i - sn x_j,
// x_j' = sn x_i + cs x_j with tangent t the smaller root of
// t^2 + 2 zeta t - 1 = 0, zeta = (b - a) / (2 g) (Rutishauser's rotation).
// Returns false (no rotation) when |g| <= tol |x_i| |x_j|, when a squared
// norm underflowed to zero (the column is negligible and its products would
// never settle), when g is not finite, or when the angle rounds to zero; so
// every rotation reported changes the columns.  tan_out receives t.
// The rotation sits on the critical path of every inner Jacobi round:
//   * squared norms within 1e+-100: two reciprocal square roots.  With
//     d = b - a and h = sqrt(d^2 + 4 g^2) (in [4e-230, 5e200] there),
//     cos 2theta = |d| / h, cs^2 = x = (1 + |d| / h) / 2,
//     sn = sign(d) g / (h cs) and t = sn / cs; cs^2 + sn^2 = 1 to rounding
//     (at d = 0 the angle is 45 degrees with t = sign(g));
//   * within 1e+-140: t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
//     cs = 1 / sqrt(1 + t^2), by the hardware reciprocal and reciprocal
//     square root with Newton steps, the threshold on squares (a b can
//     neither overflow nor underflow);
//   * other scales: the same formulas correctly rounded; for
//     |zeta| > 1e150 the tangent takes its asymptotic form 1 / (2 zeta),
//     where zeta^2 would overflow.
__device__ __forceinline__ bool jacobi_rotation(double a, double b, double g, double tol, double& cs, double& sn,
                                                double& tan_out)
{
  if (a == 0.0 || b == 0.0) return false;
  if (a > 1.0e-100 && a < 1.0e100 && b > 1.0e-100 && b < 1.0e100)
  {
    if (!(g * g > (tol * tol) * (a * b))) return false;
    const double d = b - a;
    const double r = rsqrt_newton(fma(d, d, 4.0 * (g * g)));  // 1 / h
    const double x = fma(0.5 * fabs(d), r, 0.5);                // cs^2
    const double ri = rsqrt_newton(x);                         // 1 / cs
    const double gs = d >= 0.0 ? g : -g;
    cs = x * ri;
    sn = (gs * r) * ri;
    tan_out = sn * ri;
    return sn != 0.0;
  }
  if (!(a > 1.0e-140 && a < 1.0e140 && b > 1.0e-140 && b < 1.0e140))
  {
    // Extreme scales: the correctly rounded operations throughout.
    if (!(fabs(g) > tol * sqrt(a) * sqrt(b))) return false;
    const double zeta = (b - a) / (2.0 * g);
    const double az = fabs(zeta);
    const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (az > 1.0e150 ? 2.0 * az : az + sqrt(1.0 + zeta * zeta));
    cs = 1.0 / sqrt(1.0 + t * t);
    sn = cs * t;
    tan_out = t;
    return sn != 0.0;
  }
  // Here |g| > tol sqrt(a b) > 1e-155 once past the test: no denormal g.
  if (!(g * g > (tol * tol) * (a * b))) return false;
  const double zeta = 0.5 * (b - a) * rcp_newton(g);
  const double az = fabs(zeta);
  double den;
  if (az > 1.0e150)
    den = 2.0 * az;
  else
  {
    const double y = fma(zeta, zeta, 1.0);
    den = az + y * rsqrt_newton(y);
  }
  const double t = (zeta >= 0.0 ? 1.0 : -1.0) * rcp_newton(den);
  cs = rsqrt_newton(fma(t, t, 1.0));
  sn = cs * t;
  tan_out = t;
  return sn != 0.0;
}

// The same without the tangent t = sn / cs.
__device__ __forceinline__ bool jacobi_rotation(double a, double b, double g, double tol, double& cs, double& sn)
{
  double t;
  return jacobi_rotation(a, b, g, tol, cs, sn, t);
}

// Circle-method round-robin pairing of qq (even) players: round s in
// [0, qq-1), pair p in [0, qq/2).  Every unordered pair meets exactly once
// per qq-1 rounds and the pairs of one round are disjoint.
__device__ __forceinline__ void rr_pair(int s, int p, int qq, int& i, int& j)
{
  const int mdl = qq - 1;
  if (p == 0)
  {
    i = mdl;
    j = s;
  }
  else
  {
    i = (s + p) % mdl;
    j = (s - p + mdl) % mdl;
  }
}

// ---------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi on the q columns of X (p x q, column-major,
// leading dimension ldx), accumulating the rotations into J (q x q,
// column-major, ldj): on return X_in * J = X_out with mutually orthogonal
// columns and J orthogonal.  Parallel ordering: each round applies q/2
// disjoint rotations, one wave per pair.  Convergence: a pair is rotated only
// while |x_i . x_j| > tol * ||x_i|| ||x_j||, tol = max(1e-15, sqrt(p) * u);
// stops after a sweep with no rotation or after max_sweeps.
// Returns the number of sweeps run (to thread 0 via the return value).
// ---------------------------------------------------------------------------
template <int NT>
__device__ int jacobi_sweeps(double* __restrict__ X, int ldx, int p, int q, double* __restrict__ J, int ldj,
                             int max_sweeps, int* flag_lds)
{
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  constexpr int NW = NT / kWave;
  for (int c = 0; c < q; ++c)
    for (int i = tid; i < q; i += NT) J[(size_t)c * ldj + i] = (i == c) ? 1.0 : 0.0;
  const double tol = fmax(1.0e-15, sqrt((double)p) * 2.220446049250313e-16);
  const int qq = (q + 1) & ~1;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep)
  {
    if (tid == 0) *flag_lds = 0;
    __syncthreads();
    int rotated = 0;
    for (int s = 0; s < qq - 1; ++s)
    {
      for (int pr = wid; pr < qq / 2; pr += NW)
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
        double* xi = X + (size_t)ci * ldx;
        double* xj = X + (size_t)cj * ldx;
        double a = 0.0, b = 0.0, g = 0.0;
        for (int t = lane; t < p; t += kWave)
        {
          const double u = xi[t], v = xj[t];
          a += u * u;
          b += v * v;
          g += u * v;
        }
        a = wave_sum(a);
        b = wave_sum(b);
        g = wave_sum(g);
        double cs, sn;
        if (!jacobi_rotation(a, b, g, tol, cs, sn)) continue;
        rotated = 1;
        for (int t = lane; t < p; t += kWave)
        {
          const double u = xi[t], v = xj[t];
          xi[t] = cs * u - sn * v;
          xj[t] = sn * u + cs * v;
        }
        double* ji = J + (size_t)ci * ldj;
        double* jj = J + (size_t)cj * ldj;
        for (int t = lane; t < q; t += kWave)
        {
          const double u = ji[t], v = jj[t];
          ji[t] = cs * u - sn * v;
          jj[t] = sn * u + cs * v;
        }
      }
      __syncthreads();
    }
    if (rotated && lane == 0) atomicOr(flag_lds, 1);
    __syncthreads();
    const int any = *flag_lds;
    __syncthreads();
    if (!any) break;
  }
  return sweep;
}

// Norms of the q columns of W (p x q col-major), then a stable descending
// order: perm[o] = index of the o-th largest, s[o] its norm.  `nrm` scratch
// of q doubles (global or LDS).
template <int NT>
__device__ void column_norms_sorted(const double* __restrict__ W, int ldw, int p, int q, double* nrm,
                                    double* __restrict__ s, int* __restrict__ perm)
{
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  constexpr int NW = NT / kWave;
  for (int c = wid; c < q; c += NW)
  {
    const double* w = W + (size_t)c * ldw;
    double acc = 0.0;
    for (int t = lane; t < p; t += kWave) acc += w[t] * w[t];
    acc = wave_sum(acc);
    if (lane == 0) nrm[c] = sqrt(acc);
  }
  __syncthreads();
  // A total order, so perm is always a permutation of 0..q-1 and every s[o]
  // is written: norms descending, a NaN norm (from a breakdown upstream)
  // ordered after every number (key -1; norms are >= 0), ties by index.
  for (int c = tid; c < q; c += NT)
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
    perm[rank] = c;
    s[rank] = v;
  }
  __syncthreads();
}

// truncate_eps_rank: sequential tail sum from the smallest value, strict '>'
// stops; at least 1; capped by max_rank.  Thread 0 only.
__device__ __forceinline__ int truncate_rank(const double* s, int n, double delta, int max_rank)
{
  if (n == 0) return 1;
  const double thresh = delta * delta;
  double tail2 = 0.0;
  int r_new = n;
  for (int i = n - 1; i >= 0; --i)
  {
    const double next = tail2 + s[i] * s[i];
    if (next > thresh) break;
    tail2 = next;
    r_new = i;
  }
  if (r_new < 1) r_new = 1;
  if (max_rank > 0 && r_new > max_rank) r_new = max_rank;
  return r_new;
}

}  // namespace dev
}  // namespace tta
