// Device-side generators for the batched benchmark inputs (tens of GB of
// N(0,1) core entries, too many to draw on the host).  Entries come from the
// counter-based Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw, SC'11)
// keyed by a per-TT seed and indexed by (core, element), turned into a
// normal deviate by Box-Muller.  Values differ from std::mt19937_64 +
// std::normal_distribution (the single-TT generator), with the same
// distribution; every value is reproducible from (seed, core, element).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

namespace tta
{
namespace
{

__device__ __forceinline__ void philox_round(uint32_t (&c)[4], uint32_t (&k)[2])
{
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
  const uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
  k[0] += 0x9E3779B9u;
  k[1] += 0xBB67AE85u;
}

// N(0,1) deviate for (seed, stream, index).
__device__ __forceinline__ double philox_normal(unsigned long long seed, unsigned stream, unsigned long long index)
{
  uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0x54545454u};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma unroll
  for (int r = 0; r < 10; ++r) philox_round(c, k);
  const double u1 = ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6)) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(c[2] >> 5) * 67108864.0 + (double)(c[3] >> 6)) * (1.0 / 9007199254740992.0);
  const double rr = sqrt(-2.0 * log(1.0 - u1));
  return rr * cos(6.283185307179586 * u2);
}

// One (TT, core) of c = add(u, add(scale(u, 0.5), v)) (block concatenation,
// include/core/detail/block_concat.hpp:40-156, scalars in the head core).
struct sum_desc
{
  double* dst;
  int d, k, n;
  int ul, ur, vl, vr;  // ranks of u and v around core k
  unsigned long long su, sv;
};

__global__ void sum_build_kernel(const sum_desc* ds)
{
  const sum_desc& D = ds[blockIdx.y];
  const bool head = D.k == 0, tail = D.k == D.d - 1;
  const int ol = head ? 1 : 2 * D.ul + D.vl;
  const int orr = tail ? 1 : 2 * D.ur + D.vr;
  const long total = (long)ol * D.n * orr;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int b = (int)(e % orr);
    const int i = (int)((e / orr) % D.n);
    const int a = (int)(e / ((long)orr * D.n));
    double v = 0.0;
    if (D.d == 1)
    {
      v = 1.5 * philox_normal(D.su, 0, (unsigned long long)i) + philox_normal(D.sv, 0, (unsigned long long)i);
    }
    else
    {
      // Source block: 0 = u, 1 = 0.5 u (head) / u copy, 2 = v; -1 = zero.
      int blk = -1, sa = 0, sb = 0;
      if (head)
      {
        if (b < D.ur) blk = 0, sb = b;
        else if (b < 2 * D.ur) blk = 1, sb = b - D.ur;
        else blk = 2, sb = b - 2 * D.ur;
      }
      else if (tail)
      {
        if (a < D.ul) blk = 0, sa = a;
        else if (a < 2 * D.ul) blk = 1, sa = a - D.ul;
        else blk = 2, sa = a - 2 * D.ul;
      }
      else
      {
        if (a < D.ul && b < D.ur) blk = 0, sa = a, sb = b;
        else if (a >= D.ul && a < 2 * D.ul && b >= D.ur && b < 2 * D.ur) blk = 1, sa = a - D.ul, sb = b - D.ur;
        else if (a >= 2 * D.ul && b >= 2 * D.ur) blk = 2, sa = a - 2 * D.ul, sb = b - 2 * D.ur;
      }
      if (blk == 0 || blk == 1)
      {
        const int sr = tail ? 1 : D.ur;
        const unsigned long long idx = ((unsigned long long)sa * D.n + i) * sr + sb;
        v = philox_normal(D.su, (unsigned)D.k, idx);
        if (blk == 1 && head) v *= 0.5;
      }
      else if (blk == 2)
      {
        const int sr = tail ? 1 : D.vr;
        const unsigned long long idx = ((unsigned long long)sa * D.n + i) * sr + sb;
        v = philox_normal(D.sv, (unsigned)D.k, idx);
      }
    }
    D.dst[e] = v;
  }
}

// Column-major Gaussian (m x nc) for (seed, core), FP64 (dst) or FP32
// (dst32, when set).
struct gauss_desc
{
  double* dst;
  int m, nc;
  unsigned long long seed;
  int k;
  float* dst32;
};

__global__ void gauss_kernel(const gauss_desc* ds)
{
  const gauss_desc& D = ds[blockIdx.y];
  const long total = (long)D.m * D.nc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const double v = philox_normal(D.seed, (unsigned)D.k, (unsigned long long)e);
    if (D.dst32)
      D.dst32[e] = (float)v;
    else
      D.dst[e] = v;
  }
}

// core (m x nc row-major) = Q (m x nc column-major) * diag(10^(-j s)); for
// the last core (nc = 1) the Gaussian scaled to unit norm (scale given).
struct qs_desc
{
  const double* Q;
  double* dst;
  int m, nc;
  double s;       // decay exponent per column
  double scale;   // extra factor (1 / norm for the last core)
};

__global__ void qscale_kernel(const qs_desc* ds)
{
  const qs_desc& D = ds[blockIdx.y];
  const long total = (long)D.m * D.nc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
  {
    const int row = (int)(e / D.nc), j = (int)(e % D.nc);
    D.dst[e] = D.Q[(size_t)j * D.m + row] * exp10(-(double)j * D.s) * D.scale;
  }
}

template <class T>
T* upload(tta_ctx* c, const std::vector<T>& v, std::vector<void*>& own)
{
  void* p = c->mem.get(std::max<size_t>(v.size(), 1) * sizeof(T));
  own.push_back(p);
  upload_async(c, p, v.data(), v.size() * sizeof(T));
  return static_cast<T*>(p);
}

// Rank rule of random() (tt_factory.hpp:155-190) for [n]^d, max_rank.
std::vector<int> rand_ranks(int d, int n, int mr)
{
  std::vector<int> r(d + 1, 1);
  for (int k = 0; k < d; ++k)
  {
    long long cap = (long long)r[k] * n;
    if (cap > mr && mr > 0) cap = mr;
    r[k + 1] = (int)cap;
  }
  for (int k = d - 1; k >= 0; --k)
  {
    long long cap = (long long)r[k + 1] * n;
    if (cap > mr && mr > 0) cap = mr;
    if (cap < r[k]) r[k] = (int)cap;
  }
  r[0] = 1;
  r[d] = 1;
  return r;
}

tta_tt* alloc_tt(tta_ctx* c, const std::vector<int>& modes, const std::vector<int>& ranks)
{
  auto* t = new tta_tt();
  t->modes = modes;
  t->ranks = ranks;
  for (size_t k = 0; k < modes.size(); ++k)
    t->cores.push_back(static_cast<double*>(
      c->mem.get(std::max<size_t>((size_t)ranks[k] * modes[k] * ranks[k + 1], 1) * sizeof(double))));
  return t;
}

}  // namespace

// Row-major N(0,1) blocks (m x nc) for (seed, stream k), one per descriptor.
void launch_gauss_batched(void* const* dst, const int* m, const int* nc, const unsigned long long* seed, const int* k,
                          int count, bool f32, tta_ctx* c)
{
  std::vector<gauss_desc> gd;
  long max_el = 0;
  for (int i = 0; i < count; ++i)
  {
    gd.push_back(f32 ? gauss_desc{nullptr, m[i], nc[i], seed[i], k[i], static_cast<float*>(dst[i])}
                     : gauss_desc{static_cast<double*>(dst[i]), m[i], nc[i], seed[i], k[i], nullptr});
    max_el = std::max(max_el, (long)m[i] * nc[i]);
  }
  std::vector<void*> own;
  const gauss_desc* dd = upload(c, gd, own);
  const int g = (int)std::min<long>((max_el + 255) / 256, 64);
  for (size_t off = 0; off < gd.size(); off += 65535)
  {
    const int chunk = (int)std::min<size_t>(65535, gd.size() - off);
    hipLaunchKernelGGL(gauss_kernel, dim3(g < 1 ? 1 : g, chunk), dim3(256), 0, c->stream, dd + off);
  }
  for (void* p : own) c->mem.put(p);  // stream-ordered reuse
}

}  // namespace tta

using namespace tta;

extern "C"
{

int tta_build_sum_batch(tta_ctx* c, int count, const int* d, const int* n, const int* ru, const int* rv,
                        const unsigned long long* seed_u, const unsigned long long* seed_v, tta_tt** out)
{
  if (!c || count < 0 || (count > 0 && (!d || !n || !ru || !rv || !seed_u || !seed_v || !out))) return TTA_ERR_ARG;
  std::vector<void*> own;
  try
  {
    std::vector<sum_desc> ds;
    long max_el = 0;
    for (int i = 0; i < count; ++i)
    {
      if (d[i] < 1 || n[i] < 1 || ru[i] < 1 || rv[i] < 1) fail(TTA_ERR_SHAPE, "build_sum_batch: bad shape %d", i);
      const std::vector<int> Ru = rand_ranks(d[i], n[i], ru[i]), Rv = rand_ranks(d[i], n[i], rv[i]);
      std::vector<int> modes(d[i], n[i]), ranks(d[i] + 1, 1);
      for (int k = 1; k < d[i]; ++k) ranks[k] = 2 * Ru[k] + Rv[k];
      tta_tt* t = alloc_tt(c, modes, ranks);
      out[i] = t;
      for (int k = 0; k < d[i]; ++k)
      {
        sum_desc s{t->cores[k], d[i], k, n[i], Ru[k], Ru[k + 1], Rv[k], Rv[k + 1], seed_u[i], seed_v[i]};
        ds.push_back(s);
        max_el = std::max(max_el, (long)ranks[k] * n[i] * ranks[k + 1]);
      }
    }
    if (!ds.empty())
    {
      const sum_desc* dd = upload(c, ds, own);
      const int g = (int)std::min<long>((max_el + 255) / 256, 64);
      // grid.y is limited to 65535: launch in chunks of descriptors.
      for (size_t off = 0; off < ds.size(); off += 65535)
      {
        const int chunk = (int)std::min<size_t>(65535, ds.size() - off);
        hipLaunchKernelGGL(sum_build_kernel, dim3(g < 1 ? 1 : g, chunk), dim3(256), 0, c->stream, dd + off);
      }
    }
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  catch (const tta::error& e)
  {
    for (void* p : own) c->mem.put(p);
    c->last_error = e.what();
    return e.code;
  }
  for (void* p : own) c->mem.put(p);
  return TTA_OK;
}

int tta_build_decay_batch(tta_ctx* c, int count, int d, int n, int rmax, const double* s,
                          const unsigned long long* seeds, tta_tt** out)
{
  if (!c || count < 0 || d < 2 || n < 1 || rmax < 1 || (count > 0 && (!s || !seeds || !out))) return TTA_ERR_ARG;
  std::vector<void*> own;
  try
  {
    std::vector<int> ranks(d + 1, 1), modes(d, n);
    for (int k = 1; k < d; ++k)
    {
      double a = std::pow((double)n, (double)k), b = std::pow((double)n, (double)(d - k));
      ranks[k] = (int)std::min<double>(std::min(a, b), (double)rmax);
    }
    // Gaussian column-major blocks for every (TT, core), QR of cores < d-1.
    std::vector<gauss_desc> gd;
    std::vector<qrp> qr;
    std::vector<qs_desc> qs;
    std::vector<double*> tmp;
    int max_m = 0, max_n = 0;
    long max_el = 0;
    for (int i = 0; i < count; ++i)
    {
      tta_tt* t = alloc_tt(c, modes, ranks);
      out[i] = t;
      for (int k = 0; k < d; ++k)
      {
        const int m = ranks[k] * n, nc = ranks[k + 1];
        double* G = static_cast<double*>(c->mem.get((size_t)m * nc * 8));
        own.push_back(G);
        gd.push_back(gauss_desc{G, m, nc, seeds[i], k});
        max_el = std::max(max_el, (long)m * nc);
        if (k < d - 1)
        {
          qrp p;
          p.A = G;
          p.lda = m;
          p.m = m;
          p.n = nc;
          p.tau = static_cast<double*>(c->mem.get((size_t)nc * 8));
          p.T = static_cast<double*>(c->mem.get((size_t)((nc + 15) / 16) * 256 * 8));
          p.Q = static_cast<double*>(c->mem.get((size_t)m * nc * 8));
          p.ldq = m;
          own.push_back(p.tau);
          own.push_back(p.T);
          own.push_back(p.Q);
          qr.push_back(p);
          qs.push_back(qs_desc{p.Q, t->cores[k], m, nc, s[i], 1.0});
          max_m = std::max(max_m, m);
          max_n = std::max(max_n, nc);
        }
        else
        {
          qs.push_back(qs_desc{G, t->cores[k], m, nc, 0.0, 0.0});  // scale set below
        }
      }
    }
    const gauss_desc* gdd = upload(c, gd, own);
    const int g = (int)std::min<long>((max_el + 255) / 256, 64);
    for (size_t off = 0; off < gd.size(); off += 65535)
    {
      const int chunk = (int)std::min<size_t>(65535, gd.size() - off);
      hipLaunchKernelGGL(gauss_kernel, dim3(g < 1 ? 1 : g, chunk), dim3(256), 0, c->stream, gdd + off);
    }
    // The last core is a Gaussian vector of length r_{d-1} n; its norm is
    // read back so the TT has unit-norm last core.
    std::vector<sq_desc> sq;
    double* nrm = static_cast<double*>(c->mem.get((size_t)std::max(count, 1) * 8));
    own.push_back(nrm);
    for (int i = 0; i < count; ++i)
    {
      const qs_desc& last = qs[(size_t)i * d + (d - 1)];
      sq.push_back(sq_desc{last.Q, (long)last.m * last.nc, nrm + i, nullptr, 0.0});
    }
    launch_frob_batched(upload(c, sq, own), count, c->stream);
    for (size_t off = 0; off < qr.size(); off += 65535)
    {
      qr_batch b;
      std::vector<qrp> part(qr.begin() + (long)off, qr.begin() + (long)std::min(qr.size(), off + 65535));
      b.descs = upload(c, part, own);
      b.batch = (int)part.size();
      b.max_m = max_m;
      b.max_n = max_n;
      b.form_q = true;
      launch_qr_blocked(b, c->stream);
    }
    std::vector<double> hn(count);
    hip_check(hipMemcpyAsync(hn.data(), nrm, (size_t)count * 8, hipMemcpyDeviceToHost, c->stream), "norms");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (int i = 0; i < count; ++i) qs[(size_t)i * d + (d - 1)].scale = hn[i] > 0 ? 1.0 / hn[i] : 1.0;
    const qs_desc* qsd = upload(c, qs, own);
    for (size_t off = 0; off < qs.size(); off += 65535)
    {
      const int chunk = (int)std::min<size_t>(65535, qs.size() - off);
      hipLaunchKernelGGL(qscale_kernel, dim3(g < 1 ? 1 : g, chunk), dim3(256), 0, c->stream, qsd + off);
    }
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  catch (const tta::error& e)
  {
    for (void* p : own) c->mem.put(p);
    c->last_error = e.what();
    return e.code;
  }
  for (void* p : own) c->mem.put(p);
  return TTA_OK;
}

}  // extern "C"
