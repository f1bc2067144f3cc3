// Learned-rank rounding: per-bond features of the input TTs, the MLP's
// predicted truncation ranks, and the randomized rounding with those target
// ranks (plain fallback when the discarded norm exceeds eps).  No reference
// counterpart.
//
// Features of bond k (1 <= k < d) of a TT, from the column norms c_0 >= c_1
// >= ... of the left unfolding of core k-1 (r_{k-1} n x r_k), computed on the
// device:
//   [ log2(r_k)/8, log2(n_{k-1})/3, k/d, -log10(eps)/10,
//     -log10(eps / sqrt(d-1))/10,
//     log10(c_j / c_0)/10 for j = 1, 2, 4, 8, 16, 32, 64 (j clamped to r_k-1),
//     log10(c_{r_k-1} / c_0)/10 ]                                (13 values)
// The label is the plain rounding's new rank at that bond over r_k.  The
// target rank is ceil(pred r_k (1 + slack)) + oversample, clipped to [1, r_k].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

namespace tta
{
void round_batched_plain(tta_ctx* c, std::vector<tta_tt*>& ts, double eps, int max_rank, bool mixed);
}

using namespace tta;

namespace tta
{
void mlp_predict_dev(tta_ctx* c, tta_mlp* m, const float* X_dev, int n, float* y_host);
}

namespace
{
constexpr int NF = 13;
constexpr int FMAX = 4096;  // largest bond rank handled (sort in LDS)

struct bf_desc
{
  const double* M;  // left unfolding of core k-1: rows x cols row-major
  int rows, cols;   // r_{k-1} n_{k-1}, r_k
  int n, k, d;      // n_{k-1}, bond index, TT length
};

__device__ __forceinline__ float feat_log(double x, double c0)
{
  return (float)(log10(fmax(x / c0, 1e-300)) / 10.0);
}

// One workgroup per bond: column norms (thread per column, rows walked in
// order, so the wave's loads are contiguous and the sums deterministic),
// bitonic sort descending in LDS, then the 13 features.
__global__ __launch_bounds__(256) void bond_feat_kernel(const bf_desc* ds, double eps, float* F)
{
  const bf_desc& D = ds[blockIdx.x];
  __shared__ double v[FMAX];
  const int tid = threadIdx.x;
  int P = 1;
  while (P < D.cols) P <<= 1;
  for (int c = tid; c < P; c += 256)
  {
    double acc = -1.0;  // padding sorts last
    if (c < D.cols)
    {
      acc = 0.0;
      for (int i = 0; i < D.rows; ++i)
      {
        const double x = D.M[(size_t)i * D.cols + c];
        acc = fma(x, x, acc);
      }
      acc = sqrt(acc);
    }
    v[c] = acc;
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1)
    {
      for (int i = tid; i < P; i += 256)
      {
        const int j = i ^ stride;
        if (j > i)
        {
          const bool desc = (i & size) == 0;
          const double a = v[i], b = v[j];
          if (desc ? (a < b) : (a > b))
          {
            v[i] = b;
            v[j] = a;
          }
        }
      }
      __syncthreads();
    }
  if (tid == 0)
  {
    const int r = D.cols;
    const double c0 = v[0] > 0.0 ? v[0] : 1.0;
    float* f = F + (size_t)blockIdx.x * NF;
    f[0] = (float)(log2((double)r) / 8.0);
    f[1] = (float)(log2((double)D.n) / 3.0);
    f[2] = (float)D.k / (float)D.d;
    f[3] = (float)(-log10(eps) / 10.0);
    f[4] = (float)(-log10(eps / sqrt((double)max(D.d - 1, 1))) / 10.0);
    const int js[7] = {1, 2, 4, 8, 16, 32, 64};
    for (int q = 0; q < 7; ++q) f[5 + q] = feat_log(v[min(js[q], r - 1)], c0);
    f[12] = feat_log(v[r - 1], c0);
  }
}

// Features of every bond k = 1..d-1 of every TT, into device memory
// (bonds in TT order, then bond order); returns the bond count.
int bond_features_dev(tta_ctx* c, int count, tta_tt* const* in, double eps, arena& A, float** F_out)
{
  std::vector<bf_desc> ds;
  for (int i = 0; i < count; ++i)
    for (int k = 1; k < in[i]->d(); ++k)
    {
      const tta_tt* t = in[i];
      if (t->ranks[k] > FMAX) fail(TTA_ERR_UNSUPPORTED, "bond features: rank %d above %d", t->ranks[k], FMAX);
      ds.push_back(bf_desc{t->cores[k - 1], t->ranks[k - 1] * t->modes[k - 1], t->ranks[k], t->modes[k - 1], k, t->d()});
    }
  const int nb = (int)ds.size();
  float* F = A.alloc<float>((size_t)std::max(nb, 1) * NF);
  if (nb > 0)
  {
    const bf_desc* dd = A.upload(ds);
    for (int o = 0; o < nb; o += 65535)
      hipLaunchKernelGGL(bond_feat_kernel, dim3(std::min(65535, nb - o)), dim3(256), 0, c->stream, dd + o, eps,
                         F + (size_t)o * NF);
  }
  *F_out = F;
  return nb;
}
}  // namespace

extern "C"
{

int tta_num_bond_features() { return NF; }

int tta_bond_features(tta_ctx* c, int count, tta_tt* const* in, double eps, float* feats)
{
  if (!c || count < 0 || (count > 0 && (!in || !feats)) || !(eps > 0.0)) return TTA_ERR_ARG;
  try
  {
    arena A(c, size_t(4) << 20);
    float* F = nullptr;
    const int nb = bond_features_dev(c, count, in, eps, A, &F);
    if (nb > 0)
      hip_check(hipMemcpyAsync(feats, F, (size_t)nb * NF * sizeof(float), hipMemcpyDeviceToHost, c->stream), "features");
    ctx_sync(c);
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
  return TTA_OK;
}

int tta_mlp_predict(tta_ctx* c, tta_mlp* m, const float* X, int n, float* y_out);
int tta_round_randomized_batched(tta_ctx* c, int count, tta_tt* const* in, const int* ranks, double eps,
                                 unsigned long long seed, int fallback, tta_tt** out, double* rel_err,
                                 int* fell_back);

int tta_round_learned_batched(tta_ctx* c, tta_mlp* m, int count, tta_tt* const* in, double eps, int oversample,
                              double slack, unsigned long long seed, int fallback, tta_tt** out, double* rel_err,
                              int* fell_back, int* ranks_used)
{
  if (!c || !m || count < 0 || (count > 0 && (!in || !out)) || !(eps > 0.0) || oversample < 0 || !(slack >= 0.0))
    return TTA_ERR_ARG;
  if (fallback != TTA_FALLBACK_FP64 && fallback != TTA_FALLBACK_MIXED) return TTA_ERR_ARG;
  try
  {
    std::vector<float> pred;
    {
      arena A(c, size_t(4) << 20);
      float* F = nullptr;
      const int nb = bond_features_dev(c, count, in, eps, A, &F);
      pred.assign((size_t)std::max(nb, 1), 0.f);
      if (nb > 0) mlp_predict_dev(c, m, F, nb, pred.data());
    }
    std::vector<int> ranks;
    size_t b = 0;
    for (int i = 0; i < count; ++i)
    {
      const int d = in[i]->d();
      ranks.push_back(1);
      for (int k = 1; k < d; ++k, ++b)
      {
        const int r = in[i]->ranks[k];
        const double frac = std::min(1.0, std::max(0.0, (double)pred[b]));
        int l = (int)std::ceil(frac * r * (1.0 + slack)) + oversample;
        l = std::max(1, std::min(l, r));
        ranks.push_back(l);
      }
      if (d >= 1) ranks.push_back(1);
    }
    if (ranks_used) std::copy(ranks.begin(), ranks.end(), ranks_used);
    return tta_round_randomized_batched(c, count, in, ranks.data(), eps, seed, fallback, out, rel_err, fell_back);
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
}

}  // extern "C"
