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
void round_batched_plain(tta_ctx* c, std::vector<tta_tt*>& ts, double eps, int max_rank);
}

using namespace tta;

namespace
{
constexpr int NF = 13;

void bond_features(tta_ctx* c, int count, tta_tt* const* in, double eps, std::vector<float>& F)
{
  F.clear();
  std::vector<cn_desc> cn;
  std::vector<size_t> off;
  size_t total = 0;
  int max_cols = 0;
  for (int i = 0; i < count; ++i)
    for (int k = 1; k < in[i]->d(); ++k)
    {
      off.push_back(total);
      total += (size_t)in[i]->ranks[k];
      max_cols = std::max(max_cols, in[i]->ranks[k]);
    }
  if (total == 0) return;
  double* nrm = static_cast<double*>(c->mem.get(total * 8));
  size_t b = 0;
  for (int i = 0; i < count; ++i)
    for (int k = 1; k < in[i]->d(); ++k, ++b)
      cn.push_back(cn_desc{in[i]->cores[k - 1], in[i]->ranks[k - 1] * in[i]->modes[k - 1], in[i]->ranks[k], nrm + off[b]});
  void* dd = c->mem.get(cn.size() * sizeof(cn_desc));
  upload_async(c, dd, cn.data(), cn.size() * sizeof(cn_desc));
  for (size_t o = 0; o < cn.size(); o += 65535)
    launch_col_norms_batched(static_cast<cn_desc*>(dd) + o, (int)std::min<size_t>(65535, cn.size() - o), max_cols,
                             c->stream);
  std::vector<double> h(total);
  hip_check(hipMemcpyAsync(h.data(), nrm, total * 8, hipMemcpyDeviceToHost, c->stream), "norms");
  ctx_sync(c);
  c->mem.put(dd);
  c->mem.put(nrm);
  b = 0;
  const int js[7] = {1, 2, 4, 8, 16, 32, 64};
  for (int i = 0; i < count; ++i)
  {
    const int d = in[i]->d();
    for (int k = 1; k < d; ++k, ++b)
    {
      const int r = in[i]->ranks[k];
      std::vector<double> v(h.begin() + (long)off[b], h.begin() + (long)off[b] + r);
      std::sort(v.begin(), v.end(), [](double x, double y) { return x > y; });
      const double c0 = v[0] > 0 ? v[0] : 1.0;
      auto lg = [&](double x) { return (float)(std::log10(std::max(x / c0, 1e-300)) / 10.0); };
      F.push_back((float)(std::log2((double)r) / 8.0));
      F.push_back((float)(std::log2((double)in[i]->modes[k - 1]) / 3.0));
      F.push_back((float)k / (float)d);
      F.push_back((float)(-std::log10(eps) / 10.0));
      F.push_back((float)(-std::log10(eps / std::sqrt((double)std::max(d - 1, 1))) / 10.0));
      for (int j : js) F.push_back(lg(v[std::min(j, r - 1)]));
      F.push_back(lg(v[r - 1]));
    }
  }
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
    std::vector<float> F;
    bond_features(c, count, in, eps, F);
    std::copy(F.begin(), F.end(), feats);
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
                              double slack, unsigned long long seed, tta_tt** out, double* rel_err, int* fell_back,
                              int* ranks_used)
{
  if (!c || !m || count < 0 || (count > 0 && (!in || !out)) || !(eps > 0.0) || oversample < 0 || !(slack >= 0.0))
    return TTA_ERR_ARG;
  try
  {
    std::vector<float> F;
    bond_features(c, count, in, eps, F);
    const int nb = (int)(F.size() / NF);
    std::vector<float> pred((size_t)std::max(nb, 1));
    if (nb > 0)
    {
      const int st = tta_mlp_predict(c, m, F.data(), nb, pred.data());
      if (st != TTA_OK) return st;
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
    return tta_round_randomized_batched(c, count, in, ranks.data(), eps, seed, 1, out, rel_err, fell_back);
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
}

}  // extern "C"
