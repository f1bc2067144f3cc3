// Seeded input generators of the reference (include/core/factory/
// tt_factory.hpp:141-207 random(shape, max_rank, seed), :31-78 zeros/ones,
// tt_matrix_factory.hpp:133-191 random(row_shape, col_shape, max_rank,
// seed)), generated on the host and uploaded.  Entries are
// std::normal_distribution<double>(0, 1) over std::mt19937_64(seed), drawn
// core by core in buffer order; with the same standard library (libstdc++)
// the values are bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"

namespace
{

// Rank rule of random(): forward cap r_{k+1} <= r_k N_k, backward cap
// r_k <= r_{k+1} N_k (both also <= max_rank when max_rank > 0), and only then
// r_0 = r_d = 1 -- so the last bond keeps the forward cap (a quirk kept
// on purpose: random([4]^10, 8) has a last core (8, 4, 1)).
int random_ranks(int d, const int* N, int max_rank, std::vector<int>& ranks)
{
  ranks.assign(d + 1, 1);
  for (int k = 0; k < d; ++k)
    if (N[k] <= 0) return TTA_ERR_SHAPE;
  for (int k = 0; k < d; ++k)
  {
    long long cap = (long long)ranks[k] * N[k];
    if (cap > max_rank && max_rank > 0) cap = max_rank;
    if (cap > std::numeric_limits<int>::max()) return TTA_ERR_SHAPE;
    ranks[k + 1] = (int)cap;
  }
  for (int k = d - 1; k >= 0; --k)
  {
    long long cap = (long long)ranks[k + 1] * N[k];
    if (cap > max_rank && max_rank > 0) cap = max_rank;
    if (cap > std::numeric_limits<int>::max()) return TTA_ERR_SHAPE;
    if (cap < ranks[k]) ranks[k] = (int)cap;
  }
  ranks[0] = 1;
  ranks[d] = 1;
  return TTA_OK;
}

void fill_normal(const std::vector<int>& ranks, const int* N, int d, unsigned long long seed, double* out)
{
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  size_t off = 0;
  for (int k = 0; k < d; ++k)
  {
    const size_t sz = (size_t)ranks[k] * N[k] * ranks[k + 1];
    for (size_t i = 0; i < sz; ++i) out[off + i] = nd(rng);
    off += sz;
  }
}

}  // namespace

extern "C"
{

int tta_random_host(int d, const int* modes, int max_rank, unsigned long long seed, int* ranks_out, double* data_out,
                    long long* numel_out)
{
  if (d < 0 || (d > 0 && !modes)) return TTA_ERR_ARG;
  std::vector<int> ranks;
  const int st = random_ranks(d, modes, max_rank, ranks);
  if (st != TTA_OK) return st;
  long long n = 0;
  for (int k = 0; k < d; ++k) n += (long long)ranks[k] * modes[k] * ranks[k + 1];
  if (ranks_out) std::memcpy(ranks_out, ranks.data(), sizeof(int) * (d + 1));
  if (numel_out) *numel_out = n;
  if (data_out) fill_normal(ranks, modes, d, seed, data_out);
  return TTA_OK;
}

int tta_random(tta_ctx* c, int d, const int* modes, int max_rank, unsigned long long seed, tta_tt** out)
{
  if (!c || !out) return TTA_ERR_ARG;
  std::vector<int> ranks(d + 1);
  long long n = 0;
  int st = tta_random_host(d, modes, max_rank, seed, ranks.data(), nullptr, &n);
  if (st != TTA_OK) return st;
  std::vector<double> host((size_t)n);
  tta_random_host(d, modes, max_rank, seed, nullptr, host.data(), nullptr);
  st = tta_tt_create(c, d, modes, ranks.data(), out);
  if (st != TTA_OK) return st;
  st = tta_tt_upload(c, *out, host.data());
  if (st != TTA_OK)
  {
    tta_tt_free(c, *out);
    *out = nullptr;
  }
  return st;
}

int tta_random_matrix(tta_ctx* c, int d, const int* row_modes, const int* col_modes, int max_rank,
                      unsigned long long seed, tta_tt** out)
{
  if (!c || !out || d < 0 || (d > 0 && (!row_modes || !col_modes))) return TTA_ERR_ARG;
  std::vector<int> N(d), ranks(d + 1);
  for (int k = 0; k < d; ++k) N[k] = row_modes[k] * col_modes[k];
  long long n = 0;
  int st = tta_random_host(d, N.data(), max_rank, seed, ranks.data(), nullptr, &n);
  if (st != TTA_OK) return st;
  std::vector<double> host((size_t)n);
  tta_random_host(d, N.data(), max_rank, seed, nullptr, host.data(), nullptr);
  st = tta_ttm_create(c, d, row_modes, col_modes, ranks.data(), out);
  if (st != TTA_OK) return st;
  st = tta_tt_upload(c, *out, host.data());
  if (st != TTA_OK)
  {
    tta_tt_free(c, *out);
    *out = nullptr;
  }
  return st;
}

// zeros (tt_factory.hpp:31-52) and ones (:59-78): rank-1 cores.
int tta_constant(tta_ctx* c, int d, const int* modes, double value, tta_tt** out)
{
  if (!c || !out || d < 0 || (d > 0 && !modes)) return TTA_ERR_ARG;
  std::vector<int> ranks(d + 1, 1);
  int st = tta_tt_create(c, d, modes, ranks.data(), out);
  if (st != TTA_OK || value == 0.0) return st;
  size_t n = 0;
  for (int k = 0; k < d; ++k) n += (size_t)modes[k];
  std::vector<double> host(n, value);
  st = tta_tt_upload(c, *out, host.data());
  if (st != TTA_OK)
  {
    tta_tt_free(c, *out);
    *out = nullptr;
  }
  return st;
}

}  // extern "C"
