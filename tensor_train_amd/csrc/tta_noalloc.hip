// inner(a, b) with no device allocation, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).  The environment
// sweep of inner_impl (contract_inner_step, include/core/detail/
// contract.hpp:155-205) runs with its two environment matrices and its
// intermediate in `scratch`, a file-scope device array of fixed size, and
// writes the result to a device double the caller provides.  Nothing is
// read back to the host and the stream is not synchronized, so the call can
// be captured into a graph.
//
// Every context on a device shares that device's array: calls that overlap
// on one device (two streams, or two host threads) race on it.  The array
// keeps the last call's contents; the first environment is set by a kernel
// at the start of every call, and every later value is written before it is
// read.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

namespace tta
{
// 4 Mi doubles (32 MiB) per device, held until the process ends.
constexpr size_t kNoAllocDoubles = size_t(4) << 20;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace
{
size_t round_up(size_t n) { return (n + 31) / 32 * 32; }  // keep each part 256-byte aligned
}  // namespace

void noalloc_init(tta_ctx* c)
{
  // The array's name is a host-side stand-in; kernels get its device address,
  // looked up once here with the context's device current.
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  c->noalloc_scratch = static_cast<double*>(p);
}

}  // namespace tta

extern "C" int tta_inner_noalloc(tta_ctx* c, const tta_tt* a, const tta_tt* b, double* out)
{
  using namespace tta;
  if (!c || !a || !b || !out) return TTA_ERR_ARG;
  try
  {
    if (!c->noalloc_scratch) fail(TTA_ERR_HIP, "inner_noalloc: scratch array not set up");
    const int d = a->d();
    if (b->d() != d || a->modes != b->modes)
      fail(TTA_ERR_SHAPE, "inner_noalloc: operands differ in core count or mode sizes");
    if (d == 0)
    {
      launch_set_value(out, 0.0, c->stream);
      return TTA_OK;
    }
    size_t emax = 1, tmax = 1;
    for (int k = 0; k <= d; ++k) emax = std::max(emax, (size_t)a->ranks[k] * b->ranks[k]);
    for (int k = 0; k < d; ++k) tmax = std::max(tmax, (size_t)b->ranks[k] * a->modes[k] * a->ranks[k + 1]);
    const size_t need = 2 * round_up(emax) + round_up(tmax);
    if (need > kNoAllocDoubles)
      fail(TTA_ERR_UNSUPPORTED, "inner_noalloc: needs %zu doubles of scratch, the array holds %zu", need,
           kNoAllocDoubles);
    double* E0 = c->noalloc_scratch;
    double* E1 = E0 + round_up(emax);
    double* T = E1 + round_up(emax);
    launch_set_value(E0, 1.0, c->stream);
    for (int k = 0; k < d; ++k)
    {
      const int ra = a->ranks[k], n = a->modes[k], rap = a->ranks[k + 1];
      const int rb = b->ranks[k], rbp = b->ranks[k + 1];
      double* dst = (k == d - 1) ? out : E1;  // the last environment is 1 x 1
      // T (rb x n rap) = E0^T right_unfold(a_k); E1 (rap x rbp) = T_as^T left_unfold(b_k).
      gemm_desc g;
      g.M = rb, g.N = n * rap, g.K = ra;
      g.A = E0, g.sar = 1, g.sac = rb;
      g.B = a->cores[k], g.sbr = (long)n * rap, g.sbc = 1;
      g.C = T, g.scr = (long)n * rap, g.scc = 1;
      launch_gemm(g, 1, c->stream);
      gemm_desc h;
      h.M = rap, h.N = rbp, h.K = rb * n;
      h.A = T, h.sar = 1, h.sac = rap;
      h.B = b->cores[k], h.sbr = rbp, h.sbc = 1;
      h.C = dst, h.scr = rbp, h.scc = 1;
      launch_gemm(h, 1, c->stream);
      std::swap(E0, E1);
    }
    hip_check(hipGetLastError(), "inner_noalloc launch");
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
  catch (const std::exception& e)
  {
    c->last_error = e.what();
    return TTA_ERR_HIP;
  }
  return TTA_OK;
}
