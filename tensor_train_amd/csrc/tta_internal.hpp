// Host-side state behind the C ABI: context, device tensors, pool.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace tta
{

struct error : std::runtime_error
{
  int code;
  error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...);
void hip_check(hipError_t e, const char* what);

// Size-bucketed caching allocator over hipMalloc.  All work of a context is
// ordered on its one stream, so a block returned here may be handed to the
// next request without synchronisation.
class pool
{
 public:
  void* get(size_t bytes);
  void put(void* p);
  void trim();
  ~pool();
  static size_t bucket(size_t bytes);

 private:
  std::map<size_t, std::vector<void*>> free_;
  std::unordered_map<void*, size_t> size_of_;
};

struct ws_slot
{
  double* p = nullptr;
  size_t cap = 0;
};

}  // namespace tta

struct tta_ctx
{
  int device = 0;
  hipStream_t stream = nullptr;
  tta::pool mem;
  tta::ws_slot ws[16];
  int* h_int = nullptr;     // pinned host scalars
  double* h_dbl = nullptr;
  std::string last_error;
  // Per-launch event timing (tta_prof_enable): recorded pairs awaiting
  // tta_prof_read, recycled events, totals per kernel class.
  bool prof = false;
  struct prof_rec
  {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
  };
  std::vector<prof_rec> prof_pending;
  std::vector<hipEvent_t> prof_free;
  double prof_ms[8] = {0};
  long long prof_n[8] = {0};
  double prof_flops[8] = {0};
  double prof_bytes[8] = {0};
};

// A device TT or TT-matrix.  Core k holds (ranks[k], modes[k], ranks[k+1])
// doubles row-major in its own pool block; for matrices modes[k] =
// row_modes[k] * col_modes[k].
struct tta_tt
{
  std::vector<int> modes, ranks, row_modes, col_modes;
  std::vector<double*> cores;
  bool matrix = false;
  int d() const { return (int)cores.size(); }
  bool is_matrix() const { return matrix || !row_modes.empty(); }
};

namespace tta
{
double* ctx_ws(tta_ctx* c, int slot, size_t n);
}
