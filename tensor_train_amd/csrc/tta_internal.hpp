// Host-side state behind the C ABI: context, device tensors, pool.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <set>
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

// Caching sub-allocator over hipMalloc.  Requests are carved best-fit out of
// device segments (16 MiB segments for blocks up to 1 MiB, at least 32 MiB
// for larger ones); freed blocks coalesce with free neighbours of the same
// segment and are reused without going back to the device.  All work of a
// context is ordered on its one stream, so a block returned here may be
// handed to the next request without synchronisation.
class pool
{
 public:
  void* get(size_t bytes);
  void put(void* p);
  // Returns every segment with no live block to the device.
  void trim();
  ~pool();

  size_t live_bytes = 0, cached_bytes = 0;
  unsigned long long device_allocs = 0, trims = 0;

 private:
  struct block
  {
    char* ptr;
    size_t size;
    bool free, small;
    block* prev;  // neighbours inside the same segment
    block* next;
  };
  struct by_size
  {
    bool operator()(const block* a, const block* b) const
    {
      return a->size != b->size ? a->size < b->size : a->ptr < b->ptr;
    }
  };
  block* carve(std::set<block*, by_size>& fl, size_t size, bool small);
  std::set<block*, by_size> small_free_, large_free_;
  std::unordered_map<void*, block*> live_;
  std::vector<block*> segments_;  // first block of each segment
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
  // Pinned staging for host-to-device descriptor uploads: a copy from
  // pageable memory would wait for the stream, so uploads go through this
  // buffer and stay asynchronous.  Offsets restart after each stream
  // synchronisation (tta::ctx_sync).
  char* stage = nullptr;
  size_t stage_cap = 0, stage_off = 0;
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
  // Per kernel class (tta_kernel_class; room for 16).
  double prof_ms[16] = {0};
  long long prof_n[16] = {0};
  double prof_flops[16] = {0};
  double prof_bytes[16] = {0};
  // Block-Jacobi launches (q > 64) since the last tta_prof_enable and the
  // block sweeps they ran (host-driven, so counted whether or not timing is on).
  long long jac_calls = 0, jac_sweeps = 0;
  // Sweeps the last single-problem block Jacobi needed: the next call reads
  // the convergence flag from three sweeps before that on (each read drains
  // the stream; a sweep of a converged problem costs only early-exiting
  // launches).
  int jac_sweep_hint = 0;
  // Block-Jacobi problems that had not converged after the sweep limit.
  long long jac_unconverged = 0;
  // SVD-sweep steps of the mixed batched rounding tried by the rank-revealing
  // CholQR2 certificate, and those it accepted.
  long long rr_tried = 0, rr_held = 0;
  // Block-Jacobi sweeps captured in a HIP graph and replayed (tta_set_graphs).
  bool graphs = false;
  // Block-Jacobi rounds of few problems driven by the block pair's Gram
  // matrix (bj_gram_kernel; tta_set_jacobi_gram, on by default).
  bool jac_gram = true;
  // Device address of the allocation-free entry point's scratch array on
  // this context's device (looked up at context creation).
  double* noalloc_scratch = nullptr;
  // Full-rank certificate of the SVD steps (on by default): steps where it
  // was tried, and where it held so that the Jacobi SVD was not needed.
  bool rank_certificate = true;
  long long cert_tried = 0, cert_held = 0;
  // CholQR2 for tall QR factorizations of the sweeps (on by default): tried
  // and accepted counts.
  bool cholqr = true;
  long long cholqr_tried = 0, cholqr_used = 0;
  long long cholqr_triangular = 0;  // accepted in the triangular form
  // Speculative SVD sweeps of round (on by default): CholQR2 acceptance and
  // the rank certificates are decided on the device into one flag that the
  // host reads after the sweep; a failed sweep is undone and run again with
  // host decisions.  spec marks a running speculative sweep; spec_svd_ok
  // records whether the last SVD sweep needed no fallback (speculation
  // resumes after a checked sweep that needed none).
  bool speculate = true, spec = false;
  // Chained sweep blocks in round's sweeps (on by default): blocks
  // accepted, the steps they covered, those accepted after one pass, and
  // those rejected.
  bool chain = true;
  long long chain_blocks = 0, chain_steps = 0;
  long long chain_one_pass = 0;  // blocks whose second pass was not needed
  long long chain_rejected = 0;  // blocks whose steps then ran one by one
  bool spec_svd_ok = true;
  long long spec_sweeps = 0, spec_failed = 0;
  // Certificate decisions of the running speculative sweep, made together
  // before its flag is read: their inputs (n values in mode 1, 2n in mode
  // 0), and the pool buffers that hold those inputs until then.
  struct spec_cert_rec
  {
    const double* a;
    int n, mode;
    double delta;
  };
  std::vector<spec_cert_rec> spec_certs;
  std::vector<void*> spec_cert_bufs;
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
// Looks up the device address of tta_noalloc.hip's scratch array for the
// current device (the context's) and keeps it in the context.
void noalloc_init(tta_ctx* c);
}

namespace tta
{
// Asynchronous upload of `bytes` host bytes to device memory `dev` on the
// context stream, through the pinned staging buffer (the source may be
// reused or freed on return).
void upload_async(tta_ctx* c, void* dev, const void* src, size_t bytes);
// Stream synchronisation; also recycles the staging buffer.
void ctx_sync(tta_ctx* c);
void free_stage(tta_ctx* c);

hipEvent_t prof_event(tta_ctx* c);
void hip_check(hipError_t e, const char* what);

// Algorithmic work credited to one launch (not hardware counters):
// flops of the textbook algorithm at that shape and the minimum HBM bytes
// (each operand read once, each result written once).
struct prof_scope
{
  tta_ctx* c;
  int cls;
  double flops, bytes;
  hipEvent_t a = nullptr;
  prof_scope(tta_ctx* c_, int cls_, double f = 0.0, double by = 0.0) : c(c_), cls(cls_), flops(f), bytes(by)
  {
    if (c->prof)
    {
      a = prof_event(c);
      hip_check(hipEventRecord(a, c->stream), "hipEventRecord");
    }
  }
  ~prof_scope()
  {
    if (!a) return;
    hipEvent_t b = prof_event(c);
    if (hipEventRecord(b, c->stream) == hipSuccess) c->prof_pending.push_back({cls, a, b, flops, bytes});
  }
};

// Bump allocator over large pool blocks for the many short-lived buffers
// of a batched step (one pool call per chunk instead of one per buffer).
// Buffers are 256-byte aligned and all return to the pool when the arena
// is destroyed; stream ordering makes that safe while kernels still queue.
class arena
{
 public:
  explicit arena(tta_ctx* c, size_t chunk_bytes = size_t(64) << 20) : c_(c), chunk_(chunk_bytes) {}
  arena(const arena&) = delete;
  arena& operator=(const arena&) = delete;
  ~arena()
  {
    for (void* p : chunks_) c_->mem.put(p);
  }
  template <class T>
  T* alloc(size_t n)
  {
    const size_t b = (std::max<size_t>(n, 1) * sizeof(T) + 255) / 256 * 256;
    if (b > left_)
    {
      const size_t sz = std::max(chunk_, b);
      cur_ = static_cast<char*>(c_->mem.get(sz));
      chunks_.push_back(cur_);
      left_ = sz;
    }
    T* r = reinterpret_cast<T*>(cur_);
    cur_ += b;
    left_ -= b;
    return r;
  }
  template <class T>
  T* upload(const std::vector<T>& v)
  {
    T* d = alloc<T>(v.size());
    upload_async(c_, d, v.data(), v.size() * sizeof(T));
    return d;
  }

 private:
  tta_ctx* c_;
  size_t chunk_;
  std::vector<void*> chunks_;
  char* cur_ = nullptr;
  size_t left_ = 0;
};

}  // namespace tta
