// Device memory of a context: a caching best-fit sub-allocator over
// hipMalloc'd segments.  Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"

namespace tta
{
namespace
{
constexpr size_t kAlign = 512;
constexpr size_t kSmallMax = size_t(1) << 20;      // blocks up to 1 MiB
constexpr size_t kSmallSegment = size_t(16) << 20;  // come from 16 MiB segments
constexpr size_t kLargeSegment = size_t(32) << 20;  // larger ones from >= 32 MiB
constexpr size_t kLargeRound = size_t(2) << 20;
constexpr size_t kLargeSplit = size_t(1) << 20;  // smallest remainder split off a large block

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
}  // namespace

pool::block* pool::carve(std::set<block*, by_size>& fl, size_t size, bool small)
{
  block key{nullptr, size, true, small, nullptr, nullptr};
  auto it = fl.lower_bound(&key);
  if (it == fl.end()) return nullptr;
  block* b = *it;
  fl.erase(it);
  const size_t rest = b->size - size;
  if (rest >= (small ? kAlign : kLargeSplit))
  {
    block* r = new block{b->ptr + size, rest, true, small, b, b->next};
    if (b->next) b->next->prev = r;
    b->next = r;
    b->size = size;
    fl.insert(r);
  }
  b->free = false;
  cached_bytes -= b->size;
  return b;
}

void* pool::get(size_t bytes)
{
  const size_t size = round_up(std::max<size_t>(bytes, 1), kAlign);
  const bool small = size <= kSmallMax;
  auto& fl = small ? small_free_ : large_free_;
  block* b = carve(fl, size, small);
  if (!b)
  {
    const size_t seg = small ? kSmallSegment : std::max(kLargeSegment, round_up(size, kLargeRound));
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, seg);
    if (e != hipSuccess)
    {
      (void)hipGetLastError();
      trim();
      ++trims;
      e = hipMalloc(&p, seg);
      if (e != hipSuccess)
      {
        (void)hipGetLastError();
        fail(TTA_ERR_NOMEM, "device allocation of %zu bytes failed", seg);
      }
    }
    ++device_allocs;
    block* s = new block{static_cast<char*>(p), seg, true, small, nullptr, nullptr};
    segments_.push_back(s);
    fl.insert(s);
    cached_bytes += seg;
    b = carve(fl, size, small);
  }
  live_[b->ptr] = b;
  live_bytes += b->size;
  return b->ptr;
}

void pool::put(void* p)
{
  if (!p) return;
  auto it = live_.find(p);
  if (it == live_.end()) return;
  block* b = it->second;
  live_.erase(it);
  live_bytes -= b->size;
  cached_bytes += b->size;
  auto& fl = b->small ? small_free_ : large_free_;
  b->free = true;
  if (b->next && b->next->free)
  {
    block* n = b->next;
    fl.erase(n);
    b->size += n->size;
    b->next = n->next;
    if (n->next) n->next->prev = b;
    delete n;
  }
  if (b->prev && b->prev->free)
  {
    block* q = b->prev;
    fl.erase(q);
    q->size += b->size;
    q->next = b->next;
    if (b->next) b->next->prev = q;
    delete b;
    b = q;
  }
  fl.insert(b);
}

void pool::trim()
{
  (void)hipDeviceSynchronize();
  std::vector<block*> keep;
  for (block* s : segments_)
  {
    // A segment with no live block has coalesced into its first block.
    if (s->free && !s->next)
    {
      (s->small ? small_free_ : large_free_).erase(s);
      cached_bytes -= s->size;
      (void)hipFree(s->ptr);
      delete s;
    }
    else
      keep.push_back(s);
  }
  segments_.swap(keep);
}

void upload_async(tta_ctx* c, void* dev, const void* src, size_t bytes)
{
  if (bytes == 0) return;
  const size_t need = round_up(bytes, 256);
  if (c->stage_off + need > c->stage_cap)
  {
    // Every staged copy so far must finish before the buffer is reused.
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->stage_off = 0;
    if (need > c->stage_cap)
    {
      if (c->stage) (void)hipHostFree(c->stage);
      c->stage = nullptr;
      const size_t cap = std::max(need, std::max<size_t>(2 * c->stage_cap, size_t(8) << 20));
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->stage), cap, hipHostMallocDefault), "hipHostMalloc");
      c->stage_cap = cap;
    }
  }
  char* p = c->stage + c->stage_off;
  std::memcpy(p, src, bytes);
  c->stage_off += need;
  hip_check(hipMemcpyAsync(dev, p, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
}

void ctx_sync(tta_ctx* c)
{
  hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  c->stage_off = 0;
}

void free_stage(tta_ctx* c)
{
  if (c->stage) (void)hipHostFree(c->stage);
  c->stage = nullptr;
  c->stage_cap = c->stage_off = 0;
}

pool::~pool()
{
  (void)hipDeviceSynchronize();
  for (block* s : segments_)
  {
    (void)hipFree(s->ptr);
    for (block* b = s; b;)
    {
      block* n = b->next;
      delete b;
      b = n;
    }
  }
}

}  // namespace tta
