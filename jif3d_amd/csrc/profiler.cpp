// Per-kernel HIP-event timing for a context (jgm_profile_* in the C ABI).
#include <hip/hip_runtime.h>

#include "context.hpp"

namespace jgm {

long prof_begin(jgm_context* ctx, const char* name) {
  if (ctx->prof_parent) ctx = ctx->prof_parent;
  Profiler& p = ctx->prof;
  if (!p.enabled) return -1;
  size_t idx = p.entries.size();
  for (size_t i = 0; i < p.entries.size(); ++i)
    if (p.entries[i].name == name) {
      idx = i;
      break;
    }
  if (idx == p.entries.size()) {
    Profiler::Entry e;
    e.name = name;
    p.entries.push_back(e);
  }
  Profiler::Pending pd{idx, nullptr, nullptr};
  if (hipEventCreate(&pd.start) != hipSuccess) return -1;
  if (hipEventCreate(&pd.stop) != hipSuccess) {
    (void)hipEventDestroy(pd.start);
    return -1;
  }
  if (hipEventRecord(pd.start, ctx->stream) != hipSuccess) {
    (void)hipEventDestroy(pd.start);
    (void)hipEventDestroy(pd.stop);
    return -1;
  }
  p.pending.push_back(pd);
  return (long)p.pending.size() - 1;
}

void prof_end(jgm_context* ctx, long token) {
  if (ctx->prof_parent) ctx = ctx->prof_parent;
  if (token < 0 || (size_t)token >= ctx->prof.pending.size()) return;
  (void)hipEventRecord(ctx->prof.pending[token].stop, ctx->stream);
}

hipError_t prof_harvest(jgm_context* ctx) {
  Profiler& p = ctx->prof;
  hipError_t first = hipSuccess;
  for (auto& pd : p.pending) {
    hipError_t e = hipEventSynchronize(pd.stop);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, pd.start, pd.stop);
    if (e == hipSuccess) {
      p.entries[pd.entry].total_ms += ms;
      p.entries[pd.entry].launches += 1;
    } else if (first == hipSuccess) {
      first = e;
    }
    (void)hipEventDestroy(pd.start);
    (void)hipEventDestroy(pd.stop);
  }
  p.pending.clear();
  return first;
}

void prof_clear(jgm_context* ctx) {
  (void)prof_harvest(ctx);
  ctx->prof.entries.clear();
}

}  // namespace jgm
