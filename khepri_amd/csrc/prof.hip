// Optional per-kernel timing with HIP events recorded on the launch stream,
// and a plain device copy used to calibrate hardware byte counters.
#include <mutex>
#include <string>
#include <vector>

#include "khx_internal.h"
#include "khx_prof.h"

namespace khx {

namespace {
struct Rec {
  const char* name;
  hipEvent_t a, b;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

ProfScope::ProfScope(const char* name, hipStream_t s) : name_(name), s_(s), a_(nullptr) {
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_mu);
  a_ = take_event();
  if (a_) hipEventRecord(static_cast<hipEvent_t>(a_), s_);
}

ProfScope::~ProfScope() {
  if (!a_) return;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t b = take_event();
  if (!b) return;
  hipEventRecord(b, s_);
  g_recs.push_back(Rec{name_, static_cast<hipEvent_t>(a_), b});
}

__global__ void copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace khx

extern "C" {

int khx_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(khx::g_mu);
  for (auto& r : khx::g_recs) {
    khx::g_pool.push_back(r.a);
    khx::g_pool.push_back(r.b);
  }
  khx::g_recs.clear();
  khx::g_on = on != 0;
  return 0;
}

int khx_prof_collect(const char* name, double* total_ms, long long* launches) {
  std::lock_guard<std::mutex> lk(khx::g_mu);
  double tot = 0.0;
  long long cnt = 0;
  for (auto& r : khx::g_recs) {
    if (std::string(r.name) != name) continue;
    if (hipEventSynchronize(r.b) != hipSuccess) {
      khx::set_error("khx_prof_collect: event synchronisation failed");
      return -3;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
      khx::set_error("khx_prof_collect: elapsed time failed");
      return -3;
    }
    tot += ms;
    ++cnt;
  }
  *total_ms = tot;
  *launches = cnt;
  return 0;
}

int khx_copy(const void* src, void* dst, size_t bytes, void* stream) {
  khx::set_error("");
  KHX_REQUIRE(bytes % 16 == 0, "khx_copy: size must be a multiple of 16 bytes");
  const size_t n = bytes / 16;
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) return 0;
  khx::copy_kernel<<<dim3((unsigned)blocks), 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
      static_cast<const double2*>(src), static_cast<double2*>(dst), n);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
