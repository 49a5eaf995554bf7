// Scoped launch timing: when enabled (khx_prof_enable), a ProfScope records a
// HIP event pair around the launches it encloses, on the launch stream.
#pragma once
#include <hip/hip_runtime.h>

namespace khx {
class ProfScope {
 public:
  ProfScope(const char* name, hipStream_t s);
  ~ProfScope();

 private:
  const char* name_;
  hipStream_t s_;
  void* a_;
};
}  // namespace khx
