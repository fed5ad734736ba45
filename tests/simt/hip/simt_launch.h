// Kernel launch for the CPU stand-in of <hip/hip_runtime.h> (test only).
#pragma once
#include <functional>

namespace simt {
void run_grid(const std::function<void()>& body, dim3 grid, dim3 block, bool cooperative);

template <class F, class... A>
void launch(F kernel, dim3 grid, dim3 block, bool cooperative, A... args) {
  std::function<void()> body = [=]() { kernel(args...); };
  run_grid(body, grid, block, cooperative);
}
}  // namespace simt
