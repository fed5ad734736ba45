// Kernel launch for the CPU stand-in of <hip/hip_runtime.h> (test only).
#pragma once
#include <functional>

namespace simt {
void run_grid(const std::function<void()>& body, dim3 grid, dim3 block, bool cooperative);
// Stream capture: while a capture is open, launches are recorded (with
// their arguments) instead of run; a graph launch replays them in order.
bool capturing();
void record(std::function<void()> launch_fn);

// SIMT_GUARD=1: reports any 8-byte word of a kernel argument that points
// into a device buffer freed earlier and not mapped again.
void check_arg(const void* kernel, int index, const void* arg, size_t bytes);

template <class F, class... A>
void launch(F kernel, dim3 grid, dim3 block, bool cooperative, A... args) {
  {
    int i = 0;
    (check_arg(reinterpret_cast<const void*>(kernel), i++, &args, sizeof(A)), ...);
  }
  std::function<void()> body = [=]() { kernel(args...); };
  if (capturing()) {
    record([=]() { run_grid(body, grid, block, cooperative); });
    return;
  }
  run_grid(body, grid, block, cooperative);
}
}  // namespace simt
