// SFXT without device allocation: levelize + relax + successor of a graph
// that is already on the device, for callers that cannot call hipMalloc
// (serving loops that forbid it, harnesses that check for it).
//
// Every intermediate buffer of the level loop (the fanin counters, the
// relax candidates, the level-order queue and offsets, the float suffix
// distances and the loop's control block) lives in one file-scope device
// array, `scratch`, sized for kNoallocMaxVerts vertices.  Its device
// address is taken once per device with hipGetSymbolAddress (the array's
// host-side name is not a device address) and kept per device; kernels get
// pointers into it.  Each device holds its own copy until the process
// ends.  Calls that overlap on one device race on it: callers serialise
// them (one stream, or one call at a time).  The control block is cleared
// by a kernel at the start of every call.
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace cpg {

namespace {

constexpr int kNoallocMaxVerts = 1 << 20;  // 1,048,576 vertices
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
// layout: control block, best (8 B), deps, queue, distf (4 B each), lvlp
// (n + 2 ints); best and deps carry the level loop's kSfxtPad no-op slots
constexpr size_t kOffCtl = 0;
constexpr size_t kOffBest = align_up(sizeof(SfxtCtl));
constexpr size_t kOffDeps = kOffBest + align_up(8ull * (kNoallocMaxVerts + kSfxtPad));
constexpr size_t kOffQueue = kOffDeps + align_up(4ull * (kNoallocMaxVerts + kSfxtPad));
constexpr size_t kOffDistf = kOffQueue + align_up(4ull * kNoallocMaxVerts);
constexpr size_t kOffLvlp = kOffDistf + align_up(4ull * kNoallocMaxVerts);
constexpr size_t kScratchBytes = kOffLvlp + align_up(4ull * (kNoallocMaxVerts + 2));  // ~25 MiB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kNoallocBlocks = 24;  // level-loop working blocks (levels of a few thousand vertices)
constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// The current device's address of `scratch`, taken on its first use there.
unsigned char* scratch_ptr() {
  int dev = 0;
  CPG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    CPG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("no device address for the scratch array");
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch[dev];
}

__global__ void noalloc_clear_ctl_kernel(SfxtCtl* ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctl = SfxtCtl{};
}

}  // namespace

int sfxt_noalloc_max_verts() { return kNoallocMaxVerts; }

int sfxt_noalloc(int n, const int* oadjp, const int* iadjp, const int2* iedge, const int* sinks,
                 int nsinks, int* dist, int* succ, int* level, hipStream_t stream) {
  if (n < 0 || nsinks < 0 || nsinks > n) throw std::runtime_error("bad vertex or sink count");
  if (n > kNoallocMaxVerts)
    throw std::runtime_error("graph has " + std::to_string(n) + " vertices; the no-allocation SFXT takes at most " +
                             std::to_string(kNoallocMaxVerts));
  if (n == 0) return 0;
  if (!oadjp || !iadjp || !iedge || (nsinks > 0 && !sinks) || !dist || !succ || !level)
    throw std::runtime_error("null device pointer");
  unsigned char* base = scratch_ptr();
  SfxtArgs a{};
  a.n = n;
  a.m = 0;  // list positions per lane step: 3 (the sparse case); the result does not depend on it
  a.nsinks = nsinks;
  a.oadjp = oadjp;
  a.iadjp = iadjp;
  a.iedge = iedge;
  a.sinks = sinks;
  a.ctl = reinterpret_cast<SfxtCtl*>(base + kOffCtl);
  a.best = reinterpret_cast<unsigned long long*>(base + kOffBest);
  a.deps = reinterpret_cast<int*>(base + kOffDeps);
  a.queue = reinterpret_cast<int*>(base + kOffQueue);
  a.distf = reinterpret_cast<float*>(base + kOffDistf);
  a.lvlp = reinterpret_cast<int*>(base + kOffLvlp);
  a.dist = dist;
  a.succ = succ;
  a.level = level;
  hipLaunchKernelGGL(noalloc_clear_ctl_kernel, dim3(1), dim3(64), 0, stream, a.ctl);
  CPG_HIP(hipGetLastError());
  launch_sfxt_dataflow_init(a, stream);  // counters, candidates, queue = sinks then -1
  launch_sfxt_levels_edges(a, kNoallocBlocks, stream);
  SfxtCtl c{};
  CPG_HIP(hipMemcpyAsync(&c, a.ctl, sizeof(SfxtCtl), hipMemcpyDeviceToHost, stream));
  CPG_HIP(hipStreamSynchronize(stream));
  if (c.err) throw std::runtime_error("SFXT grid barrier did not complete");
  return c.num_levels;
}

}  // namespace cpg
