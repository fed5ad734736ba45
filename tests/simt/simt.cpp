// Fiber scheduler behind tests/simt/hip/hip_runtime.h (test only).
//
// One block runs on one OS thread as blockDim.x fibers (ucontext).  A
// fiber runs until it reaches a collective: __syncthreads (all live
// fibers of the block), or a wave collective (all live fibers of its
// 64-lane wave).  When every live fiber is parked and no collective can
// complete, the launch is reported as a deadlock and the process aborts
// with the parked positions.  Plain launches run their blocks one after
// another; cooperative launches run every block on its own OS thread so
// that grid-wide barriers made of atomics can complete.
#include "hip/hip_runtime.h"

#include <dlfcn.h>
#include <execinfo.h>
#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>
#include <ucontext.h>

#include <atomic>
#include <chrono>
#include <map>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <vector>

namespace simt {
namespace {

enum State { kRun, kBar, kWave, kDone };
constexpr size_t kStack = 256 * 1024;

struct Block;
// Fiber stacks are mapped with mmap rather than taken from the heap (so
// that AddressSanitizer builds never see stack frames inside heap chunks)
// and kept in a process-wide pool for reuse by later blocks and launches.
std::mutex g_stack_mu;
std::vector<char*> g_stack_pool;
struct Stack {
  char* p = nullptr;
  Stack() = default;
  Stack(const Stack&) = delete;
  Stack& operator=(const Stack&) = delete;
  Stack(Stack&& o) noexcept : p(o.p) { o.p = nullptr; }
  ~Stack() {
    if (!p) return;  // stacks (and their guard pages) stay mapped in the pool
    std::lock_guard<std::mutex> lk(g_stack_mu);
    g_stack_pool.push_back(p);
  }
  char* get() {
    if (p) return p;
    {
      std::lock_guard<std::mutex> lk(g_stack_mu);
      if (!g_stack_pool.empty()) {
        p = g_stack_pool.back();
        g_stack_pool.pop_back();
        return p;
      }
    }
    // one inaccessible page below the stack: an overflow faults there
    void* m = mmap(nullptr, kStack + 4096, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (m == MAP_FAILED) std::abort();
    mprotect(m, 4096, PROT_NONE);
    p = static_cast<char*>(m) + 4096;
    return p;
  }
};
struct Lane {
  ucontext_t ctx;
  Stack stack;
  int tid = 0;
  State st = kRun;
  Block* blk = nullptr;
};
struct Wave {
  int live = 0;
  int arrived = 0;
  unsigned gen = 0;
  uint64_t present[2] = {0, 0};
  uint64_t slots[2][64];
};
struct Block {
  std::vector<Lane> lanes;
  std::vector<Wave> waves;
  int live = 0;
  int bar_arrived = 0;
  dim3 idx, dim, grid;
  ucontext_t sched;
  const std::function<void()>* body = nullptr;
};

thread_local Block* t_blk = nullptr;
thread_local Lane* t_lane = nullptr;

void lane_entry(int lo, int hi) {
  Lane* L = reinterpret_cast<Lane*>((static_cast<uintptr_t>(static_cast<unsigned>(hi)) << 32) |
                                    static_cast<unsigned>(lo));
  (*L->blk->body)();
  L->st = kDone;
  Block* B = L->blk;
  B->live--;
  B->waves[L->tid / 64].live--;
  swapcontext(&L->ctx, &B->sched);
}

void park(State s) {
  Lane* L = t_lane;
  L->st = s;
  swapcontext(&L->ctx, &L->blk->sched);
}

void dump_and_abort(Block& B, const char* why) {
  std::fprintf(stderr, "simt: %s in block %u:", why, B.idx.x);
  for (auto& l : B.lanes)
    if (l.st != kDone) std::fprintf(stderr, " t%d:%s", l.tid, l.st == kBar ? "bar" : l.st == kWave ? "wave" : "run");
  std::fprintf(stderr, "\n");
  std::abort();
}

// The fault reporter runs on its own per-thread stack (a fiber's stack
// may be the one that overflowed).
void ensure_altstack() {
  static thread_local bool done = false;
  if (done) return;
  done = true;
  static thread_local char alt[64 * 1024];
  stack_t ss{};
  ss.ss_sp = alt;
  ss.ss_size = sizeof alt;
  sigaltstack(&ss, nullptr);
}

void run_block(Block& B) {
  ensure_altstack();
  const int n = static_cast<int>(B.dim.x);
  B.lanes.resize(n);
  B.waves.assign((n + 63) / 64, Wave{});
  B.live = n;
  B.bar_arrived = 0;
  for (int t = 0; t < n; ++t) {
    Lane& L = B.lanes[t];
    L.tid = t;
    L.st = kRun;
    L.blk = &B;
    getcontext(&L.ctx);
    L.ctx.uc_stack.ss_sp = L.stack.get();
    L.ctx.uc_stack.ss_size = kStack;
    L.ctx.uc_link = nullptr;
    const uintptr_t p = reinterpret_cast<uintptr_t>(&L);
    makecontext(&L.ctx, reinterpret_cast<void (*)()>(lane_entry), 2,
                static_cast<int>(p & 0xffffffffu), static_cast<int>(p >> 32));
    B.waves[t / 64].live++;
  }
  t_blk = &B;
  for (;;) {
    bool ran = false;
    for (auto& L : B.lanes) {
      if (L.st != kRun) continue;
      ran = true;
      t_lane = &L;
      swapcontext(&B.sched, &L.ctx);
    }
    if (B.live == 0) break;
    bool released = false;
    if (B.bar_arrived > 0 && B.bar_arrived == B.live) {
      for (auto& L : B.lanes)
        if (L.st == kBar) L.st = kRun;
      B.bar_arrived = 0;
      released = true;
    }
    for (size_t w = 0; w < B.waves.size(); ++w) {
      Wave& W = B.waves[w];
      if (W.arrived > 0 && W.arrived == W.live) {
        for (int t = static_cast<int>(w) * 64; t < std::min(n, static_cast<int>(w) * 64 + 64); ++t)
          if (B.lanes[t].st == kWave) B.lanes[t].st = kRun;
        W.arrived = 0;
        W.gen++;
        W.present[W.gen & 1] = 0;
        released = true;
      }
    }
    if (!ran && !released) dump_and_abort(B, "deadlock (collective not reached by every live lane)");
  }
  t_blk = nullptr;
  t_lane = nullptr;
}

}  // namespace

dim3 thread_idx() { return dim3(static_cast<unsigned>(t_lane->tid), 0, 0); }
dim3 block_idx() { return t_blk->idx; }
dim3 block_dim() { return t_blk->dim; }
dim3 grid_dim() { return t_blk->grid; }
int lane_id() { return t_lane->tid & 63; }

uint64_t* wave_sync(uint64_t v) {
  Block& B = *t_blk;
  Lane& L = *t_lane;
  Wave& W = B.waves[L.tid / 64];
  // the wave cannot mix a collective with lanes parked at a block barrier
  for (int t = (L.tid / 64) * 64; t < std::min(static_cast<int>(B.dim.x), (L.tid / 64) * 64 + 64); ++t)
    if (B.lanes[t].st == kBar) dump_and_abort(B, "wave collective while lanes wait at __syncthreads");
  const unsigned g = W.gen & 1;
  W.slots[g][L.tid & 63] = v;
  W.present[g] |= 1ull << (L.tid & 63);
  W.arrived++;
  park(kWave);
  return W.slots[g];
}

uint64_t wave_present() {
  Block& B = *t_blk;
  Wave& W = B.waves[t_lane->tid / 64];
  return W.present[(W.gen & 1) ^ 1];
}

void block_sync() {
  t_blk->bar_arrived++;
  park(kBar);
}

// A polling wait (a grid barrier) sleeps briefly so that the blocks it
// waits for get the CPU: blocks are OS threads, often more than cores.
void relax() { std::this_thread::sleep_for(std::chrono::microseconds(20)); }

void run_grid(const std::function<void()>& body, dim3 grid, dim3 block, bool cooperative) {
  const unsigned nb = grid.x * grid.y * grid.z;
  // A GPU dispatches every block of a small grid at once, and the engine's
  // grid-barrier kernels rely on that (they launch ordinary grids within
  // the device's resident capacity): grids of up to SIMT_RESIDENT blocks
  // (default 16) run one OS thread per block, as cooperative launches do.
  static const unsigned resident = [] {
    const char* e = std::getenv("SIMT_RESIDENT");
    return e ? static_cast<unsigned>(std::max(1, std::atoi(e))) : 16u;
  }();
  if (!cooperative && nb <= resident) cooperative = true;
  if (!cooperative) {
    // SIMT_THREADS > 1 spreads the blocks of ordinary launches over that
    // many OS threads, so that blocks run concurrently as they do on a GPU
    static const int nthr = [] {
      const char* e = std::getenv("SIMT_THREADS");
      return e ? std::max(1, std::atoi(e)) : 1;
    }();
    auto worker = [&](unsigned first, unsigned step) {
      Block B;
      for (unsigned b = first; b < nb; b += step) {
        B.idx = dim3(b % grid.x, (b / grid.x) % grid.y, b / (grid.x * grid.y));
        B.dim = block;
        B.grid = grid;
        B.body = &body;
        run_block(B);
      }
    };
    if (nthr == 1 || nb == 1) {
      worker(0, 1);
      return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthr; ++t) th.emplace_back(worker, static_cast<unsigned>(t), static_cast<unsigned>(nthr));
    for (auto& t : th) t.join();
    return;
  }
  std::vector<std::thread> th;
  std::vector<Block> blocks(nb);
  for (unsigned b = 0; b < nb; ++b) {
    blocks[b].idx = dim3(b % grid.x, (b / grid.x) % grid.y, b / (grid.x * grid.y));
    blocks[b].dim = block;
    blocks[b].grid = grid;
    blocks[b].body = &body;
    th.emplace_back([&blocks, b]() { run_block(blocks[b]); });
  }
  for (auto& t : th) t.join();
}

}  // namespace simt

// ---------------------------------------------------------------- host API
struct simt_event { std::chrono::steady_clock::time_point t; };
struct simt_stream { int dummy; };

const char* hipGetErrorString(hipError_t) { return "simt"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
// SIMT_CUS sets the reported compute-unit count (default 2), so that
// grids sized from it match a real device's.
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  static const int cus = [] {
    const char* e = std::getenv("SIMT_CUS");
    return e ? std::max(1, std::atoi(e)) : 2;
  }();
  *v = a == hipDeviceAttributeMultiprocessorCount ? cus : 1;
  return hipSuccess;
}
// SIMT_GUARD=1 places every device buffer at the end of its own mapping,
// against an inaccessible page, so a read or write past the end of a
// buffer (by more than the 16-byte alignment slack) faults at once.
namespace {
bool guard_on() {
  static const bool on = [] { const char* e = std::getenv("SIMT_GUARD"); return e && *e == '1'; }();
  return on;
}
std::mutex g_map_mu;
std::map<void*, std::pair<void*, size_t>> g_maps;  // buffer -> (mapping, length)
std::vector<std::pair<uintptr_t, uintptr_t>> g_freed;  // [begin, end) of freed buffers
}  // namespace

void simt::check_arg(const void* kernel, int index, const void* arg, size_t bytes) {
  if (!guard_on()) return;
  std::lock_guard<std::mutex> lk(g_map_mu);
  for (size_t o = 0; o + 8 <= bytes; o += 8) {
    uintptr_t v;
    std::memcpy(&v, static_cast<const char*>(arg) + o, 8);
    if (v < 4096) continue;
    bool live = false;
    for (auto& kv : g_maps) {
      const uintptr_t m = reinterpret_cast<uintptr_t>(kv.second.first);
      if (v >= m && v < m + kv.second.second) { live = true; break; }
    }
    if (live) continue;
    for (auto& f : g_freed) {
      if (v >= f.first && v < f.second) {
        Dl_info di{};
        dladdr(kernel, &di);
        std::fprintf(stderr, "simt: kernel %s argument %d word at +%zu = %#lx points into a freed "
                     "buffer [%#lx, %#lx)\n", di.dli_sname ? di.dli_sname : "?", index, o,
                     static_cast<unsigned long>(v), static_cast<unsigned long>(f.first),
                     static_cast<unsigned long>(f.second));
        std::abort();
      }
    }
  }
}
// With SIMT_GUARD=1, a fault reports the faulting address, the buffer
// whose guard page it hit (its size) and the native backtrace.
void guard_fault(int, siginfo_t* si, void*) {
  char buf[256];
  const uintptr_t a = reinterpret_cast<uintptr_t>(si->si_addr);
  int len = std::snprintf(buf, sizeof buf, "simt: fault at %#lx (lane stack %p)\n",
                          static_cast<unsigned long>(a),
                          simt::t_lane ? static_cast<void*>(simt::t_lane->stack.p) : nullptr);
  (void)!write(2, buf, len);
  uintptr_t best = 0, best_sz = 0;
  for (auto& kv : g_maps) {
    const uintptr_t b = reinterpret_cast<uintptr_t>(kv.first);
    const uintptr_t m = reinterpret_cast<uintptr_t>(kv.second.first);
    if (b <= a && b > best) {
      best = b;
      best_sz = m + kv.second.second - 4096 - b;
    }
  }
  len = std::snprintf(buf, sizeof buf, "simt: nearest buffer below: %#lx, %lu bytes, offset %+ld\n",
                      static_cast<unsigned long>(best), static_cast<unsigned long>(best_sz),
                      static_cast<long>(a - best));
  (void)!write(2, buf, len);
  for (auto& kv : g_maps) {
    const uintptr_t b = reinterpret_cast<uintptr_t>(kv.first);
    const uintptr_t m = reinterpret_cast<uintptr_t>(kv.second.first);
    if (a >= m && a < m + kv.second.second) {
      len = std::snprintf(buf, sizeof buf, "simt: buffer %#lx, %lu bytes to its guard, offset %+ld\n",
                          static_cast<unsigned long>(b), static_cast<unsigned long>(m + kv.second.second - 4096 - b),
                          static_cast<long>(a - b));
      (void)!write(2, buf, len);
    }
  }
  for (auto& kv : g_maps) {
    len = std::snprintf(buf, sizeof buf, "simt: live buffer %#lx mapping %#lx +%lu\n",
                        static_cast<unsigned long>(reinterpret_cast<uintptr_t>(kv.first)),
                        static_cast<unsigned long>(reinterpret_cast<uintptr_t>(kv.second.first)),
                        static_cast<unsigned long>(kv.second.second));
    (void)!write(2, buf, len);
  }
  void* fr[64];
  const int nf = backtrace(fr, 64);
  backtrace_symbols_fd(fr, nf, 2);
  _exit(139);
}
struct GuardInit {
  GuardInit() {
    if (!guard_on()) return;
    struct sigaction sa {};
    sa.sa_sigaction = guard_fault;
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sigaction(SIGSEGV, &sa, nullptr);
  }
} g_guard_init;
hipError_t simt_malloc(void** p, size_t bytes) {
  if (!bytes) bytes = 1;
  if (!guard_on()) {
    *p = std::calloc(1, bytes);
    return *p ? hipSuccess : 2;
  }
  const size_t pg = 4096;
  const size_t body = (bytes + 15) / 16 * 16;
  const size_t len = (body + pg - 1) / pg * pg + pg;
  void* m = mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (m == MAP_FAILED) return 2;
  char* guard = static_cast<char*>(m) + len - pg;
  mprotect(guard, pg, PROT_NONE);
  void* b = guard - body;
  std::lock_guard<std::mutex> lk(g_map_mu);
  g_maps[b] = {m, len};
  *p = b;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  if (!guard_on()) {
    std::free(p);
    return hipSuccess;
  }
  std::lock_guard<std::mutex> lk(g_map_mu);
  auto it = g_maps.find(p);
  if (it == g_maps.end()) std::abort();
  const uintptr_t b = reinterpret_cast<uintptr_t>(p);
  const uintptr_t e = reinterpret_cast<uintptr_t>(it->second.first) + it->second.second - 4096;
  g_freed.emplace_back(b, e);
  munmap(it->second.first, it->second.second);
  g_maps.erase(it);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t) { std::memset(p, v, n); return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new simt_stream{0}; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new simt_event{}; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }

struct simt_graph {
  std::vector<std::function<void()>> launches;
  int refs = 1;
};
namespace {
thread_local simt_graph* t_capture = nullptr;
}
bool simt::capturing() { return t_capture != nullptr; }
void simt::record(std::function<void()> launch_fn) { t_capture->launches.push_back(std::move(launch_fn)); }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  t_capture = new simt_graph();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = t_capture;
  t_capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) {
  g->refs++;
  *e = g;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& f : e->launches) f();
  return hipSuccess;
}
static void graph_release(simt_graph* g) {
  if (g && --g->refs == 0) delete g;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { graph_release(e); return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { graph_release(g); return hipSuccess; }
