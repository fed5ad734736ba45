// CSR relabelling: the graph's vertices renumbered in level order (sinks
// first, then each level in ascending original id), with every vertex's
// fanout and fanin lists kept in their order.
//
// This is the reference's E_ORIENTED reorder (reorder_csr_e_oriented,
// gpucpg/gpucpg.cu:2048-2079, run at :17753-17765 after levelize): one
// thread per edge places edge e of vertex u at new_adjp[perm[u]] +
// (e - adjp[u]) and relabels its head.  The permutation itself comes from
// the host (a counting sort of the levels); these kernels move the O(M)
// edge arrays.  A level's vertices then occupy one id range, so the level
// loop's reads of best[], deps[] and the CSR offsets of a level touch a few
// contiguous lines instead of one line per vertex across the whole graph.
#include "common.hpp"
#include "kernels.hpp"

namespace cpg {

namespace {

// owner[e] = u for every edge e of u (one thread per vertex fills its range).
__global__ void __launch_bounds__(256) edge_owner_kernel(const int* adjp, int n, int* owner) {
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const int b = adjp[u], e = adjp[u + 1];
    for (int j = b; j < e; ++j) owner[j] = u;
  }
}

__global__ void __launch_bounds__(256) reorder_fanout_kernel(const int* adjp, const int* adj,
                                                              const float* w, const int* owner,
                                                              const int* perm, const int* new_adjp,
                                                              int64_t m, int* new_adj, float* new_w) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int u = owner[e];
    const int ne = new_adjp[perm[u]] + static_cast<int>(e - adjp[u]);
    new_adj[ne] = perm[adj[e]];
    new_w[ne] = w[e];
  }
}

__global__ void __launch_bounds__(256) reorder_fanin_kernel(const int* iadjp, const int2* iedge,
                                                             const int* owner, const int* perm,
                                                             const int* new_iadjp, int64_t m,
                                                             int2* new_iedge) {
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < m;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int v = owner[j];
    const int nj = new_iadjp[perm[v]] + static_cast<int>(j - iadjp[v]);
    const int2 r = iedge[j];
    new_iedge[nj] = make_int2(perm[r.x], r.y);
  }
}

__global__ void __launch_bounds__(256) gather_u8_kernel(const uint8_t* in, const int* idx, int n,
                                                         uint8_t* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = in[idx[i]];
}

int grid_for(int64_t items) {
  int dev = 0, cus = 0;
  CPG_HIP(hipGetDevice(&dev));
  CPG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t g = (items + 255) / 256;
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(g, static_cast<int64_t>(cus) * 16)));
}

}  // namespace

void launch_edge_owner(const int* adjp, int n, int* owner, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(edge_owner_kernel, dim3(grid_for(n)), dim3(256), 0, stream, adjp, n, owner);
  CPG_HIP(hipGetLastError());
}

void launch_reorder_fanout(const int* adjp, const int* adj, const float* w, const int* owner,
                           const int* perm, const int* new_adjp, int64_t m, int* new_adj,
                           float* new_w, hipStream_t stream) {
  if (m <= 0) return;
  hipLaunchKernelGGL(reorder_fanout_kernel, dim3(grid_for(m)), dim3(256), 0, stream, adjp, adj, w,
                     owner, perm, new_adjp, m, new_adj, new_w);
  CPG_HIP(hipGetLastError());
}

void launch_reorder_fanin(const int* iadjp, const int2* iedge, const int* owner, const int* perm,
                          const int* new_iadjp, int64_t m, int2* new_iedge, hipStream_t stream) {
  if (m <= 0) return;
  hipLaunchKernelGGL(reorder_fanin_kernel, dim3(grid_for(m)), dim3(256), 0, stream, iadjp, iedge,
                     owner, perm, new_iadjp, m, new_iedge);
  CPG_HIP(hipGetLastError());
}

void launch_gather_u8(const uint8_t* in, const int* idx, int n, uint8_t* out, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gather_u8_kernel, dim3(grid_for(n)), dim3(256), 0, stream, in, idx, n, out);
  CPG_HIP(hipGetLastError());
}

}  // namespace cpg
