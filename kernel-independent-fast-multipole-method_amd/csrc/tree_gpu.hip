// tree_gpu.hip -- Tree::build_device: the octree of tree.cpp with the
// per-point work on the GPU.
//
// 1. Every point gets its octant path down to depth maxlevel-1 as a key of
//    3 bits per level, from the same comparisons x >= centre as the host
//    partition (let3d.cpp:96-102, centres of let3d.cpp:309-319 evaluated
//    with the same operations, no contraction), plus the root-box check.
// 2. The source keys are radix-sorted; the host grows the node structure in
//    the same order as Tree::build (Tree::grow), reading each node's octant
//    counts off the sorted keys by binary search.
// 3. Every point walks the node structure to its leaf; a stable radix sort of
//    (leaf node, point index) gives the leaf-contiguous permutation with
//    leaves in node (BFS) order and points in index order inside a leaf --
//    the order of the host's stable partitions.
// 4. Targets take steps 1 and 3 against the source tree; per-node target
//    counts come from the sorted leaf ids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "tree.hpp"

namespace kifmm {

namespace {

#define TCHK(call)                                                                                              \
  do {                                                                                                        \
    hipError_t e_ = (call);                                                                                   \
    if (e_ != hipSuccess)                                                                                     \
      throw Error(e_ == hipErrorOutOfMemory ? KIFMM3D_ENOMEM : KIFMM3D_EDEVICE,                               \
                  std::string(#call) + ": " + hipGetErrorString(e_));                                         \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  DBuf() = default;
  explicit DBuf(size_t n) { alloc(n); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t n) { TCHK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T))); }
};

struct Box {
  double lo[3], hi[3];  // ctr -+ rad: the root-box test and the centre origin
  double w;             // 2 rad
};

// Octant keys (most significant digit = depth-0 split) and the root-box test;
// bad receives the smallest index of a point outside the box.
__global__ __launch_bounds__(256) void k_keys(const double* __restrict__ pos, int n, Box b, int D,
                                              unsigned long long* __restrict__ key, int* __restrict__ idx,
                                              int* __restrict__ bad)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double x[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
  bool in = true;
  for (int d = 0; d < 3; d++) in = in && (x[d] >= b.lo[d] && x[d] <= b.hi[d]);
  if (!in) atomicMin(bad, i);
  unsigned long long k = 0;
  int p[3] = {0, 0, 0};
  for (int l = 0; l < D; l++) {
    const double t = (double)(1 << l);
    int o = 0;
    for (int d = 0; d < 3; d++) {
      // Tree::center: (ctr - rad) + (2 rad) * (path + 0.5) / 2^depth
      const double c = b.lo[d] + b.w * ((double)p[d] + 0.5) / t;
      const int bit = x[d] >= c;
      o = 2 * o + bit;
      p[d] = 2 * p[d] + bit;
    }
    k = (k << 3) | (unsigned long long)o;
  }
  key[i] = k;
  idx[i] = i;
}

// Leaf node of every point: walk from the root by the key's octant digits.
__global__ __launch_bounds__(256) void k_leaf(const unsigned long long* __restrict__ key, int n,
                                              const int* __restrict__ chd, int D, unsigned* __restrict__ leaf,
                                              int* __restrict__ idx)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = key[i];
  int node = 0;
  for (int l = 0; l < D && chd[node] >= 0; l++) node = chd[node] + (int)((k >> (3 * (D - 1 - l))) & 7);
  leaf[i] = (unsigned)node;
  idx[i] = i;
}

inline unsigned grid(int n) { return (unsigned)((n + 255) / 256); }

template <class K>
void sort_pairs(const K* kin, K* kout, const int* vin, int* vout, int n, int bits, hipStream_t s)
{
  size_t bytes = 0;
  TCHK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0, bits, s));
  DBuf<char> tmp(bytes);
  TCHK(rocprim::radix_sort_pairs((void*)tmp.p, bytes, kin, kout, vin, vout, (size_t)n, 0, bits, s));
  TCHK(hipStreamSynchronize(s));
}

// Octant keys of n points on the device (positions uploaded into dpos).
void keys_of(const double* pos, int n, const Box& box, int D, DBuf<double>& dpos, DBuf<unsigned long long>& key,
             DBuf<int>& idx, hipStream_t s, const char* what)
{
  dpos.alloc((size_t)3 * n);
  key.alloc(n);
  idx.alloc(n);
  DBuf<int> bad(1);
  const int big = 0x7fffffff;
  TCHK(hipMemcpyAsync(dpos.p, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(bad.p, &big, sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_keys, dim3(grid(n)), dim3(256), 0, s, dpos.p, n, box, D, key.p, idx.p, bad.p);
  TCHK(hipGetLastError());
  int b = big;
  TCHK(hipMemcpyAsync(&b, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  if (b != big) throw Error(KIFMM3D_EBOX, std::string(what) + " point " + std::to_string(b) + " lies outside the root box");
}

// Leaf-contiguous permutation (leaves in node order, index order inside) and
// the sorted leaf id of every slot.
void leaf_order(const DBuf<unsigned long long>& key, int n, const DBuf<int>& dchd, int D, int nnode,
                std::vector<int>& perm, std::vector<unsigned>& leaf_sorted, hipStream_t s)
{
  DBuf<unsigned> leaf(n), leaf2(n);
  DBuf<int> idx(n), idx2(n);
  hipLaunchKernelGGL(k_leaf, dim3(grid(n)), dim3(256), 0, s, key.p, n, dchd.p, D, leaf.p, idx.p);
  TCHK(hipGetLastError());
  int bits = 1;
  while ((1LL << bits) < nnode) bits++;
  sort_pairs<unsigned>(leaf.p, leaf2.p, idx.p, idx2.p, n, bits, s);
  perm.resize(n);
  leaf_sorted.resize(n);
  TCHK(hipMemcpyAsync(perm.data(), idx2.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(leaf_sorted.data(), leaf2.p, sizeof(unsigned) * (size_t)n, hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
}

struct Stream {
  hipStream_t s = nullptr;
  Stream() { TCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

}  // namespace

void Tree::build_device(const double* spos, int ns, const double* tpos, int nt)
{
  if (maxlevel < 1 || maxlevel > 16) throw Error(KIFMM3D_EINVAL, "maxlevel must be in [1, 16]");
  if (ns == 0 || nt == 0) {  // nothing to sort
    build(spos, ns, tpos, nt);
    return;
  }
  rad = std::pow(2.0, -rootlvl);
  Box box;
  for (int d = 0; d < 3; d++) {
    box.lo[d] = ctr[d] - rad;
    box.hi[d] = ctr[d] + rad;
  }
  box.w = 2 * rad;
  const int D = maxlevel - 1;  // octant digits: nodes split only above depth maxlevel-1
  Stream st;
  const hipStream_t s = st.s;

  // sources: keys, sorted keys on the host, nodes
  DBuf<double> dspos;
  DBuf<unsigned long long> skey;
  DBuf<int> sidx;
  keys_of(spos, ns, box, D, dspos, skey, sidx, s, "source");
  std::vector<unsigned long long> sk(ns);
  {
    DBuf<unsigned long long> k2(ns);
    DBuf<int> i2(ns);
    if (D > 0) {
      sort_pairs<unsigned long long>(skey.p, k2.p, sidx.p, i2.p, ns, 3 * D, s);
      TCHK(hipMemcpy(sk.data(), k2.p, sizeof(unsigned long long) * (size_t)ns, hipMemcpyDeviceToHost));
    }
  }
  std::vector<int> beg, cnt;
  grow(ns, beg, cnt, [&](int k, const double*, int* c8) {
    // the points of node k are sk[beg, beg + cnt), sorted; octant digit at depth(k)
    const int sh = 3 * (D - 1 - depth[k]);
    const unsigned long long* b = sk.data() + beg[k];
    const unsigned long long* e = b + cnt[k];
    const unsigned long long* lo = b;
    for (int o = 0; o < 8; o++) {
      const unsigned long long* hi = std::partition_point(lo, e, [&](unsigned long long v) { return (int)((v >> sh) & 7) <= o; });
      c8[o] = (int)(hi - lo);
      lo = hi;
    }
  });
  index_sources(cnt);
  DBuf<int> dchd(nnode);
  TCHK(hipMemcpy(dchd.p, chd.data(), sizeof(int) * (size_t)nnode, hipMemcpyHostToDevice));
  std::vector<unsigned> leaf_sorted;
  leaf_order(skey, ns, dchd, D, nnode, sperm, leaf_sorted, s);

  // targets against the source tree
  std::vector<int> tcnt(nnode, 0);
  if (tpos == spos && nt == ns) {
    tperm = sperm;
    tcnt = cnt;
  } else {
    DBuf<double> dtpos;
    DBuf<unsigned long long> tkey;
    DBuf<int> tidx;
    keys_of(tpos, nt, box, D, dtpos, tkey, tidx, s, "target");
    leaf_order(tkey, nt, dchd, D, nnode, tperm, leaf_sorted, s);
    for (int i = 0; i < nt; i++) tcnt[leaf_sorted[i]]++;
    for (int g = nnode - 1; g > 0; g--) tcnt[par[g]] += tcnt[g];
  }
  index_targets(tcnt);
  build_lists();
}

}  // namespace kifmm
