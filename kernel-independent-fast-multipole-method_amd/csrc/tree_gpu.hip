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
// 5. The U/V/W/X lists (Let3d::calgnext, let3d.cpp:203-282): one thread per
//    target node counts, then writes, its entries (k_lists), and a segmented
//    radix sort puts every node's entries in ascending node order -- the
//    arrays of Tree::build_lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

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

// ---------------------------------------------------------------------------
// U/V/W/X lists on the device (Let3d::calgnext, let3d.cpp:203-282; the host
// form is Tree::build_lists).
struct NodeView {
  const int *par, *chd, *depth, *path, *tag;
  int nnode;
};

// findgnt (let3d.cpp:325-337): the deepest existing node containing cell
// (x, y, z) of depth d -- every split creates all eight children, so the
// walk down from the root by the cell's octant bits finds it.
__device__ inline int findgnt_dev(const int* __restrict__ chd, int d, int x, int y, int z)
{
  int node = 0;
  for (int l = 0; l < d; l++) {
    const int c = chd[node];
    if (c < 0) break;
    const int s = d - l - 1;
    node = c + ((x >> s) & 1) * 4 + ((y >> s) & 1) * 2 + ((z >> s) & 1);
  }
  return node;
}

// adjacent (let3d.cpp:339-352): the boxes touch (faces, edges or corners)
__device__ inline bool adjacent_dev(const NodeView& T, int me, int yo)
{
  const int md = max(T.depth[me], T.depth[yo]);
  const int merad = 1 << (md - T.depth[me]), yorad = 1 << (md - T.depth[yo]);
  const int r = merad + yorad;
  int linf = 0;
  bool ok = true;
  for (int d = 0; d < 3; d++) {
    const int dif = abs((2 * T.path[3 * me + d] + 1) * merad - (2 * T.path[3 * yo + d] + 1) * yorad);
    if (dif > r) ok = false;
    linf = max(linf, dif);
  }
  return ok && linf == r;
}

constexpr int LIST_STACK = 8 * 17;  // depth-first walk below a same-depth neighbour (7 per level + 1)

// One thread per node.  WRITE = false: cnt[l * nnode + g] = entries of list l
// of node g; WRITE = true: the entries at out[l] + ptr[l][g], unsorted.  The
// 6 x 6 x 6 cells at node g's depth around its parent's children are visited
// as the host does; a coarser node that several cells reach is taken only
// from the first such cell in that order (the corner of the cells it covers
// inside the window), so no list holds an entry twice; a node without
// sources is never entered (tag filter of the host lists).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_lists(NodeView T, int* __restrict__ cnt, const int* __restrict__ ptr,
                                               int* __restrict__ out0, int* __restrict__ out1, int* __restrict__ out2,
                                               int* __restrict__ out3, int* __restrict__ err)
{
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= T.nnode) return;
  int n[4] = {0, 0, 0, 0};
  int* out[4] = {out0, out1, out2, out3};
  auto emit = [&](int l, int node) {
    if (!(T.tag[node] & NODE_SRC)) return;
    if (WRITE) out[l][ptr[l * (T.nnode + 1) + g] + n[l]] = node;
    n[l]++;
  };
  if (T.tag[g] & NODE_TRG) {
    const bool leaf = T.chd[g] < 0;
    if (T.par[g] >= 0) {
      const int pg = T.par[g], dg = T.depth[g], maxidx = 1 << dg;
      const int p0 = 2 * T.path[3 * pg], p1 = 2 * T.path[3 * pg + 1], p2 = 2 * T.path[3 * pg + 2];
      const int w0 = max(p0 - 2, 0), w1 = max(p1 - 2, 0), w2 = max(p2 - 2, 0);  // first cell in the window
      const int g0 = T.path[3 * g], g1 = T.path[3 * g + 1], g2 = T.path[3 * g + 2];
      for (int i = -2; i < 4; i++)
        for (int j = -2; j < 4; j++)
          for (int k = -2; k < 4; k++) {
            const int t0 = p0 + i, t1 = p1 + j, t2 = p2 + k;
            if (t0 < 0 || t1 < 0 || t2 < 0 || t0 >= maxidx || t1 >= maxidx || t2 >= maxidx) continue;
            if (t0 == g0 && t1 == g1 && t2 == g2) continue;
            const int res = findgnt_dev(T.chd, dg, t0, t1, t2);
            const int dr = T.depth[res];
            const bool adj = adjacent_dev(T, res, g);
            if (dr < dg) {
              const int sh = dg - dr;
              const bool first = t0 == max(w0, T.path[3 * res] << sh) && t1 == max(w1, T.path[3 * res + 1] << sh) &&
                                 t2 == max(w2, T.path[3 * res + 2] << sh);
              if (!first) continue;
              if (adj) {
                if (leaf) emit(LIST_U, res);
              } else {
                emit(LIST_X, res);
              }
            } else if (!adj) {
              emit(LIST_V, res);
            } else if (leaf) {
              // the descendants of an adjacent same-depth node: W where not
              // adjacent, U at adjacent leaves (let3d.cpp:251-266)
              int stk[LIST_STACK];
              int sp = 0;
              stk[sp++] = res;
              while (sp > 0) {
                const int f = stk[--sp];
                if (!adjacent_dev(T, f, g)) {
                  emit(LIST_W, f);
                } else if (T.chd[f] < 0) {
                  emit(LIST_U, f);
                } else if (sp + 8 <= LIST_STACK) {
                  for (int c = 7; c >= 0; c--) stk[sp++] = T.chd[f] + c;
                } else {
                  atomicOr(err, 1);  // deeper than any tree of maxlevel <= 16
                  sp = 0;
                }
              }
            }
          }
    }
    if (leaf) emit(LIST_U, g);
  }
  if (!WRITE)
    for (int l = 0; l < 4; l++) cnt[l * T.nnode + g] = n[l];
}

struct Stream {
  hipStream_t s = nullptr;
  Stream() { TCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

// Lists of every node on the device, into the host CSR arrays.
void lists_device(Tree& t, const DBuf<int>& dchd, hipStream_t s)
{
  const int nn = t.nnode;
  DBuf<int> dpar(nn), ddep(nn), dpath((size_t)3 * nn), dtag(nn), cnt((size_t)4 * nn), err(1);
  TCHK(hipMemcpyAsync(dpar.p, t.par.data(), sizeof(int) * nn, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(ddep.p, t.depth.data(), sizeof(int) * nn, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dpath.p, t.path.data(), sizeof(int) * 3 * (size_t)nn, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dtag.p, t.tag.data(), sizeof(int) * nn, hipMemcpyHostToDevice, s));
  TCHK(hipMemsetAsync(err.p, 0, sizeof(int), s));
  const NodeView T{dpar.p, dchd.p, ddep.p, dpath.p, dtag.p, nn};
  hipLaunchKernelGGL(k_lists<false>, dim3(grid(nn)), dim3(256), 0, s, T, cnt.p, nullptr, nullptr, nullptr, nullptr,
                     nullptr, err.p);
  TCHK(hipGetLastError());
  std::vector<int> c((size_t)4 * nn);
  TCHK(hipMemcpyAsync(c.data(), cnt.p, sizeof(int) * c.size(), hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  // CSR pointers per list (host scan), [4][nnode + 1] on the device
  std::vector<int> ptr((size_t)4 * (nn + 1));
  for (int l = 0; l < 4; l++) {
    int* p = ptr.data() + (size_t)l * (nn + 1);
    p[0] = 0;
    for (int g = 0; g < nn; g++) p[g + 1] = p[g] + c[(size_t)l * nn + g];
  }
  DBuf<int> dptr(ptr.size());
  TCHK(hipMemcpyAsync(dptr.p, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, s));
  DBuf<int> raw[4], sorted[4];
  for (int l = 0; l < 4; l++) {
    const int tot = ptr[(size_t)l * (nn + 1) + nn];
    raw[l].alloc(tot);
    sorted[l].alloc(tot);
  }
  hipLaunchKernelGGL(k_lists<true>, dim3(grid(nn)), dim3(256), 0, s, T, nullptr, dptr.p, raw[0].p, raw[1].p,
                     raw[2].p, raw[3].p, err.p);
  TCHK(hipGetLastError());
  // ascending node order inside every node's entries
  int bits = 1;
  while ((1LL << bits) < nn) bits++;
  for (int l = 0; l < 4; l++) {
    t.lptr[l].assign(ptr.begin() + (size_t)l * (nn + 1), ptr.begin() + (size_t)(l + 1) * (nn + 1));
    const int tot = t.lptr[l][nn];
    t.lidx[l].resize(tot);
    if (tot == 0) continue;
    const int* seg = dptr.p + (size_t)l * (nn + 1);
    size_t bytes = 0;
    TCHK(rocprim::segmented_radix_sort_keys(nullptr, bytes, raw[l].p, sorted[l].p, (unsigned)tot, (unsigned)nn, seg,
                                            seg + 1, 0, bits, s));
    DBuf<char> tmp(bytes);
    TCHK(rocprim::segmented_radix_sort_keys((void*)tmp.p, bytes, raw[l].p, sorted[l].p, (unsigned)tot,
                                            (unsigned)nn, seg, seg + 1, 0, bits, s));
    TCHK(hipMemcpyAsync(t.lidx[l].data(), sorted[l].p, sizeof(int) * (size_t)tot, hipMemcpyDeviceToHost, s));
  }
  int e = 0;
  TCHK(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  if (e) throw Error(KIFMM3D_EINVAL, "internal: list construction exceeded its stack");
}

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
  const char* lb = std::getenv("KIFMM3D_LISTS");
  if (lb && std::string(lb) == "host")
    build_lists();
  else
    lists_device(*this, dchd, s);
}

}  // namespace kifmm
