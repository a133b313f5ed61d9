// tree.hpp -- adaptive octree (the reference's Let3d) and U/V/W/X lists in
// flat arrays ready for upload.  Node numbering, split rule, child order,
// point order inside leaves and list contents are those of fmm3d/let3d.cpp;
// the construction itself is array-based (stable partitions, a hash map for
// findgnt, threads over target nodes).
#pragma once

#include <functional>
#include <vector>

namespace kifmm {

enum { NODE_SRC = 1, NODE_TRG = 2 };  // let3d.hpp:27-30
enum { LIST_U = 0, LIST_V = 1, LIST_W = 2, LIST_X = 3 };

struct Tree {
  // parameters (let3d.hpp:78-84)
  double ctr[3] = {0, 0, 0};
  int rootlvl = 0, ptsmax = 150, maxlevel = 10;
  double rad = 1.0;  // 2^-rootlvl (let3d.hpp:98)

  // nodes in BFS order
  int nnode = 0, nlevel = 0;
  std::vector<int> par, chd, depth, tag;
  std::vector<int> path;  // 3 per node
  // source side (let3d.cpp:118-135)
  std::vector<int> sndeidx, sextbeg, sextnum;
  int sndecnt = 0, sextcnt = 0;
  std::vector<int> sperm;  // leaf-order position -> caller's source index
  // target side (let3d.cpp:183-200)
  std::vector<int> tndeidx, textbeg, textnum;
  int tndecnt = 0, textcnt = 0;
  std::vector<int> tperm;
  // lists as CSR over nodes, entries ascending by node id (let3d.cpp:272-279)
  std::vector<int> lptr[4], lidx[4];

  // Throws Error(KIFMM3D_EBOX) when a point is outside the root box.
  void build(const double* spos, int ns, const double* tpos, int nt);
  // The same tree with the per-point work on the current HIP device: octant
  // keys by the same centre comparisons, radix sorts, the node structure from
  // the sorted keys; lists on the host.  Identical arrays to build().
  void build_device(const double* spos, int ns, const double* tpos, int nt);

  void center(int g, double* c) const;  // let3d.cpp:309-319
  double radius(int g) const;           // let3d.cpp:320-323
  bool terminal(int g) const { return chd[g] < 0; }
  int child(int g, int a, int b, int c) const { return chd[g] + a * 4 + b * 2 + c; }

 private:
  void grow(int ns, std::vector<int>& beg, std::vector<int>& cnt,
            const std::function<void(int, const double*, int*)>& split);
  void index_sources(const std::vector<int>& cnt);
  void index_targets(const std::vector<int>& cnt);
  void check_box(const double* pos, int n, const char* what) const;
  void build_sources(const double* spos, int ns);
  void insert_targets(const double* tpos, int nt);
  void build_lists();
};

}  // namespace kifmm
