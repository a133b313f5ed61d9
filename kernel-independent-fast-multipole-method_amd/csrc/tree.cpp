// tree.cpp -- octree construction and U/V/W/X lists (see tree.hpp).
#include "tree.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <numeric>
#include <unordered_map>

#include "common.hpp"

namespace kifmm {

namespace {

inline uint64_t cell_key(int d, int x, int y, int z)
{
  return ((uint64_t)d << 51) | ((uint64_t)x << 34) | ((uint64_t)y << 17) | (uint64_t)z;
}

bool in_box(const Tree& t, const double* p)
{
  for (int d = 0; d < 3; d++)
    if (!(p[d] >= t.ctr[d] - t.rad && p[d] <= t.ctr[d] + t.rad)) return false;
  return true;
}

// Stable partition of perm[b, b+n) into the 8 octants of centre cc, child order
// 4a + 2b + c with a = (x >= cx) etc. (let3d.cpp:96-102, :304-308).
void split_range(const double* pos, int* perm, int b, int n, const double* cc, int* cnt8, std::vector<int>& tmp,
                 std::vector<unsigned char>& oct)
{
  for (int o = 0; o < 8; o++) cnt8[o] = 0;
  oct.resize(n);
  for (int i = 0; i < n; i++) {
    const double* p = pos + 3 * (size_t)perm[b + i];
    int o = 4 * (p[0] >= cc[0]) + 2 * (p[1] >= cc[1]) + (p[2] >= cc[2]);
    oct[i] = (unsigned char)o;
    cnt8[o]++;
  }
  int off[8];
  off[0] = 0;
  for (int o = 1; o < 8; o++) off[o] = off[o - 1] + cnt8[o - 1];
  tmp.resize(n);
  for (int i = 0; i < n; i++) tmp[off[oct[i]]++] = perm[b + i];
  std::copy(tmp.begin(), tmp.begin() + n, perm + b);
}

}  // namespace

void Tree::center(int g, double* c) const
{
  int tmp = 1 << depth[g];
  for (int d = 0; d < 3; d++) c[d] = (ctr[d] - rad) + (2 * rad) * (path[3 * g + d] + 0.5) / double(tmp);
}

double Tree::radius(int g) const { return rad / double(1 << depth[g]); }

void Tree::build(const double* spos, int ns, const double* tpos, int nt)
{
  if (maxlevel < 1 || maxlevel > 16) throw Error(KIFMM3D_EINVAL, "maxlevel must be in [1, 16]");
  rad = std::pow(2.0, -rootlvl);
  const bool trace = std::getenv("KIFMM3D_TRACE") != nullptr;
  auto clk = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    auto now = std::chrono::steady_clock::now();
    if (trace) std::fprintf(stderr, "[kifmm3d]   tree %-10s %8.3f s\n", what, std::chrono::duration<double>(now - clk).count());
    clk = now;
  };
  build_sources(spos, ns);
  lap("sources");
  insert_targets(tpos, nt);
  lap("targets");
  build_lists();
  lap("lists");
}

// Nodes of the source tree, level by level (Let3d::srcdata, let3d.cpp:56-137):
// a node with more than ptsmax points above depth maxlevel-1 gets 8 children
// numbered after the nodes already made, in child order 4a + 2b + c.
// split(k, cc, c8) counts node k's points per octant of its centre cc and
// leaves them octant-ordered in [beg[k], beg[k] + cnt[k]).
void Tree::grow(int ns, std::vector<int>& beg, std::vector<int>& cnt,
                const std::function<void(int, const double*, int*)>& split)
{
  par.assign(1, -1); chd.assign(1, -1); depth.assign(1, 0); tag.assign(1, 0); path.assign(3, 0);
  beg.assign(1, 0);
  cnt.assign(1, ns);
  nnode = 1;
  int level = 0, arrbeg = 0, arrend = 1;
  while (arrbeg < arrend) {
    int arrcnt = arrend;
    for (int k = arrbeg; k < arrend; k++) {
      if (cnt[k] > ptsmax && level < maxlevel - 1) {
        chd[k] = arrcnt;
        arrcnt += 8;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
              par.push_back(k); chd.push_back(-1); depth.push_back(depth[k] + 1); tag.push_back(0);
              path.push_back(2 * path[3 * k] + a); path.push_back(2 * path[3 * k + 1] + b);
              path.push_back(2 * path[3 * k + 2] + c);
              beg.push_back(0); cnt.push_back(0);
              nnode++;
            }
        double cc[3];
        center(k, cc);
        int c8[8];
        split(k, cc, c8);
        int off = beg[k];
        for (int o = 0; o < 8; o++) {
          beg[chd[k] + o] = off;
          cnt[chd[k] + o] = c8[o];
          off += c8[o];
        }
      }
    }
    level++;
    arrbeg = arrend;
    arrend = arrcnt;
  }
  nlevel = level;
}

// Source tags, node indices and leaf ranges from per-node point counts
// (let3d.cpp:118-135).
void Tree::index_sources(const std::vector<int>& cnt)
{
  sndeidx.assign(nnode, 0); sextbeg.assign(nnode, 0); sextnum.assign(nnode, 0);
  int c = 0, sum = 0;
  for (int g = 0; g < nnode; g++) {
    if (cnt[g] > 0) {
      tag[g] |= NODE_SRC;
      sndeidx[g] = c++;
      if (chd[g] < 0) {
        sextbeg[g] = sum;
        sextnum[g] = cnt[g];
        sum += cnt[g];
      }
    }
  }
  sndecnt = c;
  sextcnt = sum;
}

// The same for targets (let3d.cpp:183-200).
void Tree::index_targets(const std::vector<int>& cnt)
{
  tndeidx.assign(nnode, 0); textbeg.assign(nnode, 0); textnum.assign(nnode, 0);
  int c = 0, sum = 0;
  for (int g = 0; g < nnode; g++) {
    if (cnt[g] > 0) {
      tag[g] |= NODE_TRG;
      tndeidx[g] = c++;
      if (chd[g] < 0) {
        textbeg[g] = sum;
        textnum[g] = cnt[g];
        sum += cnt[g];
      }
    }
  }
  tndecnt = c;
  textcnt = sum;
}

void Tree::check_box(const double* pos, int n, const char* what) const
{
  for (int k = 0; k < n; k++)
    if (!in_box(*this, pos + 3 * (size_t)k))
      throw Error(KIFMM3D_EBOX, std::string(what) + " point " + std::to_string(k) + " lies outside the root box");
}

// Let3d::srcdata, let3d.cpp:56-137
void Tree::build_sources(const double* pos, int ns)
{
  check_box(pos, ns, "source");
  std::vector<int> perm(ns), beg, cnt;
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<int> tmp;
  std::vector<unsigned char> oct;
  grow(ns, beg, cnt, [&](int k, const double* cc, int* c8) { split_range(pos, perm.data(), beg[k], cnt[k], cc, c8, tmp, oct); });
  index_sources(cnt);
  sperm.resize(ns);
  for (int g = 0; g < nnode; g++)
    if (cnt[g] > 0 && chd[g] < 0) std::copy(perm.begin() + beg[g], perm.begin() + beg[g] + cnt[g], sperm.begin() + sextbeg[g]);
}

// Let3d::trgdata, let3d.cpp:140-200
void Tree::insert_targets(const double* pos, int nt)
{
  check_box(pos, nt, "target");
  std::vector<int> perm(nt), beg(nnode, 0), cnt(nnode, 0);
  std::iota(perm.begin(), perm.end(), 0);
  cnt[0] = nt;
  std::vector<int> tmp;
  std::vector<unsigned char> oct;
  for (int g = 0; g < nnode; g++) {
    if (chd[g] < 0) continue;
    double cc[3];
    center(g, cc);
    int c8[8];
    split_range(pos, perm.data(), beg[g], cnt[g], cc, c8, tmp, oct);
    int off = beg[g];
    for (int o = 0; o < 8; o++) {
      beg[chd[g] + o] = off;
      cnt[chd[g] + o] = c8[o];
      off += c8[o];
    }
  }
  index_targets(cnt);
  tperm.resize(nt);
  for (int g = 0; g < nnode; g++)
    if (cnt[g] > 0 && chd[g] < 0) std::copy(perm.begin() + beg[g], perm.begin() + beg[g] + cnt[g], tperm.begin() + textbeg[g]);
}

// Let3d::calgnext for every target node, let3d.cpp:203-282
void Tree::build_lists()
{
  std::unordered_map<uint64_t, int> cells;
  cells.reserve((size_t)nnode * 2);
  for (int g = 0; g < nnode; g++) cells[cell_key(depth[g], path[3 * g], path[3 * g + 1], path[3 * g + 2])] = g;

  // findgnt (let3d.cpp:325-337): the deepest existing node on the way to the
  // cell; every split creates all eight children, so that is the deepest
  // existing ancestor-or-self cell.
  auto findgnt = [&](int d, const int* p) {
    for (int dd = d; dd > 0; dd--) {
      int s = d - dd;
      auto it = cells.find(cell_key(dd, p[0] >> s, p[1] >> s, p[2] >> s));
      if (it != cells.end()) return it->second;
    }
    return 0;
  };
  // adjacent (let3d.cpp:339-352)
  auto adjacent = [&](int me, int yo) {
    int md = std::max(depth[me], depth[yo]);
    int merad = 1 << (md - depth[me]), yorad = 1 << (md - depth[yo]);
    int r = merad + yorad, linf = 0;
    bool ok = true;
    for (int d = 0; d < 3; d++) {
      int dif = std::abs((2 * path[3 * me + d] + 1) * merad - (2 * path[3 * yo + d] + 1) * yorad);
      if (dif > r) ok = false;
      linf = std::max(linf, dif);
    }
    return ok && linf == r;
  };

  std::vector<int> trg;
  for (int g = 0; g < nnode; g++)
    if (tag[g] & NODE_TRG) trg.push_back(g);
  std::vector<std::vector<int>> L[4];
  for (int l = 0; l < 4; l++) L[l].assign(nnode, {});

  const int nchunk = std::max(1, std::min((int)trg.size(), 256));
  parallel_for(nchunk, [&](int ch) {
    size_t b = trg.size() * (size_t)ch / nchunk, e = trg.size() * (size_t)(ch + 1) / nchunk;
    std::vector<int> U, V, W, X, q;
    for (size_t ti = b; ti < e; ti++) {
      int g = trg[ti];
      U.clear(); V.clear(); W.clear(); X.clear();
      if (par[g] >= 0) {
        int pg = par[g];
        int maxidx = 1 << depth[g];
        for (int i = -2; i < 4; i++)
          for (int j = -2; j < 4; j++)
            for (int k = -2; k < 4; k++) {
              int tp[3] = {2 * path[3 * pg] + i, 2 * path[3 * pg + 1] + j, 2 * path[3 * pg + 2] + k};
              if (tp[0] < 0 || tp[1] < 0 || tp[2] < 0 || tp[0] >= maxidx || tp[1] >= maxidx || tp[2] >= maxidx) continue;
              if (tp[0] == path[3 * g] && tp[1] == path[3 * g + 1] && tp[2] == path[3 * g + 2]) continue;
              int res = findgnt(depth[g], tp);
              bool adj = adjacent(res, g);
              if (depth[res] < depth[g]) {
                if (adj) {
                  if (chd[g] < 0) U.push_back(res);
                } else {
                  X.push_back(res);
                }
              }
              if (depth[res] == depth[g]) {
                if (!adj) {
                  V.push_back(res);
                } else if (chd[g] < 0) {
                  q.clear();
                  q.push_back(res);
                  for (size_t h = 0; h < q.size(); h++) {
                    int f = q[h];
                    if (!adjacent(f, g))
                      W.push_back(f);
                    else if (chd[f] < 0)
                      U.push_back(f);
                    else
                      for (int c = 0; c < 8; c++) q.push_back(chd[f] + c);
                  }
                }
              }
            }
      }
      if (chd[g] < 0) U.push_back(g);
      std::vector<int>* src[4] = {&U, &V, &W, &X};
      for (int l = 0; l < 4; l++) {
        std::vector<int>& v = *src[l];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        auto& out = L[l][g];
        for (int x : v)
          if (tag[x] & NODE_SRC) out.push_back(x);
      }
    }
  });
  for (int l = 0; l < 4; l++) {
    lptr[l].assign(nnode + 1, 0);
    for (int g = 0; g < nnode; g++) lptr[l][g + 1] = lptr[l][g] + (int)L[l][g].size();
    lidx[l].resize(lptr[l][nnode]);
    for (int g = 0; g < nnode; g++) std::copy(L[l][g].begin(), L[l][g].end(), lidx[l].begin() + lptr[l][g]);
  }
}

}  // namespace kifmm
