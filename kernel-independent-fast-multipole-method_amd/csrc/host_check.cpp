// host_check.cpp -- the library's host work (octree, U/V/W/X lists, evaluate
// plan, operators and their disk cache, argument checks) run through the C
// ABI on many point sets, built with AddressSanitizer and
// UndefinedBehaviorSanitizer (make sanitize).  The int indices these host
// products hold feed every GPU kernel, so an out-of-range or overflowing index
// here is found on the CPU first.  On a machine without a GPU setup stops
// after the host phase (KIFMM3D_EDEVICE) with the host products complete; with
// one it also uploads them.  Exits non-zero on the first failed check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "kifmm3d.h"

namespace {

int g_fail = 0;

#define CHECK(cond, ...)                                    \
  do {                                                      \
    if (!(cond)) {                                          \
      std::fprintf(stderr, "host_check: FAIL %s: ", #cond); \
      std::fprintf(stderr, __VA_ARGS__);                    \
      std::fprintf(stderr, "\n");                           \
      g_fail++;                                             \
    }                                                       \
  } while (0)

struct Set {
  std::string name;
  std::vector<double> src, nor, trg;  // 3 x N column-major
  double ctr[3];
  int rootlvl;
};

Set uniform(const char* name, int n, unsigned seed, double lo, double hi)
{
  Set s{name, {}, {}, {}, {0.5 * (lo + hi), 0.5 * (lo + hi), 0.5 * (lo + hi)}, lo == 0.0 ? 1 : 0};
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(lo, hi);
  for (int i = 0; i < 3 * n; i++) s.src.push_back(u(g));
  s.trg = s.src;
  s.nor = s.src;
  return s;
}

Set clusters(int n, unsigned seed)
{
  Set s{"clusters", {}, {}, {}, {0.5, 0.5, 0.5}, 1};
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(0.15, 0.85);
  std::normal_distribution<double> nd(0.0, 1.0);
  double c[4][3], sig[4] = {0.002, 0.01, 0.03, 0.08};
  for (auto& cc : c)
    for (double& v : cc) v = u(g);
  while ((int)s.src.size() < 3 * n) {
    const int k = (int)(g() % 4);
    double p[3];
    bool in = true;
    for (int d = 0; d < 3; d++) {
      p[d] = c[k][d] + sig[k] * nd(g);
      in = in && p[d] >= 0.0 && p[d] < 1.0;
    }
    if (in)
      for (double v : p) s.src.push_back(v);
  }
  s.trg = s.src;
  s.nor = s.src;
  return s;
}

Set sphere(int n, unsigned seed, double rad)
{
  Set s{"sphere", {}, {}, {}, {0.0, 0.0, 0.0}, 0};
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> uc(-1.0, 1.0), up(0.0, 2.0 * M_PI);
  for (int i = 0; i < n; i++) {
    const double ct = uc(g), ph = up(g), st = std::sqrt(1.0 - ct * ct);
    const double v[3] = {st * std::cos(ph), st * std::sin(ph), ct};
    for (double x : v) {
      s.nor.push_back(x);
      s.src.push_back(rad * x);
    }
  }
  s.trg = s.src;
  return s;
}

// points on the splitting planes and faces of the root box [0, 1]^3, repeated points
Set planes(int n, unsigned seed)
{
  Set s{"planes", {}, {}, {}, {0.5, 0.5, 0.5}, 1};
  std::mt19937_64 g(seed);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      const unsigned r = (unsigned)(g() % 64);
      s.src.push_back(r < 32 ? (double)(r % 17) / 16.0 * (r % 17 == 16 ? 0.9999999 : 1.0) : (double)(g() % 1000) / 1000.0);
    }
  for (int i = 0; i < 200 && 3 * (i + 1) <= (int)s.src.size(); i++)
    for (int d = 0; d < 3; d++) s.src.push_back(s.src[3 * i + d]);
  s.trg = s.src;
  s.nor = s.src;
  return s;
}

long tree_len(kifmm3d* h, int what) { return kifmm3d_get_tree(h, what, nullptr, 0); }

std::vector<int> tree(kifmm3d* h, int what)
{
  const long n = tree_len(h, what);
  std::vector<int> v(n > 0 ? n : 0);
  if (n > 0) kifmm3d_get_tree(h, what, v.data(), n);
  return v;
}

void check_products(kifmm3d* h, const std::string& tag, int ns, int nt)
{
  int info[12] = {0};
  CHECK(kifmm3d_info(h, info, 12) == KIFMM3D_OK, "%s: info", tag.c_str());
  const int nnode = info[0];
  CHECK(info[3] == ns && info[5] == nt, "%s: point counts %d %d", tag.c_str(), info[3], info[5]);
  const std::vector<int> nodes = tree(h, 0);
  CHECK((int)nodes.size() == 13 * nnode, "%s: node table size", tag.c_str());
  for (int l = 0; l < 4; l++) {
    const std::vector<int> ptr = tree(h, 1 + l), idx = tree(h, 5 + l);
    CHECK((int)ptr.size() == nnode + 1, "%s: list %d pointer size", tag.c_str(), l);
    if ((int)ptr.size() != nnode + 1) continue;
    CHECK(ptr[0] == 0 && ptr[nnode] == (int)idx.size(), "%s: list %d bounds", tag.c_str(), l);
    for (int g = 0; g < nnode; g++) CHECK(ptr[g] <= ptr[g + 1], "%s: list %d not monotone at %d", tag.c_str(), l, g);
    for (int v : idx) CHECK(v >= 0 && v < nnode, "%s: list %d entry %d out of range", tag.c_str(), l, v);
  }
  for (int w = 0; w < 2; w++) {
    const std::vector<int> p = tree(h, 9 + w);
    const int n = w == 0 ? ns : nt;
    std::vector<char> seen(n, 0);
    CHECK((int)p.size() == n, "%s: permutation %d size", tag.c_str(), w);
    for (int v : p)
      if (v >= 0 && v < n) seen[v]++;
      else CHECK(false, "%s: permutation %d entry %d", tag.c_str(), w, v);
    for (int i = 0; i < n; i++) CHECK(seen[i] == 1, "%s: permutation %d misses %d", tag.c_str(), w, i);
  }
  double st[11] = {0};
  CHECK(kifmm3d_plan_stats(h, st, 11) == KIFMM3D_OK, "%s: plan stats", tag.c_str());
  CHECK(st[0] >= 0 && st[2] >= 0, "%s: plan counts", tag.c_str());
  for (int w = 0; w < 2; w++) {
    const long n = kifmm3d_get_operator(h, w, nullptr, 0);
    CHECK(n > 0, "%s: operator %d", tag.c_str(), w);
    std::vector<double> a(n > 0 ? n : 1);
    kifmm3d_get_operator(h, w, a.data(), n);
    for (double v : a) CHECK(std::isfinite(v), "%s: operator %d not finite", tag.c_str(), w);
  }
}

void run(const Set& s, int kt, const std::vector<double>& coefs, int np, int ptsmax, int maxlevel, bool sep_targets)
{
  const int ns = (int)s.src.size() / 3;
  std::vector<double> trg = s.trg;
  if (sep_targets) {  // targets: every other source, shifted inside the box
    trg.clear();
    for (int i = 0; i + 1 < ns; i += 2)
      for (int d = 0; d < 3; d++) trg.push_back(0.5 * (s.src[3 * i + d] + s.src[3 * (i + 1) + d]));
  }
  const int nt = (int)trg.size() / 3;
  const std::string tag = s.name + " kt " + std::to_string(kt) + " np " + std::to_string(np) + " ptsmax " +
                          std::to_string(ptsmax) + (sep_targets ? " separate targets" : "");
  kifmm3d* h = nullptr;
  int rc = kifmm3d_create(&h, kt, coefs.data(), (int)coefs.size(), np, ptsmax, maxlevel, s.ctr, s.rootlvl);
  CHECK(rc == KIFMM3D_OK && h, "%s: create %d", tag.c_str(), rc);
  if (!h) return;
  rc = kifmm3d_setup(h, s.src.data(), s.nor.data(), ns, trg.data(), nt);
  CHECK(rc == KIFMM3D_OK || rc == KIFMM3D_EDEVICE, "%s: setup %d %s", tag.c_str(), rc, kifmm3d_last_error(h));
  if (rc == KIFMM3D_OK || rc == KIFMM3D_EDEVICE) check_products(h, tag, ns, nt);
  std::printf("host_check: %-48s nodes %6ld  setup rc %d\n", tag.c_str(), tree_len(h, 0) / 13, rc);
  kifmm3d_destroy(h);
}

void error_paths()
{
  kifmm3d* h = nullptr;
  const double c1[1] = {1.0}, ctr[3] = {0, 0, 0};
  CHECK(kifmm3d_create(&h, 999, c1, 1, 4, 50, 10, ctr, 0) == KIFMM3D_EINVAL, "bad kernel");
  CHECK(kifmm3d_create(&h, 111, c1, 1, 5, 50, 10, ctr, 0) == KIFMM3D_EINVAL, "bad np");
  CHECK(kifmm3d_create(&h, 111, c1, 1, 4, 0, 10, ctr, 0) == KIFMM3D_EINVAL, "bad ptsmax");
  CHECK(kifmm3d_create(&h, 111, c1, 1, 4, 50, 0, ctr, 0) == KIFMM3D_EINVAL, "bad maxlevel");
  CHECK(kifmm3d_create(&h, 111, c1, 1, 4, 50, 10, ctr, 0) == KIFMM3D_OK, "create");
  const double out[6] = {0.0, 0.0, 0.0, 1.5, 0.0, 0.0};
  CHECK(kifmm3d_setup(h, out, nullptr, 2, out, 2) == KIFMM3D_EBOX, "point outside the box");
  CHECK(kifmm3d_setup(h, nullptr, nullptr, 2, out, 2) == KIFMM3D_EINVAL, "null sources");
  CHECK(kifmm3d_setup(h, out, nullptr, -1, out, 2) == KIFMM3D_EINVAL, "negative count");
  double v = 0;
  CHECK(kifmm3d_evaluate(h, &v, &v, 1, 0) == KIFMM3D_ESTATE, "evaluate before setup");
  kifmm3d_destroy(h);
  kifmm3d* k = nullptr;
  CHECK(kifmm3d_create(&k, 321, c1, 1, 4, 50, 10, ctr, 0) == KIFMM3D_OK, "create 321");
  CHECK(kifmm3d_setup(k, out, nullptr, 1, out, 1) == KIFMM3D_EINVAL, "321 without normals");
  kifmm3d_destroy(k);
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc > 1) kifmm3d_set_operator_cache(argv[1]);  // exercise the disk cache writer and reader
  const std::vector<double> lap = {1.0}, stk = {1.0}, nav = {1.0, 0.3};
  error_paths();
  const Set u = uniform("uniform", 20000, 1, 0.0, 1.0), c = clusters(20000, 2), p = planes(6000, 3);
  const Set sp = sphere(12000, 4, 0.8);
  run(u, 111, lap, 4, 60, 10, false);
  run(u, 111, lap, 6, 150, 10, true);
  run(c, 111, lap, 4, 30, 10, false);
  run(c, 511, nav, 4, 40, 10, false);
  run(c, 511, nav, 4, 40, 3, true);
  run(p, 111, lap, 4, 20, 10, false);
  run(p, 111, lap, 4, 1, 16, false);
  run(sp, 321, stk, 4, 40, 10, false);
  run(sp, 121, lap, 6, 60, 10, true);
  run(uniform("one point", 1, 5, -1.0, 1.0), 111, lap, 4, 10, 10, false);
  run(uniform("small", 30, 6, -1.0, 1.0), 311, stk, 4, 100, 10, false);
  if (argc > 1) {  // the same operators again after the in-memory sets are freed: read from disk
    kifmm3d_release_operators();
    run(u, 111, lap, 4, 60, 10, false);
  }
  int built = 0, loaded = 0;
  kifmm3d_operator_cache_stats(&built, &loaded);
  std::printf("host_check: operator sets built %d, loaded %d; %d failed checks\n", built, loaded, g_fail);
  if (argc > 1) CHECK(loaded >= 1, "operator sets were not read back from the disk cache");
  return g_fail ? 1 : 0;
}
