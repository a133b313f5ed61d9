// operators.cpp -- build and cache the translation operators.
#include "operators.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "linalg.hpp"

namespace kifmm {

std::vector<double> surface_grid(int n, double R)
{
  std::vector<double> pos;
  const double step = 2.0 / (n - 1), init = -1.0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++)
        if (i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 || k == n - 1) {
          pos.push_back(R * (init + i * step));
          pos.push_back(R * (init + j * step));
          pos.push_back(R * (init + k * step));
        }
  return pos;
}

namespace {

std::shared_ptr<Operators> build(const KernelInfo& mm, int np)
{
  auto op = std::make_shared<Operators>();
  Operators& o = *op;
  o.mm = mm;
  o.np = np;
  const int sm = mm.sdof, tm = mm.tdof;
  // MatMgnt3d::setup, matmgnt3d.cpp:51-56
  const int gn[4] = {np, np + 2, np, np};
  const double gR[4] = {1.0, 3.0, 3.0, 1.0};
  for (int t = 0; t < 4; t++) {
    o.unit[t] = surface_grid(gn[t], gR[t]);
    o.nsp[t] = (int)o.unit[t].size() / 3;
  }
  o.pln[SURF_UE] = o.nsp[SURF_UE] * sm;
  o.pln[SURF_DE] = o.nsp[SURF_DE] * sm;
  o.pln[SURF_UC] = o.nsp[SURF_UC] * tm;
  o.pln[SURF_DC] = o.nsp[SURF_DC] * tm;
  o.n = 2 * np;
  o.G = o.n * o.n * o.n;
  o.C = (np + 1) * o.n * o.n;
  const double eps = std::pow(0.1, np + 1);  // matmgnt3d.cpp:38-41

  // UC2UE = pinv(K(UC <- UE)), DC2DE = pinv(K(DC <- DE)) at level 0
  // (matmgnt3d.cpp:96-105, 163-171; homogeneous kernels, R = 1)
  {
    const int M = o.pln[SURF_UC], N = o.pln[SURF_UE];
    std::vector<double> K((size_t)M * N);
    kernel_matrix(mm, o.nsp[SURF_UE], o.unit[SURF_UE].data(), nullptr, o.nsp[SURF_UC], o.unit[SURF_UC].data(), K.data());
    o.rank_uc2ue = pinv_truncated(M, N, K.data(), eps, o.uc2ue);
  }
  {
    const int M = o.pln[SURF_DC], N = o.pln[SURF_DE];
    std::vector<double> K((size_t)M * N);
    kernel_matrix(mm, o.nsp[SURF_DE], o.unit[SURF_DE].data(), nullptr, o.nsp[SURF_DC], o.unit[SURF_DC].data(), K.data());
    o.rank_dc2de = pinv_truncated(M, N, K.data(), eps, o.dc2de);
  }
  // M2M: K(parent UC <- child UE), child UE shifted by (2c - 1) R, parent UC at 2 R
  // (matmgnt3d.cpp:132-139).  L2L: K(child DC <- parent DE), child DC at R / 2
  // shifted by (c - 1/2) R (matmgnt3d.cpp:198-205).
  {
    const size_t bs = (size_t)o.pln[SURF_UC] * o.pln[SURF_UE];
    const size_t bd = (size_t)o.pln[SURF_DC] * o.pln[SURF_DE];
    o.ue2uc.resize(8 * bs);
    o.de2dc.resize(8 * bd);
    for (int c = 0; c < 8; c++) {
      const int idx[3] = {c >> 2, (c >> 1) & 1, c & 1};
      std::vector<double> chk(o.unit[SURF_UC]), den(o.unit[SURF_UE]);
      for (size_t i = 0; i < chk.size(); i++) chk[i] = 2.0 * chk[i];
      for (int j = 0; j < o.nsp[SURF_UE]; j++)
        for (int d = 0; d < 3; d++) den[3 * j + d] = den[3 * j + d] + (2 * idx[d] - 1) * 1.0;
      kernel_matrix(mm, o.nsp[SURF_UE], den.data(), nullptr, o.nsp[SURF_UC], chk.data(), &o.ue2uc[c * bs]);
      std::vector<double> dden(o.unit[SURF_DE]), dchk(o.unit[SURF_DC]);
      for (int j = 0; j < o.nsp[SURF_DC]; j++)
        for (int d = 0; d < 3; d++) dchk[3 * j + d] = 0.5 * dchk[3 * j + d] + (double(idx[d]) - 0.5) * 1.0;
      kernel_matrix(mm, o.nsp[SURF_DE], dden.data(), nullptr, o.nsp[SURF_DC], dchk.data(), &o.de2dc[c * bd]);
    }
  }
  // M2L: T_idx = r2c(K(reg - 2 idx)) / G per (source dof j, target dof i)
  // (matmgnt3d.cpp:307-325, 344-359), regular grid of matmgnt3d.cpp:400-419.
  {
    const int n = o.n, G = o.G, C = o.C;
    std::vector<double> reg((size_t)3 * G);
    const double step = 2.0 / (np - 1);
    for (int k = 0, cnt = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++, cnt++) {
          int gi = (i < np) ? i : i - n, gj = (j < np) ? j : j - n, gk = (k < np) ? k : k - n;
          reg[3 * cnt] = gi * step;
          reg[3 * cnt + 1] = gj * step;
          reg[3 * cnt + 2] = gk * step;
        }
    const size_t per = (size_t)sm * tm * C * 2;
    o.m2l.assign(343 * per, 0.0);
    const KernelConsts kc = kernel_consts(mm);
    parallel_for(343, [&](int oi) {
      const int a = oi / 49 - 3, b = (oi / 7) % 7 - 3, c = oi % 7 - 3;
      if (std::max(std::abs(a), std::max(std::abs(b), std::abs(c))) <= 1) return;
      const double src[3] = {a * 2.0, b * 2.0, c * 2.0};
      std::vector<double> grid((size_t)G * sm * tm), spec((size_t)2 * C), one(G);
      double blk[12];
      for (int g = 0; g < G; g++) {
        kernel_block(mm, kc, reg[3 * g] - src[0], reg[3 * g + 1] - src[1], reg[3 * g + 2] - src[2], nullptr, blk);
        for (int t = 0; t < tm; t++)
          for (int s = 0; s < sm; s++) grid[(size_t)(s * tm + t) * G + g] = blk[t * sm + s];
      }
      const double inv = 1.0 / double(G);
      for (int s = 0; s < sm; s++)
        for (int t = 0; t < tm; t++) {
          rdft3(n, &grid[(size_t)(s * tm + t) * G], spec.data());
          double* dst = &o.m2l[oi * per + (size_t)(s * tm + t) * C * 2];
          for (int q = 0; q < 2 * C; q++) dst[q] = spec[q] * inv;
        }
    });
  }
  // cube coordinate -> surface index (the count order of matmgnt3d.cpp:251-261)
  o.cube2surf.assign((size_t)np * np * np, -1);
  for (int i = 0, cnt = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1)
          o.cube2surf[((size_t)i * np + j) * np + k] = (short)cnt++;
  return op;
}

}  // namespace

// ---------------------------------------------------------------------------
// On-disk cache.  One file per (translation kernel, coefficients, np):
//   "KIFMMOP1", int32 {version, kt, np, sdof, tdof, 0}, double {c0, c1},
//   int32 {nsp[4], pln[4], n, G, C, rank_uc2ue, rank_dc2de},
//   arrays (int64 count + data): unit[0..3], uc2ue, dc2de, ue2uc, de2dc, m2l
//   (double), cube2surf (int16),
//   uint64 FNV-1a hash of everything before it.
// A file whose header, sizes or hash do not match is ignored and rewritten.

namespace {

constexpr int OPS_VERSION = 1;

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n)
  {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  }
};

struct Writer {
  std::FILE* f;
  Fnv fnv;
  bool ok = true;
  void put(const void* p, size_t n)
  {
    fnv.add(p, n);
    ok = ok && std::fwrite(p, 1, n, f) == n;
  }
  template <class T>
  void vec(const std::vector<T>& v)
  {
    int64_t n = (int64_t)v.size();
    put(&n, sizeof n);
    put(v.data(), v.size() * sizeof(T));
  }
};

struct Reader {
  std::FILE* f;
  Fnv fnv;
  bool ok = true;
  void get(void* p, size_t n)
  {
    ok = ok && std::fread(p, 1, n, f) == n;
    if (ok) fnv.add(p, n);
  }
  template <class T>
  void vec(std::vector<T>& v, size_t expect)
  {
    int64_t n = -1;
    get(&n, sizeof n);
    if (!ok || n != (int64_t)expect) {
      ok = false;
      return;
    }
    v.resize(expect);
    get(v.data(), expect * sizeof(T));
  }
};

std::mutex g_disk_mu;
std::string g_dir;
bool g_dir_init = false;
int g_built = 0, g_loaded = 0;

std::string cache_dir()
{
  if (!g_dir_init) {
    g_dir_init = true;
    if (const char* e = std::getenv("KIFMM3D_OPS_CACHE")) g_dir = e;
  }
  return g_dir;
}

std::string cache_path(const std::string& dir, const KernelInfo& mm, int np, double c0, double c1)
{
  uint64_t b0, b1;
  std::memcpy(&b0, &c0, 8);
  std::memcpy(&b1, &c1, 8);
  char name[160];
  std::snprintf(name, sizeof name, "/kifmm3d_ops_v%d_kt%d_np%d_%016llx_%016llx.bin", OPS_VERSION, mm.kt, np,
                (unsigned long long)b0, (unsigned long long)b1);
  return dir + name;
}

void save(const std::string& path, const Operators& o, double c0, double c1)
{
  // a temporary name unique per process and call, created exclusively: two
  // writers never share one file, and the rename publishes a whole file
  static std::atomic<unsigned long> serial{0};
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid()) + "_" + std::to_string(serial++);
  const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_EXCL, 0644);
  if (fd < 0) return;  // an unwritable cache directory only costs the rebuild
  std::FILE* f = ::fdopen(fd, "wb");
  if (!f) {
    ::close(fd);
    std::remove(tmp.c_str());
    return;
  }
  Writer w{f};
  w.put("KIFMMOP1", 8);
  const int32_t hdr[6] = {OPS_VERSION, o.mm.kt, o.np, o.mm.sdof, o.mm.tdof, 0};
  w.put(hdr, sizeof hdr);
  const double cc[2] = {c0, c1};
  w.put(cc, sizeof cc);
  const int32_t sz[13] = {o.nsp[0], o.nsp[1], o.nsp[2], o.nsp[3], o.pln[0], o.pln[1], o.pln[2], o.pln[3],
                          o.n, o.G, o.C, o.rank_uc2ue, o.rank_dc2de};
  w.put(sz, sizeof sz);
  for (int t = 0; t < 4; t++) w.vec(o.unit[t]);
  w.vec(o.uc2ue);
  w.vec(o.dc2de);
  w.vec(o.ue2uc);
  w.vec(o.de2dc);
  w.vec(o.m2l);
  w.vec(o.cube2surf);
  const uint64_t h = w.fnv.h;
  w.ok = w.ok && std::fwrite(&h, 1, sizeof h, f) == sizeof h;
  const bool ok = (std::fclose(f) == 0) && w.ok;
  if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

std::shared_ptr<Operators> load(const std::string& path, const KernelInfo& mm, int np, double c0, double c1)
{
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return nullptr;
  auto op = std::make_shared<Operators>();
  Operators& o = *op;
  Reader r{f};
  char magic[8];
  r.get(magic, 8);
  int32_t hdr[6];
  r.get(hdr, sizeof hdr);
  double cc[2];
  r.get(cc, sizeof cc);
  int32_t sz[13];
  r.get(sz, sizeof sz);
  bool ok = r.ok && std::memcmp(magic, "KIFMMOP1", 8) == 0 && hdr[0] == OPS_VERSION && hdr[1] == mm.kt &&
            hdr[2] == np && hdr[3] == mm.sdof && hdr[4] == mm.tdof && cc[0] == c0 && cc[1] == c1;
  // sizes implied by np (the same formulas as build())
  const int gn[4] = {np, np + 2, np, np};
  for (int t = 0; ok && t < 4; t++) {
    const int n = gn[t], nsp = n * n * n - (n - 2) * (n - 2) * (n - 2);
    ok = sz[t] == nsp;
  }
  if (ok) {
    o.mm = mm;
    o.np = np;
    for (int t = 0; t < 4; t++) o.nsp[t] = sz[t], o.pln[t] = sz[4 + t];
    o.n = sz[8];
    o.G = sz[9];
    o.C = sz[10];
    o.rank_uc2ue = sz[11];
    o.rank_dc2de = sz[12];
    ok = o.pln[SURF_UE] == o.nsp[SURF_UE] * mm.sdof && o.pln[SURF_DE] == o.nsp[SURF_DE] * mm.sdof &&
         o.pln[SURF_UC] == o.nsp[SURF_UC] * mm.tdof && o.pln[SURF_DC] == o.nsp[SURF_DC] * mm.tdof && o.n == 2 * np &&
         o.G == o.n * o.n * o.n && o.C == (np + 1) * o.n * o.n;
  }
  if (ok) {
    for (int t = 0; t < 4; t++) r.vec(o.unit[t], (size_t)3 * o.nsp[t]);
    r.vec(o.uc2ue, (size_t)o.pln[SURF_UE] * o.pln[SURF_UC]);
    r.vec(o.dc2de, (size_t)o.pln[SURF_DE] * o.pln[SURF_DC]);
    r.vec(o.ue2uc, (size_t)8 * o.pln[SURF_UC] * o.pln[SURF_UE]);
    r.vec(o.de2dc, (size_t)8 * o.pln[SURF_DC] * o.pln[SURF_DE]);
    r.vec(o.m2l, (size_t)343 * mm.sdof * mm.tdof * o.C * 2);
    r.vec(o.cube2surf, (size_t)np * np * np);
    const uint64_t want = r.fnv.h;
    uint64_t h = 0;
    ok = r.ok && std::fread(&h, 1, sizeof h, f) == sizeof h && h == want;
  }
  std::fclose(f);
  return ok ? op : nullptr;
}

}  // namespace

void set_operator_cache_dir(const std::string& dir)
{
  std::lock_guard<std::mutex> lk(g_disk_mu);
  g_dir_init = true;
  g_dir = dir;
}

void operator_cache_stats(int* built, int* loaded)
{
  std::lock_guard<std::mutex> lk(g_disk_mu);
  *built = g_built;
  *loaded = g_loaded;
}

namespace {
std::mutex g_mem_mu;
std::map<std::tuple<int, double, double, int>, std::shared_ptr<const Operators>> g_mem;
}  // namespace

int release_operators()
{
  std::lock_guard<std::mutex> lk(g_mem_mu);
  int n = 0;
  for (auto it = g_mem.begin(); it != g_mem.end();) {
    if (it->second.use_count() == 1) {
      it = g_mem.erase(it);
      n++;
    } else {
      ++it;
    }
  }
  return n;
}

std::shared_ptr<const Operators> get_operators(const KernelInfo& mm, int np)
{
  const double c0 = mm.kt == 111 ? 0.0 : mm.coefs[0];
  const double c1 = mm.kt == 511 ? mm.coefs[1] : 0.0;
  auto key = std::make_tuple(mm.kt, c0, c1, np);
  std::lock_guard<std::mutex> lk(g_mem_mu);
  auto it = g_mem.find(key);
  if (it != g_mem.end()) return it->second;
  std::string dir;
  {
    std::lock_guard<std::mutex> dk(g_disk_mu);
    dir = cache_dir();
  }
  std::shared_ptr<Operators> op;
  if (!dir.empty()) op = load(cache_path(dir, mm, np, c0, c1), mm, np, c0, c1);
  if (op) {
    std::lock_guard<std::mutex> dk(g_disk_mu);
    g_loaded++;
  } else {
    op = build(mm, np);
    if (!dir.empty()) save(cache_path(dir, mm, np, c0, c1), *op, c0, c1);
    std::lock_guard<std::mutex> dk(g_disk_mu);
    g_built++;
  }
  g_mem[key] = op;
  return op;
}

}  // namespace kifmm
