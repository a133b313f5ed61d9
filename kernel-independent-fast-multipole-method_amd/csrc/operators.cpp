// operators.cpp -- build and cache the translation operators.
#include "operators.hpp"

#include <cmath>
#include <map>
#include <mutex>
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

std::shared_ptr<const Operators> get_operators(const KernelInfo& mm, int np)
{
  static std::mutex mu;
  static std::map<std::tuple<int, double, double, int>, std::shared_ptr<const Operators>> cache;
  const double c0 = mm.kt == 111 ? 0.0 : mm.coefs[0];
  const double c1 = mm.kt == 511 ? mm.coefs[1] : 0.0;
  auto key = std::make_tuple(mm.kt, c0, c1, np);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  auto op = build(mm, np);
  cache[key] = op;
  return op;
}

}  // namespace kifmm
