// kernels_host.cpp -- host evaluation of the kernels (operator precompute only;
// every evaluate-pass kernel evaluation runs on the GPU in device.hip).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

#include "common.hpp"

namespace kifmm {

KernelInfo kernel_info(int kt, const double* coefs, int ncoefs)
{
  KernelInfo k;
  k.kt = kt;
  if (ncoefs > 0 && coefs) k.coefs[0] = coefs[0];
  if (ncoefs > 1 && coefs) k.coefs[1] = coefs[1];
  switch (kt) {
  case 111:  // KNL_LAP_S_U
    k.sdof = 1; k.tdof = 1; k.deg[0] = 1;
    break;
  case 121:  // KNL_LAP_D_U: Laplace double layer
    k.sdof = 1; k.tdof = 1; k.needs_normal = true;
    break;
  case 311:  // KNL_STK_S_U: Stokes single layer velocity (Stokeslet)
    k.sdof = 3; k.tdof = 3;
    break;
  case 312:  // KNL_STK_S_P: Stokes single layer pressure
    k.sdof = 3; k.tdof = 1;
    break;
  case 322:  // KNL_STK_D_P: Stokes double layer pressure
    k.sdof = 3; k.tdof = 1; k.needs_normal = true;
    break;
  case 521:  // KNL_NAV_D_U: Navier double layer displacement
    k.sdof = 3; k.tdof = 3; k.needs_normal = true;
    if (ncoefs < 2) throw Error(KIFMM3D_EINVAL, "kernel 521 needs coefs {mu, nu}");
    if (k.coefs[1] == 1.0) throw Error(KIFMM3D_EINVAL, "kernel 521 needs nu != 1");
    break;
  case 301:  // KNL_STK_F_U: Stokeslet plus potential source, mm kernel of 321
    k.sdof = 4; k.tdof = 3; k.deg[0] = k.deg[1] = k.deg[2] = 1; k.deg[3] = 2;
    break;
  case 321:  // KNL_STK_D_U: Stokes double layer velocity
    k.sdof = 3; k.tdof = 3; k.needs_normal = true;
    break;
  case 511:  // KNL_NAV_S_U: Navier (Kelvin) single layer displacement
    k.sdof = 3; k.tdof = 3; k.deg[0] = k.deg[1] = k.deg[2] = 1;
    if (ncoefs < 2) throw Error(KIFMM3D_EINVAL, "kernel 511 needs coefs {mu, nu}");
    if (k.coefs[1] == 1.0) throw Error(KIFMM3D_EINVAL, "kernel 511 needs nu != 1");
    break;
  default:
    throw Error(KIFMM3D_EINVAL, "unsupported kernel type " + std::to_string(kt) +
                                    " (supported: 111, 121, 311, 312, 321, 322, 511, 521)");
  }
  if ((kt == 301 || kt == 311 || kt == 321 || kt == 511 || kt == 521) && k.coefs[0] == 0.0)
    throw Error(KIFMM3D_EINVAL, "mu must be nonzero");
  return k;
}

int mm_kernel_type(int kt)
{
  switch (kt) {
  case 111: return 111;
  case 121: return 111;
  case 311: return 301;
  case 312: return 111;
  case 321: return 301;
  case 322: return 111;
  case 511: return 511;
  case 521: return 511;
  default: return -1;
  }
}

KernelConsts kernel_consts(const KernelInfo& k)
{
  KernelConsts c;
  switch (k.kt) {
  case 111: c.a = 1.0 / (4.0 * M_PI); break;
  case 301: {
    double oomu = 1.0 / k.coefs[0];
    c.a = oomu * (1.0 / (8.0 * M_PI));  // G, H numerators
    c.b = 1.0 / (4.0 * M_PI);           // A numerator
    break;
  }
  case 321: c.a = -(6.0 / (8.0 * M_PI)); break;
  case 121: c.a = -(1.0 / (4.0 * M_PI)); break;
  case 311: c.a = (1.0 / k.coefs[0]) * (1.0 / (8.0 * M_PI)); break;
  case 312: c.a = 1.0 / (4.0 * M_PI); break;
  case 322: c.a = k.coefs[0] * (1.0 / (2.0 * M_PI)); break;
  case 521: {
    double ve = k.coefs[1];
    c.a = -(1 - 2.0 * ve) / (8.0 * M_PI * (1.0 - ve));
    c.b = (1 - 2.0 * ve) / (8.0 * M_PI * (1.0 - ve));
    c.c = -3.0 / (8.0 * M_PI * (1.0 - ve));
    break;
  }
  case 511: {
    double ve = k.coefs[1];
    c.a = (3.0 - 4.0 * ve) / (16.0 * M_PI * (1.0 - ve));
    c.b = 1.0 / (16.0 * M_PI * (1.0 - ve));
    c.c = 1.0 / k.coefs[0];
    break;
  }
  }
  return c;
}

void kernel_block(const KernelInfo& k, const KernelConsts& kc, double x, double y, double z, const double* nrm,
                  double* b)
{
  const int n = k.sdof * k.tdof;
  double r2 = x * x + y * y + z * z;
  double r = std::sqrt(r2);
  if (r < 1e-8) {
    for (int i = 0; i < n; i++) b[i] = 0.0;
    return;
  }
  const double d[3] = {x, y, z};
  switch (k.kt) {
  case 111: b[0] = kc.a / r; break;
  case 301: {
    double r3 = r2 * r;
    double G = kc.a / r, H = kc.a / r3, A = kc.b / r3;
    for (int t = 0; t < 3; t++) {
      for (int s = 0; s < 3; s++) b[t * 4 + s] = (t == s ? G : 0.0) + H * d[t] * d[s];
      b[t * 4 + 3] = A * d[t];
    }
    break;
  }
  case 321: {
    double rn = x * nrm[0] + y * nrm[1] + z * nrm[2];
    double r5 = r2 * r2 * r;
    double C = kc.a / r5;
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++) b[t * 3 + s] = C * rn * d[t] * d[s];
    break;
  }
  case 511: {
    double r3 = r2 * r;
    double G = kc.a / r, H = kc.b / r3;
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++) b[t * 3 + s] = kc.c * ((t == s ? G : 0.0) + H * d[t] * d[s]);
    break;
  }
  case 121: b[0] = kc.a / (r2 * r) * (x * nrm[0] + y * nrm[1] + z * nrm[2]); break;
  case 311: {
    double G = kc.a / r, H = kc.a / (r2 * r);
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++) b[t * 3 + s] = (t == s ? G : 0.0) + H * d[t] * d[s];
    break;
  }
  case 312:
    for (int s = 0; s < 3; s++) b[s] = kc.a * d[s] / (r2 * r);
    break;
  case 322: {
    double rn = x * nrm[0] + y * nrm[1] + z * nrm[2];
    double r3 = r2 * r, r5 = r3 * r2;
    for (int s = 0; s < 3; s++) b[s] = kc.a * (nrm[s] / r3 - 3 * rn * d[s] / r5);
    break;
  }
  case 521: {
    double rn = x * nrm[0] + y * nrm[1] + z * nrm[2];
    double r3 = r2 * r, r5 = r3 * r2;
    double A = kc.a / r3, B = kc.b / r3, C = kc.c / r5;
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++)
        b[t * 3 + s] = A * ((t == s ? rn : 0.0) + nrm[t] * d[s]) + B * (d[t] * nrm[s]) + C * rn * d[t] * d[s];
    break;
  }
  }
}

void kernel_matrix(const KernelInfo& k, int ns, const double* spos, const double* snor, int nt, const double* tpos,
                   double* mat)
{
  KernelConsts kc = kernel_consts(k);
  const long M = (long)nt * k.tdof;
  double b[12];
  for (int j = 0; j < ns; j++)
    for (int i = 0; i < nt; i++) {
      kernel_block(k, kc, tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
                   tpos[3 * i + 2] - spos[3 * j + 2], snor ? snor + 3 * j : nullptr, b);
      for (int t = 0; t < k.tdof; t++)
        for (int s = 0; s < k.sdof; s++) mat[((long)j * k.sdof + s) * M + (long)i * k.tdof + t] = b[t * k.sdof + s];
    }
}

int host_threads()
{
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return (int)std::min(n, 16u);
}

void parallel_for(int n, const std::function<void(int)>& fn)
{
  int nt = std::min(host_threads(), n);
  if (nt <= 1) {
    for (int i = 0; i < n; i++) fn(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&]() {
      for (;;) {
        int i = next.fetch_add(1);
        if (i >= n) break;
        fn(i);
      }
    });
  for (auto& x : th) x.join();
}

}  // namespace kifmm
