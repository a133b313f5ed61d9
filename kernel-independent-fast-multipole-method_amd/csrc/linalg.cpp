// linalg.cpp -- host FP64 numerics for operator precompute:
//  * truncated pseudoinverse by one-sided (Hestenes) Jacobi SVD, the
//    truncation rule of common/svdrep.cpp:52-58 and the product
//    V_k S_k^-1 U_k^T of common/vecmatop.cpp:124-159;
//  * real 3-D DFT in the layout of FFTW 2's rfftwnd (row-major, last
//    dimension halved, sign -1, unnormalised), used to transform the M2L
//    kernel grids (matmgnt3d.cpp:307-325).
#include "linalg.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <thread>

#include "common.hpp"

namespace kifmm {

namespace {

// Sense-reversing spin barrier for the Jacobi worker threads.
class SpinBarrier {
 public:
  explicit SpinBarrier(int n) : n_(n), count_(n), sense_(false) {}
  void wait(bool& local)
  {
    local = !local;
    if (count_.fetch_sub(1) == 1) {
      count_.store(n_);
      sense_.store(local);
    } else {
      int spins = 0;
      while (sense_.load() != local)
        if (++spins > 256) std::this_thread::yield();
    }
  }

 private:
  int n_;
  std::atomic<int> count_;
  std::atomic<bool> sense_;
};

}  // namespace

int pinv_truncated(int m, int n, const double* A, double eps, std::vector<double>& R)
{
  const bool tr = m < n;
  const int M = tr ? n : m, N = tr ? m : n;
  std::vector<double> W((size_t)M * N), V((size_t)N * N, 0.0);
  for (int j = 0; j < N; j++)
    for (int i = 0; i < M; i++) W[(size_t)j * M + i] = tr ? A[(size_t)i * m + j] : A[(size_t)j * m + i];
  for (int j = 0; j < N; j++) V[(size_t)j * N + j] = 1.0;

  const int P = N + (N & 1);
  const int npair = P / 2;
  const double tol = 1e-15 * std::sqrt((double)M);
  const int nth = std::max(1, std::min(host_threads(), npair / 8));

  auto rotate = [&](int p, int q) -> bool {
    double* wp = &W[(size_t)p * M];
    double* wq = &W[(size_t)q * M];
    double al = 0, be = 0, ga = 0;
    for (int i = 0; i < M; i++) {
      al += wp[i] * wp[i];
      be += wq[i] * wq[i];
      ga += wp[i] * wq[i];
    }
    if (ga == 0.0 || al == 0.0 || be == 0.0) return false;
    if (std::fabs(ga) <= tol * std::sqrt(al) * std::sqrt(be)) return false;
    double zeta = (be - al) / (2.0 * ga);
    double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::hypot(1.0, zeta));
    double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
    for (int i = 0; i < M; i++) {
      double a = wp[i], b = wq[i];
      wp[i] = c * a - s * b;
      wq[i] = s * a + c * b;
    }
    double* vp = &V[(size_t)p * N];
    double* vq = &V[(size_t)q * N];
    for (int i = 0; i < N; i++) {
      double a = vp[i], b = vq[i];
      vp[i] = c * a - s * b;
      vq[i] = s * a + c * b;
    }
    return true;
  };

  for (int sweep = 0; sweep < 60; sweep++) {
    std::atomic<long> nrot(0);
    SpinBarrier bar(nth);
    auto worker = [&](int tid) {
      bool sense = false;
      long local = 0;
      for (int step = 0; step < P - 1; step++) {
        // round-robin (circle method) pairing: every pair once per sweep
        for (int k = tid; k < npair; k += nth) {
          int a, b;
          if (k == 0) {
            a = P - 1;
            b = step;
          } else {
            a = (step + k) % (P - 1);
            b = (step - k + (P - 1)) % (P - 1);
          }
          if (a >= N || b >= N) continue;
          if (rotate(std::min(a, b), std::max(a, b))) local++;
        }
        bar.wait(sense);
      }
      nrot += local;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; t++) th.emplace_back(worker, t);
    worker(0);
    for (auto& x : th) x.join();
    if (nrot.load() == 0) break;
  }

  std::vector<double> sig(N);
  for (int j = 0; j < N; j++) {
    double s = 0;
    for (int i = 0; i < M; i++) s += W[(size_t)j * M + i] * W[(size_t)j * M + i];
    sig[j] = std::sqrt(s);
  }
  std::vector<int> ord(N);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sig[a] > sig[b]; });
  // keep sigma_i >= eps * sigma_0, scanning in descending order (svdrep.cpp:52-58)
  int rank = 0;
  const double cutoff = eps * sig[ord[0]];
  while (rank < N && sig[ord[rank]] >= cutoff && sig[ord[rank]] > 0) rank++;
  std::vector<double> wk((size_t)rank * M), vk((size_t)rank * N);
  for (int k = 0; k < rank; k++) {
    int j = ord[k];
    double s2 = 1.0 / (sig[j] * sig[j]);
    for (int i = 0; i < M; i++) wk[(size_t)k * M + i] = W[(size_t)j * M + i];
    for (int i = 0; i < N; i++) vk[(size_t)k * N + i] = V[(size_t)j * N + i] * s2;
  }
  // R is n x m column-major.
  R.assign((size_t)n * m, 0.0);
  parallel_for(m, [&](int j) {
    double* out = &R[(size_t)j * n];
    for (int k = 0; k < rank; k++) {
      if (!tr) {  // R(i, j) = sum_k V(i,k) W(j,k) / s_k^2
        double w = wk[(size_t)k * M + j];
        const double* v = &vk[(size_t)k * N];
        for (int i = 0; i < n; i++) out[i] += v[i] * w;
      } else {  // R(i, j) = sum_k W(i,k) V(j,k) / s_k^2
        double v = vk[(size_t)k * N + j];
        const double* w = &wk[(size_t)k * M];
        for (int i = 0; i < n; i++) out[i] += w[i] * v;
      }
    }
  });
  return rank;
}

void rdft3(int n, const double* in, double* out)
{
  const int h = n / 2 + 1;
  std::vector<double> c(n), s(n);
  for (int m = 0; m < n; m++) {
    c[m] = std::cos(2.0 * M_PI * m / n);
    s[m] = std::sin(2.0 * M_PI * m / n);
  }
  std::vector<double> a((size_t)2 * n * n * h), b((size_t)2 * n * n * h);
  // along x (fastest), real -> half complex
  for (int zy = 0; zy < n * n; zy++)
    for (int kx = 0; kx < h; kx++) {
      double re = 0, im = 0;
      for (int x = 0; x < n; x++) {
        double v = in[(size_t)zy * n + x];
        int m = (kx * x) % n;
        re += v * c[m];
        im -= v * s[m];
      }
      a[2 * ((size_t)zy * h + kx)] = re;
      a[2 * ((size_t)zy * h + kx) + 1] = im;
    }
  // along y
  for (int z = 0; z < n; z++)
    for (int ky = 0; ky < n; ky++)
      for (int kx = 0; kx < h; kx++) {
        double re = 0, im = 0;
        for (int y = 0; y < n; y++) {
          size_t o = 2 * (((size_t)z * n + y) * h + kx);
          int m = (ky * y) % n;
          re += a[o] * c[m] + a[o + 1] * s[m];
          im += a[o + 1] * c[m] - a[o] * s[m];
        }
        size_t o = 2 * (((size_t)z * n + ky) * h + kx);
        b[o] = re;
        b[o + 1] = im;
      }
  // along z
  for (int kz = 0; kz < n; kz++)
    for (int ky = 0; ky < n; ky++)
      for (int kx = 0; kx < h; kx++) {
        double re = 0, im = 0;
        for (int z = 0; z < n; z++) {
          size_t o = 2 * (((size_t)z * n + ky) * h + kx);
          int m = (kz * z) % n;
          re += b[o] * c[m] + b[o + 1] * s[m];
          im += b[o + 1] * c[m] - b[o] * s[m];
        }
        size_t o = 2 * (((size_t)kz * n + ky) * h + kx);
        out[o] = re;
        out[o + 1] = im;
      }
}

}  // namespace kifmm
