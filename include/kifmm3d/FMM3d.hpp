// FMM3d.hpp -- header-only C++ drop-in for the reference's FMM3d class
// (fmm3d/fmm3d.hpp:27-115, fmm3d/knlmat3d.hpp:26-49, fmm3d/kernel3d.hpp:56-80)
// over the C ABI of libkifmm3d.so.  Code written against the reference's
//
//   FMM3d* fmm = new FMM3d("fmm3d_");
//   fmm->srcpos() = &srcpos; fmm->srcnor() = &srcnor; fmm->trgpos() = &trgpos;
//   fmm->ctr() = Point3(0,0,0); fmm->rootlvl() = 0; fmm->knl() = knl;
//   fmm->setup(opts); fmm->eval(srcden, trgval); fmm->check(srcden, trgval, 20);
//
// compiles unchanged against this header (fmm3d/tt.cpp:85-105).  DblNumMat is
// column-major 3 x N as in common/nummat.hpp:77-84.  The reference's methods
// return 0 and assert on errors; these return the library's status code (0 on
// success) and keep the message in last_error().
#ifndef KIFMM3D_FMM3D_HPP
#define KIFMM3D_FMM3D_HPP

#include <cmath>
#include <cstdio>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../kifmm3d.h"

namespace kifmm3d_cxx {

// Minimal counterparts of common/numvec.hpp, nummat.hpp and vec3t.hpp.
class DblNumVec {
 public:
  explicit DblNumVec(int m = 0) : v_(m, 0.0) {}
  int m() const { return (int)v_.size(); }
  void resize(int m) { v_.assign(m, 0.0); }
  double& operator()(int i) { return v_[i]; }
  double operator()(int i) const { return v_[i]; }
  double* data() { return v_.data(); }
  const double* data() const { return v_.data(); }

 private:
  std::vector<double> v_;
};

class DblNumMat {
 public:
  DblNumMat(int m = 0, int n = 0) : m_(m), n_(n), v_((size_t)m * n, 0.0) {}
  int m() const { return m_; }
  int n() const { return n_; }
  void resize(int m, int n) { m_ = m; n_ = n; v_.assign((size_t)m * n, 0.0); }
  double& operator()(int i, int j) { return v_[i + (size_t)j * m_]; }
  double operator()(int i, int j) const { return v_[i + (size_t)j * m_]; }
  double* data() { return v_.data(); }
  const double* data() const { return v_.data(); }

 private:
  int m_, n_;
  std::vector<double> v_;
};

struct Point3 {
  double v[3];
  Point3(double a = 0, double b = 0, double c = 0) : v{a, b, c} {}
  double& operator()(int i) { return v[i]; }
  double operator()(int i) const { return v[i]; }
};

enum {  // fmm3d/kernel3d.hpp:29-53 (the kernel types this evaluator supports)
  KNL_LAP_S_U = 111,
  KNL_LAP_D_U = 121,
  KNL_STK_S_U = 311,
  KNL_STK_S_P = 312,
  KNL_STK_D_U = 321,
  KNL_STK_D_P = 322,
  KNL_NAV_S_U = 511,
  KNL_NAV_D_U = 521,
};

class Kernel3d {
 public:
  Kernel3d() : kt_(-1) {}
  Kernel3d(int kt, const std::vector<double>& coefs) : kt_(kt), coefs_(coefs) {}
  int& kt() { return kt_; }
  const int& kt() const { return kt_; }
  std::vector<double>& coefs() { return coefs_; }
  const std::vector<double>& coefs() const { return coefs_; }
  // kernel3d.cpp:24-85
  int sdof() const { return (kt_ == 111 || kt_ == 121) ? 1 : 3; }
  int tdof() const { return (kt_ == 111 || kt_ == 121 || kt_ == 312 || kt_ == 322) ? 1 : 3; }

 private:
  int kt_;
  std::vector<double> coefs_;
};

class FMM3d {
 public:
  explicit FMM3d(const std::string& prefix) : prefix_(prefix) {}
  ~FMM3d() { kifmm3d_destroy(h_); }
  FMM3d(const FMM3d&) = delete;
  FMM3d& operator=(const FMM3d&) = delete;

  DblNumMat*& srcpos() { return srcpos_; }
  DblNumMat*& srcnor() { return srcnor_; }
  DblNumMat*& trgpos() { return trgpos_; }
  Kernel3d& knl() { return knl_; }
  Point3& ctr() { return ctr_; }
  int& rootlvl() { return rootlvl_; }
  int& np() { return np_; }
  const std::string& prefix() const { return prefix_; }
  int sdof() const { return knl_.sdof(); }
  int tdof() const { return knl_.tdof(); }
  const char* last_error() const { return kifmm3d_last_error(h_); }

  // FMM3d::setup (fmm3d_setup.cpp:26-63): reads -<prefix>np,
  // -<prefix>let3d_ptsmax and -<prefix>let3d_maxlevel.
  int setup(std::map<std::string, std::string>& opts)
  {
    int ptsmax = 0, maxlevel = 0;
    if (!read(opts, "np", np_) || !read(opts, "let3d_ptsmax", ptsmax) || !read(opts, "let3d_maxlevel", maxlevel))
      return KIFMM3D_EINVAL;
    if (!srcpos_ || !trgpos_ || srcpos_->m() != 3 || trgpos_->m() != 3) return KIFMM3D_EINVAL;
    kifmm3d_destroy(h_);
    h_ = nullptr;
    int rc = kifmm3d_create(&h_, knl_.kt(), knl_.coefs().data(), (int)knl_.coefs().size(), np_, ptsmax, maxlevel,
                            ctr_.v, rootlvl_);
    if (rc != KIFMM3D_OK) return rc;
    return kifmm3d_setup(h_, srcpos_->data(), srcnor_ ? srcnor_->data() : nullptr, srcpos_->n(), trgpos_->data(),
                         trgpos_->n());
  }

  // FMM3d::eval (fmm3d_eval.cpp:24-243): trgval is overwritten.
  int eval(const DblNumVec& srcden, DblNumVec& trgval)
  {
    if (srcden.m() != sdof() * srcpos_->n() || trgval.m() != tdof() * trgpos_->n()) return KIFMM3D_EINVAL;
    return kifmm3d_evaluate(h_, srcden.data(), trgval.data(), 1, KIFMM3D_MODE_PLAIN);
  }

  // FMM3d::check (fmm3d_check.cpp:25-74): prints the relative error at numchk
  // sampled targets against the FP64 direct sum (computed on the GPU).
  int check(const DblNumVec& srcden, DblNumVec& trgval, int numchk)
  {
    const int nt = trgpos_->n(), td = tdof();
    std::mt19937_64 rng(12345);
    std::vector<int> idx(numchk);
    for (int k = 0; k < numchk; k++) idx[k] = (int)(rng() % (unsigned long long)nt);
    std::vector<double> dir((size_t)numchk * td);
    int rc = kifmm3d_direct(h_, srcden.data(), 1, idx.data(), numchk, dir.data());
    if (rc != KIFMM3D_OK) return rc;
    double vn = 0, en = 0;
    for (int k = 0; k < numchk; k++)
      for (int i = 0; i < td; i++) {
        double v = trgval(idx[k] * td + i), e = dir[(size_t)k * td + i] - v;
        vn += v * v;
        en += e * e;
      }
    std::printf("relative error: %g\nerror norm    : %g\nsolution norm : %g\n", std::sqrt(en / vn), std::sqrt(en),
                std::sqrt(vn));
    return KIFMM3D_OK;
  }

 private:
  bool read(std::map<std::string, std::string>& opts, const char* key, int& out)
  {
    auto it = opts.find("-" + prefix_ + key);
    if (it == opts.end()) return false;
    std::istringstream ss(it->second);
    return (bool)(ss >> out);
  }

  std::string prefix_;
  DblNumMat *srcpos_ = nullptr, *srcnor_ = nullptr, *trgpos_ = nullptr;
  Kernel3d knl_;
  Point3 ctr_;
  int rootlvl_ = 0, np_ = 6;
  kifmm3d* h_ = nullptr;
};

}  // namespace kifmm3d_cxx

#endif
