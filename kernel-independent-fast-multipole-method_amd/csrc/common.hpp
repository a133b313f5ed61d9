// common.hpp -- shared host types: errors, kernel descriptors, thread helper.
#pragma once

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "kifmm3d.h"

namespace kifmm {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Kernel descriptor.  Kernel ids, dof counts and coefficient meaning follow
// fmm3d/kernel3d.hpp:29-53 and fmm3d/kernel3d.cpp:24-113.
struct KernelInfo {
  int kt = 0;
  int sdof = 0, tdof = 0;
  bool needs_normal = false;
  double coefs[2] = {1.0, 0.0};  // {mu} (301, 321) or {mu, nu} (511)
  int deg[4] = {0, 0, 0, 0};     // homogeneity degree per source dof (mm kernels)
};

// Constants of the closed forms, precomputed exactly as the reference forms
// them (kernel3d.cpp:127, :165-167, :226, :320-322) so host and device share
// one definition.
struct KernelConsts {
  double a = 0, b = 0, c = 0;
};

KernelInfo kernel_info(int kt, const double* coefs, int ncoefs);  // throws Error(EINVAL)
int mm_kernel_type(int kt);                                        // fmm3d_setup.cpp:41-54
KernelConsts kernel_consts(const KernelInfo& k);

// One tdof x sdof block K(x = trg - src, n), blk[t * sdof + s]; zero when
// |x| < 1e-8 (kernel3d.cpp:21).
void kernel_block(const KernelInfo& k, const KernelConsts& kc, double x, double y, double z, const double* nrm,
                  double* blk);

// Dense (nt*tdof) x (ns*sdof) column-major kernel matrix; positions 3 x n
// column-major (Kernel3d::kernel, kernel3d.cpp:116-123).
void kernel_matrix(const KernelInfo& k, int ns, const double* spos, const double* snor, int nt, const double* tpos,
                   double* mat);

// Run fn(i) for i in [0, n) on up to hardware_concurrency threads (capped at 16).
void parallel_for(int n, const std::function<void(int)>& fn);
int host_threads();

}  // namespace kifmm
