// operators.hpp -- FP64 translation operators of one (mm kernel, np) pair,
// the counterpart of the reference's MatMgnt3d (fmm3d/matmgnt3d.hpp).  Built
// once per process and shared read-only by every handle.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace kifmm {

enum { SURF_UE = 0, SURF_UC = 1, SURF_DE = 2, SURF_DC = 3 };

struct Operators {
  KernelInfo mm;
  int np = 0;
  int nsp[4] = {0, 0, 0, 0};   // surface point counts
  std::vector<double> unit[4];  // unit surfaces 3 x n: UE (np, 1), UC (np+2, 3), DE (np, 3), DC (np, 1)
  int pln[4] = {0, 0, 0, 0};   // plndatasize: UE/DE n*sdof, UC/DC n*tdof (matmgnt3d.cpp:72-78)
  int n = 0, G = 0, C = 0;     // n = 2np, G = n^3, C = (np+1) n^2 complex per dof
  std::vector<double> uc2ue;   // plnUE x plnUC, column-major
  std::vector<double> dc2de;   // plnDE x plnDC
  int rank_uc2ue = 0, rank_dc2de = 0;
  std::vector<double> ue2uc;   // 8 blocks of plnUC x plnUE (child index 4a+2b+c)
  std::vector<double> de2dc;   // 8 blocks of plnDC x plnDE
  // M2L operators, already divided by G: [343 offsets][sdof][tdof][C] complex
  // (re, im); offset index (i+3)*49 + (j+3)*7 + (k+3); zero for |offset|_inf <= 1.
  std::vector<double> m2l;
  std::vector<short> cube2surf;  // np^3 -> surface point index (x-major order) or -1
};

// Operators for (mm kernel, np): from the in-memory cache, else from the
// on-disk cache directory (set_operator_cache_dir or KIFMM3D_OPS_CACHE) when
// a valid file is there, else built (and written to the directory).
std::shared_ptr<const Operators> get_operators(const KernelInfo& mm, int np);
void set_operator_cache_dir(const std::string& dir);  // "" disables the disk cache
void operator_cache_stats(int* built, int* loaded);     // counts since process start
int release_operators();                                // drop in-memory sets no handle uses

// Unit surface grid of matmgnt3d.cpp:375-397 (boundary points of an n^3 grid on [-R, R]^3).
std::vector<double> surface_grid(int n, double R);

}  // namespace kifmm
