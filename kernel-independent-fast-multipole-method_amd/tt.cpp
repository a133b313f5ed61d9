// tt.cpp -- command-line driver with the behaviour of the reference's
// fmm3d/tt.cpp:40-111, written against the drop-in header
// include/kifmm3d/FMM3d.hpp: reads an options file (-numsrc, -numtrg, -kt and
// the -fmm3d_* keys), makes random sources/targets in [-1,1]^3 (normals =
// positions, coefs {1, 0.25}), runs setup, three evals and a check.
// Unlike the reference the random data are seeded (reproducible).
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>

#include "kifmm3d/FMM3d.hpp"

using namespace kifmm3d_cxx;

static int options_create(const char* file, std::map<std::string, std::string>& opts)
{  // tt.cpp:23-38: first token is the key, the rest of the line the value
  std::ifstream fin(file);
  if (!fin.good()) {
    std::cerr << "wrong option file" << std::endl;
    return 1;
  }
  std::string line;
  while (std::getline(fin, line)) {
    std::istringstream ss(line);
    std::string key;
    if (!(ss >> key)) continue;
    std::string rest;
    std::getline(ss, rest);
    opts[key] = rest;
  }
  return 0;
}

static int get_int(std::map<std::string, std::string>& o, const char* k, int def)
{
  auto it = o.find(k);
  if (it == o.end()) return def;
  std::istringstream ss(it->second);
  int v = def;
  ss >> v;
  return v;
}

int main(int argc, char** argv)
{
  if (argc != 2) {
    std::cerr << "usage: tt <optionsfile>" << std::endl;
    return 1;
  }
  std::map<std::string, std::string> opts;
  if (options_create(argv[1], opts)) return 1;
  const int numsrc = get_int(opts, "-numsrc", 10000), numtrg = get_int(opts, "-numtrg", 20000);
  const int kt = get_int(opts, "-kt", 111);
  Kernel3d knl(kt, {1.0, 0.25});
  std::mt19937_64 rng(20041);
  std::uniform_real_distribution<double> u(-1.0, 1.0), u01(0.0, 1.0);
  DblNumMat srcpos(3, numsrc), trgpos(3, numtrg);
  for (int i = 0; i < numsrc; i++)
    for (int d = 0; d < 3; d++) srcpos(d, i) = u(rng);
  for (int i = 0; i < numtrg; i++)
    for (int d = 0; d < 3; d++) trgpos(d, i) = u(rng);
  DblNumVec srcden(knl.sdof() * numsrc), trgval(knl.tdof() * numtrg);
  for (int i = 0; i < srcden.m(); i++) srcden(i) = u01(rng);

  FMM3d* fmm = new FMM3d("fmm3d_");
  fmm->srcpos() = &srcpos;
  fmm->srcnor() = &srcpos;
  fmm->trgpos() = &trgpos;
  fmm->ctr() = Point3(0, 0, 0);
  fmm->rootlvl() = 0;
  fmm->knl() = knl;
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  if (fmm->setup(opts)) {
    std::cerr << "setup failed: " << fmm->last_error() << std::endl;
    return 1;
  }
  std::cout << "fmm setup used " << std::chrono::duration<double>(clk::now() - t0).count() << "secs " << std::endl;
  for (int i = 0; i < 3; i++) {
    t0 = clk::now();
    if (fmm->eval(srcden, trgval)) {
      std::cerr << "eval failed: " << fmm->last_error() << std::endl;
      return 1;
    }
    std::cout << "fmm eval used " << std::chrono::duration<double>(clk::now() - t0).count() << "secs " << std::endl;
  }
  t0 = clk::now();
  if (fmm->check(srcden, trgval, 20)) return 1;
  std::cout << "fmm check used " << std::chrono::duration<double>(clk::now() - t0).count() << "sec " << std::endl;
  delete fmm;
  return 0;
}
