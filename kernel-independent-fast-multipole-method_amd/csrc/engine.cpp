// engine.cpp -- plan construction, device memory and the evaluate pipeline.
//
// Stage order and the W/X direct-versus-expansion decisions follow
// fmm3d/fmm3d_eval.cpp:24-243; each stage runs as batched launches over
// whole levels (up/down GEMMs, V-list FFTs) or over all leaves at once
// (near field, L2T/M2T, S2M, S2L).
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>

namespace kifmm {

#define HIPCHK(call)                                                                                        \
  do {                                                                                                      \
    hipError_t e_ = (call);                                                                                 \
    if (e_ != hipSuccess)                                                                                   \
      throw Error(e_ == hipErrorOutOfMemory ? KIFMM3D_ENOMEM : KIFMM3D_EDEVICE,                             \
                  std::string(#call) + ": " + hipGetErrorString(e_));                                       \
  } while (0)

template <class T>
void DVec<T>::reset()
{
  if (p && own) (void)hipFree(p);
  p = nullptr;
  n = 0;
  own = true;
}
template <class T>
void DVec<T>::alloc(size_t count)
{
  reset();
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void**)&p, count * sizeof(T)));
  n = count;
}
template <class T>
void DVec<T>::upload(const std::vector<T>& v)
{
  alloc(v.size());
  if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}
template <class T>
bool DVec<T>::ensure(size_t count)
{
  if (count <= n && p != nullptr) return false;
  alloc(count);
  return true;
}
template struct DVec<double>;
template struct DVec<float>;
template struct DVec<int>;
template struct DVec<short>;

namespace {

size_t pool_i(std::vector<int>& I, const std::vector<int>& v)
{
  while (I.size() % 4) I.push_back(0);
  size_t off = I.size();
  I.insert(I.end(), v.begin(), v.end());
  return off;
}
size_t pool_d(std::vector<double>& D, const std::vector<double>& v)
{
  while (D.size() % 4) D.push_back(0.0);
  size_t off = D.size();
  D.insert(D.end(), v.begin(), v.end());
  return off;
}

}  // namespace

Engine::Engine(int kt, const double* coefs, int ncoefs, int np, int ptsmax, int maxlevel, const double* ctr,
               int rootlvl)
{
  user_ = kernel_info(kt, coefs, ncoefs);
  mm_ = kernel_info(mm_kernel_type(kt), user_.coefs, 2);
  kcu_ = kernel_consts(user_);
  kcm_ = kernel_consts(mm_);
  if (np != 4 && np != 6 && np != 8) throw Error(KIFMM3D_EINVAL, "np must be 4, 6 or 8");
  if (ptsmax < 1) throw Error(KIFMM3D_EINVAL, "ptsmax must be positive");
  if (maxlevel < 1 || maxlevel > 16) throw Error(KIFMM3D_EINVAL, "maxlevel must be in [1, 16]");
  np_ = np;
  rootlvl_ = rootlvl;
  for (int d = 0; d < 3; d++) ctr_[d] = ctr ? ctr[d] : 0.0;
  tree_.ptsmax = ptsmax;
  tree_.maxlevel = maxlevel;
  tree_.rootlvl = rootlvl;
  for (int d = 0; d < 3; d++) tree_.ctr[d] = ctr_[d];
  if (const char* rb = std::getenv("KIFMM3D_M2L_RB")) m2l_rb_ = std::atoi(rb);
  if (const char* bs = std::getenv("KIFMM3D_M2L_BS")) m2l_bs_ = std::atoi(bs);
  if (const char* v = std::getenv("KIFMM3D_P2P")) p2p_variant_ = std::atoi(v);
  if (const char* g = std::getenv("KIFMM3D_GEMM")) gemm_mfma_ = std::string(g) == "valu" ? 0 : std::string(g) == "tile64" ? 1 : 2;
  if (const char* g = std::getenv("KIFMM3D_GEMM_SPLIT")) split_blocks_ = std::max(1, std::atoi(g));
  if (const char* m = std::getenv("KIFMM3D_M2L")) {
    const std::string v(m);
    m2l_tile_ = v != "group" && v != "pair" && v != "matrix";
    m2l_matrix_ = v != "group" && v != "pair";
    m2l_wave_ = v == "wave";
  }
  if (const char* d = std::getenv("KIFMM3D_M2L_DEPTH")) m2l_depth_ = std::atoi(d);
  if (const char* d = std::getenv("KIFMM3D_M2L_FPART")) m2l_fpart_ = std::atoi(d);
  if (const char* gph = std::getenv("KIFMM3D_GRAPH")) use_graph_ = std::atoi(gph) != 0;
  if (const char* ov = std::getenv("KIFMM3D_OVERLAP")) overlap_ = std::atoi(ov) != 0;
  if (const char* v = std::getenv("KIFMM3D_S2M32")) s2m32_ = std::atoi(v) != 0;
}

Engine::~Engine()
{
  drop_graph();
  for (hipEvent_t e : evpool_) (void)hipEventDestroy(e);
  if (stream_) (void)hipStreamDestroy(stream_);
  if (side_) (void)hipStreamDestroy(side_);
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (join_ev_) (void)hipEventDestroy(join_ev_);
}

void Engine::setup(const double* spos, const double* snor, int ns, const double* tpos, int nt)
{
  if (ns < 0 || nt < 0 || (ns > 0 && !spos) || (nt > 0 && !tpos)) throw Error(KIFMM3D_EINVAL, "bad point arrays");
  if (user_.needs_normal && ns > 0 && !snor) throw Error(KIFMM3D_EINVAL, "this kernel needs source normals");
  // every device array below is replaced: a graph of the previous set-up
  // must never be launched again
  drop_graph();
  gen_++;
  setup_done_ = false;
  host_done_ = false;
  const bool trace = std::getenv("KIFMM3D_TRACE") != nullptr;
  auto clk = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    auto now = std::chrono::steady_clock::now();
    if (trace) std::fprintf(stderr, "[kifmm3d] setup %-10s %8.3f s\n", what, std::chrono::duration<double>(now - clk).count());
    clk = now;
  };
  ops_ = get_operators(mm_, np_);
  lap("operators");
  Tree t;
  t.ptsmax = tree_.ptsmax;
  t.maxlevel = tree_.maxlevel;
  t.rootlvl = tree_.rootlvl;
  for (int d = 0; d < 3; d++) t.ctr[d] = tree_.ctr[d];
  // the per-point tree work runs on the GPU when one is visible (identical
  // tree; KIFMM3D_TREE=host keeps it on the host)
  int ndev = 0;
  const bool have_dev = hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0;
  const char* tb = std::getenv("KIFMM3D_TREE");
  if (have_dev && !(tb && std::string(tb) == "host"))
    t.build_device(spos, ns, tpos, nt);
  else
    t.build(spos, ns, tpos, nt);
  tree_ = std::move(t);
  ns_ = ns;
  nt_ = nt;
  tpos_host_.assign(tpos, tpos + 3 * (size_t)nt);
  lap("tree");
  build_plan();
  lap("plan");
  host_done_ = true;
  // device phase
  if (!have_dev) throw Error(KIFMM3D_EDEVICE, "no HIP device is visible (tree, lists and operators are built)");
  if (!stream_) HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  if (!side_) {
    HIPCHK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
  }
  // leaf-ordered coordinates (fmm3d_setup.cpp:65-123) as SoA
  std::vector<double> ax(ns), ay(ns), az(ns);
  for (int k = 0; k < tree_.sextcnt; k++) {
    const double* p = spos + 3 * (size_t)tree_.sperm[k];
    ax[k] = p[0]; ay[k] = p[1]; az[k] = p[2];
  }
  sx_.upload(ax); sy_.upload(ay); sz_.upload(az);
  if (user_.needs_normal) {
    for (int k = 0; k < tree_.sextcnt; k++) {
      const double* p = snor + 3 * (size_t)tree_.sperm[k];
      ax[k] = p[0]; ay[k] = p[1]; az[k] = p[2];
    }
    snx_.upload(ax); sny_.upload(ay); snz_.upload(az);
  }
  std::vector<double> bx(nt), by(nt), bz(nt);
  for (int k = 0; k < tree_.textcnt; k++) {
    const double* p = tpos + 3 * (size_t)tree_.tperm[k];
    bx[k] = p[0]; by[k] = p[1]; bz[k] = p[2];
  }
  tx_.upload(bx); ty_.upload(by); tz_.upload(bz);
  {
    std::vector<int> inv(tree_.sperm.size());
    for (size_t k = 0; k < inv.size(); k++) inv[tree_.sperm[k]] = (int)k;
    siperm_.upload(inv);
  }
  tperm_.upload(tree_.tperm);
  {
    std::vector<int> inv(tree_.tperm.size());
    for (size_t k = 0; k < inv.size(); k++) inv[tree_.tperm[k]] = (int)k;
    tiperm_.upload(inv);
  }
  upload_static();
  near_flags();
  HIPCHK(hipDeviceSynchronize());
  lap("upload");
  gen_++;
  setup_done_ = true;
}

// Which near-field entries the mixed form's FP32 near fields must leave to
// the FP64 pass.  Their displacements come from FP32 offsets relative to the
// chunk's first target, each rounded to 2^-24 of at most span (the largest
// distance of the entry's points from that origin), so a pair at distance r
// gets 1/r with a relative error of up to ~2 sqrt(3) 2^-24 span / r.  An
// entry whose closest pair of distinct points (FP64, measured on the device
// at set-up) is nearer than span / 128 -- a relative error above 2.6e-5 for
// that pair, which dominates its target -- or nearer than 1e-7, so that the
// reference's r < 1e-8 exclusion (kernel3d.cpp:21, :132) cannot be decided
// from the offsets, is flagged 2: the FP32 kernels skip it and the FP64
// kernel adds it (near_fix).  Coincident points, which fall in one leaf,
// give equal offsets and drop out of the own-leaf entry (flag 1) exactly.
// Uniform clouds flag almost nothing; tight clusters flag the entries of
// their densest leaves.
void Engine::near_flags()
{
  Plan& P = plan_;
  P.near_fix = ChunkList{};
  fix_.reset();
  nflag_ = 0;
  const bool fp32_near = user_.kt == 111 || near32_ok();
  if (!fp32_near || P.near.n == 0) return;
  std::vector<int> eptr(P.near.n + 1), tbeg(P.near.n), tnum(P.near.n);
  HIPCHK(hipMemcpy(eptr.data(), I_.p + P.near.eptr, sizeof(int) * eptr.size(), hipMemcpyDeviceToHost));
  const int nent = eptr[P.near.n];
  std::copy(P.I.begin() + P.near.tbeg, P.I.begin() + P.near.tbeg + P.near.n, tbeg.begin());
  std::copy(P.I.begin() + P.near.tnum, P.I.begin() + P.near.tnum + P.near.n, tnum.begin());
  DVec<double> ext, text;
  ext.alloc((size_t)2 * std::max(nent, 1));
  text.alloc((size_t)P.near.n);
  dev::Points src{sx_.p, sy_.p, sz_.p, snx_.p, sny_.p, snz_.p, nullptr, 0, 0};
  dev::Targets trg{tx_.p, ty_.p, tz_.p, nullptr, 0};
  dev::near_extent(stream_, P.near.n, I_.p + P.near.tbeg, I_.p + P.near.tnum, I_.p + P.near.eptr,
                   (const int2*)(I_.p + P.near.ent), src, trg, (double2*)ext.p, text.p);
  HIPCHK(hipGetLastError());
  std::vector<double> e2((size_t)2 * nent), t2(P.near.n);
  HIPCHK(hipStreamSynchronize(stream_));
  if (nent) HIPCHK(hipMemcpy(e2.data(), ext.p, sizeof(double) * e2.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(t2.data(), text.p, sizeof(double) * t2.size(), hipMemcpyDeviceToHost));
  std::vector<int> chk(P.I.begin() + P.near.chk, P.I.begin() + P.near.chk + nent);
  const int* ent = P.I.data() + P.near.ent;
  std::vector<int> F, ftb, ftn, fep(1, 0), fent;
  for (int c = 0; c < P.near.n; c++) {
    bool any = false;
    for (int e = eptr[c]; e < eptr[c + 1]; e++) {
      const double span2 = std::max(t2[c], e2[2 * (size_t)e + 1]);
      const double rmin2 = e2[2 * (size_t)e];
      if (rmin2 < 1e-14 || rmin2 < span2 * (1.0 / 16384.0)) {
        chk[e] = 2;
        nflag_++;
        if (!any) {
          ftb.push_back(tbeg[c]);
          ftn.push_back(tnum[c]);
          any = true;
        }
        fent.push_back(ent[2 * e]);
        fent.push_back(ent[2 * e + 1]);
      }
    }
    if (any) fep.push_back((int)fent.size() / 2);
  }
  if (std::getenv("KIFMM3D_TRACE"))
    std::fprintf(stderr, "[kifmm3d] near field: %ld of %d entries left to the FP64 pass of the mixed form\n", nflag_,
                 nent);
  if (nent) HIPCHK(hipMemcpy(I_.p + P.near.chk, chk.data(), sizeof(int) * nent, hipMemcpyHostToDevice));
  std::copy(chk.begin(), chk.end(), P.I.begin() + P.near.chk);
  if (ftb.empty()) return;
  P.near_fix.n = (int)ftb.size();
  P.near_fix.tbeg = pool_i(F, ftb);
  P.near_fix.tnum = pool_i(F, ftn);
  P.near_fix.eptr = pool_i(F, fep);
  P.near_fix.ent = pool_i(F, fent);
  fix_.upload(F);
}

void Engine::upload_static()
{
  const Operators& O = *ops_;
  for (int t = 0; t < 4; t++) unit_[t].upload(O.unit[t]);
  uc2ue_.upload(O.uc2ue);
  ue2uc_.upload(O.ue2uc);
  dc2de_.upload(O.dc2de);
  de2dc_.upload(O.de2dc);
  m2l_.upload(O.m2l);
  {  // frequency-major copies for the matrix-core M2L: [C][343][sdof][tdof] complex
    const size_t st = (size_t)mm_.sdof * mm_.tdof;
    std::vector<double> t((size_t)2 * 343 * st * O.C);
    for (int o = 0; o < 343; o++)
      for (size_t e = 0; e < st; e++)
        for (int f = 0; f < O.C; f++)
          for (int c = 0; c < 2; c++)
            t[(((size_t)f * 343 + o) * st + e) * 2 + c] = O.m2l[(((size_t)o * st + e) * O.C + f) * 2 + c];
    opsT_.upload(t);
    std::vector<float> t32(t.begin(), t.end());
    opsT32_.upload(t32);
    if (st == 1) {  // tiled form: A-operand tables [C][26 x 32][2], expanded from the optab of the plan
      const int* optab = plan_.I.data() + plan_.optab;
      std::vector<double> a((size_t)O.C * 832 * 2 * 2, 0.0);
      for (int f = 0; f < O.C; f++)
        for (int e = 0; e < 832; e++)
          for (int q = 0; q < 2; q++) {
            const int n = e >> 5, k = (e >> 3) & 3, c = e & 7;
            const int o = optab[(n * 8 + (q ? 4 + k : k)) * 8 + c];
            if (o < 0) continue;
            for (int z = 0; z < 2; z++)
              a[(((size_t)f * 832 + e) * 2 + q) * 2 + z] = t[((size_t)f * 343 + o) * 2 + z];
          }
      atab_.upload(a);
      std::vector<float> a32(a.begin(), a.end());
      atab32_.upload(a32);
    }
  }
  c2s_.upload(O.cube2surf);
  std::vector<int> s2c(O.nsp[SURF_DC], 0);
  for (int c = 0; c < np_ * np_ * np_; c++)
    if (O.cube2surf[c] >= 0) s2c[O.cube2surf[c]] = c;
  s2c_.upload(s2c);
  std::vector<double> tw(2 * O.n);
  for (int m = 0; m < O.n; m++) {
    tw[2 * m] = std::cos(2.0 * M_PI * m / O.n);
    tw[2 * m + 1] = -std::sin(2.0 * M_PI * m / O.n);
  }
  tw_.upload(tw);
  I_.upload(plan_.I);
  D_.upload(plan_.D);
  {  // FP32 A operands of the parent-grouped M2L, expanded per (frequency, direction)
    const size_t na = (size_t)mm_.tdof * 4 * mm_.sdof * 64;
    avec32_.alloc((size_t)O.C * 26 * na);
    dev::expand_m2l_vec(stream_, O.C, mm_.sdof, mm_.tdof, I_.p + plan_.aidx[1], opsT32_.p, avec32_.p);
  }
  {  // fused operator of the mixed up pass: m2me_ = UC2UE x [UE2UC_0 .. UE2UC_7],
     // compensated (Dot2) like the pseudoinverse applies
    const int pUE = O.pln[SURF_UE], pUC = O.pln[SURF_UC];
    m2me_.alloc((size_t)8 * pUE * pUE);
    dev::GemmArgs a{};
    a.M = pUE; a.Kseg = pUC; a.nseg = 1; a.Rc = 1; a.ngroup = 1; a.maxcols = a.ncol = 8 * pUE;
    a.A = uc2ue_.p; a.a_gstride = 0;
    a.B = ue2uc_.p; a.bnode = I_.p + plan_.fuse_idx;
    a.C = m2me_.p; a.cnode = I_.p + plan_.fuse_idx; a.gbeg = I_.p + plan_.fuse_idx_g;
    a.compensated = 1;
    a.mfma = gemm_mfma_;
    a.ctot = a.ncol;
    dev::gemm(stream_, a);
    HIPCHK(hipGetLastError());
  }
  if (!plan_.down_mixed.empty()) {  // l2de_ = [DC2DE | DC2DE x DE2DC_c], c = 0..7
    const int pDE = O.pln[SURF_DE], pDC = O.pln[SURF_DC];
    l2de_.alloc((size_t)8 * pDE * 2 * pDC);
    for (int k = 0; k < 8; k++)
      HIPCHK(hipMemcpyAsync(l2de_.p + (size_t)k * pDE * 2 * pDC, dc2de_.p, sizeof(double) * pDE * pDC,
                            hipMemcpyDeviceToDevice, stream_));
    dev::GemmArgs a{};
    a.M = pDE; a.Kseg = pDC; a.nseg = 1; a.Rc = 1; a.ngroup = 1; a.maxcols = a.ncol = 8 * pDE;
    a.A = dc2de_.p;
    a.B = de2dc_.p; a.bnode = I_.p + plan_.fdn_b;
    a.C = l2de_.p; a.cnode = I_.p + plan_.fdn_c; a.gbeg = I_.p + plan_.fdn_g;
    a.compensated = 1;
    a.mfma = gemm_mfma_;
    a.ctot = a.ncol;
    dev::gemm(stream_, a);
    HIPCHK(hipGetLastError());
  }
}

void Engine::build_plan()
{
  const Tree& T = tree_;
  const Operators& O = *ops_;
  Plan P;
  const int sdof = user_.sdof, tdof = user_.tdof, sm = mm_.sdof;
  const int pUE = O.pln[SURF_UE], pUC = O.pln[SURF_UC], pDE = O.pln[SURF_DE], pDC = O.pln[SURF_DC];
  auto geom = [&](int g, std::vector<double>& D) {
    double c[3];
    T.center(g, c);
    D.push_back(c[0]); D.push_back(c[1]); D.push_back(c[2]); D.push_back(T.radius(g));
  };
  auto scales = [&](int l, int sign, int len) {
    std::vector<double> v(len);
    for (int i = 0; i < len; i++) v[i] = std::pow(2.0, sign * l * mm_.deg[i % sm]);
    return v;
  };
  std::vector<std::vector<int>> by_depth(std::max(T.nlevel, 1));
  for (int g = 0; g < T.nnode; g++) by_depth[T.depth[g]].push_back(g);
  for (int g = 0; g < T.nnode; g++) {
    if ((T.tag[g] & NODE_SRC) && T.depth[g] < 2) P.nlow_s = std::max(P.nlow_s, T.sndeidx[g] + 1);
    if ((T.tag[g] & NODE_TRG) && T.depth[g] < 2) P.nlow_t = std::max(P.nlow_t, T.tndeidx[g] + 1);
  }
  auto list = [&](int l, int g) {
    return std::make_pair(T.lidx[l].begin() + T.lptr[l][g], T.lidx[l].begin() + T.lptr[l][g + 1]);
  };

  // S2M items: source leaves at depth >= 2 (fmm3d_eval.cpp:72-77)
  {
    std::vector<int> out, eptr(1, 0), ent;
    std::vector<double> gm;
    for (int g = 0; g < T.nnode; g++)
      if ((T.tag[g] & NODE_SRC) && T.terminal(g) && T.depth[g] >= 2) {
        out.push_back(T.sndeidx[g]);
        geom(g, gm);
        ent.push_back(T.sextbeg[g]);
        ent.push_back(T.sextnum[g]);
        eptr.push_back((int)ent.size() / 2);
      }
    for (size_t i = 1; i < ent.size(); i += 2) P.s2m_pairs += (long)ent[i] * O.nsp[SURF_UC];
    P.s2m.n = (int)out.size();
    P.s2m.out = pool_i(P.I, out);
    P.s2m.eptr = pool_i(P.I, eptr);
    P.s2m.ent = pool_i(P.I, ent);
    P.s2m.geom = pool_d(P.D, gm);
  }
  // up pass, deepest level first (fmm3d_eval.cpp:69-92)
  for (int L = T.nlevel - 1; L >= 2; L--) {
    std::vector<int> all, inner;
    for (int g : by_depth[L])
      if (T.tag[g] & NODE_SRC) {
        all.push_back(g);
        if (!T.terminal(g)) inner.push_back(g);
      }
    if (!inner.empty()) {
      GemmOp op;
      op.M = pUC; op.Kseg = pUE; op.nseg = 8; op.ngroup = 1;
      op.ncol = op.maxcols = (int)inner.size();
      op.a_which = 1; op.b_buf = 1; op.c_buf = 0;
      std::vector<int> bnode(8 * inner.size()), cnode(inner.size());
      for (size_t n = 0; n < inner.size(); n++) {
        int g = inner[n];
        cnode[n] = T.sndeidx[g];
        for (int c = 0; c < 8; c++) {
          int ch = T.chd[g] + c;
          bnode[c * inner.size() + n] = (T.tag[ch] & NODE_SRC) ? T.sndeidx[ch] : -1;
        }
      }
      op.bnode = pool_i(P.I, bnode);
      op.cnode = pool_i(P.I, cnode);
      op.gbeg = pool_i(P.I, {0, (int)inner.size()});
      op.bscale = (long)pool_d(P.D, scales(L + 1 + rootlvl_, +1, pUE));  // child level (fmm3d_eval.cpp:84)
      P.up_flops += 2.0 * op.M * op.Kseg * op.nseg * op.ncol;
      P.up.push_back(op);
      GemmOp f = op;  // mixed form: child densities -> parent densities in one product
      f.M = pUE;
      f.a_which = 4; f.c_buf = 1;
      f.cscale = (long)pool_d(P.D, scales(L + rootlvl_, -1, pUE));
      P.up_flops_mixed += 2.0 * f.M * f.Kseg * f.nseg * f.ncol;
      P.up_mixed.push_back(f);
    }
    if (!all.empty()) {
      GemmOp op;
      op.M = pUE; op.Kseg = pUC; op.nseg = 1; op.ngroup = 1;
      op.ncol = op.maxcols = (int)all.size();
      op.a_which = 0; op.b_buf = 0; op.c_buf = 1; op.compensated = 1;
      std::vector<int> idx(all.size());
      for (size_t n = 0; n < all.size(); n++) idx[n] = T.sndeidx[all[n]];
      op.bnode = op.cnode = pool_i(P.I, idx);
      op.gbeg = pool_i(P.I, {0, (int)all.size()});
      op.cscale = (long)pool_d(P.D, scales(L + rootlvl_, -1, pUE));
      P.up_flops += 2.0 * op.M * op.Kseg * op.ncol;
      P.up.push_back(op);
      std::vector<int> leaves;
      for (int g : all)
        if (T.terminal(g)) leaves.push_back(T.sndeidx[g]);
      if (!leaves.empty()) {
        GemmOp u = op;
        u.ncol = u.maxcols = (int)leaves.size();
        u.bnode = u.cnode = pool_i(P.I, leaves);
        u.gbeg = pool_i(P.I, {0, (int)leaves.size()});
        P.up_flops_mixed += 2.0 * u.M * u.Kseg * u.ncol;
        P.up_mixed.push_back(u);
      }
    }
  }
  {
    std::vector<int> id(8 * (size_t)pUE);
    for (size_t i = 0; i < id.size(); i++) id[i] = (int)i;
    P.fuse_idx = pool_i(P.I, id);
    P.fuse_idx_g = pool_i(P.I, {0, (int)id.size()});
  }
  // near field, S2L items and L2T/M2T chunks, per target leaf
  {
    std::vector<int> tbeg, tnum, eptr(1, 0), ent, echk;
    std::vector<int> ftbeg, ftnum, feptr(1, 0), fnode, fkind;
    std::vector<double> fgeom;
    std::vector<int> xout, xeptr(1, 0), xent;
    std::vector<double> xgeom;
    std::vector<int> dent, dchk, sent, sknd;
    std::vector<double> sgeom;
    for (int g = 0; g < T.nnode; g++) {
      if (!(T.tag[g] & NODE_TRG)) continue;
      const bool leaf = T.terminal(g);
      const bool xdirect = leaf && T.textnum[g] * tdof < pDC;  // fmm3d_eval.cpp:183
      auto X = list(LIST_X, g);
      if (!xdirect && X.first != X.second) {
        xout.push_back(T.tndeidx[g]);
        geom(g, xgeom);
        for (auto it = X.first; it != X.second; ++it) {
          xent.push_back(T.sextbeg[*it]);
          xent.push_back(T.sextnum[*it]);
          P.s2l_pairs += (long)T.sextnum[*it] * O.nsp[SURF_DC];
        }
        xeptr.push_back((int)xent.size() / 2);
      }
      if (!leaf) continue;
      dent.clear(); dchk.clear(); sent.clear(); sknd.clear(); sgeom.clear();
      long npair = 0, fpair = 0;
      auto U = list(LIST_U, g);
      for (auto it = U.first; it != U.second; ++it) {
        // 1: the target's own leaf (coincident points drop out by the
        // exclusion test); entries are flagged 2 at the device set-up
        // (near_flags) where the FP32 offsets cannot resolve a pair
        dchk.push_back(*it == g ? 1 : 0);
        dent.push_back(T.sextbeg[*it]); dent.push_back(T.sextnum[*it]);
        npair += T.sextnum[*it];
      }
      auto W = list(LIST_W, g);
      for (auto it = W.first; it != W.second; ++it) {
        int v = *it;
        if (T.terminal(v) && T.sextnum[v] * sdof < pUE) {  // fmm3d_eval.cpp:161
          dchk.push_back(0);  // W and X boxes are not adjacent to the target leaf
          dent.push_back(T.sextbeg[v]); dent.push_back(T.sextnum[v]);
          npair += T.sextnum[v];
        } else {
          sent.push_back(T.sndeidx[v]); sknd.push_back(1); geom(v, sgeom);
          fpair += O.nsp[SURF_UE];
        }
      }
      if (xdirect)
        for (auto it = X.first; it != X.second; ++it) {
          dchk.push_back(0);
          dent.push_back(T.sextbeg[*it]); dent.push_back(T.sextnum[*it]);
          npair += T.sextnum[*it];
        }
      if (T.depth[g] >= 2) {  // L2T from DE (dneden is zero above depth 2)
        sent.push_back(T.tndeidx[g]); sknd.push_back(0); geom(g, sgeom);
        fpair += O.nsp[SURF_DE];
      }
      P.near_pairs += npair * T.textnum[g];
      P.far_pairs += fpair * T.textnum[g];
      for (int off = 0; off < T.textnum[g]; off += 128) {
        int n = std::min(128, T.textnum[g] - off);
        tbeg.push_back(T.textbeg[g] + off); tnum.push_back(n);
        ent.insert(ent.end(), dent.begin(), dent.end());
        echk.insert(echk.end(), dchk.begin(), dchk.end());
        eptr.push_back((int)ent.size() / 2);
        ftbeg.push_back(T.textbeg[g] + off); ftnum.push_back(n);
        fnode.insert(fnode.end(), sent.begin(), sent.end());
        fkind.insert(fkind.end(), sknd.begin(), sknd.end());
        fgeom.insert(fgeom.end(), sgeom.begin(), sgeom.end());
        feptr.push_back((int)fnode.size());
      }
    }
    P.near.n = (int)tbeg.size();
    P.near.tbeg = pool_i(P.I, tbeg); P.near.tnum = pool_i(P.I, tnum);
    P.near.eptr = pool_i(P.I, eptr); P.near.ent = pool_i(P.I, ent); P.near.chk = pool_i(P.I, echk);
    P.far.n = (int)ftbeg.size();
    P.far.tbeg = pool_i(P.I, ftbeg); P.far.tnum = pool_i(P.I, ftnum);
    P.far.eptr = pool_i(P.I, feptr); P.far.enode = pool_i(P.I, fnode);
    P.far.ekind = pool_i(P.I, fkind); P.far.egeom = pool_d(P.D, fgeom);
    P.s2l.n = (int)xout.size();
    P.s2l.out = pool_i(P.I, xout); P.s2l.eptr = pool_i(P.I, xeptr);
    P.s2l.ent = pool_i(P.I, xent); P.s2l.geom = pool_d(P.D, xgeom);
  }
  // V list per level (fmm3d_eval.cpp:115-150).  The V list of a node c is the set
  // of source children s of the 26 neighbours Q of its parent P that are not
  // adjacent to c (let3d.cpp:228-248), i.e. offset o = 2 (Q - P) + s - c with
  // |o|_inf >= 2.  The grouped plan stores, per target parent, the spectrum
  // slots of its 8 children and of the 8 children of each neighbour; the plan
  // is checked against the lists pair by pair.
  {
    const char* m2l_env = std::getenv("KIFMM3D_M2L");
    P.m2l_pairs = m2l_env && std::string(m2l_env) == "pair";
    std::vector<int> optab(26 * 64, -1);
    int dirs[26][3];
    {
      int n = 0;
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dz = -1; dz <= 1; dz++)
            if (dx || dy || dz) {
              dirs[n][0] = dx; dirs[n][1] = dy; dirs[n][2] = dz;
              for (int sc = 0; sc < 8; sc++)
                for (int cc = 0; cc < 8; cc++) {
                  const int sv[3] = {sc >> 2, (sc >> 1) & 1, sc & 1}, cv[3] = {cc >> 2, (cc >> 1) & 1, cc & 1};
                  int o[3], linf = 0;
                  for (int k = 0; k < 3; k++) {
                    o[k] = 2 * dirs[n][k] + sv[k] - cv[k];
                    linf = std::max(linf, std::abs(o[k]));
                  }
                  if (linf >= 2) optab[(n * 8 + sc) * 8 + cc] = (o[0] + 3) * 49 + (o[1] + 3) * 7 + (o[2] + 3);
                }
              n++;
            }
    }
    P.optab = pool_i(P.I, optab);
    // A-operand tables of the parent-grouped matrix-core M2L: entry (n, rt,
    // ks, lane) of the real embedding [[Mr, -Mi], [Mi, Mr]] of the direction-n
    // operator block, for the FP64 row layout (row r16 = output r16 % 8, part
    // r16 / 8) and the FP32 one (output r16 / 2, part r16 % 2)
    for (int prec = 0; prec < 2; prec++) {
      const int tmv = mm_.tdof, nks = 4 * sm;
      std::vector<int> aidx((size_t)26 * tmv * nks * 64, -1);
      for (int n = 0; n < 26; n++)
        for (int rt = 0; rt < tmv; rt++)
          for (int ks = 0; ks < nks; ks++)
            for (int ll = 0; ll < 64; ll++) {
              const int r16 = ll & 15, kk = ll >> 4;
              const int ro = prec == 0 ? (r16 & 7) : (r16 >> 1);
              const bool imrow = prec == 0 ? r16 >= 8 : (r16 & 1) == 1;
              const int qo = rt * 8 + ro, c = qo / tmv, ti = qo % tmv;
              const int kap = ks * 4 + kk, pin = kap / (8 * sm), cs = kap % (8 * sm), sc = cs / sm, j = cs % sm;
              const int o = optab[(n * 8 + sc) * 8 + c];
              if (o < 0) continue;
              const int part = imrow ? (pin == 0 ? 1 : 0) : pin;  // Yr = Mr Xr - Mi Xi, Yi = Mi Xr + Mr Xi
              const int neg = (!imrow && pin == 1) ? 1 : 0;
              const int vidx = ((o * sm + j) * tmv + ti) * 2 + part;
              aidx[(((size_t)n * tmv + rt) * nks + ks) * 64 + ll] = vidx * 2 + neg;
            }
      P.aidx[prec] = pool_i(P.I, aidx);
    }
    std::vector<int> sslot_of(T.nnode, -1), tslot_of(T.nnode, -1);
    for (int L = 2; L < T.nlevel; L++) {
      std::vector<int> trg, src, vptr(1, 0), vsrc, vop;
      long npair = 0;
      for (int g : by_depth[L]) {
        if (!(T.tag[g] & NODE_TRG)) continue;
        auto V = list(LIST_V, g);
        if (V.first == V.second) continue;
        tslot_of[g] = (int)trg.size();
        trg.push_back(T.tndeidx[g]);
        for (auto it = V.first; it != V.second; ++it) {
          int v = *it;
          if (sslot_of[v] < 0) {
            sslot_of[v] = (int)src.size();
            src.push_back(T.sndeidx[v]);
          }
          int d[3], linf = 0;
          for (int k = 0; k < 3; k++) {
            d[k] = T.path[3 * v + k] - T.path[3 * g + k];
            linf = std::max(linf, std::abs(d[k]));
          }
          if (linf < 2 || linf > 3) throw Error(KIFMM3D_EINVAL, "internal: V-list offset out of range");
          npair++;
          if (P.m2l_pairs) {
            vsrc.push_back(sslot_of[v]);
            vop.push_back((d[0] + 3) * 49 + (d[1] + 3) * 7 + (d[2] + 3));
          }
        }
        if (P.m2l_pairs) vptr.push_back((int)vsrc.size());
      }
      if (trg.empty()) continue;
      VLevel vl;
      vl.depth = L;
      vl.nsrc = (int)src.size();
      vl.ntrg = (int)trg.size();
      vl.npair = (int)npair;
      vl.src_node = pool_i(P.I, src);
      vl.trg_node = pool_i(P.I, trg);
      if (P.m2l_pairs) {
        vl.vptr = pool_i(P.I, vptr);
        vl.vsrc = pool_i(P.I, vsrc);
        vl.vop = pool_i(P.I, vop);
      } else {
        // parents at depth L-1 with a child that has a V list
        std::unordered_map<uint64_t, int> cell;
        for (int g : by_depth[L - 1])
          cell[((uint64_t)T.path[3 * g] << 40) | ((uint64_t)T.path[3 * g + 1] << 20) | (uint64_t)T.path[3 * g + 2]] = g;
        std::vector<int> tsl, ssl, gpar;
        long checked = 0;
        std::vector<int> mine;
        for (int pg : by_depth[L - 1]) {
          if (T.terminal(pg)) continue;
          bool any = false;
          for (int c = 0; c < 8; c++) any = any || tslot_of[T.chd[pg] + c] >= 0;
          if (!any) continue;
          gpar.push_back(pg);
          for (int c = 0; c < 8; c++) tsl.push_back(tslot_of[T.chd[pg] + c]);
          const int lim = 1 << (L - 1);
          for (int n = 0; n < 26; n++) {
            int q[3];
            bool in = true;
            for (int k = 0; k < 3; k++) {
              q[k] = T.path[3 * pg + k] + dirs[n][k];
              in = in && q[k] >= 0 && q[k] < lim;
            }
            int qg = -1;
            if (in) {
              auto it = cell.find(((uint64_t)q[0] << 40) | ((uint64_t)q[1] << 20) | (uint64_t)q[2]);
              if (it != cell.end() && !T.terminal(it->second)) qg = it->second;
            }
            for (int sc = 0; sc < 8; sc++) ssl.push_back(qg >= 0 ? sslot_of[T.chd[qg] + sc] : -1);
          }
          // the grouped pairs of each target child must be exactly its V list
          const int* sp = &ssl[ssl.size() - 26 * 8];
          for (int c = 0; c < 8; c++) {
            int g = T.chd[pg] + c;
            if (tslot_of[g] < 0) continue;
            mine.clear();
            for (int n = 0; n < 26; n++)
              for (int sc = 0; sc < 8; sc++)
                if (sp[n * 8 + sc] >= 0 && optab[(n * 8 + sc) * 8 + c] >= 0) mine.push_back(sp[n * 8 + sc]);
            auto V = list(LIST_V, g);
            std::vector<int> ref;
            for (auto it = V.first; it != V.second; ++it) ref.push_back(sslot_of[*it]);
            std::sort(mine.begin(), mine.end());
            std::sort(ref.begin(), ref.end());
            if (mine != ref) throw Error(KIFMM3D_EINVAL, "internal: grouped M2L plan differs from the V list");
            checked += (long)ref.size();
          }
        }
        if (checked != npair) throw Error(KIFMM3D_EINVAL, "internal: grouped M2L plan misses V pairs");
        vl.npar = (int)(tsl.size() / 8);
        std::vector<int> pm(vl.npar, 0);  // bit n: the parent has a source child in direction n
        for (int q = 0; q < vl.npar; q++)
          for (int n = 0; n < 26; n++)
            for (int sc = 0; sc < 8; sc++)
              if (ssl[((size_t)q * 26 + n) * 8 + sc] >= 0) pm[q] |= 1 << n;
        // the parent-grouped form's columns: parents in decreasing order of
        // their direction masks (stable), so that the 16 parents of a
        // matrix-core tile share their directions and a tile skips more of
        // them (1M Navier clusters: tiles 51 -> 60 % occupied, M2L 20.4 ->
        // 19.9 ms); the tiled form below keeps the tree order
        std::vector<int> ord(vl.npar);
        for (int q = 0; q < vl.npar; q++) ord[q] = q;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return pm[a] > pm[b]; });
        std::vector<int> gtsl(tsl.size()), gssl(ssl.size()), gpm(pm.size());
        for (int q = 0; q < vl.npar; q++) {
          const int o = ord[q];
          std::copy(tsl.begin() + (size_t)o * 8, tsl.begin() + (size_t)o * 8 + 8, gtsl.begin() + (size_t)q * 8);
          std::copy(ssl.begin() + (size_t)o * 208, ssl.begin() + (size_t)o * 208 + 208, gssl.begin() + (size_t)q * 208);
          gpm[q] = pm[o];
        }
        vl.tslot = pool_i(P.I, gtsl);
        vl.sslot = pool_i(P.I, gssl);
        vl.pmask = pool_i(P.I, gpm);
        if (std::getenv("KIFMM3D_TRACE")) {
          // occupancy of the matrix-core M2L tiles (16 parents x 26 directions x
          // 8 x 8 child pairs) by V pairs, for one right-hand side
          long used = 0, run = 0;
          for (int q0 = 0; q0 < vl.npar; q0 += 16)
            for (int n = 0; n < 26; n++) {
              bool any = false;
              for (int q = q0; q < std::min(q0 + 16, vl.npar); q++) {
                any = any || ((gpm[q] >> n) & 1);
                for (int sc = 0; sc < 8; sc++)
                  for (int c = 0; c < 8; c++)
                    used += gssl[((size_t)q * 26 + n) * 8 + sc] >= 0 && gtsl[(size_t)q * 8 + c] >= 0 &&
                            optab[(n * 8 + sc) * 8 + c] >= 0;
              }
              run += any ? 16 * 64 : 0;
            }
          std::fprintf(stderr, "[kifmm3d] V level %2d: %7d parents %8d sources %8d targets %10ld pairs, "
                       "matrix tiles %5.1f%% occupied\n", L, vl.npar, (int)src.size(), (int)trg.size(), npair,
                       run ? 100.0 * used / run : 0.0);
        }
        // tiled form: the grouped parents under one grandparent form a tile; the
        // neighbour n of the parent at position (x, y, z) of the tile is halo
        // entry ((x+1+dx)*4 + (y+1+dy))*4 + (z+1+dz)
        if (sm == 1 && mm_.tdof == 1) {
          std::unordered_map<int, int> tile_of;
          std::vector<int> tt, ht;
          for (int q = 0; q < vl.npar; q++) {
            const int pg = gpar[q], gp = T.par[pg];
            auto ins = tile_of.emplace(gp, (int)(tt.size() / 64));
            if (ins.second) {
              tt.resize(tt.size() + 64, -1);
              ht.resize(ht.size() + 512, -1);
            }
            const int tile = ins.first->second;
            int pv[3];
            for (int k = 0; k < 3; k++) pv[k] = T.path[3 * pg + k] - 2 * T.path[3 * gp + k];
            const int pos = pv[0] * 4 + pv[1] * 2 + pv[2];
            for (int c = 0; c < 8; c++) tt[((size_t)tile * 8 + pos) * 8 + c] = tsl[(size_t)q * 8 + c];
            for (int n = 0; n < 26; n++) {
              const int h = ((pv[0] + 1 + dirs[n][0]) * 4 + (pv[1] + 1 + dirs[n][1])) * 4 + (pv[2] + 1 + dirs[n][2]);
              for (int sc = 0; sc < 8; sc++) {
                const int v = ssl[((size_t)q * 26 + n) * 8 + sc];
                int& e = ht[((size_t)tile * 64 + h) * 8 + sc];
                if (e >= 0 && v >= 0 && e != v) throw Error(KIFMM3D_EINVAL, "internal: inconsistent M2L tile halo");
                if (v >= 0) e = v;
              }
            }
          }
          vl.ntile = (int)(tt.size() / 64);
          vl.ttile = pool_i(P.I, tt);
          vl.htile = pool_i(P.I, ht);
        }
      }
      for (int g : by_depth[L]) tslot_of[g] = -1;
      vl.dscale = (long)pool_d(P.D, scales(L + rootlvl_, +1, sm));  // plnden2effden (matmgnt3d.cpp:226-235)
      P.max_vsrc = std::max(P.max_vsrc, vl.nsrc);
      P.max_vtrg = std::max(P.max_vtrg, vl.ntrg);
      P.v_pairs += vl.npair;
      P.fft_fwd += (long)vl.nsrc * mm_.sdof;
      P.fft_inv += (long)vl.ntrg * mm_.tdof;
      P.vlev.push_back(vl);
    }
  }
  {  // the inverse FFT of a level assigns the DC rows of its V-list targets
     // (each node is a target on its own level only); the other rows are zeroed
    std::vector<char> covered(T.tndecnt, 0);
    for (const VLevel& vl : P.vlev)
      for (int i = 0; i < vl.ntrg; i++) covered[P.I[vl.trg_node + i]] = 1;
    std::vector<int> rows;
    for (int n = 0; n < T.tndecnt; n++)
      if (!covered[n]) rows.push_back(n);
    P.n_dc_zero = (int)rows.size();
    P.dc_zero = pool_i(P.I, rows);
  }
  // down pass, top level first (fmm3d_eval.cpp:195-209)
  const bool fuse_down = pDC == pDE && pDC % 8 == 0 && gemm_mfma_ == 2;
  for (int L = 2; L < T.nlevel; L++) {
    std::vector<int> all;
    for (int g : by_depth[L])
      if (T.tag[g] & NODE_TRG) all.push_back(g);
    if (all.empty()) continue;
    if (L >= 3) {
      GemmOp op;
      op.M = pDC; op.Kseg = pDE; op.nseg = 1; op.ngroup = 8;
      op.a_which = 3; op.a_gstride = (long)pDC * pDE; op.b_buf = 3; op.c_buf = 2; op.accumulate = 1;
      std::vector<int> cnode, bnode, gbeg(9, 0);
      for (int c = 0; c < 8; c++) {
        gbeg[c] = (int)cnode.size();
        for (int g : all) {
          int pg = T.par[g];
          int ci = (T.path[3 * g] - 2 * T.path[3 * pg]) * 4 + (T.path[3 * g + 1] - 2 * T.path[3 * pg + 1]) * 2 +
                   (T.path[3 * g + 2] - 2 * T.path[3 * pg + 2]);
          if (ci != c) continue;
          cnode.push_back(T.tndeidx[g]);
          bnode.push_back(T.tndeidx[pg]);
        }
        op.maxcols = std::max(op.maxcols, (int)cnode.size() - gbeg[c]);
      }
      gbeg[8] = (int)cnode.size();
      op.ncol = (int)cnode.size();
      op.cnode = pool_i(P.I, cnode);
      op.bnode = pool_i(P.I, bnode);
      op.gbeg = pool_i(P.I, gbeg);
      op.bscale = (long)pool_d(P.D, scales(L - 1 + rootlvl_, +1, pDE));  // parent level (fmm3d_eval.cpp:203)
      P.down_flops += 2.0 * op.M * op.Kseg * op.ncol;
      P.down.push_back(op);
      if (fuse_down) {
        GemmOp f = op;  // [own check values; parent density] -> own density
        f.M = pDE; f.nseg = 2;
        f.a_which = 5; f.a_gstride = (long)pDE * 2 * pDC;
        f.b_buf = 2; f.b2_buf = 3; f.bscale_seg = 1; f.c_buf = 3; f.accumulate = 0;
        std::vector<int> b2(cnode);
        b2.insert(b2.end(), bnode.begin(), bnode.end());
        f.bnode = pool_i(P.I, b2);
        f.cscale = (long)pool_d(P.D, scales(L + rootlvl_, -1, pDE));
        P.down_mixed.push_back(f);
      }
    }
    GemmOp op;
    op.M = pDE; op.Kseg = pDC; op.nseg = 1; op.ngroup = 1;
    op.ncol = op.maxcols = (int)all.size();
    op.a_which = 2; op.b_buf = 2; op.c_buf = 3; op.compensated = 1;
    std::vector<int> idx(all.size());
    for (size_t n = 0; n < all.size(); n++) idx[n] = T.tndeidx[all[n]];
    op.bnode = op.cnode = pool_i(P.I, idx);
    op.gbeg = pool_i(P.I, {0, (int)all.size()});
    op.cscale = (long)pool_d(P.D, scales(L + rootlvl_, -1, pDE));
    P.down_flops += 2.0 * op.M * op.Kseg * op.ncol;
    P.down.push_back(op);
    if (fuse_down && L < 3) P.down_mixed.push_back(op);
  }
  if (fuse_down) {
    std::vector<int> b(8 * (size_t)pDE), c(8 * (size_t)pDE);
    for (int k = 0; k < 8; k++)
      for (int j = 0; j < pDE; j++) {
        b[(size_t)k * pDE + j] = k * pDE + j;                // column j of DE2DC_k
        c[(size_t)k * pDE + j] = k * 2 * pDE + pDE + j;      // column pDE + j of block k
      }
    P.fdn_b = pool_i(P.I, b);
    P.fdn_c = pool_i(P.I, c);
    P.fdn_g = pool_i(P.I, {0, 8 * pDE});
  }
  plan_ = std::move(P);
}

int Engine::choose_rhs_chunk(int nrhs)
{
  const Operators& O = *ops_;
  const Tree& T = tree_;
  double per = 8.0 * ((double)T.sextcnt * user_.sdof * 1.5 + 2.0 * T.textcnt * user_.tdof +
                      (double)T.sndecnt * (O.pln[SURF_UC] + O.pln[SURF_UE]) +
                      (double)T.tndecnt * (O.pln[SURF_DC] + O.pln[SURF_DE])) +
               16.0 * O.C * ((double)plan_.max_vsrc * mm_.sdof + (double)plan_.max_vtrg * mm_.tdof);
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  double budget = 0.6 * (double)fr + 4.0 * (double)sden32_.n +
                  8.0 * (double)(sden_.n + tval_.n + tvalT_.n + upcval_.n + upeden_.n + dncval_.n + dneden_.n) +
                  8.0 * (double)(effden_.n + effval_.n);
  int rc = std::min((int)std::max(1.0, std::floor(budget / std::max(per, 1.0))), nrhs);
  if (rc >= 16) rc -= rc % 16;  // chunks of 16k right-hand sides take the matrix-core kernels
  return rc;
}

void Engine::work_counts(int rc, size_t cnt[NWORK]) const
{
  const Operators& O = *ops_;
  const Tree& T = tree_;
  cnt[0] = (size_t)rc * T.sextcnt * user_.sdof;
  cnt[1] = (size_t)rc * T.textcnt * user_.tdof;
  cnt[2] = (size_t)rc * T.textcnt * user_.tdof;
  cnt[3] = (size_t)rc * T.sndecnt * O.pln[SURF_UC];
  cnt[4] = (size_t)rc * T.sndecnt * O.pln[SURF_UE];
  cnt[5] = (size_t)rc * T.tndecnt * O.pln[SURF_DC];
  cnt[6] = (size_t)rc * T.tndecnt * O.pln[SURF_DE];
  cnt[7] = (size_t)2 * rc * (plan_.max_vsrc + 1) * mm_.sdof * O.C;  // + the zero slot of the wave-per-tile M2L
  cnt[8] = (size_t)2 * rc * plan_.max_vtrg * mm_.tdof * O.C;
  // split-K partials of the matrix-core GEMMs with few columns, for every
  // chunk size up to rc (the last chunk of an evaluate may be smaller)
  size_t gw = 0;
  for (const std::vector<GemmOp>* ops : {&plan_.up, &plan_.up_mixed, &plan_.down, &plan_.down_mixed})
    for (const GemmOp& op : *ops)
      for (int r = 1; r <= rc; r++) {
        dev::GemmArgs a{};
        a.M = op.M; a.Kseg = op.Kseg; a.nseg = op.nseg; a.Rc = r; a.ngroup = op.ngroup; a.maxcols = op.maxcols;
        a.compensated = 0;
        a.mfma = gemm_mfma_;
        a.split_blocks = split_blocks_;
        const int ks = dev::gemm_split(a);
        if (ks > 1) gw = std::max(gw, (size_t)ks * op.ncol * r * op.M);
      }
  cnt[9] = gw;
  cnt[10] = user_.kt == 111 ? (size_t)rc * T.sextcnt : 0;
}

void Engine::ensure_work(int rc)
{
  size_t cnt[NWORK];
  work_counts(rc, cnt);
  DVec<double>* bufs[NWORK - 1] = {&sden_,   &tval_,   &tvalT_, &upcval_, &upeden_,
                                   &dncval_, &dneden_, &effden_, &effval_, &gwork_};
  bool grew = false;
  for (int b = 0; b < NWORK - 1; b++) grew |= bufs[b]->ensure(cnt[b]);
  if (user_.kt == 111) grew |= sden32_.ensure(cnt[10]);
  // graphs captured before hold the freed buffers: they are dropped here and
  // the batch graphs that recorded this engine see the new generation
  if (grew) {
    drop_graph();
    gen_++;
  }
}

void Engine::check_eval_args(const double* d_srcden, const double* d_trgval, int nrhs, int mode) const
{
  if (!setup_done_) throw Error(KIFMM3D_ESTATE, "evaluate called before setup");
  if (nrhs < 1) throw Error(KIFMM3D_EINVAL, "nrhs must be >= 1");
  if (mode != KIFMM3D_MODE_PLAIN && mode != KIFMM3D_MODE_MIXED) throw Error(KIFMM3D_EINVAL, "unsupported evaluation mode");
  if ((ns_ > 0 && !d_srcden) || (nt_ > 0 && !d_trgval)) throw Error(KIFMM3D_EINVAL, "null density or value pointer");
}

int Engine::prepare(int nrhs)
{
  const int rc = choose_rhs_chunk(nrhs);
  ensure_work(rc);
  last_in_scratch_ = false;
  return rc;
}

void Engine::record(const double* d_srcden, double* d_trgval, int nrhs, int rc, int mode, hipStream_t s)
{
  if (nt_ == 0) return;
  for (int r0 = 0; r0 < nrhs; r0 += rc) {
    const int n = std::min(rc, nrhs - r0);
    run_chunk(d_srcden + (size_t)r0 * ns_ * user_.sdof, d_trgval + (size_t)r0 * nt_ * user_.tdof, n, mode, s, nullptr);
    last_rc_ = n;
  }
}

void Engine::drop_graph()
{
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  gexec_ = nullptr;
}

void Engine::evaluate_device(const double* d_srcden, double* d_trgval, int nrhs, int mode)
{
  check_eval_args(d_srcden, d_trgval, nrhs, mode);
  last_nrhs_ = nrhs;
  for (double& m : stage_ms) m = 0.0;
  for (double& m : kernel_ms) m = 0.0;
  for (int& c : kernel_calls) c = 0;
  if (nt_ == 0) return;
  int rc = prepare(nrhs);
  if (use_graph_) {
    // one captured graph per (buffers, nrhs, mode); replayed while they stay the same
    if (!gexec_ || gsrc_ != d_srcden || gtrg_ != d_trgval || gnrhs_ != nrhs || gmode_ != mode || grc_ != rc ||
        ggen_ != gen_) {
      drop_graph();
      HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      hipGraph_t g = nullptr;
      try {
        record(d_srcden, d_trgval, nrhs, rc, mode, stream_);
      } catch (...) {
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHK(hipStreamEndCapture(stream_, &g));
      hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        gexec_ = nullptr;
        throw Error(KIFMM3D_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      gsrc_ = d_srcden; gtrg_ = d_trgval; gnrhs_ = nrhs; gmode_ = mode; grc_ = rc; ggen_ = gen_;
    }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    float ms = 0;
    hipError_t e = hipEventRecord(a, stream_);
    if (e == hipSuccess) e = hipGraphLaunch(gexec_, stream_);
    if (e == hipSuccess) e = hipEventRecord(b, stream_);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) throw Error(KIFMM3D_EDEVICE, std::string("graph launch: ") + hipGetErrorString(e));
    stage_ms[KIFMM3D_STAGE_TOTAL] = ms;
    last_rc_ = std::min(rc, nrhs - (nrhs - 1) / rc * rc);
    return;
  }
  vlev_ms_.assign(plan_.vlev.size(), 0.0);
  hipEvent_t ev[9];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  try {
    for (int r0 = 0; r0 < nrhs; r0 += rc) {
      int n = std::min(rc, nrhs - r0);
      run_chunk(d_srcden + (size_t)r0 * ns_ * user_.sdof, d_trgval + (size_t)r0 * nt_ * user_.tdof, n, mode, stream_,
                ev);
      HIPCHK(hipEventSynchronize(ev[8]));
      for (int s = 0; s < 8; s++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
        stage_ms[s] += ms;
      }
      for (size_t e = 0; e + 1 < nev_used_; e += 2) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, evpool_[e], evpool_[e + 1]));
        kernel_ms[evcls_[e / 2]] += ms;
        kernel_calls[evcls_[e / 2]] += 1;
        if (evlvl_[e / 2] >= 0) vlev_ms_[evlvl_[e / 2]] += ms;
      }
      last_rc_ = n;
    }
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  double tot = 0;
  for (int s = 0; s < 8; s++) tot += stage_ms[s];
  stage_ms[KIFMM3D_STAGE_TOTAL] = tot;
}

void Engine::evaluate_device_noalloc(const double* d_srcden, double* d_trgval, int nrhs, int mode)
{
  check_eval_args(d_srcden, d_trgval, nrhs, mode);
  void* base = nullptr;
  HIPCHK(dev::scratch_ptr(&base));
  for (double& m : stage_ms) m = 0.0;
  for (double& m : kernel_ms) m = 0.0;
  for (int& c : kernel_calls) c = 0;
  if (nt_ == 0) return;
  // bytes of the work buffers for chunks of r right-hand sides, each piece
  // starting on a 256-byte boundary of the array (debugging aid: with
  // KIFMM3D_SCRATCH_GAP = g bytes, g-byte gaps filled with 0xff bytes follow
  // every piece and are checked afterwards; KIFMM3D_SCRATCH_POISON = mask
  // fills the pieces of the set bits with 0xff bytes (NaN) first)
  const char* gap_env = std::getenv("KIFMM3D_SCRATCH_GAP");
  const char* poison_env = std::getenv("KIFMM3D_SCRATCH_POISON");
  const size_t gap = gap_env ? (size_t)std::atol(gap_env) / 256 * 256 : 0;
  const unsigned poison = poison_env ? (unsigned)std::atoi(poison_env) : 0u;
  size_t cnt[NWORK], off[NWORK], len[NWORK];
  auto layout = [&](int r) {
    work_counts(r, cnt);
    size_t at = 0;
    for (int b = 0; b < NWORK; b++) {
      off[b] = at;
      len[b] = (cnt[b] * (b == NWORK - 1 ? sizeof(float) : sizeof(double)) + 255) / 256 * 256;
      at += len[b] + gap;
    }
    return at;
  };
  int rc = choose_rhs_chunk(nrhs);
  while (rc > 1 && layout(rc) > dev::SCRATCH_BYTES) rc = rc > 16 ? rc - 16 : rc - 1;
  const size_t need = layout(rc);
  if (need > dev::SCRATCH_BYTES)
    throw Error(KIFMM3D_ENOMEM, "the no-allocation evaluate needs " + std::to_string(need) +
                                    " bytes of work memory for one right-hand side of this problem; its scratch "
                                    "array holds " + std::to_string(dev::SCRATCH_BYTES));
  // point the work buffers at their pieces of the array for this call, then
  // give the handle its own buffers back
  DVec<double>* bufs[NWORK - 1] = {&sden_,   &tval_,   &tvalT_, &upcval_, &upeden_,
                                   &dncval_, &dneden_, &effden_, &effval_, &gwork_};
  struct Saved {
    double* p;
    size_t n;
    bool own;
  } saved[NWORK - 1];
  float* const p32 = sden32_.p;
  const size_t n32 = sden32_.n;
  const bool own32 = sden32_.own;
  unsigned char* b0 = static_cast<unsigned char*>(base);
  for (int b = 0; b < NWORK - 1; b++) {
    saved[b] = {bufs[b]->p, bufs[b]->n, bufs[b]->own};
    bufs[b]->p = reinterpret_cast<double*>(b0 + off[b]);
    bufs[b]->n = cnt[b];
    bufs[b]->own = false;
  }
  sden32_.p = reinterpret_cast<float*>(b0 + off[NWORK - 1]);
  sden32_.n = cnt[NWORK - 1];
  sden32_.own = false;
  auto restore = [&] {
    for (int b = 0; b < NWORK - 1; b++) {
      bufs[b]->p = saved[b].p;
      bufs[b]->n = saved[b].n;
      bufs[b]->own = saved[b].own;
    }
    sden32_.p = p32;
    sden32_.n = n32;
    sden32_.own = own32;
  };
  try {
    for (int b = 0; b < NWORK; b++) {
      if (poison >> b & 1u) HIPCHK(hipMemset(b0 + off[b], 0xff, len[b]));
      if (gap) HIPCHK(hipMemset(b0 + off[b] + len[b], 0xff, gap));
    }
    record(d_srcden, d_trgval, nrhs, rc, mode, stream_);
    HIPCHK(hipStreamSynchronize(stream_));
    if (gap) {
      std::vector<unsigned char> g(gap);
      for (int b = 0; b < NWORK; b++) {
        HIPCHK(hipMemcpy(g.data(), b0 + off[b] + len[b], gap, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < gap; q++)
          if (g[q] != 0xff)
            throw Error(KIFMM3D_EDEVICE, "work buffer " + std::to_string(b) + " (" + std::to_string(cnt[b]) +
                                             " elements) was written " + std::to_string(q) + " bytes past its end");
      }
    }
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    restore();
    throw;
  }
  restore();
  last_nrhs_ = nrhs;
  last_in_scratch_ = true;
}

void Engine::evaluate_host(const double* srcden, double* trgval, int nrhs, int mode)
{
  if (!setup_done_) throw Error(KIFMM3D_ESTATE, "evaluate called before setup");
  if (nrhs < 1) throw Error(KIFMM3D_EINVAL, "nrhs must be >= 1");
  size_t ns = (size_t)nrhs * ns_ * user_.sdof, nt = (size_t)nrhs * nt_ * user_.tdof;
  hsrc_.ensure(ns);
  htrg_.ensure(nt);
  if (ns) HIPCHK(hipMemcpy(hsrc_.p, srcden, ns * sizeof(double), hipMemcpyHostToDevice));
  evaluate_device(hsrc_.p, htrg_.p, nrhs, mode);
  if (nt) HIPCHK(hipMemcpy(trgval, htrg_.p, nt * sizeof(double), hipMemcpyDeviceToHost));
}

int Engine::near_variant(bool mixed) const
{
  const int t = p2p_tuned_[mixed ? 1 : 0];
  // the mixed form: FP32 matrix cores for Laplace, FP32 vector kernels for the
  // kernels without normals (with normals, r.n of nearly tangent pairs cancels
  // beyond what FP32 coordinates resolve: FP64 there)
  const int v = t >= 0 ? t : mixed ? (user_.kt == 111 ? 3 : near32_ok() ? 4 : 0) : p2p_variant_;
  // the matrix-core near-field kernels are Laplace's; the other kernel types
  // run kernels 2 and 3 as the LDS-broadcast form 0
  return user_.kt != 111 && (v == 2 || v == 3) ? 0 : v;
}

// M2L form of V-list level `level` for a chunk of rc right-hand sides: the
// tuned one where set and applicable (the tiled form needs a scalar
// translation kernel and 16k right-hand sides), else tiled, parent-grouped
// matrix-core or grouped VALU as the switches allow; the per-pair form when
// the plan was built for it.
int Engine::m2l_form(bool mixed, size_t level, int rc) const
{
  if (plan_.m2l_pairs) return 3;
  const bool tile_ok = mm_.sdof == 1 && mm_.tdof == 1 && rc % 16 == 0;
  const std::vector<int>& tf = m2l_form_[mixed ? 1 : 0];
  if (level < tf.size() && tf[level] >= 0 && ((tf[level] != 0 && tf[level] != 4) || tile_ok))
    return tf[level];
  return m2l_tile_ && tile_ok ? (m2l_wave_ ? 4 : 0) : m2l_matrix_ ? 1 : 2;
}

int Engine::tune(int nrhs, int mode)
{
  if (!setup_done_) throw Error(KIFMM3D_ESTATE, "tune called before setup");
  if (nrhs < 1) throw Error(KIFMM3D_EINVAL, "nrhs must be >= 1");
  if (mode != KIFMM3D_MODE_PLAIN && mode != KIFMM3D_MODE_MIXED) throw Error(KIFMM3D_EINVAL, "unsupported evaluation mode");
  if (nt_ == 0) return 0;
  const bool mixed = mode == KIFMM3D_MODE_MIXED;
  const int mi = mixed ? 1 : 0;
  const size_t nlev = plan_.vlev.size();
  const int rc = choose_rhs_chunk(nrhs);
  // candidates: the mixed form keeps the Laplace near field on the FP32
  // matrix cores (its precision choice, not a schedule); the other kernel
  // types have the LDS-broadcast and scalar-cache forms, and in the mixed
  // form the FP32 LDS-broadcast one
  std::vector<int> nf, mf;
  if (user_.kt == 111)
    nf = mixed ? std::vector<int>{3} : std::vector<int>{2, 0, 1};
  else
    nf = mixed && near32_ok() ? std::vector<int>{4, 0, 1} : std::vector<int>{0, 1};
  if (!plan_.m2l_pairs) {
    if (mm_.sdof == 1 && mm_.tdof == 1 && rc % 16 == 0) {
      mf.push_back(0);
      mf.push_back(4);
    }
    mf.push_back(1);
    mf.push_back(2);
  }
  DVec<double> den, val;
  den.ensure((size_t)nrhs * ns_ * user_.sdof);
  val.ensure((size_t)nrhs * nt_ * user_.tdof);
  if (den.n) HIPCHK(hipMemset(den.p, 0, den.n * sizeof(double)));
  const bool graph = use_graph_;
  use_graph_ = false;
  drop_graph();
  const size_t rounds = std::max(nf.size(), std::max(mf.size(), (size_t)1));
  std::vector<double> best_lvl(nlev, 1e300);
  std::vector<int> best_form(nlev, -1);
  double best_near = 1e300;
  int best_v = -1, runs = 0;
  try {
    for (size_t r = 0; r < rounds; r++) {
      p2p_tuned_[mi] = nf[std::min(r, nf.size() - 1)];
      const int f = mf.empty() ? -1 : mf[std::min(r, mf.size() - 1)];
      m2l_form_[mi].assign(nlev, f);
      std::vector<double> lvl(nlev, 1e300);
      double near = 1e300;
      for (int k = 0; k < 3; k++) {  // one warm-up, the faster of two timed
        evaluate_device(den.p, val.p, nrhs, mode);
        runs++;
        if (k == 0) continue;
        near = std::min(near, kernel_ms[KIFMM3D_K_P2P]);
        for (size_t l = 0; l < nlev; l++) lvl[l] = std::min(lvl[l], vlev_ms_[l]);
      }
      if (near < best_near) {
        best_near = near;
        best_v = p2p_tuned_[mi];
      }
      for (size_t l = 0; l < nlev; l++)
        if (lvl[l] < best_lvl[l]) {
          best_lvl[l] = lvl[l];
          best_form[l] = f;
        }
    }
  } catch (...) {
    p2p_tuned_[mi] = -1;
    m2l_form_[mi].clear();
    use_graph_ = graph;
    throw;
  }
  p2p_tuned_[mi] = best_v;
  m2l_form_[mi] = best_form;
  use_graph_ = graph;
  gen_++;  // batch graphs recorded with the previous settings are recaptured
  return runs;
}

void Engine::set_settings(int mode, const int* in, int n)
{
  if (!host_done_) throw Error(KIFMM3D_ESTATE, "settings before setup");
  if (mode != KIFMM3D_MODE_PLAIN && mode != KIFMM3D_MODE_MIXED) throw Error(KIFMM3D_EINVAL, "unsupported evaluation mode");
  if (!in || n < 1) throw Error(KIFMM3D_EINVAL, "settings: empty array");
  const bool mixed = mode == KIFMM3D_MODE_MIXED;
  const int v = in[0];
  // kernel 3 (FP32 matrix cores) is the mixed form's Laplace near field
  // only, kernel 4 (FP32 vector) the mixed form's for kernels without normals
  if (v < -1 || v > 4 || (v == 3 && !(mixed && user_.kt == 111)) || (v == 4 && !(mixed && near32_ok())))
    throw Error(KIFMM3D_EINVAL, "settings: near-field kernel " + std::to_string(v) + " not available here");
  const size_t nlev = plan_.vlev.size();
  std::vector<int> forms(nlev, -1);
  for (size_t l = 0; l < nlev && (long)l + 1 < n; l++) {
    const int f = in[l + 1];
    if (f < -1 || f > 4 || f == 3 || ((f == 0 || f == 4) && !(mm_.sdof == 1 && mm_.tdof == 1)))
      throw Error(KIFMM3D_EINVAL, "settings: M2L form " + std::to_string(f) + " not available here");
    forms[l] = f;
  }
  p2p_tuned_[mixed ? 1 : 0] = v;
  m2l_form_[mixed ? 1 : 0] = forms;
  drop_graph();
  gen_++;
}

int Engine::settings(int mode, int* out, int n) const
{
  const bool mixed = mode == KIFMM3D_MODE_MIXED;
  const int rc = last_rc_ > 0 ? last_rc_ : 16;
  const int cnt = 1 + (int)plan_.vlev.size();
  if (out && n > 0) out[0] = near_variant(mixed);
  for (int l = 0; out && l + 1 < std::min(n, cnt); l++) out[l + 1] = m2l_form(mixed, (size_t)l, rc);
  return cnt;
}

void Engine::run_chunk(const double* srcden, double* trgval, int rc, int mode, hipStream_t s, hipEvent_t* ev)
{
  const bool mixed = mode == KIFMM3D_MODE_MIXED;
  const Operators& O = *ops_;
  const Tree& T = tree_;
  const Plan& P = plan_;
  const int* I = I_.p;
  const double* D = D_.p;
  const int sdof = user_.sdof, tdof = user_.tdof, sm = mm_.sdof, tm = mm_.tdof;
  const int pUE = O.pln[SURF_UE], pUC = O.pln[SURF_UC], pDE = O.pln[SURF_DE], pDC = O.pln[SURF_DC];
  const dev::KC kcu{kcu_.a, kcu_.b, kcu_.c}, kcm{kcm_.a, kcm_.b, kcm_.c};
  double* bufs[4] = {upcval_.p, upeden_.p, dncval_.p, dneden_.p};
  const double* opsA[6] = {uc2ue_.p, ue2uc_.p, dc2de_.p, de2dc_.p, m2me_.p, l2de_.p};
  const long trs = (long)T.textcnt * tdof;
  dev::Points src{sx_.p, sy_.p, sz_.p, snx_.p, sny_.p, snz_.p, sden_.p, (long)rc * sdof, sdof};
  // the mixed form's Laplace kernels stage FP32 densities (half the bytes, no conversions);
  // with S2M / S2L and the near field in FP32 too and every RHS block on those
  // kernels (a last block of 1-3 RHS takes the FP64 vector kernels) nothing
  // reads the FP64 copy
  float* den32 = mixed && user_.kt == 111 ? sden32_.p : nullptr;
  src.den32 = den32;
  const int nvar = near_variant(mixed);
  // the FP32 near fields skip the entries flagged for FP64 exclusion tests
  // (set-up: a pair may be closer than the FP32 offsets resolve); the FP64
  // LDS-broadcast kernel adds those entries after them
  const bool near_fix = mixed && (nvar == 3 || nvar == 4) && P.near_fix.n > 0;
  // (the FP64 pass over the flagged entries reads the FP32 copy when there is
  // no FP64 one: its positions and sums are FP64)
  double* den64 = den32 && s2m32_ && nvar == 3 && (rc % 16 == 0 || rc % 16 >= dev::PAD16_MIN) ? nullptr : sden_.p;
  dev::Targets trg{tx_.p, ty_.p, tz_.p, tval_.p, trs};

  auto run_gemm = [&](const GemmOp& op) {
    dev::GemmArgs a{};
    a.M = op.M; a.Kseg = op.Kseg; a.nseg = op.nseg; a.Rc = rc; a.ngroup = op.ngroup; a.maxcols = op.maxcols;
    a.A = opsA[op.a_which]; a.a_gstride = op.a_gstride;
    a.B = bufs[op.b_buf]; a.bnode = I + op.bnode; a.ncol = op.ncol;
    a.B2 = op.b2_buf >= 0 ? bufs[op.b2_buf] : nullptr; a.bscale_seg = op.bscale_seg;
    a.C = bufs[op.c_buf]; a.cnode = I + op.cnode; a.gbeg = I + op.gbeg;
    a.bscale = op.bscale >= 0 ? D + op.bscale : nullptr;
    a.cscale = op.cscale >= 0 ? D + op.cscale : nullptr;
    a.accumulate = op.accumulate;
    a.compensated = mixed ? 0 : op.compensated;
    a.mfma = gemm_mfma_;
    a.ctot = (long)op.ncol * rc;
    a.work = gwork_.p;
    a.work_n = (long)gwork_.n;
    a.split_blocks = split_blocks_;
    dev::gemm(s, a);
  };

  // stage and kernel-class timing: events on this stream (none when ev is
  // null, e.g. while the sequence is captured into a graph)
  auto mark = [&](int k) {
    if (ev) HIPCHK(hipEventRecord(ev[k], s));
  };
  size_t nev = 0;
  int cur_lvl = -1;  // V-list level being enqueued (for the per-level times)
  auto timed_on = [&](hipStream_t st, int cls, auto&& fn) {
    if (!ev) {
      fn();
      return;
    }
    if (evpool_.size() < nev + 2) {
      hipEvent_t a, b;
      HIPCHK(hipEventCreate(&a));
      HIPCHK(hipEventCreate(&b));
      evpool_.push_back(a);
      evpool_.push_back(b);
    }
    HIPCHK(hipEventRecord(evpool_[nev], st));
    fn();
    HIPCHK(hipEventRecord(evpool_[nev + 1], st));
    evcls_.resize(nev / 2 + 1);
    evcls_[nev / 2] = cls;
    evlvl_.resize(nev / 2 + 1);
    evlvl_[nev / 2] = cur_lvl;
    nev += 2;
  };
  auto timed = [&](int cls, auto&& fn) { timed_on(s, cls, fn); };

  mark(0);
  // zero the accumulated buffers (fmm3d_eval.cpp:33-42); upward check values
  // at depth >= 2 are overwritten by S2M (leaves) and M2M (inner nodes), the
  // downward check values of V-list targets by the inverse FFT of their level
  dev::zero(s, tval_.p, (long)rc * trs);
  dev::zero_rows(s, dncval_.p, I + P.dc_zero, P.n_dc_zero, (long)rc * pDC);
  if (P.nlow_s) dev::zero(s, upcval_.p, (long)rc * P.nlow_s * pUC);
  if (P.nlow_s) dev::zero(s, upeden_.p, (long)rc * P.nlow_s * pUE);
  if (P.nlow_t) dev::zero(s, dneden_.p, (long)rc * P.nlow_t * pDE);
  // load (fmm3d_eval.cpp:50-64)
  timed(KIFMM3D_K_GATHER,
        [&] { dev::gather(s, T.sextcnt, sdof, rc, siperm_.p, srcden, (long)ns_ * sdof, den64, den32); });
  mark(1);
  // up: S2M, then M2M + UC2UE level by level
  timed(KIFMM3D_K_S2M, [&] {
    dev::p2s(s, user_.kt, P.s2m.n, I + P.s2m.out, (const double4*)(D + P.s2m.geom), I + P.s2m.eptr,
             (const int2*)(I + P.s2m.ent), unit_[SURF_UC].p, O.nsp[SURF_UC], src, upcval_.p, rc, pUC, rc, kcu,
             mixed && s2m32_ ? 3 : std::min(p2p_variant_, 2), 0);
  });
  for (const GemmOp& op : mixed ? P.up_mixed : P.up) timed(KIFMM3D_K_GEMM_UP, [&] { run_gemm(op); });
  mark(2);
  // U list and direct W/X.  It depends only on the gathered densities, so
  // with overlap on it runs on a side stream concurrently with the up pass,
  // the V list and the down pass, and joins before L2T/M2T adds to the values.
  const bool side = overlap_ && side_;
  if (side) {
    HIPCHK(hipEventRecord(fork_ev_, s));
    HIPCHK(hipStreamWaitEvent(side_, fork_ev_, 0));
  }
  timed_on(side ? side_ : s, KIFMM3D_K_P2P, [&] {
    const int n32 = dev::p2p(side ? side_ : s, user_.kt, P.near.n, I + P.near.tbeg, I + P.near.tnum,
                             I + P.near.eptr, (const int2*)(I + P.near.ent), I + P.near.chk, src, trg, rc, kcu, nvar);
    if (near_fix && n32 > 0) {
      const int* F = fix_.p;
      dev::Points sfix = src;
      sfix.den = den64;  // null when only the FP32 copy was gathered
      dev::p2p(side ? side_ : s, user_.kt, P.near_fix.n, F + P.near_fix.tbeg, F + P.near_fix.tnum,
               F + P.near_fix.eptr, (const int2*)(F + P.near_fix.ent), nullptr, sfix, trg, n32, kcu, 0);
    }
  });
  if (side) HIPCHK(hipEventRecord(join_ev_, side_));
  mark(3);
  // V list: tiled matrix-core M2L for scalar kernels on chunks of 16k
  // right-hand sides, the parent-grouped matrix-core form otherwise (VALU
  // grouped and per-pair forms for comparison)
  for (size_t li = 0; li < P.vlev.size(); li++) {
    const VLevel& vl = P.vlev[li];
    const int form = m2l_form(mixed, li, rc);
    const bool tiled = form == 0 || form == 4;
    const int ri = tiled ? 1 : 0, single = mixed && (form <= 1 || form == 4) ? 1 : 0;
    cur_lvl = (int)li;
    timed(KIFMM3D_K_FFT_FWD, [&] {
      dev::fft_fwd(s, np_, vl.nsrc, I + vl.src_node, rc, upeden_.p, pUE, sm, D + vl.dscale, c2s_.p, effden_.p, single,
                   ri, (const double2*)tw_.p);
    });
    timed(KIFMM3D_K_M2L, [&] {
      if (form == 4) {
        // slot vl.nsrc (C x rc float2 or double2: rc or 2 rc doubles per
        // frequency) is the zero slot that absent halo children read
        const long zs = (long)O.C * rc * (single ? 1 : 2);
        dev::zero(s, effden_.p + (size_t)vl.nsrc * zs, zs);
        dev::m2l_wave(s, O.C, vl.ntile, rc, vl.nsrc, I + vl.ttile, I + vl.htile,
                      single ? (const void*)atab32_.p : (const void*)atab_.p, effden_.p, effval_.p, m2l_depth_,
                      single);
      } else if (tiled)
        dev::m2l_tiled(s, O.C, vl.ntile, rc, I + vl.ttile, I + vl.htile,
                       single ? (const void*)atab32_.p : (const void*)atab_.p, effden_.p, effval_.p, single,
                       m2l_fpart_);
      else if (form == 3)
        dev::m2l(s, sm, tm, O.C, vl.ntrg, rc, I + vl.vptr, I + vl.vsrc, I + vl.vop, (const double2*)m2l_.p,
                 (const double2*)effden_.p, (double2*)effval_.p);
      else if (form == 1)
        dev::m2l_matrix(s, sm, tm, O.C, vl.npar, rc, I + vl.tslot, I + vl.sslot, I + vl.pmask, I + P.aidx[single],
                        single ? (const void*)opsT32_.p : (const void*)opsT_.p, avec32_.p, effden_.p, effval_.p,
                        single);
      else
        dev::m2l_grouped(s, sm, tm, O.C, vl.npar, rc, m2l_rb_, m2l_bs_, I + vl.tslot, I + vl.sslot, I + P.optab,
                         (const double2*)m2l_.p, (const double2*)effden_.p, (double2*)effval_.p);
    });
    timed(KIFMM3D_K_FFT_INV, [&] {
      dev::fft_inv(s, np_, vl.ntrg, I + vl.trg_node, rc, effval_.p, single, ri, tm, s2c_.p, O.nsp[SURF_DC],
                   dncval_.p, pDC, (const double2*)tw_.p);
    });
    cur_lvl = -1;
  }
  mark(4);
  // X list S2L
  timed(KIFMM3D_K_S2L, [&] {
    dev::p2s(s, user_.kt, P.s2l.n, I + P.s2l.out, (const double4*)(D + P.s2l.geom), I + P.s2l.eptr,
             (const int2*)(I + P.s2l.ent), unit_[SURF_DC].p, O.nsp[SURF_DC], src, dncval_.p, rc, pDC, rc, kcu,
             mixed && s2m32_ ? 3 : std::min(p2p_variant_, 2));
  });
  mark(5);
  // down: L2L + DC2DE level by level
  for (const GemmOp& op : mixed && !P.down_mixed.empty() ? P.down_mixed : P.down)
    timed(KIFMM3D_K_GEMM_DOWN, [&] { run_gemm(op); });
  mark(6);
  if (side) HIPCHK(hipStreamWaitEvent(s, join_ev_, 0));
  // L2T + M2T
  timed(KIFMM3D_K_S2P, [&] {
    dev::s2p(s, mm_.kt, P.far.n, I + P.far.tbeg, I + P.far.tnum, I + P.far.eptr, I + P.far.enode,
             I + P.far.ekind, (const double4*)(D + P.far.egeom), unit_[SURF_DE].p, unit_[SURF_UE].p,
             O.nsp[SURF_DE], dneden_.p, upeden_.p, rc, pDE, trg, rc, kcm, mixed ? 3 : p2p_variant_);
  });
  mark(7);
  // save (fmm3d_eval.cpp:221-235)
  timed(KIFMM3D_K_SCATTER, [&] {
    if (tdof <= 3)
      dev::scatter_rows(s, T.textcnt, tdof, rc, tiperm_.p, tval_.p, trs, tvalT_.p, trgval, (long)nt_ * tdof);
    else
      dev::scatter(s, T.textcnt, tdof, rc, tperm_.p, tval_.p, trs, trgval, (long)nt_ * tdof);
  });
  HIPCHK(hipGetLastError());
  mark(8);
  if (ev) nev_used_ = nev;
}

long Engine::get_buffer(int what, double* buf, long n)
{
  if (!setup_done_ || last_rc_ == 0 || last_in_scratch_) return KIFMM3D_ESTATE;
  const Operators& O = *ops_;
  const DVec<double>* src[4] = {&upeden_, &upcval_, &dneden_, &dncval_};
  const int pln[4] = {O.pln[SURF_UE], O.pln[SURF_UC], O.pln[SURF_DE], O.pln[SURF_DC]};
  const int cnt[4] = {tree_.sndecnt, tree_.sndecnt, tree_.tndecnt, tree_.tndecnt};
  if (what < 0 || what > 3) return KIFMM3D_EINVAL;
  long tot = (long)cnt[what] * pln[what];
  if (!buf) return tot;
  if (n < tot) return KIFMM3D_EINVAL;
  HIPCHK(hipStreamSynchronize(stream_));
  for (int k = 0; k < cnt[what]; k++)
    HIPCHK(hipMemcpy(buf + (size_t)k * pln[what], src[what]->p + (size_t)k * last_rc_ * pln[what],
                     sizeof(double) * pln[what], hipMemcpyDeviceToHost));
  return tot;
}

void Engine::direct(const double* srcden, bool on_device, int nrhs, const int* idx, int nchk, double* out)
{
  if (!setup_done_) throw Error(KIFMM3D_ESTATE, "direct called before setup");
  if (nrhs < 1 || nchk < 0 || (nchk > 0 && (!idx || !out)) || !srcden) throw Error(KIFMM3D_EINVAL, "bad arguments");
  if (nchk == 0) return;
  const int sdof = user_.sdof, tdof = user_.tdof;
  std::vector<double> c((size_t)3 * nchk);
  for (int k = 0; k < nchk; k++) {
    if (idx[k] < 0 || idx[k] >= nt_) throw Error(KIFMM3D_EINVAL, "target index out of range");
    for (int d = 0; d < 3; d++) c[(size_t)d * nchk + k] = tpos_host_[3 * (size_t)idx[k] + d];
  }
  const double* dsrc = srcden;
  size_t nsd = (size_t)nrhs * ns_ * sdof;
  if (!on_device) {
    hsrc_.ensure(nsd);
    HIPCHK(hipMemcpy(hsrc_.p, srcden, nsd * sizeof(double), hipMemcpyHostToDevice));
    dsrc = hsrc_.p;
  }
  chk_.upload(c);
  const int nseg = std::max(1, std::min(512, ns_ / 512));
  const long srs = (long)tree_.sextcnt * sdof;
  // keep the evaluate work buffers intact: the gathered densities get their own buffer
  DVec<double> sden;
  sden.alloc((size_t)nrhs * srs);
  part_.ensure((size_t)nrhs * nseg * nchk * tdof);
  dout_.ensure((size_t)nrhs * nchk * tdof);
  dev::gather(stream_, tree_.sextcnt, sdof, nrhs, siperm_.p, dsrc, (long)ns_ * sdof, sden.p);
  dev::Points src{sx_.p, sy_.p, sz_.p, snx_.p, sny_.p, snz_.p, sden.p, (long)nrhs * sdof, sdof};
  const dev::KC kcu{kcu_.a, kcu_.b, kcu_.c};
  dev::direct(stream_, user_.kt, tdof, nchk, chk_.p, chk_.p + nchk, chk_.p + 2 * nchk, ns_, nseg, src, part_.p,
              dout_.p, nrhs, kcu);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream_));
  HIPCHK(hipMemcpy(out, dout_.p, sizeof(double) * (size_t)nrhs * nchk * tdof, hipMemcpyDeviceToHost));
}

BatchRunner::BatchRunner(const std::vector<Engine*>& engines) : eng_(engines)
{
  if (const char* b = std::getenv("KIFMM3D_BATCH_BALANCE")) balance_ = std::atoi(b);
  const int nb = (int)std::min<size_t>(8, eng_.size());
  br_.resize(nb, nullptr);
  for (auto& s : br_) HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  fork_join_.resize(2 * nb, nullptr);
  for (auto& e : fork_join_) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

BatchRunner::~BatchRunner()
{
  drop();
  for (auto s : br_)
    if (s) (void)hipStreamDestroy(s);
  for (auto e : fork_join_)
    if (e) (void)hipEventDestroy(e);
}

void BatchRunner::drop()
{
  if (exec_) (void)hipGraphExecDestroy(exec_);
  exec_ = nullptr;
}

void BatchRunner::evaluate(const double* const* d_srcden, double* const* d_trgval, const int* nrhs, int mode)
{
  const size_t n = eng_.size();
  std::vector<int> rc(n);
  for (size_t i = 0; i < n; i++) {
    eng_[i]->check_eval_args(d_srcden[i], d_trgval[i], nrhs[i], mode);
    rc[i] = eng_[i]->prepare(nrhs[i]);  // allocations happen here, outside the capture
  }
  // the graph holds every engine's work buffers and settings as they were at
  // capture: it is replayed only while none of them changed
  bool same = exec_ && kmode_ == mode;
  for (size_t i = 0; same && i < n; i++)
    same = ksrc_[i] == d_srcden[i] && ktrg_[i] == d_trgval[i] && knrhs_[i] == nrhs[i] && krc_[i] == rc[i] &&
           kgen_[i] == eng_[i]->generation();
  hipStream_t s0 = br_[0];
  if (!same) {
    drop();
    HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
      const int nb = (int)br_.size();
      HIPCHK(hipEventRecord(fork_join_[0], s0));
      for (int k = 1; k < nb; k++) HIPCHK(hipStreamWaitEvent(br_[k], fork_join_[0], 0));
      // branches balanced by cost, largest problem first onto the least loaded
      // branch; the cost is each problem's last timed evaluate per right-hand
      // side, or, unless every problem has one, its target count
      std::vector<double> cost(n);
      bool timed = true;
      for (size_t i = 0; i < n; i++) timed = timed && eng_[i]->stage_ms[KIFMM3D_STAGE_TOTAL] > 0 && eng_[i]->last_nrhs() > 0;
      for (size_t i = 0; i < n; i++)
        cost[i] = timed ? eng_[i]->stage_ms[KIFMM3D_STAGE_TOTAL] / eng_[i]->last_nrhs() * nrhs[i]
                        : (double)eng_[i]->tree().textcnt * nrhs[i];
      std::vector<size_t> order(n);
      for (size_t i = 0; i < n; i++) order[i] = i;
      const bool rr = balance_ == 0;
      if (!rr) std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
      std::vector<double> load(nb, 0.0);
      for (size_t k = 0; k < n; k++) {
        const size_t i = order[k];
        int br = (int)(k % nb);
        if (!rr) br = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[br] += cost[i];
        eng_[i]->record(d_srcden[i], d_trgval[i], nrhs[i], rc[i], mode, br_[br]);
      }
      for (int k = 1; k < nb; k++) {
        HIPCHK(hipEventRecord(fork_join_[nb + k], br_[k]));
        HIPCHK(hipStreamWaitEvent(s0, fork_join_[nb + k], 0));
      }
    } catch (...) {
      (void)hipStreamEndCapture(s0, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(s0, &g));
    hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      exec_ = nullptr;
      throw Error(KIFMM3D_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    ksrc_.assign(d_srcden, d_srcden + n);
    ktrg_.assign(d_trgval, d_trgval + n);
    knrhs_.assign(nrhs, nrhs + n);
    krc_ = rc;
    kgen_.resize(n);
    for (size_t i = 0; i < n; i++) kgen_[i] = eng_[i]->generation();
    kmode_ = mode;
  }
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  float ms = 0;
  hipError_t e = hipEventRecord(a, s0);
  if (e == hipSuccess) e = hipGraphLaunch(exec_, s0);
  if (e == hipSuccess) e = hipEventRecord(b, s0);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) throw Error(KIFMM3D_EDEVICE, std::string("batch graph launch: ") + hipGetErrorString(e));
  last_ms_ = ms;
}

}  // namespace kifmm
