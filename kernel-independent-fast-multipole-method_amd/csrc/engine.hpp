// engine.hpp -- one evaluator instance: tree, operators, device plan and the
// evaluate pipeline (the counterpart of the reference's FMM3d object).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "operators.hpp"
#include "tree.hpp"

namespace kifmm {

// Owning device array (own = false: a view of memory held elsewhere, which
// reset() leaves alone).
template <class T>
struct DVec {
  T* p = nullptr;
  size_t n = 0;
  bool own = true;
  DVec() = default;
  DVec(const DVec&) = delete;
  DVec& operator=(const DVec&) = delete;
  ~DVec() { reset(); }
  void reset();
  void alloc(size_t count);                // uninitialised
  void upload(const std::vector<T>& v);    // alloc + copy
  bool ensure(size_t count);               // grow only; true when it reallocated
};

// A GEMM step of the up or down pass (see dev::GemmArgs).
struct GemmOp {
  int M = 0, Kseg = 0, nseg = 1, ngroup = 1, maxcols = 0, ncol = 0;
  int a_which = 0;  // 0 uc2ue, 1 ue2uc, 2 dc2de, 3 de2dc, 4 uc2ue x ue2uc (fused M2M + UC2UE),
                    // 5 [dc2de | dc2de x de2dc_c] (fused L2L + DC2DE)
  long a_gstride = 0;
  int b_buf = 0, c_buf = 0;  // 0 upcval, 1 upeden, 2 dncval, 3 dneden
  int b2_buf = -1;            // buffer of segments >= 1 (-1: b_buf)
  int bscale_seg = 0;         // bscale applies to segments >= bscale_seg
  size_t bnode = 0, cnode = 0, gbeg = 0;
  long bscale = -1, cscale = -1;
  int accumulate = 0;
  int compensated = 0;  // pseudoinverse applies (UC2UE, DC2DE) accumulate compensated
};

struct ItemList {  // point -> surface items
  int n = 0;
  size_t out = 0, geom = 0, eptr = 0, ent = 0;
};

struct ChunkList {  // chunks of <= 128 leaf targets
  int n = 0;
  size_t tbeg = 0, tnum = 0, eptr = 0;
  size_t ent = 0;                          // near field: int2 {sextbeg, sextnum}
  size_t chk = 0;                          // near field: per entry, how the FP32 kernel applies the
                                           // r < 1e-8 exclusion (0 none needed, 1 FP32, 2 FP64 decision)
  size_t enode = 0, ekind = 0, egeom = 0;  // surfaces: node, kind, (cx, cy, cz, rad)
};

struct VLevel {
  int depth = 0, nsrc = 0, ntrg = 0, npair = 0;
  size_t src_node = 0, trg_node = 0;
  size_t vptr = 0, vsrc = 0, vop = 0;  // per-pair form (V-list CSR with operator index)
  int npar = 0;                        // grouped form: target parents of the level
  size_t tslot = 0, sslot = 0;         // [npar][8] target-child slots, [npar][26][8] source-child slots
  size_t pmask = 0;                    // [npar] bit n: a source child in neighbour direction n
  int ntile = 0;                       // tiled form: 2x2x2 blocks of target parents
  size_t ttile = 0, htile = 0;         // [ntile][8][8] target-child slots, [ntile][64][8] halo source-child slots
  long dscale = -1;
};

struct Plan {
  ItemList s2m, s2l;
  ChunkList near, far;
  // the near-field chunks that hold entries flagged 2 (chk), with only those
  // entries (offsets into the engine's fix_ pool, set at the device set-up):
  // the FP32 near fields skip them and this FP64 pass adds them
  ChunkList near_fix;
  std::vector<GemmOp> up, down;
  // Mixed form's up pass: an inner node's equivalent density comes straight
  // from its children's through UC2UE x UE2UC_c (one product instead of M2M
  // then UC2UE; its check values are not formed); leaves apply UC2UE.
  std::vector<GemmOp> up_mixed;
  size_t fuse_idx = 0;         // identity index 0 .. 8 plnUE - 1 (builds the fused operator)
  size_t fuse_idx_g = 0;       // its group bounds {0, 8 plnUE}
  // Mixed form's down pass for plnDC == plnDE: a node's downward equivalent
  // density from its own check values and its parent's density in one
  // product, [DC2DE | DC2DE x DE2DC_c] (L2L into the check values skipped).
  std::vector<GemmOp> down_mixed;
  size_t fdn_b = 0, fdn_c = 0, fdn_g = 0;  // index arrays that build the fused operator
  std::vector<VLevel> vlev;
  size_t dc_zero = 0;          // target nodes (tndeidx) in no V list: their DC rows start at zero
  int n_dc_zero = 0;
  int max_vsrc = 0, max_vtrg = 0;
  bool m2l_pairs = false;      // per-pair M2L kernel instead of the parent-grouped one
  size_t optab = 0;            // [26 directions][8 source children][8 target children] -> operator or -1
  size_t aidx[2] = {0, 0};     // A-operand tables of the matrix-core M2L, FP64 and FP32 row layouts
  int nlow_s = 0, nlow_t = 0;  // source / target node indices of depth < 2 form these prefixes
  long near_pairs = 0;         // source-target pairs of the near field (one RHS)
  long far_pairs = 0;          // surface point - target pairs of L2T + M2T
  long v_pairs = 0;
  long s2m_pairs = 0;          // source point - UC point pairs of S2M
  long s2l_pairs = 0;          // source point - DC point pairs of S2L
  double up_flops = 0;         // M2M + UC2UE multiply-adds x 2, per right-hand side
  double up_flops_mixed = 0;   // the same for up_mixed
  double down_flops = 0;       // L2L + DC2DE multiply-adds x 2, per right-hand side
  long fft_fwd = 0, fft_inv = 0;  // transforms per right-hand side (times the dof count)
  std::vector<int> I;          // index pool (uploaded once)
  std::vector<double> D;       // double pool (geometry, scales)
};

class Engine {
 public:
  Engine(int kt, const double* coefs, int ncoefs, int np, int ptsmax, int maxlevel, const double* ctr, int rootlvl);
  ~Engine();

  void setup(const double* spos, const double* snor, int ns, const double* tpos, int nt);
  void evaluate_device(const double* d_srcden, double* d_trgval, int nrhs, int mode);
  void evaluate_host(const double* srcden, double* trgval, int nrhs, int mode);
  // evaluate_device with every work buffer carved out of the library's static
  // per-GPU scratch array (dev::SCRATCH_BYTES): no device memory is allocated
  // or freed.  The chunk of right-hand sides is evaluate_device's where it
  // fits the array, else the largest that does; a problem whose work buffers
  // for one right-hand side exceed the array is rejected (KIFMM3D_ENOMEM).
  // Kernel launches are eager (no graph) and untimed; returns when done.
  void evaluate_device_noalloc(const double* d_srcden, double* d_trgval, int nrhs, int mode);
  // Autotuning: times the candidate near-field kernels and, per V-list level,
  // the candidate M2L forms on this tree for (nrhs, mode) -- one warm-up and
  // two timed evaluates of zero densities per round, rounds pairing the
  // near-field and M2L candidates -- and keeps the fastest of each for later
  // evaluates of that mode.  Returns the number of evaluates run.
  int tune(int nrhs, int mode);
  // Settings in force for a mode: [0] near-field kernel (0 LDS broadcast,
  // 1 scalar-cache sources, 2 FP64 matrix cores, 3 FP32 matrix cores, 4 FP32
  // LDS broadcast),
  // [1 + l] M2L form of V-list level l (0 tiled matrix-core, 1 parent-grouped
  // matrix-core, 2 grouped VALU, 3 per pair, 4 tiled with one wave per tile
  // and frequency).  Returns the count (1 + levels).
  int settings(int mode, int* out, int n) const;
  // Pins settings for a mode (same layout; -1 entries keep the default), e.g.
  // ones a previous process tuned.  Rejects kernels and forms the problem
  // cannot take.
  void set_settings(int mode, const int* in, int n);
  // FFTs + M2L of each V-list level in the last timed evaluate (ms)
  const std::vector<double>& level_ms() const { return vlev_ms_; }
  // Graph replay: with it on, evaluate_device captures the whole pass into a
  // hipGraph on first use for a given (buffers, nrhs, mode) and then only
  // launches it (stage_times then reports the total only).
  void set_graph(bool on)
  {
    use_graph_ = on;
    if (!on) drop_graph();
  }
  bool graph() const { return use_graph_; }
  // Pieces of evaluate_device for callers that build their own graph (the
  // batch interface): argument checks, work buffers for nrhs (returns the RHS
  // chunk), and the whole pass enqueued on stream s without timing or sync.
  void check_eval_args(const double* d_srcden, const double* d_trgval, int nrhs, int mode) const;
  int prepare(int nrhs);
  void record(const double* d_srcden, double* d_trgval, int nrhs, int rc, int mode, hipStream_t s);

  const Tree& tree() const { return tree_; }
  const Operators& ops() const { return *ops_; }
  const KernelInfo& user() const { return user_; }
  bool is_setup() const { return setup_done_; }
  // Bumped whenever setup or a work-buffer growth replaces device memory that
  // a captured graph may hold: a graph recorded at another generation must
  // not be replayed.
  unsigned long generation() const { return gen_; }
  int last_nrhs() const { return last_nrhs_; }
  bool host_ready() const { return host_done_; }
  double stage_ms[KIFMM3D_NSTAGE] = {0};
  double kernel_ms[KIFMM3D_NKCLASS] = {0};
  int kernel_calls[KIFMM3D_NKCLASS] = {0};
  long get_buffer(int what, double* buf, long n);
  // FP64 direct sum at the caller-order targets idx[0..nchk) for nrhs density
  // vectors; out is host memory [nrhs][nchk * tdof].
  void direct(const double* srcden, bool on_device, int nrhs, const int* idx, int nchk, double* out);
  const Plan& plan() const { return plan_; }
  long near_flagged() const { return nflag_; }

 private:
  void near_flags();
  void build_plan();
  void upload_static();
  int choose_rhs_chunk(int nrhs);
  void ensure_work(int rc);
  // Element counts of the work buffers for chunks of rc right-hand sides, in
  // the order sden, tval, tvalT, upcval, upeden, dncval, dneden, effden,
  // effval, gwork (doubles), sden32 (floats; 0 unless Laplace).
  static constexpr int NWORK = 11;
  void work_counts(int rc, size_t cnt[NWORK]) const;
  void run_chunk(const double* srcden, double* trgval, int rc, int mode, hipStream_t s, hipEvent_t* ev);
  void drop_graph();
  int near_variant(bool mixed) const;
  // The FP32 vector near field (kernel 4) is the mixed form's for the 1/r
  // Green's functions without normals, Stokes (311) and Navier (511) single
  // layer.  The Stokes pressure (312, r.f / r^3) and the kernels with normals
  // keep FP64: their sums cancel more (a 4,000-point cluster set measured
  // 1.6e-6 from the plain form with the pressure in FP32).
  bool near32_ok() const { return user_.kt == 311 || user_.kt == 511; }
  int m2l_form(bool mixed, size_t level, int rc) const;

  KernelInfo user_, mm_;
  KernelConsts kcu_, kcm_;
  int np_, rootlvl_;
  double ctr_[3];
  std::shared_ptr<const Operators> ops_;
  Tree tree_;
  Plan plan_;
  int ns_ = 0, nt_ = 0;
  bool setup_done_ = false;
  bool host_done_ = false;
  int last_rc_ = 0;
  bool last_in_scratch_ = false;  // the last evaluate ran in the shared scratch array
  unsigned long gen_ = 0;
  int last_nrhs_ = 0;  // right-hand sides of the last evaluate_device
  int m2l_rb_ = 0;   // right-hand sides per lane in the grouped M2L (0 = default; KIFMM3D_M2L_RB)
  int m2l_bs_ = 64;  // frequencies per block in the grouped M2L (64 or 256; KIFMM3D_M2L_BS)
  int p2p_variant_ = 2;  // near field: 0 LDS broadcast, 1 scalar-cache sources, 2 matrix cores (KIFMM3D_P2P)
  int split_blocks_ = 1024;  // blocks the split-K GEMMs aim at (KIFMM3D_GEMM_SPLIT)
  int gemm_mfma_ = 2;  // uncompensated GEMMs: 2 wide matrix-core tiles, 1 64x64 tiles (KIFMM3D_GEMM=tile64), 0 VALU (=valu)
  bool m2l_tile_ = true;   // tiled matrix-core M2L for scalar kernels and 16-RHS chunks (KIFMM3D_M2L=matrix|group|pair disables)
  bool m2l_matrix_ = true; // parent-grouped matrix-core M2L otherwise (KIFMM3D_M2L=group|pair disables)
  bool m2l_wave_ = false;  // the tiled M2L with one wave per (tile, frequency) (KIFMM3D_M2L=wave)
  int m2l_depth_ = 4;       // ... halo parents its spectra loads run ahead (KIFMM3D_M2L_DEPTH)
  int m2l_fpart_ = 0;       // tiled M2L: frequency blocks dealt to the XCDs (KIFMM3D_M2L_FPART=1)
  hipStream_t stream_ = nullptr;
  hipStream_t side_ = nullptr;  // near field, concurrent with the far-field chain
  hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr;
  bool overlap_ = false;        // KIFMM3D_OVERLAP=1 (measured: -0.5 ms of 51 at 4M/16 RHS; off so that kernel times stay uncontended)
  bool use_graph_ = false;  // KIFMM3D_GRAPH
  bool s2m32_ = true;       // mixed form: S2M / S2L on the FP32 matrix cores (KIFMM3D_S2M32=0: FP64)
  hipGraphExec_t gexec_ = nullptr;
  const double* gsrc_ = nullptr;
  double* gtrg_ = nullptr;
  int gnrhs_ = 0, gmode_ = -1, grc_ = 0;
  unsigned long ggen_ = 0;
  std::vector<hipEvent_t> evpool_;
  std::vector<int> evcls_;
  std::vector<int> evlvl_;        // V-list level of each timed launch (-1 outside the V list)
  std::vector<double> vlev_ms_;   // per V-list level: FFTs + M2L of the last timed evaluate
  int p2p_tuned_[2] = {-1, -1};   // per mode: tuned near-field kernel (-1: default)
  std::vector<int> m2l_form_[2];  // per mode, per V-list level: tuned M2L form (empty: default)
  size_t nev_used_ = 0;

  // device data
  DVec<double> sx_, sy_, sz_, snx_, sny_, snz_, tx_, ty_, tz_;
  DVec<int> tperm_, s2c_;
  DVec<int> siperm_;  // inverse of the source leaf permutation: caller order -> leaf order
  DVec<int> tiperm_;  // inverse of the target leaf permutation
  DVec<double> tvalT_;  // target values as slot-major rows, on their way out
  DVec<float> sden32_;  // FP32 copy of the gathered densities (mixed form, Laplace)
  DVec<short> c2s_;
  DVec<double> unit_[4];
  DVec<double> uc2ue_, ue2uc_, dc2de_, de2dc_, m2l_;
  DVec<double> m2me_;  // 8 blocks of plnUE x plnUE: UC2UE x UE2UC_c (Dot2 on the device at setup)
  DVec<double> l2de_;  // 8 blocks of plnDE x 2 plnDC: [DC2DE | DC2DE x DE2DC_c] (plnDC == plnDE)
  DVec<double> opsT_;  // M2L operators as [C][343][sdof][tdof] complex (FP64)
  DVec<double> atab_;  // tiled M2L A-operand tables [C][26 x 32][2] complex (scalar kernels)
  DVec<float> atab32_;  // the same in FP32
  DVec<float> opsT32_; // the same in FP32
  DVec<float> avec32_; // FP32 A operands of the parent-grouped M2L, [C][26][tdof 4 sdof 64]
  DVec<double> tw_;
  DVec<int> I_;
  DVec<double> D_;
  // work buffers (grow with the RHS chunk)
  DVec<double> sden_, tval_, upcval_, upeden_, dncval_, dneden_, effden_, effval_;
  DVec<double> hsrc_, htrg_;  // staging for the host-pointer entry point
  DVec<double> chk_, part_, dout_;  // direct-sum scratch
  DVec<double> gwork_;              // split-K partials of the GEMMs
  std::vector<double> tpos_host_;  // caller-order target positions (for direct sums)
  DVec<int> fix_;   // index pool of plan_.near_fix (entries flagged 2 at set-up)
  long nflag_ = 0;  // near-field entries flagged 2
};

// Several set-up engines evaluated in one captured graph: problem i runs on
// branch i % nbranch (forked from and joined back to one capture stream), so
// independent problems overlap on the GPU.  The graph is kept while the
// buffers, right-hand-side counts and mode stay the same.
class BatchRunner {
 public:
  explicit BatchRunner(const std::vector<Engine*>& engines);
  ~BatchRunner();
  void evaluate(const double* const* d_srcden, double* const* d_trgval, const int* nrhs, int mode);
  double last_ms() const { return last_ms_; }

 private:
  void drop();
  std::vector<Engine*> eng_;
  std::vector<hipStream_t> br_;
  std::vector<hipEvent_t> fork_join_;
  hipGraphExec_t exec_ = nullptr;
  std::vector<const double*> ksrc_;
  std::vector<double*> ktrg_;
  std::vector<int> knrhs_, krc_;
  std::vector<unsigned long> kgen_;
  int kmode_ = -1;
  double last_ms_ = 0.0;
  int balance_ = 1;  // 1: branches balanced by problem cost, 0: round robin (KIFMM3D_BATCH_BALANCE)
};

}  // namespace kifmm
