// device.hpp -- host-callable launchers of the gfx950 kernels in device.hip.
// All pointers are device pointers; every launch goes to the given stream.
#pragma once

#include <hip/hip_runtime.h>

namespace kifmm {
namespace dev {

struct KC {  // kernel constants (see kernel_consts in kernels_host.cpp)
  double a, b, c;
};

// Work memory of the no-allocation evaluate: one static device array of
// SCRATCH_BYTES per GPU, which that GPU holds until the process ends.  Every
// call on a GPU shares it, so calls that overlap on one GPU race.
constexpr size_t SCRATCH_BYTES = size_t(256) << 20;
// Device address of the array on the current device, resolved with
// hipGetSymbolAddress on the first call made with that device current.
hipError_t scratch_ptr(void** out);

// p[0 .. n) = 0
void zero(hipStream_t s, double* p, long n);
// p[rows[i] * rowlen .. + rowlen) = 0 for i < nrows
void zero_rows(hipStream_t s, double* p, const int* rows, int nrows, long rowlen);
// dst[0 .. n) = src[0 .. n)
void copy(hipStream_t s, const double* src, double* dst, long n);
// Source-major gather through the inverse permutation (iperm[perm[k]] = k):
// dst[(iperm[i]*nr + r)*dof + d] = src[r*src_rs + i*dof + d] for i < n
void gather(hipStream_t s, int n, int dof, int nr, const int* iperm, const double* src, long src_rs, double* dst,
            float* dst32 = nullptr);  // dst32: also (dst null: only) an FP32 copy
// The same through slot-major rows (tmp holds n * nr * dof doubles; iperm is
// the inverse of perm; dof <= 3): a transpose, then one caller point per thread.
void scatter_rows(hipStream_t s, int n, int dof, int nr, const int* iperm, const double* src, long src_rs,
                  double* tmp, double* dst, long dst_rs);
// dst[r*dst_rs + perm[k]*dof + d] = src[r*src_rs + k*dof + d]
void scatter(hipStream_t s, int n, int dof, int nr, const int* perm, const double* src, long src_rs, double* dst,
             long dst_rs);

// Leaf-ordered source points (SoA) and densities: density d of source j for
// right-hand side r is den[j*den_js + r*den_rs + d].  The evaluate pass keeps
// them source-major (den_js = Rc*sdof, den_rs = sdof) so that the densities of
// one source for a block of right-hand sides are contiguous.
// The FP32 matrix-core Laplace kernels run a last block of fewer than 16
// right-hand sides zero-padded to 16 when it has at least this many; smaller
// remainders take the FP64 vector kernels.
constexpr int PAD16_MIN = 4;

struct Points {
  const double *x, *y, *z, *nx, *ny, *nz;
  const double* den;
  long den_js, den_rs;
  const float* den32 = nullptr;  // FP32 copy of den (same layout), read by the FP32 matrix-core kernels
};
struct Targets {  // leaf-ordered targets (SoA) and values [rhs][n*tdof]
  const double *x, *y, *z;
  double* val;
  long val_rs;
};

// Near field: for every chunk of <= 128 targets, sum the user kernel over the
// source leaves ent[eptr[c] .. eptr[c+1]) (each ent = {sextbeg, sextnum}).
// variant 0: sources staged in LDS, 128-thread blocks; variant 1: sources read
// through the scalar cache (needs den_rs == sdof), one wave per 64 targets;
// variant 2: Laplace blocks of 16 right-hand sides on the FP64 matrix cores;
// variant 3: the same on the FP32 matrix cores (mixed precision).  Variants
// 2 and 3 fall back to variant 0 for other kernels and smaller blocks.
// variant 4: the kernels without normals in FP32 (mixed precision).
// echk[e] (variants 2, 3 and 4; without it variant 3 takes the FP64
// matrix-core kernel): 0 no pair of entry e can be closer than 1e-8 or than
// its FP32 offsets resolve (the FP64 matrix-core kernel skips the exclusion
// test there), 1 the target's own leaf (exclusion tested), 2 a pair may be
// closer than the FP32 offsets resolve: the FP32 kernels skip such entries.  Returns n: right-hand sides [0, n) skipped the entries flagged 2,
// which a second call with variant 0 over those entries alone (and nr = n)
// then adds in FP64.
int p2p(hipStream_t s, int kt, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int2* ent,
        const int* echk, Points src, Targets trg, int nr, KC kc, int variant);

// Set-up pass for the mixed near field, per entry e of the near-field chunks:
// out[e].x = the smallest squared distance between a chunk target and an
// entry source that are distinct points (DBL_MAX if none), out[e].y = the
// largest squared distance of an entry source from the chunk's first target;
// tout[c] = the largest squared distance of chunk c's targets from it.
void near_extent(hipStream_t s, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int2* ent,
                 Points src, Targets trg, double2* out, double* tout);

// FP64 direct sum of the user kernel over all nsrc sources at nchk points
// (cx, cy, cz): out[r][k*tdof + t]; part holds nr * nseg * nchk * tdof doubles.
void direct(hipStream_t s, int kt, int tdof, int nchk, const double* cx, const double* cy, const double* cz,
            int nsrc, int nseg, Points src, double* part, double* out, int nr, KC kc);

// Points to check surface (S2M onto UC, S2L onto DC): out[(node*Rc + r)*pln + p*tdof + t] +=
// sum over the source leaves of item i of K(ctr_i + rad_i * unit_p <- y) den(y);
// with accum = 0 the sum overwrites out (items own distinct nodes).
void p2s(hipStream_t s, int kt, int nitem, const int* out_node, const double4* geom, const int* eptr,
         const int2* ent, const double* unit, int npt, Points src, double* out, int Rc, int pln, int nr, KC kc,
         int variant, int accum = 1);

// Equivalent surfaces to targets with the mm kernel (L2T from DE, M2T from UE).
// Entry e: node enode[e] of buffer ekind[e] (0 = dneden, 1 = upeden), surface
// ekind 0 -> unit_de, 1 -> unit_ue, placed at geom[e] = (cx, cy, cz, rad).
// variant (Laplace, 16-RHS blocks): 2 FP64 matrix cores, 3 FP32 matrix cores
// (mixed form), otherwise the FP64 vector kernels.
void s2p(hipStream_t s, int kt, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int* enode,
         const int* ekind, const double4* egeom, const double* unit_de, const double* unit_ue, int nsp,
         const double* dneden, const double* upeden, int Rc, int pln, Targets trg, int nr, KC kc, int variant);

// Batched, indexed FP64 GEMM over (node, rhs) columns:
// C[:, cnode[n]*Rc + r] (+)= cscale .* sum_seg A_{grp,seg} (bscale .* B[:, bnode[seg*ncol + n]*Rc + r])
// for the nodes n in [gbeg[grp], gbeg[grp+1]); bnode < 0 contributes nothing.
struct GemmArgs {
  int M, Kseg, nseg, Rc, ngroup, maxcols;
  const double* A;
  long a_gstride;
  const double* B;
  const int* bnode;
  int ncol;
  double* C;
  const int* cnode;
  const int* gbeg;
  const double* bscale;
  const double* cscale;
  int accumulate;
  int compensated;  // 1: TwoProd/TwoSum accumulation (pseudoinverse applies), 0: plain FMA
  int mfma;         // plain form on the FP64 matrix cores (ignored when compensated):
                    // 2 wide tiles (up to 160 rows x 128 columns), 1 64 x 64 tiles, 0 VALU
  long ctot;        // all columns of the product (ncol x Rc)
  double* work;     // split-K partials, gemm_split(a) x ctot x M doubles (null or too small: no split)
  long work_n;      // capacity of work in doubles
  int split_blocks; // blocks the split-K aims at (0: 1024); lower when other streams share the GPU
  const double* B2; // segments >= 1 read this buffer instead of B (wide-tile plain form only)
  int bscale_seg;   // bscale applies to segments >= bscale_seg
};
void gemm(hipStream_t s, const GemmArgs& a);
// K splits the wide-tile product takes for these sizes (1: none).
int gemm_split(const GemmArgs& a);

// r2c of the scaled upward equivalent densities on the (2np)^3 grid into
// complex spectra, effden[slot][Rc][sm][C] (rhs_inner 0) or
// effden[slot][sm][C][Rc] (rhs_inner 1), stored as double2 or, with single,
// as float2.
void fft_fwd(hipStream_t s, int np, int nslot, const int* slot_node, int Rc, const double* upeden, int pln, int sm,
             const double* dscale, const short* cube2surf, void* effden, int single, int rhs_inner,
             const double2* tw);
// M2L products summed over V pairs -> effval[trg][Rc][tm][C] (overwritten).
void m2l(hipStream_t s, int sm, int tm, int C, int ntrg, int Rc, const int* vptr, const int* vsrc, const int* vop,
         const double2* ops, const double2* effden, double2* effval);
// M2L grouped by target parent: for parent p, effval[tslot[p][c]] =
// sum over the 26 neighbour directions n and source children s of
// T_{optab[n][s][c]} effden[sslot[p][n][s]] (entries < 0 are skipped).
// rb: right-hand sides per lane (0 = default), bs: frequencies per block (64 or 256).
void m2l_grouped(hipStream_t s, int sm, int tm, int C, int npar, int Rc, int rb, int bs, const int* tslot,
                 const int* sslot, const int* optab, const double2* ops, const double2* effden, double2* effval);
// The same sums on the matrix cores: columns (parent, right-hand side) in
// tiles of 16, operators from the frequency-major copy opsT [C][343][sm][tm];
// pmask[p] bit n set if parent p has a source child in direction n; aidx
// [26][tm][4 sm][64] the A-operand table of the precision's row layout
// (2 x index of the real value among one frequency's operators + 1 if
// negated, -1 for zero).  single selects FP32 spectra and operators (float2,
// all directions accumulated in FP32), otherwise FP64 (double2).  atabv: the
// FP32 A operands expanded by expand_m2l_vec, used by the FP32 form when not
// null (the FP64 form resolves aidx against opsT).
void m2l_matrix(hipStream_t s, int sm, int tm, int C, int npar, int Rc, const int* tslot, const int* sslot,
                const int* pmask, const int* aidx, const void* opsT, const void* atabv, const void* effden,
                void* effval, int single);
// atabv [C][26][tm 4 sm 64] = the FP32 A-operand table aidx32 resolved
// against the FP32 operators opsT32 of every frequency
void expand_m2l_vec(hipStream_t s, int C, int sm, int tm, const int* aidx32, const float* opsT32, float* atabv);
// c2r and gather onto the DC surface: dncval[(node*Rc + r)*pln + p*tm + f] = ...
// (assigned: every DC value of the listed nodes; effval in the layout and
// precision described for fft_fwd)
void fft_inv(hipStream_t s, int np, int nslot, const int* slot_node, int Rc, const void* effval, int single,
             int rhs_inner, int tm, const int* surf2cube, int nsp, double* dncval, int pln, const double2* tw);
// Tiled M2L on the matrix cores for scalar kernels, spectra with the
// right-hand sides innermost and Rc a multiple of 16: for every tile (2x2x2
// target parents, ttile [ntile][8][8] child slots) and its 4x4x4 halo of
// neighbour parents (htile [ntile][64][8] child slots), effval of each target
// child = sum over its V list of T_o effden.  atab is the A-operand table
// [C][26 directions x 32][2] complex: entry n*32 + 8k + c holds the operators
// of (source child k, target child c) and (4 + k, c) in direction n (zero for
// pairs outside the V list); single selects FP32 spectra and operators.
// fpart: deal frequency blocks rather than tile ranges to the XCDs.
void m2l_tiled(hipStream_t s, int C, int ntile, int Rc, const int* ttile, const int* htile, const void* atab,
               const void* effden, void* effval, int single, int fpart = 0);
// The products of m2l_tiled with one wave per (tile, frequency): the A
// operands of all directions in LDS, halo spectra loaded straight into the
// matrix-core operands (atab: the table of the precision), depth halo parents
// ahead (1 or 4); nslot: the spectra slots of the level, followed by one
// slot of zeros (the caller's, effden[nslot]), which absent halo children
// read.  single: FP32 spectra and operators, otherwise FP64.
void m2l_wave(hipStream_t s, int C, int ntile, int Rc, int nslot, const int* ttile, const int* htile,
              const void* atab, const void* effden, void* effval, int depth, int single);

}  // namespace dev
}  // namespace kifmm
