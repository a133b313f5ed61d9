/*
 * kifmm3d.h -- C ABI of the MI355X KIFMM3d evaluator (libkifmm3d.so).
 *
 * The library is a drop-in for the reference's FMM3d setup/eval pair
 * (fmm3d/fmm3d.hpp:27-115, fmm3d/knlmat3d.hpp:26-49).  Host C++ builds the
 * octree, the U/V/W/X lists and the FP64 translation operators; the evaluate
 * pass runs as hand-written gfx950 HIP kernels on one GPU.
 *
 * Conventions
 *   - Positions and normals are 3 x N column-major (x, y, z interleaved per
 *     point), as the reference's DblNumMat 3 x N (fmm3d/knlmat3d.hpp:30-32).
 *   - Densities are [nrhs][N * sdof], values [nrhs][N * tdof], dof fastest
 *     (fmm3d/fmm3d_eval.cpp:59, :230); nrhs = 1 is exactly the reference's eval.
 *   - Every call returns 0 on success and a negative KIFMM3D_E* code on
 *     failure; kifmm3d_last_error() then describes the failure.  The library
 *     never aborts the process (the reference asserts instead,
 *     common/commoninc.hpp:59-60).
 *   - One handle per host thread; handles are independent.  Operators are
 *     shared read-only between handles through a process-wide cache.
 */
#ifndef KIFMM3D_H
#define KIFMM3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kifmm3d kifmm3d;

/* Status codes */
#define KIFMM3D_OK 0
#define KIFMM3D_EINVAL (-1)   /* bad argument (kernel type, np, sizes, options) */
#define KIFMM3D_EBOX (-2)     /* a point lies outside the root box (let3d.cpp:73, :155) */
#define KIFMM3D_ESTATE (-3)   /* call out of order (evaluate before setup, ...) */
#define KIFMM3D_EDEVICE (-4)  /* HIP runtime error or no HIP device */
#define KIFMM3D_ENOMEM (-5)   /* device or host allocation failed */

/* Evaluation forms */
#define KIFMM3D_MODE_PLAIN 0  /* FP64 throughout: the drop-in form */
#define KIFMM3D_MODE_MIXED 1  /* mixed precision: Laplace near field, S2M / S2L and L2T / M2T as FP32 matrix-core
                                 products summed in FP64 per 128-source tile; the near field of the other
                                 kernels without normals in FP32 summed in FP64 per 64-source tile; FP32 M2L
                                 spectra, FFT arithmetic,
                                 operators and products; FP64 pseudoinverse applies without compensation; FP64
                                 elsewhere.  No reference counterpart; error against the FP64 direct sum is
                                 reported by the benchmark and bounded in the tests. */

/* Per-stage timer slots filled by kifmm3d_evaluate*() (milliseconds, HIP events) */
enum {
  KIFMM3D_STAGE_LOAD = 0,   /* density gather into leaf order           (fmm3d_eval.cpp:50-64)  */
  KIFMM3D_STAGE_UP = 1,     /* S2M, M2M, UC2UE                          (fmm3d_eval.cpp:69-92)  */
  KIFMM3D_STAGE_NEAR = 2,   /* U list + direct W/X fallbacks            (fmm3d_eval.cpp:98-111) */
  KIFMM3D_STAGE_VLIST = 3,  /* FFT, M2L products, inverse FFT           (fmm3d_eval.cpp:115-150)*/
  KIFMM3D_STAGE_XLIST = 4,  /* S2L                                      (fmm3d_eval.cpp:176-191)*/
  KIFMM3D_STAGE_DOWN = 5,   /* L2L, DC2DE                               (fmm3d_eval.cpp:195-209)*/
  KIFMM3D_STAGE_EVAL = 6,   /* L2T + M2T (W list)                       (fmm3d_eval.cpp:154-172, 210-214) */
  KIFMM3D_STAGE_SAVE = 7,   /* scatter to caller order                  (fmm3d_eval.cpp:221-235)*/
  KIFMM3D_STAGE_TOTAL = 8,
  KIFMM3D_NSTAGE = 9
};

/* Kernel classes timed by kifmm3d_kernel_times() (HIP events around each launch group) */
enum {
  KIFMM3D_K_GATHER = 0,     /* density permutation into leaf order       */
  KIFMM3D_K_S2M = 1,        /* leaf sources -> UC check surface          */
  KIFMM3D_K_GEMM_UP = 2,    /* M2M and UC2UE batched GEMMs               */
  KIFMM3D_K_P2P = 3,        /* near field (U list, direct W/X)           */
  KIFMM3D_K_FFT_FWD = 4,    /* UE densities -> spectra                   */
  KIFMM3D_K_M2L = 5,        /* frequency-domain M2L products             */
  KIFMM3D_K_FFT_INV = 6,    /* spectra -> DC check values                */
  KIFMM3D_K_S2L = 7,        /* X-list sources -> DC check surface        */
  KIFMM3D_K_GEMM_DOWN = 8,  /* L2L and DC2DE batched GEMMs               */
  KIFMM3D_K_S2P = 9,        /* L2T and M2T                               */
  KIFMM3D_K_SCATTER = 10,   /* values back to caller order               */
  KIFMM3D_NKCLASS = 11
};

/* Create a handle.  Replaces FMM3d::FMM3d + the knl()/ctr()/rootlvl() setters
 * (fmm3d/fmm3d.cpp:23-27, fmm3d/tt.cpp:85-90) and reads the three options the
 * reference's setup requires (-fmm3d_np, -fmm3d_let3d_ptsmax,
 * -fmm3d_let3d_maxlevel; fmm3d_setup.cpp:30, let3d.cpp:45-48).
 * kt (fmm3d/kernel3d.hpp:29-53): 111 Laplace single layer, 121 Laplace double
 * layer, 311 / 312 Stokes single layer velocity / pressure, 321 / 322 Stokes
 * double layer velocity / pressure, 511 / 521 Navier single / double layer
 * displacement.  coefs: {mu} for 311, 312, 321, 322 and {mu, nu} for 511, 521
 * (kernel3d.cpp:187-189, :224-226, :318-319, :341-342); ignored for 111, 121.
 * Kernels 121, 321, 322 and 521 need source normals. */
int kifmm3d_create(kifmm3d** out, int kt, const double* coefs, int ncoefs, int np, int ptsmax,
                   int maxlevel, const double ctr[3], int rootlvl);

/* Build tree, lists and operators and upload them.  Replaces FMM3d::setup
 * (fmm3d/fmm3d_setup.cpp:26-63).  srcnor may be NULL when the kernel does not
 * use normals.  The library copies the inputs; the caller keeps ownership. */
int kifmm3d_setup(kifmm3d* h, const double* srcpos, const double* srcnor, int nsrc,
                  const double* trgpos, int ntrg);

/* Evaluate nrhs density vectors.  Replaces FMM3d::eval
 * (fmm3d/fmm3d_eval.cpp:24-243): trgval is overwritten.  Host pointers; the
 * call copies srcden to the device and trgval back, and returns when trgval
 * is complete. */
int kifmm3d_evaluate(kifmm3d* h, const double* srcden, double* trgval, int nrhs, int mode);

/* Same as kifmm3d_evaluate with device-resident srcden / trgval (HBM
 * pointers from kifmm3d_device_alloc).  Returns when the work is complete. */
int kifmm3d_evaluate_device(kifmm3d* h, const double* d_srcden, double* d_trgval, int nrhs, int mode);

/* kifmm3d_evaluate_device that allocates no device memory.  No counterpart in
 * the reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Every work
 * buffer of the pass (gathered densities, surface vectors, spectra, split-K
 * partials, value staging) is carved out of one static device array of
 * 256 MiB per GPU, which each GPU the library runs on holds until the process
 * ends; the handle's set-up data and the caller's srcden / trgval are used as
 * they are.  Right-hand sides run in chunks of kifmm3d_evaluate_device's size
 * where they fit the array, else in the largest chunks that do; a problem
 * whose work for one right-hand side exceeds 256 MiB fails with
 * KIFMM3D_ENOMEM.  Every call on one GPU shares the array (each call
 * overwrites what it needs before reading it), so calls that overlap on one
 * GPU -- from several threads or handles at once -- race.  Launches are eager
 * (no graph) and untimed (stage and kernel times read 0); kifmm3d_get_buffer
 * reports KIFMM3D_ESTATE after such a call.  Returns when the work is
 * complete. */
int kifmm3d_evaluate_device_noalloc(kifmm3d* h, const double* d_srcden, double* d_trgval, int nrhs, int mode);

/* Stage times (ms) of the last evaluate; n <= KIFMM3D_NSTAGE slots. */
int kifmm3d_stage_times(const kifmm3d* h, double* ms, int n);

/* Per-class kernel time (ms) and launch-group count of the last evaluate;
 * n <= KIFMM3D_NKCLASS slots; calls may be NULL. */
int kifmm3d_kernel_times(const kifmm3d* h, double* ms, int* calls, int n);

/* Tree and list sizes after setup: out[0..11] = nnode, nlevel, sndecnt,
 * sextcnt, tndecnt, textcnt, plnUE, plnUC, plnDE, plnDC, sdof, tdof. */
int kifmm3d_info(const kifmm3d* h, int* out, int n);

/* Copy of a setup product, for verification and for the reference-style
 * accessors (let3d.hpp:102-136).  what: 0 node table (13 ints per node:
 * par, chd, path x3, depth, tag, sndeidx, sextbeg, sextnum, tndeidx, textbeg,
 * textnum), 1..4 U/V/W/X list CSR pointers (nnode+1 ints), 5..8 U/V/W/X list
 * indices, 9 source leaf permutation (sextcnt ints), 10 target leaf
 * permutation (textcnt ints).  Returns the element count; buf may be NULL. */
long kifmm3d_get_tree(const kifmm3d* h, int what, int* buf, long n);

/* Copy of an FP64 operator: what 0 UC2UE (plnUE x plnUC), 1 DC2DE
 * (plnDE x plnDC), both column-major as the reference stores them
 * (matmgnt3d.cpp:103, :169).  Returns the element count; buf may be NULL. */
long kifmm3d_get_operator(const kifmm3d* h, int what, double* buf, long n);

/* Copy of a per-node buffer of the last evaluate (first right-hand side):
 * what 0 upward equivalent densities (sndecnt x plnUE), 1 upward check values
 * (sndecnt x plnUC), 2 downward equivalent densities (tndecnt x plnDE),
 * 3 downward check values (tndecnt x plnDC) -- the reference's _supeden,
 * _supcval, _tdneden, _tdncval (fmm3d/fmm3d.hpp:61-64).  Returns the element
 * count; buf may be NULL.  After a mixed-mode evaluate the upward check values
 * of inner nodes are not formed (their equivalent densities come from the
 * children's directly) and hold no meaningful data. */
long kifmm3d_get_buffer(kifmm3d* h, int what, double* buf, long n);

/* FP64 direct sum at selected targets: out[r][k*tdof + t] = sum over all
 * sources of K(trgpos[trgidx[k]] <- src) srcden[r] -- the exact sum of
 * FMM3d::check (fmm3d/fmm3d_check.cpp:25-74) and Dense3d::eval
 * (fmm3d/dense3d.cpp:38-58), computed on the GPU.  srcden is [nrhs][Nsrc*sdof]
 * host memory; trgidx and out are host memory. */
int kifmm3d_direct(kifmm3d* h, const double* srcden, int nrhs, const int* trgidx, int nchk, double* out);
/* Same with srcden resident in device memory. */
int kifmm3d_direct_device(kifmm3d* h, const double* d_srcden, int nrhs, const int* trgidx, int nchk, double* out);

/* Work counts of the plan after setup, per right-hand side: out[0] near-field
 * source-target pairs, [1] L2T + M2T surface-target pairs, [2] V-list pairs,
 * [3] S2M pairs, [4] S2L pairs, [5] up-pass GEMM flops, [6] down-pass GEMM
 * flops, [7] forward FFTs, [8] inverse FFTs, [9] complex spectrum length C,
 * [10] up-pass GEMM flops of the mixed form (inner nodes by the fused
 * UC2UE x UE2UC product), [11] near-field entries the mixed form's FP32 near
 * field leaves to its FP64 pass (closest pair below 1/128 of the FP32 offsets'
 * span or below 1e-7; set by the device set-up, 0 before it). */
int kifmm3d_plan_stats(const kifmm3d* h, double* out, int n);

const char* kifmm3d_last_error(const kifmm3d* h);
void kifmm3d_destroy(kifmm3d* h);

/* Operator cache (no reference counterpart; the reference rebuilds its
 * pseudoinverses and M2L spectra lazily in the first eval of every process,
 * matmgnt3d.cpp:96-105, 163-171, 307-325).  With a directory set (or
 * KIFMM3D_OPS_CACHE in the environment), setup reads the operators of its
 * (translation kernel, coefficients, np) from a validated binary file there
 * and writes the file after building them; dir NULL or "" turns it off.
 * Files whose header, sizes or checksum do not match are rebuilt. */
int kifmm3d_set_operator_cache(const char* dir);
/* Operator sets built and loaded from disk since the process started. */
int kifmm3d_operator_cache_stats(int* built, int* loaded);
/* Free the in-memory operator sets that no handle uses; returns how many. */
int kifmm3d_release_operators(void);

/* Graph replay for repeated evaluations (no reference counterpart; for
 * callers that evaluate the same buffers many times): with on != 0,
 * kifmm3d_evaluate_device captures the whole pass into one hipGraph the first
 * time it sees a (buffers, nrhs, mode) combination and afterwards only
 * launches it.  Stage and kernel times then report the total only. */
int kifmm3d_set_graph(kifmm3d* h, int on);

/* Autotuner (no reference counterpart; the reference has one kernel per
 * stage).  Times, on this handle's tree and for (nrhs, mode), the candidate
 * near-field kernels and, per V-list level, the candidate M2L forms over
 * warm evaluates of zero densities (one warm-up and two timed evaluates per
 * round; rounds = the larger candidate count, at most 3), and keeps the
 * fastest of each for later evaluates of that mode.  Every candidate computes
 * the same sums (the plain form's candidates agree to rounding, 1e-12).
 * settings (may be NULL) receives up to nsettings entries: [0] near-field
 * kernel (0 LDS broadcast, 1 scalar-cache sources, 2 FP64 matrix cores,
 * 3 FP32 matrix cores, 4 FP32 LDS broadcast), [1 + l] M2L form of V-list level l, top level first
 * (0 tiled matrix-core, block per tile, 1 parent-grouped matrix-core,
 * 2 grouped VALU, 3 per pair, 4 tiled matrix-core, wave per tile and
 * frequency); nsettings = 17 always suffices.  *nrun (may be NULL) receives
 * the number of evaluates run.  Requires setup; allocates scratch densities
 * and values of nrhs vectors for its duration. */
int kifmm3d_tune(kifmm3d* h, int nrhs, int mode, int* settings, int nsettings, int* nrun);
/* The settings in force for a mode (same layout as kifmm3d_tune; the tiled
 * M2L form applies to chunks of 16k right-hand sides, judged on the last
 * evaluate's chunk).  Returns the number of entries (1 + V-list levels) or a
 * negative status. */
int kifmm3d_settings(const kifmm3d* h, int mode, int* settings, int nsettings);
/* Pin settings for a mode (same layout; -1 keeps an entry's default), e.g.
 * the ones a previous process tuned for the same problem.  EINVAL for a
 * kernel or form the problem cannot take (kernel 3 outside the mixed Laplace
 * form, the tiled M2L for vector kernels).  Requires setup. */
int kifmm3d_set_settings(kifmm3d* h, int mode, const int* settings, int nsettings);

/* Batched evaluation of independent problems (no reference counterpart; the
 * reference evaluates one FMM3d at a time).  A batch groups set-up handles;
 * kifmm3d_batch_evaluate_device evaluates handle i on d_srcden[i] ->
 * d_trgval[i] with nrhs[i] right-hand sides, all in one hipGraph whose
 * problems run on up to 8 concurrent branches.  The graph is captured on the
 * first call and replayed while the pointers, nrhs and mode stay the same.
 * The handles must outlive the batch and must not be evaluated on their own
 * while a batch call runs. */
typedef struct kifmm3d_batch kifmm3d_batch;
int kifmm3d_batch_create(kifmm3d_batch** out, kifmm3d* const* handles, int n);
int kifmm3d_batch_evaluate_device(kifmm3d_batch* b, const double* const* d_srcden, double* const* d_trgval,
                                  const int* nrhs, int mode);
/* Wall time (ms, HIP events) of the last batch launch. */
int kifmm3d_batch_time(const kifmm3d_batch* b, double* ms);
const char* kifmm3d_batch_last_error(const kifmm3d_batch* b);
void kifmm3d_batch_destroy(kifmm3d_batch* b);

/* Device memory helpers for callers without a HIP runtime of their own. */
int kifmm3d_device_alloc(void** ptr, long bytes);
int kifmm3d_device_free(void* ptr);
int kifmm3d_memcpy_h2d(void* dst, const void* src, long bytes);
int kifmm3d_memcpy_d2h(void* dst, const void* src, long bytes);
int kifmm3d_device_synchronize(void);
/* Device-to-device copy by a plain streaming kernel (bytes a multiple of 8);
 * it reads and writes exactly `bytes` each, which makes it the reference for
 * scaling hardware memory-traffic counters. */
int kifmm3d_device_copy(void* dst, const void* src, long bytes);
int kifmm3d_device_count(int* n);

#ifdef __cplusplus
}
#endif
#endif
