/*
 * kifmm_oracle.c -- CPU restatement of the reference KIFMM3d evaluate pass.
 *
 * TEST INFRASTRUCTURE ONLY.  The product never links, loads or calls this
 * file.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only as the checker.
 *
 * What it restates (paths relative to the reference repository root):
 *   fmm3d/kernel3d.cpp      kernels 111, 301, 321, 511 and the r < 1e-8 rule
 *   fmm3d/let3d.cpp         octree, BFS numbering, U/V/W/X lists
 *   fmm3d/matmgnt3d.cpp     surfaces, translation operators, FFT-based M2L
 *   fmm3d/fmm3d.cpp         point <-> surface applies
 *   fmm3d/fmm3d_setup.cpp   kernel map, leaf-contiguous gathers, V degrees
 *   fmm3d/fmm3d_eval.cpp    stage order, W/X direct fallbacks
 *   fmm3d/fmm3d_check.cpp, fmm3d/dense3d.cpp   FP64 direct sum
 * The truncated-SVD pseudoinverses (common/svdrep.cpp:31-74,
 * common/vecmatop.cpp:124-159) are computed by oracle/oracle.py with LAPACK
 * dgesvd (through scipy) -- the routine the reference itself calls -- and
 * handed in with orc_set_pinv().
 *
 * FFTs follow FFTW 2.1.5's rfftwnd conventions (row-major, last dimension
 * halved, forward sign -1, unnormalised) with a separable direct DFT.
 *
 * Parity pins (tests/test_oracle_*.py): kernels against PDE identities and
 * closed-form limits; surfaces against closed-form counts; FFT against
 * numpy.fft; FFT-M2L against the dense kernel product; tree/lists against
 * brute-force coverage invariants; end-to-end against the direct sum.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OK 0
#define ERR (-1)

/* ------------------------------------------------------------------ */
/* small growable int vector                                           */
typedef struct { int n, cap; int* a; } ivec;

static void iv_push(ivec* v, int x)
{
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 8;
    v->a = (int*)realloc(v->a, sizeof(int) * (size_t)v->cap);
  }
  v->a[v->n++] = x;
}
static void iv_free(ivec* v) { free(v->a); v->a = NULL; v->n = v->cap = 0; }
static int icmp(const void* a, const void* b)
{
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}
/* std::set<int> iteration order: ascending, unique */
static void iv_sort_unique(ivec* v)
{
  if (v->n < 2) return;
  qsort(v->a, (size_t)v->n, sizeof(int), icmp);
  int m = 1;
  for (int i = 1; i < v->n; i++)
    if (v->a[i] != v->a[m - 1]) v->a[m++] = v->a[i];
  v->n = m;
}

/* ------------------------------------------------------------------ */
/* kernels: fmm3d/kernel3d.cpp                                          */
static const double MINDIF = 1e-8; /* kernel3d.cpp:21 */

/* Summation-order switches (bit mask) used to measure the FP64 rounding floor
 * of the algorithm: the reference's sums run in BLAS/FFTW order, which is not
 * specified, so reversing them gives an equally valid evaluation.
 *   1: kernel sums over sources (Kernel3d::kernel + dgemv)
 *   2: dense matrix-vector products (pinv applies, M2M, L2L)
 *   4: order of the V-list accumulation
 *   8: DFT sums */
static int g_order = 0;
void orc_set_order(int mask) { g_order = mask; }

/* Accumulation of the dense matrix-vector products (the UC2UE / DC2DE
 * pseudoinverse applies, M2M and L2L; vecmatop.cpp:97-113).  The
 * pseudoinverses are ill-conditioned (singular values kept down to
 * 0.1^(np+1) sigma_0), so the rounding of y += A x in plain FP64 depends on
 * the summation order at the 1e-11 (np 6) / 1e-9 (np 8) level of the final
 * potential, and the reference's own result is only defined to that level
 * (its BLAS fixes the order).  With accurate = 1 (the default) these sums are
 * accumulated in extended precision (long double) and rounded once, which
 * makes the result independent of the summation order to 1e-14 (np 6) /
 * 4e-13 (np 8).  accurate = 0 restores plain FP64 accumulation. */
static int g_accurate = 1;
void orc_set_accurate(int on) { g_accurate = on; }

typedef struct { int kt; double c[2]; } oknl;

/* kernel3d.cpp:24-85 (sdof/tdof) */
int orc_kernel_dofs(int kt, int* sdof, int* tdof)
{
  switch (kt) {
  case 111: *sdof = 1; *tdof = 1; return OK;
  case 121: *sdof = 1; *tdof = 1; return OK;
  case 301: *sdof = 4; *tdof = 3; return OK;
  case 311: *sdof = 3; *tdof = 3; return OK;
  case 312: *sdof = 3; *tdof = 1; return OK;
  case 321: *sdof = 3; *tdof = 3; return OK;
  case 322: *sdof = 3; *tdof = 1; return OK;
  case 511: *sdof = 3; *tdof = 3; return OK;
  case 521: *sdof = 3; *tdof = 3; return OK;
  default: return ERR;
  }
}
/* fmm3d_setup.cpp:41-54: user kernel -> kernel used by the translation operators */
int orc_kernel_mm(int kt)
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
  default: return ERR;
  }
}
/* kernel3d.cpp:104-113 (homogeneity degree per source dof) */
static int kernel_homdeg(int kt, int* deg)
{
  switch (kt) {
  case 111: deg[0] = 1; return OK;
  case 301: deg[0] = 1; deg[1] = 1; deg[2] = 1; deg[3] = 2; return OK;
  case 511: deg[0] = 1; deg[1] = 1; deg[2] = 1; return OK;
  default: return ERR;
  }
}

/* One tdof x sdof block K(x = trg - src, n = src normal); b[t*sdof+s].
 * kernel3d.cpp:124-138 (111), :139-154 (121), :160-184 (301), :185-203 (311),
 * :204-221 (312), :222-242 (321), :243-265 (322), :315-337 (511),
 * :338-369 (521). */
static void kblock(const oknl* k, double x, double y, double z, const double* nrm, double* b)
{
  double r2 = x * x + y * y + z * z;
  double r = sqrt(r2);
  if (k->kt == 111) {
    b[0] = (r < MINDIF) ? 0.0 : (1.0 / (4.0 * M_PI)) / r;
    return;
  }
  if (k->kt == 301) {
    if (r < MINDIF) { for (int i = 0; i < 12; i++) b[i] = 0.0; return; }
    double mu = k->c[0];
    double OOFP = 1.0 / (4.0 * M_PI), OOEP = 1.0 / (8.0 * M_PI), oomu = 1.0 / mu;
    double r3 = r2 * r;
    double G = oomu * OOEP / r, H = oomu * OOEP / r3, A = OOFP / r3;
    b[0] = G + H * x * x; b[1] = H * x * y;     b[2] = H * x * z;      b[3] = A * x;
    b[4] = H * y * x;     b[5] = G + H * y * y; b[6] = H * y * z;      b[7] = A * y;
    b[8] = H * z * x;     b[9] = H * z * y;     b[10] = G + H * z * z; b[11] = A * z;
    return;
  }
  if (k->kt == 321) {
    if (r < MINDIF) { for (int i = 0; i < 9; i++) b[i] = 0.0; return; }
    double SOEP = 6.0 / (8.0 * M_PI);
    double rn = x * nrm[0] + y * nrm[1] + z * nrm[2];
    double r5 = r2 * r2 * r;
    double C = -SOEP / r5;
    double d[3] = {x, y, z};
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++) b[t * 3 + s] = C * rn * d[t] * d[s];
    return;
  }
  if (k->kt == 121) {
    if (r < MINDIF) { b[0] = 0.0; return; }
    double rn = x * nrm[0] + y * nrm[1] + z * nrm[2];
    b[0] = -(1.0 / (4.0 * M_PI)) / (r2 * r) * rn;
    return;
  }
  if (k->kt == 311) {
    if (r < MINDIF) { for (int i = 0; i < 9; i++) b[i] = 0.0; return; }
    double oomu = 1.0 / k->c[0], OOEP = 1.0 / (8.0 * M_PI);
    double G = OOEP / r, H = OOEP / (r2 * r);
    double d[3] = {x, y, z};
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++) b[t * 3 + s] = oomu * ((t == s ? G : 0.0) + H * d[t] * d[s]);
    return;
  }
  if (k->kt == 312) {
    if (r < MINDIF) { for (int i = 0; i < 3; i++) b[i] = 0.0; return; }
    double OOFP = 1.0 / (4.0 * M_PI), r3 = r2 * r;
    b[0] = OOFP * x / r3; b[1] = OOFP * y / r3; b[2] = OOFP * z / r3;
    return;
  }
  if (k->kt == 322) {
    if (r < MINDIF) { for (int i = 0; i < 3; i++) b[i] = 0.0; return; }
    double coef = k->c[0] * (1.0 / (2.0 * M_PI));
    double rn = x * nrm[0] + y * nrm[1] + z * nrm[2];
    double r3 = r2 * r, r5 = r3 * r2;
    b[0] = coef * (nrm[0] / r3 - 3 * rn * x / r5);
    b[1] = coef * (nrm[1] / r3 - 3 * rn * y / r5);
    b[2] = coef * (nrm[2] / r3 - 3 * rn * z / r5);
    return;
  }
  if (k->kt == 521) {
    if (r < MINDIF) { for (int i = 0; i < 9; i++) b[i] = 0.0; return; }
    double ve = k->c[1];
    double dc1 = -(1 - 2.0 * ve) / (8.0 * M_PI * (1.0 - ve));
    double dc2 = (1 - 2.0 * ve) / (8.0 * M_PI * (1.0 - ve));
    double dc3 = -3.0 / (8.0 * M_PI * (1.0 - ve));
    double rn = x * nrm[0] + y * nrm[1] + z * nrm[2];
    double r3 = r2 * r, r5 = r3 * r2;
    double A = dc1 / r3, B = dc2 / r3, C = dc3 / r5;
    double d[3] = {x, y, z};
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++) b[t * 3 + s] = A * ((t == s ? rn : 0.0) + nrm[t] * d[s]) + B * (d[t] * nrm[s]) + C * rn * d[t] * d[s];
    return;
  }
  if (k->kt == 511) {
    if (r < MINDIF) { for (int i = 0; i < 9; i++) b[i] = 0.0; return; }
    double mu = k->c[0], ve = k->c[1];
    double sc1 = (3.0 - 4.0 * ve) / (16.0 * M_PI * (1.0 - ve));
    double sc2 = 1.0 / (16.0 * M_PI * (1.0 - ve));
    double oomu = 1.0 / mu;
    double r3 = r2 * r;
    double G = sc1 / r, H = sc2 / r3;
    double d[3] = {x, y, z};
    for (int t = 0; t < 3; t++)
      for (int s = 0; s < 3; s++) b[t * 3 + s] = oomu * ((t == s ? G : 0.0) + H * d[t] * d[s]);
    return;
  }
}

static int knl_sdof(const oknl* k) { int s, t; orc_kernel_dofs(k->kt, &s, &t); return s; }
static int knl_tdof(const oknl* k) { int s, t; orc_kernel_dofs(k->kt, &s, &t); return t; }

/* Kernel3d::kernel (kernel3d.cpp:116-123): dense (nt*tdof) x (ns*sdof) block,
 * column-major.  Positions and normals are 3 x n column-major. */
int orc_kernel(int kt, const double* coefs, int ncoefs, int ns, const double* spos,
               const double* snor, int nt, const double* tpos, double* mat)
{
  oknl k; int sdof, tdof;
  if (orc_kernel_dofs(kt, &sdof, &tdof) != OK) return ERR;
  k.kt = kt; k.c[0] = ncoefs > 0 ? coefs[0] : 1.0; k.c[1] = ncoefs > 1 ? coefs[1] : 0.0;
  long M = (long)nt * tdof;
  double b[12];
  for (int j = 0; j < ns; j++)
    for (int i = 0; i < nt; i++) {
      kblock(&k, tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
             tpos[3 * i + 2] - spos[3 * j + 2], snor ? snor + 3 * j : NULL, b);
      for (int t = 0; t < tdof; t++)
        for (int s = 0; s < sdof; s++) mat[((long)j * sdof + s) * M + (long)i * tdof + t] = b[t * sdof + s];
    }
  return OK;
}

/* val += K(tpos <- spos) den, i.e. Kernel3d::kernel followed by dgemv(1, K, den, 1, val)
 * (fmm3d.cpp:35-60 and friends), without materialising K. */
static void kapply(const oknl* k, int ns, const double* spos, const double* snor, const double* den,
                   int nt, const double* tpos, double* val)
{
  int sdof = knl_sdof(k), tdof = knl_tdof(k);
  double b[12];
  for (int i = 0; i < nt; i++) {
    double acc[3] = {0, 0, 0};
    for (int jj = 0; jj < ns; jj++) {
      int j = (g_order & 1) ? ns - 1 - jj : jj;
      kblock(k, tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
             tpos[3 * i + 2] - spos[3 * j + 2], snor ? snor + 3 * j : NULL, b);
      for (int t = 0; t < tdof; t++) {
        double s_ = 0;
        for (int s = 0; s < sdof; s++) s_ += b[t * sdof + s] * den[j * sdof + s];
        acc[t] += s_;
      }
    }
    for (int t = 0; t < tdof; t++) val[i * tdof + t] += acc[t];
  }
}

/* Dense3d::eval / FMM3d::check direct sum (dense3d.cpp:38-58, fmm3d_check.cpp:48-54):
 * tval = K(tpos <- spos) sden (overwritten). */
int orc_direct(int kt, const double* coefs, int ncoefs, int ns, const double* spos, const double* snor,
               const double* sden, int nt, const double* tpos, double* tval)
{
  oknl k; int sdof, tdof;
  if (orc_kernel_dofs(kt, &sdof, &tdof) != OK) return ERR;
  k.kt = kt; k.c[0] = ncoefs > 0 ? coefs[0] : 1.0; k.c[1] = ncoefs > 1 ? coefs[1] : 0.0;
  memset(tval, 0, sizeof(double) * (size_t)nt * tdof);
  /* targets in blocks over the threads; each target sums its sources in order */
#pragma omp parallel for schedule(dynamic, 1)
  for (int i0 = 0; i0 < nt; i0 += 16) {
    int m = nt - i0 < 16 ? nt - i0 : 16;
    kapply(&k, ns, spos, snor, sden, m, tpos + 3 * (size_t)i0, tval + (size_t)i0 * tdof);
  }
  return OK;
}

/* ------------------------------------------------------------------ */
/* surfaces: matmgnt3d.cpp:375-419                                      */
int orc_splpos(int np, double R, double* pos)
{
  int n = np * np * np - (np - 2) * (np - 2) * (np - 2);
  double step = 2.0 / (np - 1), init = -1.0;
  int cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
          double x = init + i * step, y = init + j * step, z = init + k * step;
          if (pos) { pos[3 * cnt] = R * x; pos[3 * cnt + 1] = R * y; pos[3 * cnt + 2] = R * z; }
          cnt++;
        }
  return cnt == n ? n : ERR;
}
int orc_regpos(int np, double R, double* pos)
{
  int n = 2 * np * 2 * np * 2 * np;
  double step = 2.0 / (np - 1);
  int cnt = 0;
  for (int k = 0; k < 2 * np; k++)
    for (int j = 0; j < 2 * np; j++)
      for (int i = 0; i < 2 * np; i++) {
        int gi = (i < np) ? i : i - 2 * np, gj = (j < np) ? j : j - 2 * np, gk = (k < np) ? k : k - 2 * np;
        if (pos) { pos[3 * cnt] = R * gi * step; pos[3 * cnt + 1] = R * gj * step; pos[3 * cnt + 2] = R * gk * step; }
        cnt++;
      }
  return n;
}

/* ------------------------------------------------------------------ */
/* FFT in FFTW2 rfftwnd layout: real [z][y][x] (x fastest), complex    */
/* [z][y][x'] with x' in [0, n/2]; hm interleaved transforms with      */
/* stride hm and distance 1 (matmgnt3d.cpp:239, 269, 324).             */
static void twiddles(int n, double* c, double* s)
{
  for (int m = 0; m < n; m++) { c[m] = cos(2.0 * M_PI * m / n); s[m] = sin(2.0 * M_PI * m / n); }
}

/* forward r2c, sign -1, unnormalised */
int orc_rfft3(int n, int hm, const double* in, double* out)
{
  int h = n / 2 + 1;
  double c[64], s[64];
  if (n > 64) return ERR;
  twiddles(n, c, s);
  size_t nc = (size_t)n * n * h;
  double* a = (double*)calloc(2 * nc, sizeof(double));
  double* b = (double*)calloc(2 * nc, sizeof(double));
  for (int f = 0; f < hm; f++) {
    /* x: real -> half complex */
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int kx = 0; kx < h; kx++) {
          double re = 0, im = 0;
          for (int xx = 0; xx < n; xx++) {
            int x = (g_order & 8) ? n - 1 - xx : xx;
            double v = in[((size_t)(z * n + y) * n + x) * hm + f];
            int m = (kx * x) % n;
            re += v * c[m]; im -= v * s[m];
          }
          size_t o = (size_t)(z * n + y) * h + kx;
          a[2 * o] = re; a[2 * o + 1] = im;
        }
    /* y */
    for (int z = 0; z < n; z++)
      for (int ky = 0; ky < n; ky++)
        for (int kx = 0; kx < h; kx++) {
          double re = 0, im = 0;
          for (int yy = 0; yy < n; yy++) {
            int y = (g_order & 8) ? n - 1 - yy : yy;
            size_t o = (size_t)(z * n + y) * h + kx;
            int m = (ky * y) % n;
            re += a[2 * o] * c[m] + a[2 * o + 1] * s[m];
            im += a[2 * o + 1] * c[m] - a[2 * o] * s[m];
          }
          size_t o = (size_t)(z * n + ky) * h + kx;
          b[2 * o] = re; b[2 * o + 1] = im;
        }
    /* z */
    for (int kz = 0; kz < n; kz++)
      for (int ky = 0; ky < n; ky++)
        for (int kx = 0; kx < h; kx++) {
          double re = 0, im = 0;
          for (int zz = 0; zz < n; zz++) {
            int z = (g_order & 8) ? n - 1 - zz : zz;
            size_t o = (size_t)(z * n + ky) * h + kx;
            int m = (kz * z) % n;
            re += b[2 * o] * c[m] + b[2 * o + 1] * s[m];
            im += b[2 * o + 1] * c[m] - b[2 * o] * s[m];
          }
          size_t o = (size_t)(kz * n + ky) * h + kx;
          out[2 * (o * hm + f)] = re; out[2 * (o * hm + f) + 1] = im;
        }
  }
  free(a); free(b);
  return OK;
}

/* inverse c2r, sign +1, unnormalised; the imaginary parts of the
 * self-conjugate x' = 0 and x' = n/2 planes are ignored. */
int orc_irfft3(int n, int hm, const double* in, double* out)
{
  int h = n / 2 + 1;
  double c[64], s[64];
  if (n > 64) return ERR;
  twiddles(n, c, s);
  size_t nc = (size_t)n * n * h;
  double* a = (double*)calloc(2 * nc, sizeof(double));
  double* b = (double*)calloc(2 * nc, sizeof(double));
  for (int f = 0; f < hm; f++) {
    /* z */
    for (int z = 0; z < n; z++)
      for (int ky = 0; ky < n; ky++)
        for (int kx = 0; kx < h; kx++) {
          double re = 0, im = 0;
          for (int kk = 0; kk < n; kk++) {
            int kz = (g_order & 8) ? n - 1 - kk : kk;
            size_t o = ((size_t)(kz * n + ky) * h + kx) * hm + f;
            int m = (kz * z) % n;
            re += in[2 * o] * c[m] - in[2 * o + 1] * s[m];
            im += in[2 * o + 1] * c[m] + in[2 * o] * s[m];
          }
          size_t o = (size_t)(z * n + ky) * h + kx;
          a[2 * o] = re; a[2 * o + 1] = im;
        }
    /* y */
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int kx = 0; kx < h; kx++) {
          double re = 0, im = 0;
          for (int kk = 0; kk < n; kk++) {
            int ky = (g_order & 8) ? n - 1 - kk : kk;
            size_t o = (size_t)(z * n + ky) * h + kx;
            int m = (ky * y) % n;
            re += a[2 * o] * c[m] - a[2 * o + 1] * s[m];
            im += a[2 * o + 1] * c[m] + a[2 * o] * s[m];
          }
          size_t o = (size_t)(z * n + y) * h + kx;
          b[2 * o] = re; b[2 * o + 1] = im;
        }
    /* x: half complex -> real */
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          size_t o = (size_t)(z * n + y) * h;
          double v = b[2 * o];
          if (n % 2 == 0) v += ((x % 2) ? -1.0 : 1.0) * b[2 * (o + n / 2)];
          for (int kx = 1; kx < (n + 1) / 2; kx++) {
            int m = (kx * x) % n;
            v += 2.0 * (b[2 * (o + kx)] * c[m] - b[2 * (o + kx) + 1] * s[m]);
          }
          out[((size_t)(z * n + y) * n + x) * hm + f] = v;
        }
  }
  free(a); free(b);
  return OK;
}

/* ------------------------------------------------------------------ */
/* octree / LET: fmm3d/let3d.cpp                                        */
enum { SRCNODE = 1, TRGNODE = 2 }; /* let3d.hpp:27-30 */

typedef struct {
  int par, chd, path[3], depth, tag;
  int sndeidx, sextbeg, sextnum;
  int tndeidx, textbeg, textnum;
  ivec sown, town;
  ivec U, V, W, X;
} onode;

typedef struct orc_fmm {
  oknl knl, mm;
  int sdof, tdof, sm, tm, deg[4];
  int np, ptsmax, maxlevel, rootlvl;
  double ctr[3];
  int ns, nt;
  double *spos, *snor, *tpos; /* owned copies, 3 x n column-major */
  int nnode, capnode, level;
  onode* nd;
  int sndecnt, sextcnt, tndecnt, textcnt;
  /* operators (matmgnt3d) */
  int nsp[4];     /* UE, UC, DE, DC point counts */
  double* sp[4];  /* unit surfaces: UE (np,1), UC (np+2,3), DE (np,3), DC (np,1) */
  double* reg;    /* regular grid, R = 1 */
  int G, C;       /* (2np)^3 real, (np+1)(2np)^2 complex per dof */
  int pln[4];     /* plndatasize */
  double *uc2ue, *dc2de;
  double *ue2uc[8], *de2dc[8];
  double* ue2dc[343];
  /* eval data (fmm3d_setup.cpp:65-140) */
  double *sextpos, *sextnor, *sextden, *supeden, *supcval;
  double *textpos, *textval, *tdneden, *tdncval;
  int *vinnum, *votnum, *vincnt, *votcnt;
  double **effden, **effval;
} orc_fmm;

enum { UE = 0, UC = 1, DE = 2, DC = 3 };

static double rad(const orc_fmm* h) { return pow(2.0, -h->rootlvl); } /* let3d.hpp:98 */
static int terminal(const orc_fmm* h, int g) { return h->nd[g].chd == -1; }
static int child(const orc_fmm* h, int g, int a, int b, int c) { return h->nd[g].chd + (a * 4 + b * 2 + c); }

/* let3d.cpp:309-323 */
static void center(const orc_fmm* h, int g, double* res)
{
  double r = rad(h);
  int tmp = 1 << h->nd[g].depth;
  for (int d = 0; d < 3; d++) res[d] = (h->ctr[d] - r) + (2 * r) * (h->nd[g].path[d] + 0.5) / (double)tmp;
}
static double radius(const orc_fmm* h, int g) { return rad(h) / (double)(1 << h->nd[g].depth); }

static int push_node(orc_fmm* h, int par, int path0, int path1, int path2, int depth)
{
  if (h->nnode == h->capnode) {
    h->capnode = h->capnode ? 2 * h->capnode : 64;
    h->nd = (onode*)realloc(h->nd, sizeof(onode) * (size_t)h->capnode);
  }
  onode* n = &h->nd[h->nnode];
  memset(n, 0, sizeof(onode));
  n->par = par; n->chd = -1; n->path[0] = path0; n->path[1] = path1; n->path[2] = path2; n->depth = depth;
  return h->nnode++;
}

static int inbox(const orc_fmm* h, const double* p)
{
  double r = rad(h);
  for (int d = 0; d < 3; d++)
    if (!(p[d] >= h->ctr[d] - r && p[d] <= h->ctr[d] + r)) return 0;
  return 1;
}

/* Let3d::srcdata, let3d.cpp:56-137 */
static int tree_src(orc_fmm* h)
{
  h->nnode = 0;
  push_node(h, -1, 0, 0, 0, 0);
  ivec* vis = (ivec*)calloc(1, sizeof(ivec));
  int* lsm = (int*)calloc(1, sizeof(int));
  int cap = 1;
  for (int k = 0; k < h->ns; k++) {
    if (!inbox(h, h->spos + 3 * k)) { free(vis[0].a); free(vis); free(lsm); return ERR; }
    iv_push(&vis[0], k);
  }
  lsm[0] = vis[0].n;
  int level = 0, arrbeg = 0, arrend = 1;
  while (arrbeg < arrend) {
    int arrcnt = arrend;
    for (int k = arrbeg; k < arrend; k++) {
      if (lsm[k] > h->ptsmax && level < h->maxlevel - 1) {
        h->nd[k].chd = arrcnt;
        arrcnt += 8;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++)
              push_node(h, k, 2 * h->nd[k].path[0] + a, 2 * h->nd[k].path[1] + b, 2 * h->nd[k].path[2] + c,
                        h->nd[k].depth + 1);
        if (h->nnode > cap) {
          int nc = 2 * h->nnode;
          vis = (ivec*)realloc(vis, sizeof(ivec) * (size_t)nc);
          lsm = (int*)realloc(lsm, sizeof(int) * (size_t)nc);
          memset(vis + cap, 0, sizeof(ivec) * (size_t)(nc - cap));
          memset(lsm + cap, 0, sizeof(int) * (size_t)(nc - cap));
          cap = nc;
        }
        double cc[3];
        center(h, k, cc);
        for (int q = 0; q < vis[k].n; q++) {
          int pi = vis[k].a[q];
          const double* p = h->spos + 3 * pi;
          int chd = child(h, k, p[0] >= cc[0], p[1] >= cc[1], p[2] >= cc[2]);
          iv_push(&vis[chd], pi);
        }
        iv_free(&vis[k]);
        for (int c = 0; c < 8; c++) lsm[h->nd[k].chd + c] = vis[h->nd[k].chd + c].n;
      }
    }
    level++;
    arrbeg = arrend;
    arrend = arrcnt;
  }
  h->level = level;
  int cnt = 0, sum = 0;
  for (int g = 0; g < h->nnode; g++) {
    if (lsm[g] > 0) {
      h->nd[g].tag |= SRCNODE;
      h->nd[g].sndeidx = cnt++;
      if (h->nd[g].chd == -1) {
        h->nd[g].sextbeg = sum;
        h->nd[g].sextnum = lsm[g];
        sum += lsm[g];
        h->nd[g].sown = vis[g];
        vis[g].a = NULL; vis[g].n = vis[g].cap = 0;
      }
    }
  }
  h->sndecnt = cnt; h->sextcnt = sum;
  for (int g = 0; g < cap; g++) iv_free(&vis[g]);
  free(vis); free(lsm);
  return OK;
}

/* Let3d::findgnt, let3d.cpp:325-337 */
static int findgnt(const orc_fmm* h, int wntdepth, const int* wntpath)
{
  int cur = 0;
  int left[3] = {wntpath[0], wntpath[1], wntpath[2]};
  while (h->nd[cur].depth < wntdepth && !terminal(h, cur)) {
    int dif = wntdepth - h->nd[cur].depth;
    int tmp = 1 << (dif - 1);
    int ch[3];
    for (int d = 0; d < 3; d++) { ch[d] = left[d] / tmp; left[d] -= ch[d] * tmp; }
    cur = child(h, cur, ch[0], ch[1], ch[2]);
  }
  return cur;
}

/* Let3d::adjacent, let3d.cpp:339-352 */
static int adjacent(const orc_fmm* h, int me, int yo)
{
  int dm = h->nd[me].depth, dy = h->nd[yo].depth;
  int md = dm > dy ? dm : dy;
  int merad = 1 << (md - dm), yorad = 1 << (md - dy);
  int r = merad + yorad, linf = 0, ok = 1;
  for (int d = 0; d < 3; d++) {
    int mc = (2 * h->nd[me].path[d] + 1) * merad;
    int yc = (2 * h->nd[yo].path[d] + 1) * yorad;
    int dif = abs(mc - yc);
    if (dif > r) ok = 0;
    if (dif > linf) linf = dif;
  }
  return ok && linf == r;
}

/* Let3d::calgnext, let3d.cpp:214-282 */
static void calgnext(orc_fmm* h, int g)
{
  ivec U = {0}, V = {0}, W = {0}, X = {0};
  onode* cur = &h->nd[g];
  if (cur->par != -1) {
    int par = cur->par;
    int maxidx = 1 << cur->depth;
    for (int i = -2; i < 4; i++)
      for (int j = -2; j < 4; j++)
        for (int k = -2; k < 4; k++) {
          int tp[3] = {2 * h->nd[par].path[0] + i, 2 * h->nd[par].path[1] + j, 2 * h->nd[par].path[2] + k};
          int ok = 1;
          for (int d = 0; d < 3; d++)
            if (!(tp[d] >= 0 && tp[d] < maxidx)) ok = 0;
          if (!ok) continue;
          if (tp[0] == cur->path[0] && tp[1] == cur->path[1] && tp[2] == cur->path[2]) continue;
          int res = findgnt(h, cur->depth, tp);
          int adj = adjacent(h, res, g);
          if (h->nd[res].depth < cur->depth) {
            if (adj) {
              if (terminal(h, g)) iv_push(&U, res);
            } else {
              iv_push(&X, res);
            }
          }
          if (h->nd[res].depth == cur->depth) {
            if (!adj) {
              iv_push(&V, res);
            } else if (terminal(h, g)) {
              /* breadth-first walk below the same-depth neighbour */
              ivec q = {0};
              int head = 0;
              iv_push(&q, res);
              while (head < q.n) {
                int f = q.a[head++];
                if (!adjacent(h, f, g))
                  iv_push(&W, f);
                else if (terminal(h, f))
                  iv_push(&U, f);
                else
                  for (int c = 0; c < 8; c++) iv_push(&q, h->nd[f].chd + c);
              }
              iv_free(&q);
            }
          }
        }
  }
  if (terminal(h, g)) iv_push(&U, g);
  ivec* lists[4] = {&U, &V, &W, &X};
  ivec* outs[4] = {&h->nd[g].U, &h->nd[g].V, &h->nd[g].W, &h->nd[g].X};
  for (int l = 0; l < 4; l++) {
    iv_sort_unique(lists[l]);
    for (int q = 0; q < lists[l]->n; q++)
      if (h->nd[lists[l]->a[q]].tag & SRCNODE) iv_push(outs[l], lists[l]->a[q]);
    iv_free(lists[l]);
  }
}

/* Let3d::trgdata, let3d.cpp:140-211 */
static int tree_trg(orc_fmm* h)
{
  int nn = h->nnode;
  ivec* vis = (ivec*)calloc((size_t)nn, sizeof(ivec));
  int* lsm = (int*)calloc((size_t)nn, sizeof(int));
  for (int k = 0; k < h->nt; k++) {
    if (!inbox(h, h->tpos + 3 * k)) { free(vis[0].a); free(vis); free(lsm); return ERR; }
    iv_push(&vis[0], k);
  }
  lsm[0] = vis[0].n;
  for (int g = 0; g < nn; g++) {
    if (h->nd[g].chd != -1) {
      double cc[3];
      center(h, g, cc);
      for (int q = 0; q < vis[g].n; q++) {
        int pi = vis[g].a[q];
        const double* p = h->tpos + 3 * pi;
        iv_push(&vis[child(h, g, p[0] >= cc[0], p[1] >= cc[1], p[2] >= cc[2])], pi);
      }
      iv_free(&vis[g]);
      for (int c = 0; c < 8; c++) lsm[h->nd[g].chd + c] = vis[h->nd[g].chd + c].n;
    }
  }
  int cnt = 0, sum = 0;
  for (int g = 0; g < nn; g++) {
    if (lsm[g] > 0) {
      h->nd[g].tag |= TRGNODE;
      h->nd[g].tndeidx = cnt++;
      if (h->nd[g].chd == -1) {
        h->nd[g].textbeg = sum;
        h->nd[g].textnum = lsm[g];
        sum += lsm[g];
        h->nd[g].town = vis[g];
        vis[g].a = NULL; vis[g].n = vis[g].cap = 0;
      }
    }
  }
  h->tndecnt = cnt; h->textcnt = sum;
  for (int g = 0; g < nn; g++) iv_free(&vis[g]);
  free(vis); free(lsm);
  for (int g = 0; g < nn; g++)
    if (h->nd[g].tag & TRGNODE) calgnext(h, g);
  return OK;
}

/* ------------------------------------------------------------------ */
/* FMM object                                                           */
orc_fmm* orc_create(int kt, const double* coefs, int ncoefs, int np, int ptsmax, int maxlevel,
                    const double* ctr, int rootlvl)
{
  int s, t, mmkt = orc_kernel_mm(kt);
  if (mmkt == ERR || orc_kernel_dofs(kt, &s, &t) != OK) return NULL;
  if (np < 2 || np > 16 || maxlevel < 1 || maxlevel > 16) return NULL;
  orc_fmm* h = (orc_fmm*)calloc(1, sizeof(orc_fmm));
  h->knl.kt = kt;
  h->knl.c[0] = ncoefs > 0 ? coefs[0] : 1.0;
  h->knl.c[1] = ncoefs > 1 ? coefs[1] : 0.0;
  h->mm = h->knl;
  h->mm.kt = mmkt;
  h->sdof = s; h->tdof = t;
  orc_kernel_dofs(mmkt, &h->sm, &h->tm);
  kernel_homdeg(mmkt, h->deg);
  h->np = np; h->ptsmax = ptsmax; h->maxlevel = maxlevel; h->rootlvl = rootlvl;
  for (int d = 0; d < 3; d++) h->ctr[d] = ctr[d];
  /* MatMgnt3d::setup, matmgnt3d.cpp:44-62 */
  int nps[4] = {np, np + 2, np, np};
  double Rs[4] = {1.0, 3.0, 3.0, 1.0};
  for (int q = 0; q < 4; q++) {
    h->nsp[q] = orc_splpos(nps[q], Rs[q], NULL);
    h->sp[q] = (double*)malloc(sizeof(double) * 3 * (size_t)h->nsp[q]);
    orc_splpos(nps[q], Rs[q], h->sp[q]);
  }
  h->G = 8 * np * np * np;
  h->C = (np + 1) * 4 * np * np;
  h->reg = (double*)malloc(sizeof(double) * 3 * (size_t)h->G);
  orc_regpos(np, 1.0, h->reg);
  /* plndatasize, matmgnt3d.cpp:72-78 */
  h->pln[UE] = h->nsp[UE] * h->sm; h->pln[DE] = h->nsp[DE] * h->sm;
  h->pln[UC] = h->nsp[UC] * h->tm; h->pln[DC] = h->nsp[DC] * h->tm;
  return h;
}

static void free_tree(orc_fmm* h)
{
  for (int g = 0; g < h->nnode; g++) {
    onode* n = &h->nd[g];
    iv_free(&n->sown); iv_free(&n->town);
    iv_free(&n->U); iv_free(&n->V); iv_free(&n->W); iv_free(&n->X);
  }
  free(h->nd); h->nd = NULL; h->nnode = h->capnode = 0;
}

static void free_eval(orc_fmm* h)
{
  free(h->sextpos); free(h->sextnor); free(h->sextden); free(h->supeden); free(h->supcval);
  free(h->textpos); free(h->textval); free(h->tdneden); free(h->tdncval);
  free(h->vinnum); free(h->votnum); free(h->vincnt); free(h->votcnt);
  if (h->effden)
    for (int g = 0; g < h->nnode; g++) { free(h->effden[g]); free(h->effval[g]); }
  free(h->effden); free(h->effval);
  h->sextpos = h->sextnor = h->sextden = h->supeden = h->supcval = NULL;
  h->textpos = h->textval = h->tdneden = h->tdncval = NULL;
  h->vinnum = h->votnum = h->vincnt = h->votcnt = NULL;
  h->effden = h->effval = NULL;
}

void orc_destroy(orc_fmm* h)
{
  if (!h) return;
  free_eval(h);
  free_tree(h);
  for (int q = 0; q < 4; q++) free(h->sp[q]);
  free(h->reg); free(h->uc2ue); free(h->dc2de);
  for (int c = 0; c < 8; c++) { free(h->ue2uc[c]); free(h->de2dc[c]); }
  for (int o = 0; o < 343; o++) free(h->ue2dc[o]);
  free(h->spos); free(h->snor); free(h->tpos);
  free(h);
}

/* FMM3d::setup + srcdata + trgdata, fmm3d_setup.cpp:26-140 */
int orc_setup(orc_fmm* h, int ns, const double* spos, const double* snor, int nt, const double* tpos)
{
  free_eval(h);
  free_tree(h);
  free(h->spos); free(h->snor); free(h->tpos);
  h->ns = ns; h->nt = nt;
  h->spos = (double*)malloc(sizeof(double) * 3 * (size_t)(ns ? ns : 1));
  h->snor = (double*)malloc(sizeof(double) * 3 * (size_t)(ns ? ns : 1));
  h->tpos = (double*)malloc(sizeof(double) * 3 * (size_t)(nt ? nt : 1));
  memcpy(h->spos, spos, sizeof(double) * 3 * (size_t)ns);
  memcpy(h->snor, snor ? snor : spos, sizeof(double) * 3 * (size_t)ns);
  memcpy(h->tpos, tpos, sizeof(double) * 3 * (size_t)nt);
  if (tree_src(h) != OK) return ERR;
  if (tree_trg(h) != OK) return ERR;
  h->sextpos = (double*)calloc(3 * (size_t)h->sextcnt + 1, sizeof(double));
  h->sextnor = (double*)calloc(3 * (size_t)h->sextcnt + 1, sizeof(double));
  h->sextden = (double*)calloc((size_t)h->sextcnt * h->sdof + 1, sizeof(double));
  h->supeden = (double*)calloc((size_t)h->sndecnt * h->pln[UE] + 1, sizeof(double));
  h->supcval = (double*)calloc((size_t)h->sndecnt * h->pln[UC] + 1, sizeof(double));
  h->textpos = (double*)calloc(3 * (size_t)h->textcnt + 1, sizeof(double));
  h->textval = (double*)calloc((size_t)h->textcnt * h->tdof + 1, sizeof(double));
  h->tdneden = (double*)calloc((size_t)h->tndecnt * h->pln[DE] + 1, sizeof(double));
  h->tdncval = (double*)calloc((size_t)h->tndecnt * h->pln[DC] + 1, sizeof(double));
  for (int g = 0; g < h->nnode; g++) {
    onode* n = &h->nd[g];
    if ((n->tag & SRCNODE) && n->chd == -1)
      for (int k = 0; k < n->sown.n; k++)
        for (int d = 0; d < 3; d++) {
          h->sextpos[3 * (n->sextbeg + k) + d] = h->spos[3 * n->sown.a[k] + d];
          h->sextnor[3 * (n->sextbeg + k) + d] = h->snor[3 * n->sown.a[k] + d];
        }
    if ((n->tag & TRGNODE) && n->chd == -1)
      for (int k = 0; k < n->town.n; k++)
        for (int d = 0; d < 3; d++) h->textpos[3 * (n->textbeg + k) + d] = h->tpos[3 * n->town.a[k] + d];
  }
  h->vinnum = (int*)calloc((size_t)h->nnode, sizeof(int));
  h->votnum = (int*)calloc((size_t)h->nnode, sizeof(int));
  h->vincnt = (int*)calloc((size_t)h->nnode, sizeof(int));
  h->votcnt = (int*)calloc((size_t)h->nnode, sizeof(int));
  h->effden = (double**)calloc((size_t)h->nnode, sizeof(double*));
  h->effval = (double**)calloc((size_t)h->nnode, sizeof(double*));
  for (int g = 0; g < h->nnode; g++)
    if (h->nd[g].tag & TRGNODE) {
      h->vinnum[g] = h->nd[g].V.n;
      for (int q = 0; q < h->nd[g].V.n; q++) h->votnum[h->nd[g].V.a[q]]++;
    }
  return OK;
}

/* The two check-to-equivalent kernel matrices that the pseudoinverses invert
 * (matmgnt3d.cpp:98-102 and :164-168, homogeneous kernels, level 0):
 * kuc = K(UC <- UE), plnUC x plnUE; kdc = K(DC <- DE), plnDC x plnDE. */
int orc_check_matrices(const orc_fmm* h, double* kuc, double* kdc)
{
  double cf[2] = {h->mm.c[0], h->mm.c[1]};
  if (kuc && orc_kernel(h->mm.kt, cf, 2, h->nsp[UE], h->sp[UE], h->sp[UE], h->nsp[UC], h->sp[UC], kuc) != OK) return ERR;
  if (kdc && orc_kernel(h->mm.kt, cf, 2, h->nsp[DE], h->sp[DE], h->sp[DE], h->nsp[DC], h->sp[DC], kdc) != OK) return ERR;
  return OK;
}

/* uc2ue: plnUE x plnUC, dc2de: plnDE x plnDC, column-major */
int orc_set_pinv(orc_fmm* h, const double* uc2ue, const double* dc2de)
{
  size_t a = (size_t)h->pln[UE] * h->pln[UC], b = (size_t)h->pln[DE] * h->pln[DC];
  free(h->uc2ue); free(h->dc2de);
  h->uc2ue = (double*)malloc(sizeof(double) * a);
  h->dc2de = (double*)malloc(sizeof(double) * b);
  memcpy(h->uc2ue, uc2ue, sizeof(double) * a);
  memcpy(h->dc2de, dc2de, sizeof(double) * b);
  return OK;
}

int orc_info(const orc_fmm* h, int* out /* 16 ints */)
{
  out[0] = h->nnode; out[1] = h->level; out[2] = h->sndecnt; out[3] = h->sextcnt;
  out[4] = h->tndecnt; out[5] = h->textcnt; out[6] = h->pln[UE]; out[7] = h->pln[UC];
  out[8] = h->pln[DE]; out[9] = h->pln[DC]; out[10] = h->sdof; out[11] = h->tdof;
  out[12] = h->sm; out[13] = h->tm; out[14] = h->C; out[15] = h->G;
  return OK;
}

/* per-node table: [par, chd, p0, p1, p2, depth, tag, sndeidx, sextbeg, sextnum,
 * tndeidx, textbeg, textnum] x nnode */
int orc_nodes(const orc_fmm* h, int* out)
{
  for (int g = 0; g < h->nnode; g++) {
    const onode* n = &h->nd[g];
    int* o = out + 13 * g;
    o[0] = n->par; o[1] = n->chd; o[2] = n->path[0]; o[3] = n->path[1]; o[4] = n->path[2];
    o[5] = n->depth; o[6] = n->tag; o[7] = n->sndeidx; o[8] = n->sextbeg; o[9] = n->sextnum;
    o[10] = n->tndeidx; o[11] = n->textbeg; o[12] = n->textnum;
  }
  return OK;
}

/* CSR of one list over all nodes: which = 0 U, 1 V, 2 W, 3 X; ptr has nnode+1
 * entries; idx may be NULL to query the size (returned). */
int orc_list(const orc_fmm* h, int which, int* ptr, int* idx)
{
  int tot = 0;
  for (int g = 0; g < h->nnode; g++) {
    const ivec* l = which == 0 ? &h->nd[g].U : which == 1 ? &h->nd[g].V : which == 2 ? &h->nd[g].W : &h->nd[g].X;
    if (ptr) ptr[g] = tot;
    if (idx) memcpy(idx + tot, l->a, sizeof(int) * (size_t)l->n);
    tot += l->n;
  }
  if (ptr) ptr[h->nnode] = tot;
  return tot;
}

/* leaf-contiguous permutations sownvis / townvis concatenated in BFS order */
int orc_perm(const orc_fmm* h, int* sperm, int* tperm)
{
  for (int g = 0; g < h->nnode; g++) {
    const onode* n = &h->nd[g];
    if (sperm && (n->tag & SRCNODE) && n->chd == -1) memcpy(sperm + n->sextbeg, n->sown.a, sizeof(int) * (size_t)n->sown.n);
    if (tperm && (n->tag & TRGNODE) && n->chd == -1) memcpy(tperm + n->textbeg, n->town.a, sizeof(int) * (size_t)n->town.n);
  }
  return OK;
}

/* ------------------------------------------------------------------ */
/* translation operators: matmgnt3d.cpp                                 */
static void locpos(const orc_fmm* h, int tp, const double* c, double r, double* pos)
{ /* matmgnt3d.cpp:364-372 */
  for (int j = 0; j < h->nsp[tp]; j++)
    for (int d = 0; d < 3; d++) pos[3 * j + d] = c[d] + r * h->sp[tp][3 * j + d];
}

static void dgemv_acc(int m, int n, const double* A, const double* x, double* y)
{ /* y += A x, column-major (vecmatop.cpp:97-113) */
  if (!g_accurate) {
    for (int jj = 0; jj < n; jj++) {
      int j = (g_order & 2) ? n - 1 - jj : jj;
      double xj = x[j];
      const double* a = A + (size_t)j * m;
      for (int i = 0; i < m; i++) y[i] += a[i] * xj;
    }
    return;
  }
  long double* t = (long double*)malloc(sizeof(long double) * (size_t)m);
  for (int i = 0; i < m; i++) t[i] = y[i];
  /* four columns per pass over t (each t[i] still takes them in order) */
  int jj = 0;
  for (; jj + 4 <= n; jj += 4) {
    int j0 = (g_order & 2) ? n - 1 - jj : jj, dj = (g_order & 2) ? -1 : 1;
    long double x0 = x[j0], x1 = x[j0 + dj], x2 = x[j0 + 2 * dj], x3 = x[j0 + 3 * dj];
    const double *a0 = A + (size_t)j0 * m, *a1 = A + (size_t)(j0 + dj) * m, *a2 = A + (size_t)(j0 + 2 * dj) * m,
                 *a3 = A + (size_t)(j0 + 3 * dj) * m;
    for (int i = 0; i < m; i++) {
      long double u = t[i];
      u += (long double)a0[i] * x0;
      u += (long double)a1[i] * x1;
      u += (long double)a2[i] * x2;
      u += (long double)a3[i] * x3;
      t[i] = u;
    }
  }
  for (; jj < n; jj++) {
    int j = (g_order & 2) ? n - 1 - jj : jj;
    long double xj = x[j];
    const double* a = A + (size_t)j * m;
    for (int i = 0; i < m; i++) t[i] += (long double)a[i] * xj;
  }
  for (int i = 0; i < m; i++) y[i] = (double)t[i];
  free(t);
}

/* UE2UC_dgemv, matmgnt3d.cpp:125-155 (homogeneous branch, R = 1) */
static void ue2uc(orc_fmm* h, int l, const int* idx, const double* den, double* chk)
{
  int ci = idx[0] * 4 + idx[1] * 2 + idx[2];
  int M = h->pln[UC], N = h->pln[UE];
  if (!h->ue2uc[ci]) {
    double R = 1.0;
    double* chkpos = (double*)malloc(sizeof(double) * 3 * h->nsp[UC]);
    double* denpos = (double*)malloc(sizeof(double) * 3 * h->nsp[UE]);
    for (int j = 0; j < h->nsp[UC]; j++)
      for (int d = 0; d < 3; d++) chkpos[3 * j + d] = 2.0 * R * h->sp[UC][3 * j + d];
    for (int j = 0; j < h->nsp[UE]; j++)
      for (int d = 0; d < 3; d++) denpos[3 * j + d] = R * h->sp[UE][3 * j + d] + (2 * idx[d] - 1) * R;
    h->ue2uc[ci] = (double*)malloc(sizeof(double) * (size_t)M * N);
    double cf[2] = {h->mm.c[0], h->mm.c[1]};
    orc_kernel(h->mm.kt, cf, 2, h->nsp[UE], denpos, denpos, h->nsp[UC], chkpos, h->ue2uc[ci]);
    free(chkpos); free(denpos);
  }
  double* tmp = (double*)malloc(sizeof(double) * N);
  for (int i = 0, cnt = 0; i < h->nsp[UE]; i++)
    for (int s = 0; s < h->sm; s++, cnt++) tmp[cnt] = den[cnt] * pow(2.0, l * h->deg[s]);
  dgemv_acc(M, N, h->ue2uc[ci], tmp, chk);
  free(tmp);
}

/* UC2UE_dgemv / DC2DE_dgemv, matmgnt3d.cpp:91-123 and :158-188 (homogeneous) */
static void c2e(orc_fmm* h, const double* P, int nout, int nin, int npts, int l, const double* chk, double* den)
{
  double* tmp = (double*)calloc((size_t)nout, sizeof(double));
  dgemv_acc(nout, nin, P, chk, tmp);
  for (int i = 0, cnt = 0; i < npts; i++)
    for (int s = 0; s < h->sm; s++, cnt++) den[cnt] = den[cnt] + tmp[cnt] * pow(2.0, -l * h->deg[s]);
  free(tmp);
}

/* DE2DC_dgemv, matmgnt3d.cpp:190-221 (homogeneous branch, R = 1) */
static void de2dc(orc_fmm* h, int l, const int* idx, const double* den, double* chk)
{
  int ci = idx[0] * 4 + idx[1] * 2 + idx[2];
  int M = h->pln[DC], N = h->pln[DE];
  if (!h->de2dc[ci]) {
    double R = 1.0;
    double* denpos = (double*)malloc(sizeof(double) * 3 * h->nsp[DE]);
    double* chkpos = (double*)malloc(sizeof(double) * 3 * h->nsp[DC]);
    for (int j = 0; j < h->nsp[DE]; j++)
      for (int d = 0; d < 3; d++) denpos[3 * j + d] = R * h->sp[DE][3 * j + d];
    for (int j = 0; j < h->nsp[DC]; j++)
      for (int d = 0; d < 3; d++) chkpos[3 * j + d] = 0.5 * R * h->sp[DC][3 * j + d] + ((double)idx[d] - 0.5) * R;
    h->de2dc[ci] = (double*)malloc(sizeof(double) * (size_t)M * N);
    double cf[2] = {h->mm.c[0], h->mm.c[1]};
    orc_kernel(h->mm.kt, cf, 2, h->nsp[DE], denpos, denpos, h->nsp[DC], chkpos, h->de2dc[ci]);
    free(chkpos); free(denpos);
  }
  double* tmp = (double*)malloc(sizeof(double) * N);
  for (int i = 0, cnt = 0; i < h->nsp[DE]; i++)
    for (int s = 0; s < h->sm; s++, cnt++) tmp[cnt] = den[cnt] * pow(2.0, l * h->deg[s]);
  dgemv_acc(M, N, h->de2dc[ci], tmp, chk);
  free(tmp);
}

/* plnden2effden + splden2regden, matmgnt3d.cpp:223-264 */
static void plnden2effden(const orc_fmm* h, int l, const double* plnden, double* effden)
{
  int np = h->np, rg = 2 * np, sm = h->sm;
  double* regden = (double*)calloc((size_t)h->G * sm, sizeof(double));
  int cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
          int rgoff = (k + np / 2) * rg * rg + (j + np / 2) * rg + (i + np / 2);
          for (int f = 0; f < sm; f++) regden[sm * rgoff + f] += plnden[sm * cnt + f] * pow(2.0, l * h->deg[f]);
          cnt++;
        }
  orc_rfft3(rg, sm, regden, effden);
  free(regden);
}

/* effval2plnval + regval2splval, matmgnt3d.cpp:266-294 */
static void effval2plnval(const orc_fmm* h, const double* effval, double* plnval)
{
  int np = h->np, rg = 2 * np, tm = h->tm;
  double* regval = (double*)malloc(sizeof(double) * (size_t)h->G * tm);
  orc_irfft3(rg, tm, effval, regval);
  int cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
          int rgoff = (k + np / 2) * rg * rg + (j + np / 2) * rg + (i + np / 2);
          for (int f = 0; f < tm; f++) plnval[tm * cnt + f] += regval[tm * rgoff + f];
          cnt++;
        }
  free(regval);
}

/* UE2DC_dgemv + cptwvv, matmgnt3d.cpp:296-361, 422-432 (R = 1) */
static int ue2dc(orc_fmm* h, const int* idx, const double* effden, double* effval)
{
  int o = (idx[0] + 3) * 49 + (idx[1] + 3) * 7 + (idx[2] + 3);
  int sm = h->sm, tm = h->tm, G = h->G, C = h->C;
  int linf = 0;
  for (int d = 0; d < 3; d++) { int a = abs(idx[d]); if (a > 3) return ERR; if (a > linf) linf = a; }
  if (linf <= 1) return ERR; /* matmgnt3d.cpp:309 */
  if (!h->ue2dc[o]) {
    double R = 1.0;
    double denpos[3] = {idx[0] * 2.0 * R, idx[1] * 2.0 * R, idx[2] * 2.0 * R};
    double* tt = (double*)malloc(sizeof(double) * (size_t)G * tm * sm);
    double cf[2] = {h->mm.c[0], h->mm.c[1]};
    orc_kernel(h->mm.kt, cf, 2, 1, denpos, denpos, G, h->reg, tt); /* (G*tm) x sm */
    double* tmp = (double*)malloc(sizeof(double) * (size_t)G * tm * sm);
    for (int k = 0; k < G; k++)
      for (int i = 0; i < tm; i++)
        for (int j = 0; j < sm; j++) tmp[(size_t)(j + k * sm) * tm + i] = tt[(size_t)j * G * tm + (size_t)i + (size_t)k * tm];
    h->ue2dc[o] = (double*)malloc(sizeof(double) * 2 * (size_t)C * sm * tm);
    orc_rfft3(2 * h->np, sm * tm, tmp, h->ue2dc[o]);
    free(tt); free(tmp);
  }
  double nrm = 1.0 / (double)G;
  const double* mat = h->ue2dc[o];
  /* per output (q, i) the terms are added in ascending j, as the reference's
   * loop nest (i, j, q) does; q outermost walks the arrays contiguously */
  for (int q = 0; q < C; q++)
    for (int i = 0; i < tm; i++)
      for (int j = 0; j < sm; j++) {
        const double* x = mat + 2 * ((size_t)q * sm * tm + j * tm + i);
        const double* y = effden + 2 * ((size_t)q * sm + j);
        double* z = effval + 2 * ((size_t)q * tm + i);
        z[0] += nrm * (x[0] * y[0] - x[1] * y[1]);
        z[1] += nrm * (x[0] * y[1] + x[1] * y[0]);
      }
  return OK;
}

/* Exposed for the FFT-M2L vs dense-M2L test: val (plnDC) += M2L(idx) den (plnUE) at level l. */
int orc_m2l_fft(orc_fmm* h, int l, const int* idx, const double* den, double* val)
{
  double* ed = (double*)calloc(2 * (size_t)h->C * h->sm, sizeof(double));
  double* ev = (double*)calloc(2 * (size_t)h->C * h->tm, sizeof(double));
  plnden2effden(h, l, den, ed);
  int rc = ue2dc(h, idx, ed, ev);
  if (rc == OK) effval2plnval(h, ev, val);
  free(ed); free(ev);
  return rc;
}

/* ------------------------------------------------------------------ */
/* FMM3d::eval, fmm3d_eval.cpp:24-243 (one right-hand side)             */
/* Threads of the evaluation (OpenMP when the oracle is built with it).
 * Every output value is summed by one thread in the order of the sequential
 * restatement (per node, its list order), so the result does not depend on
 * the thread count; n <= 0 restores the OpenMP default. */
void orc_set_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
  (void)n;
#endif
}
int orc_get_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* The operators the sequential code builds lazily on first use (M2M, L2L and
 * the M2L spectra of the 316 offsets, matmgnt3d.cpp:130-138, 195-203, 307-325)
 * are built before the parallel stages, which then only read them. */
static void prebuild_operators(orc_fmm* h)
{
  double* z = (double*)calloc((size_t)h->pln[UC] + h->pln[DC] + h->pln[UE] + h->pln[DE] + 2 * (size_t)h->C * (h->sm + h->tm), sizeof(double));
  double* chk = z;
  double* den = chk + h->pln[UC] + h->pln[DC];
  double* ed = den + h->pln[UE] + h->pln[DE];
  double* ev = ed + 2 * (size_t)h->C * h->sm;
  for (int c = 0; c < 8; c++) {
    int idx[3] = {c >> 2, (c >> 1) & 1, c & 1};
    if (!h->ue2uc[c]) ue2uc(h, 0, idx, den, chk);
    if (!h->de2dc[c]) de2dc(h, 0, idx, den, chk);
  }
  for (int o = 0; o < 343; o++) {
    int idx[3] = {o / 49 - 3, (o / 7) % 7 - 3, o % 7 - 3};
    int linf = 0;
    for (int d = 0; d < 3; d++) if (abs(idx[d]) > linf) linf = abs(idx[d]);
    if (linf >= 2 && !h->ue2dc[o]) ue2dc(h, idx, ed, ev);
  }
  free(z);
}

#include <time.h>
static double wall_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
/* ORC_TRACE in the environment: per-stage wall times of each eval on stderr */
static void trace_stage(const char* what, double* t)
{
  static int on = -1;
  if (on < 0) on = getenv("ORC_TRACE") != NULL;
  double now = wall_s();
  if (on) fprintf(stderr, "[oracle] %-8s %8.3f s\n", what, now - *t);
  *t = now;
}

int orc_eval(orc_fmm* h, const double* srcden, double* trgval)
{
  if (!h->uc2ue || !h->dc2de || !h->nd) return ERR;
  double tt = wall_s();
  int sdof = h->sdof, tdof = h->tdof;
  int pUE = h->pln[UE], pUC = h->pln[UC], pDE = h->pln[DE], pDC = h->pln[DC];
  prebuild_operators(h);
  /* 1. zero (:33-42) */
  memset(trgval, 0, sizeof(double) * (size_t)h->nt * tdof);
  memset(h->sextden, 0, sizeof(double) * (size_t)h->sextcnt * sdof);
  memset(h->supeden, 0, sizeof(double) * (size_t)h->sndecnt * pUE);
  memset(h->supcval, 0, sizeof(double) * (size_t)h->sndecnt * pUC);
  memset(h->textval, 0, sizeof(double) * (size_t)h->textcnt * tdof);
  memset(h->tdneden, 0, sizeof(double) * (size_t)h->tndecnt * pDE);
  memset(h->tdncval, 0, sizeof(double) * (size_t)h->tndecnt * pDC);
  int nn = h->nnode;
  /* nodes are in BFS order: depth d occupies [dbeg[d], dbeg[d+1]) */
  int nd = 0;
  for (int g = 0; g < nn; g++) if (h->nd[g].depth + 1 > nd) nd = h->nd[g].depth + 1;
  int* dbeg = (int*)calloc((size_t)nd + 1, sizeof(int));
  for (int g = 0; g < nn; g++) dbeg[h->nd[g].depth + 1] = g + 1;
  for (int d = 1; d <= nd; d++) if (dbeg[d] < dbeg[d - 1]) dbeg[d] = dbeg[d - 1];
  /* 2. load (:50-64) */
  for (int g = nn - 1; g >= 0; g--) {
    onode* n = &h->nd[g];
    if ((n->tag & SRCNODE) && n->chd == -1)
      for (int k = 0; k < n->sown.n; k++)
        for (int d = 0; d < sdof; d++) h->sextden[(n->sextbeg + k) * sdof + d] = srcden[n->sown.a[k] * sdof + d];
  }
  /* 3. up (:69-92), deepest level first; the nodes of one level are independent */
  for (int dep = nd - 1; dep >= 2; dep--) {
#pragma omp parallel
    {
      double* pos = (double*)malloc(sizeof(double) * 3 * h->nsp[UC]);
#pragma omp for schedule(dynamic, 4)
      for (int g = dbeg[dep]; g < dbeg[dep + 1]; g++) {
        onode* n = &h->nd[g];
        if (!(n->tag & SRCNODE)) continue;
        double* cval = h->supcval + (size_t)n->sndeidx * pUC;
        double* eden = h->supeden + (size_t)n->sndeidx * pUE;
        if (n->chd == -1) {
          double c[3];
          center(h, g, c);
          locpos(h, UC, c, radius(h, g), pos);
          kapply(&h->knl, n->sextnum, h->sextpos + 3 * n->sextbeg, h->sextnor + 3 * n->sextbeg,
                 h->sextden + (size_t)n->sextbeg * sdof, h->nsp[UC], pos, cval);
        } else {
          for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
              for (int c = 0; c < 2; c++) {
                int chi = child(h, g, a, b, c);
                if (h->nd[chi].tag & SRCNODE) {
                  int idx[3] = {a, b, c};
                  ue2uc(h, h->nd[chi].depth + h->rootlvl, idx, h->supeden + (size_t)h->nd[chi].sndeidx * pUE, cval);
                }
              }
        }
        c2e(h, h->uc2ue, pUE, pUC, h->nsp[UE], n->depth + h->rootlvl, cval, eden);
      }
      free(pos);
    }
  }
  trace_stage("up", &tt);
  /* U (:98-111): each target leaf sums its U list in list order */
#pragma omp parallel for schedule(dynamic, 4)
  for (int g = 0; g < nn; g++) {
    onode* n = &h->nd[g];
    if (!(n->tag & TRGNODE) || n->chd != -1) continue;
    for (int q = 0; q < n->U.n; q++) {
      onode* s = &h->nd[n->U.a[q]];
      kapply(&h->knl, s->sextnum, h->sextpos + 3 * s->sextbeg, h->sextnor + 3 * s->sextbeg,
             h->sextden + (size_t)s->sextbeg * sdof, n->textnum, h->textpos + 3 * n->textbeg,
             h->textval + (size_t)n->textbeg * tdof);
    }
  }
  trace_stage("U", &tt);
  /* V (:115-150).  The reference transforms a source's density on first use
   * and a target's spectrum after its last V source; here targets are taken
   * in BFS chunks: the spectra their V lists need are formed first (a source
   * keeps its spectrum until its last V target is done), then each target
   * accumulates its products in its V-list order and is transformed back --
   * the same sums in the same order per target. */
  size_t effn = 2 * (size_t)h->C;
  int* vleft = (int*)malloc(sizeof(int) * (size_t)(nn ? nn : 1));
  int* need = (int*)malloc(sizeof(int) * (size_t)(nn ? nn : 1));
  memcpy(vleft, h->votnum, sizeof(int) * (size_t)nn);
  int verr = 0;
  const int CH = 1024;
  for (int c0 = 0; c0 < nn; c0 += CH) {
    int c1 = c0 + CH < nn ? c0 + CH : nn, nneed = 0;
    for (int g = c0; g < c1; g++) {
      onode* n = &h->nd[g];
      if (!(n->tag & TRGNODE)) continue;
      for (int q = 0; q < n->V.n; q++) {
        int vi = n->V.a[q];
        if (!h->effden[vi]) {
          h->effden[vi] = (double*)calloc(effn * h->sm, sizeof(double));
          need[nneed++] = vi;
        }
      }
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nneed; k++) {
      int vi = need[k];
      plnden2effden(h, h->nd[vi].depth + h->rootlvl, h->supeden + (size_t)h->nd[vi].sndeidx * pUE, h->effden[vi]);
    }
#pragma omp parallel for schedule(dynamic, 4) reduction(| : verr)
    for (int g = c0; g < c1; g++) {
      onode* n = &h->nd[g];
      if (!(n->tag & TRGNODE) || n->V.n == 0) continue;
      double gc[3];
      center(h, g, gc);
      double D = 2.0 * radius(h, g);
      double* ev = (double*)calloc(effn * h->tm, sizeof(double));
      for (int qq = 0; qq < n->V.n; qq++) {
        int q = (g_order & 4) ? n->V.n - 1 - qq : qq;
        int vi = n->V.a[q];
        double vc[3];
        center(h, vi, vc);
        int idx[3];
        for (int d = 0; d < 3; d++) idx[d] = (int)round((vc[d] - gc[d]) / D);
        if (ue2dc(h, idx, h->effden[vi], ev) != OK) verr = 1;
      }
      effval2plnval(h, ev, h->tdncval + (size_t)n->tndeidx * pDC);
      free(ev);
    }
    for (int g = c0; g < c1; g++) {
      onode* n = &h->nd[g];
      if (!(n->tag & TRGNODE)) continue;
      for (int q = 0; q < n->V.n; q++) {
        int vi = n->V.a[q];
        if (--vleft[vi] == 0) { free(h->effden[vi]); h->effden[vi] = NULL; }
      }
    }
  }
  for (int g = 0; g < nn; g++)  /* none is left, but no spectrum may outlive the eval */
    if (h->effden[g]) { free(h->effden[g]); h->effden[g] = NULL; }
  free(vleft); free(need);
  if (verr) { free(dbeg); return ERR; }
  trace_stage("V", &tt);
  /* W (:154-172) */
#pragma omp parallel
  {
    double* uepos = (double*)malloc(sizeof(double) * 3 * h->nsp[UE]);
#pragma omp for schedule(dynamic, 4)
    for (int g = 0; g < nn; g++) {
      onode* n = &h->nd[g];
      if (!(n->tag & TRGNODE) || n->chd != -1) continue;
      for (int q = 0; q < n->W.n; q++) {
        int vi = n->W.a[q];
        onode* s = &h->nd[vi];
        if (s->chd == -1 && s->sextnum * sdof < pUE) {
          kapply(&h->knl, s->sextnum, h->sextpos + 3 * s->sextbeg, h->sextnor + 3 * s->sextbeg,
                 h->sextden + (size_t)s->sextbeg * sdof, n->textnum, h->textpos + 3 * n->textbeg,
                 h->textval + (size_t)n->textbeg * tdof);
        } else {
          double c[3];
          center(h, vi, c);
          locpos(h, UE, c, radius(h, vi), uepos);
          kapply(&h->mm, h->nsp[UE], uepos, uepos, h->supeden + (size_t)s->sndeidx * pUE, n->textnum,
                 h->textpos + 3 * n->textbeg, h->textval + (size_t)n->textbeg * tdof);
        }
      }
    }
    free(uepos);
  }
  trace_stage("W", &tt);
  /* X (:176-191) */
#pragma omp parallel
  {
    double* pos = (double*)malloc(sizeof(double) * 3 * h->nsp[DC]);
#pragma omp for schedule(dynamic, 4)
    for (int g = 0; g < nn; g++) {
      onode* n = &h->nd[g];
      if (!(n->tag & TRGNODE)) continue;
      for (int q = 0; q < n->X.n; q++) {
        onode* s = &h->nd[n->X.a[q]];
        if (n->chd == -1 && n->textnum * tdof < pDC) {
          kapply(&h->knl, s->sextnum, h->sextpos + 3 * s->sextbeg, h->sextnor + 3 * s->sextbeg,
                 h->sextden + (size_t)s->sextbeg * sdof, n->textnum, h->textpos + 3 * n->textbeg,
                 h->textval + (size_t)n->textbeg * tdof);
        } else {
          double c[3];
          center(h, g, c);
          locpos(h, DC, c, radius(h, g), pos);
          kapply(&h->knl, s->sextnum, h->sextpos + 3 * s->sextbeg, h->sextnor + 3 * s->sextbeg,
                 h->sextden + (size_t)s->sextbeg * sdof, h->nsp[DC], pos, h->tdncval + (size_t)n->tndeidx * pDC);
        }
      }
    }
    free(pos);
  }
  trace_stage("X", &tt);
  /* combine (:195-216), top level first; the nodes of one level are independent */
  for (int dep = 0; dep < nd; dep++) {
#pragma omp parallel
    {
      double* depos = (double*)malloc(sizeof(double) * 3 * h->nsp[DE]);
#pragma omp for schedule(dynamic, 4)
      for (int g = dbeg[dep]; g < dbeg[dep + 1]; g++) {
        onode* n = &h->nd[g];
        if (!(n->tag & TRGNODE)) continue;
        if (n->depth >= 3) {
          int par = n->par;
          int idx[3];
          for (int d = 0; d < 3; d++) idx[d] = n->path[d] - 2 * h->nd[par].path[d];
          de2dc(h, h->nd[par].depth + h->rootlvl, idx, h->tdneden + (size_t)h->nd[par].tndeidx * pDE,
                h->tdncval + (size_t)n->tndeidx * pDC);
        }
        if (n->depth >= 2)
          c2e(h, h->dc2de, pDE, pDC, h->nsp[DE], n->depth + h->rootlvl, h->tdncval + (size_t)n->tndeidx * pDC,
              h->tdneden + (size_t)n->tndeidx * pDE);
        if (n->chd == -1) {
          double c[3];
          center(h, g, c);
          locpos(h, DE, c, radius(h, g), depos);
          kapply(&h->mm, h->nsp[DE], depos, depos, h->tdneden + (size_t)n->tndeidx * pDE, n->textnum,
                 h->textpos + 3 * n->textbeg, h->textval + (size_t)n->textbeg * tdof);
        }
      }
      free(depos);
    }
  }
  trace_stage("down", &tt);
  free(dbeg);
  /* save (:221-235) */
  for (int g = 0; g < nn; g++) {
    onode* n = &h->nd[g];
    if ((n->tag & TRGNODE) && n->chd == -1)
      for (int k = 0; k < n->town.n; k++)
        for (int d = 0; d < tdof; d++) trgval[n->town.a[k] * tdof + d] = h->textval[(n->textbeg + k) * tdof + d];
  }
  return OK;
}

/* Intermediate per-node buffers of the last eval (for stage-by-stage checks):
 * which = 0 supeden, 1 supcval, 2 tdneden, 3 tdncval, 4 textval. Returns the
 * element count; out may be NULL. */
long orc_buffer(const orc_fmm* h, int which, double* out)
{
  const double* p = NULL;
  long n = 0;
  switch (which) {
  case 0: p = h->supeden; n = (long)h->sndecnt * h->pln[UE]; break;
  case 1: p = h->supcval; n = (long)h->sndecnt * h->pln[UC]; break;
  case 2: p = h->tdneden; n = (long)h->tndecnt * h->pln[DE]; break;
  case 3: p = h->tdncval; n = (long)h->tndecnt * h->pln[DC]; break;
  case 4: p = h->textval; n = (long)h->textcnt * h->tdof; break;
  default: return ERR;
  }
  if (out && p) memcpy(out, p, sizeof(double) * (size_t)n);
  return n;
}
