// device.hip -- gfx950 kernels of the FP64 evaluate pass.
//
// Thread mapping is written for 64-wide wavefronts: blocks of 128/256
// threads, sources or surface points staged in LDS and read as broadcasts,
// one target (or surface point, or frequency) per lane.  Kernel closed forms
// follow fmm3d/kernel3d.cpp (111: :124-138, 301: :160-184, 321: :222-242,
// 511: :315-337) including the r < 1e-8 exclusion (:21).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdlib>
#include <mutex>
#include <type_traits>
#include <utility>

#include "device.hpp"

namespace kifmm {
namespace dev {

// ---------------------------------------------------------------------------
// kernels: accumulate u[q][t] += sum_s K_ts(d, n) f[q*S + s]
//
// 1/r comes from the hardware estimate v_rsq_f64 refined by one third-order
// step (about 1 ulp), which replaces the reference's sqrt and divisions; the
// r < 1e-8 exclusion (kernel3d.cpp:21) is applied as r^2 < 1e-16 by zeroing
// 1/r.  The few-ulp differences this makes in single kernel values are far
// below the parity bar (checked with the oracle: 1-ulp kernel perturbations
// move the result by < 1e-14 at np 6).
//
// The estimate y0 has relative error below 2^-24 on gfx950 (measured over 4M
// inputs spanning 120 binades); with e = 1 - x y0^2 the step
// y = y0 (1 + e/2 + 3e^2/8) leaves a truncation error of (5/16) e^3 < 2^-70.
__device__ inline double rsqrt_nr(double x)
{
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  double p = fma(e, 0.375, 0.5);
  return fma(y * e, p, y);
}

// 1/r from r^2, or 0 inside the exclusion radius: every kernel term carries a
// factor of 1/r, so a zero here drops the pair without a branch.
__device__ inline double rinv(double r2)
{
  double ri = rsqrt_nr(r2);
  return r2 < 1e-16 ? 0.0 : ri;
}

// r.n in the reference's order, (x nx + y ny) + z nz (kernel3d.cpp:150, :235,
// :257, :353), with no fused multiply-add.  For a target on the source's own
// smooth surface the displacement is nearly tangent, r.n ~ |r|^2 / (2 R), so
// the three products cancel to about 1 part in 2 R / |r|; a fused evaluation
// rounds differently by up to that factor (1e-11 of the pair's value for a
// 1e-5 neighbour gap on the unit sphere), which is the largest single
// difference from the reference's own FP64 sums at full problem sizes.
__device__ inline double rdotn(double dx, double dy, double dz, const double* n)
{
#pragma clang fp contract(off)
  return (dx * n[0] + dy * n[1]) + dz * n[2];
}

// FP32 forms (mixed-precision near field of the vector kernels): 1/r from
// v_rsq_f32 with the same exclusion, r.n unfused.
__device__ inline float rinv(float r2) { return r2 < 1e-16f ? 0.0f : __builtin_amdgcn_rsqf(r2); }
__device__ inline float rdotn(float dx, float dy, float dz, const float* n)
{
#pragma clang fp contract(off)
  return (dx * n[0] + dy * n[1]) + dz * n[2];
}

// the normal pointer may be null (kernels without normals): F comes from the
// coordinates
template <class F>
struct NoDeduceT {
  using type = F;
};
template <class F>
using NoDeduce = typename NoDeduceT<F>::type;

template <int KT>
struct Knl;

template <>
struct Knl<111> {
  static constexpr int S = 1, T = 1;
  static constexpr bool NRM = false;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>*, const F* f, F (&u)[NR][T], const KC& k)
  {
    F g = (F)k.a * rinv(dx * dx + dy * dy + dz * dz);
#pragma unroll
    for (int q = 0; q < NR; q++) u[q][0] += g * f[q];
  }
};

template <>
struct Knl<301> {
  static constexpr int S = 4, T = 3;
  static constexpr bool NRM = false;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>*, const F* f, F (&u)[NR][T], const KC& k)
  {
    F ri = rinv(dx * dx + dy * dy + dz * dz), ri3 = ri * ri * ri;
    F G = (F)k.a * ri, H = (F)k.a * ri3, A = (F)k.b * ri3;
#pragma unroll
    for (int q = 0; q < NR; q++) {
      const F* g = f + 4 * q;
      F df = dx * g[0] + dy * g[1] + dz * g[2];
      F h = H * df + A * g[3];
      u[q][0] += G * g[0] + h * dx;
      u[q][1] += G * g[1] + h * dy;
      u[q][2] += G * g[2] + h * dz;
    }
  }
};

template <>
struct Knl<321> {
  static constexpr int S = 3, T = 3;
  static constexpr bool NRM = true;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>* n, const F* f, F (&u)[NR][T], const KC& k)
  {
    F rn = rdotn(dx, dy, dz, n);
    F ri = rinv(dx * dx + dy * dy + dz * dz), ri2 = ri * ri;
    F c = ((F)k.a * (ri2 * ri2 * ri)) * rn;
#pragma unroll
    for (int q = 0; q < NR; q++) {
      const F* g = f + 3 * q;
      F w = c * (dx * g[0] + dy * g[1] + dz * g[2]);
      u[q][0] += w * dx;
      u[q][1] += w * dy;
      u[q][2] += w * dz;
    }
  }
};

template <>
struct Knl<511> {
  static constexpr int S = 3, T = 3;
  static constexpr bool NRM = false;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>*, const F* f, F (&u)[NR][T], const KC& k)
  {
    F ri = rinv(dx * dx + dy * dy + dz * dz);
    F G = (F)k.a * ri, H = (F)k.b * (ri * ri * ri);
#pragma unroll
    for (int q = 0; q < NR; q++) {
      const F* g = f + 3 * q;
      F h = H * (dx * g[0] + dy * g[1] + dz * g[2]);
      u[q][0] += (F)k.c * (G * g[0] + h * dx);
      u[q][1] += (F)k.c * (G * g[1] + h * dy);
      u[q][2] += (F)k.c * (G * g[2] + h * dz);
    }
  }
};

// kernel3d.cpp:139-154: -(1/4 pi) (r.n) / r^3
template <>
struct Knl<121> {
  static constexpr int S = 1, T = 1;
  static constexpr bool NRM = true;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>* n, const F* f, F (&u)[NR][T], const KC& k)
  {
    F ri = rinv(dx * dx + dy * dy + dz * dz);
    F g = (F)k.a * (ri * ri * ri) * rdotn(dx, dy, dz, n);
#pragma unroll
    for (int q = 0; q < NR; q++) u[q][0] += g * f[q];
  }
};

// kernel3d.cpp:185-203: (1/mu)(1/8 pi) (I / r + r r^T / r^3)
template <>
struct Knl<311> {
  static constexpr int S = 3, T = 3;
  static constexpr bool NRM = false;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>*, const F* f, F (&u)[NR][T], const KC& k)
  {
    F ri = rinv(dx * dx + dy * dy + dz * dz);
    F G = (F)k.a * ri, H = (F)k.a * (ri * ri * ri);
#pragma unroll
    for (int q = 0; q < NR; q++) {
      const F* g = f + 3 * q;
      F h = H * (dx * g[0] + dy * g[1] + dz * g[2]);
      u[q][0] += G * g[0] + h * dx;
      u[q][1] += G * g[1] + h * dy;
      u[q][2] += G * g[2] + h * dz;
    }
  }
};

// kernel3d.cpp:204-221: (1/4 pi) r^T / r^3
template <>
struct Knl<312> {
  static constexpr int S = 3, T = 1;
  static constexpr bool NRM = false;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>*, const F* f, F (&u)[NR][T], const KC& k)
  {
    F ri = rinv(dx * dx + dy * dy + dz * dz);
    F H = (F)k.a * (ri * ri * ri);
#pragma unroll
    for (int q = 0; q < NR; q++) {
      const F* g = f + 3 * q;
      u[q][0] += H * (dx * g[0] + dy * g[1] + dz * g[2]);
    }
  }
};

// kernel3d.cpp:243-265: (mu/2 pi) (n^T / r^3 - 3 (r.n) r^T / r^5)
template <>
struct Knl<322> {
  static constexpr int S = 3, T = 1;
  static constexpr bool NRM = true;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>* n, const F* f, F (&u)[NR][T], const KC& k)
  {
    F ri = rinv(dx * dx + dy * dy + dz * dz), ri2 = ri * ri, ri3 = ri2 * ri;
    F rn = rdotn(dx, dy, dz, n);
    F A = (F)k.a * ri3, B = (F)-3 * (F)k.a * (ri3 * ri2) * rn;
#pragma unroll
    for (int q = 0; q < NR; q++) {
      const F* g = f + 3 * q;
      u[q][0] += A * (n[0] * g[0] + n[1] * g[1] + n[2] * g[2]) + B * (dx * g[0] + dy * g[1] + dz * g[2]);
    }
  }
};

// kernel3d.cpp:338-369: A ((r.n) I + n r^T) + B r n^T + C (r.n) r r^T with
// A = dc1 / r^3, B = dc2 / r^3, C = dc3 / r^5
template <>
struct Knl<521> {
  static constexpr int S = 3, T = 3;
  static constexpr bool NRM = true;
  template <int NR, class F>
  __device__ static inline void add(F dx, F dy, F dz, const NoDeduce<F>* n, const F* f, F (&u)[NR][T], const KC& k)
  {
    F ri = rinv(dx * dx + dy * dy + dz * dz), ri2 = ri * ri, ri3 = ri2 * ri;
    F rn = rdotn(dx, dy, dz, n);
    F A = (F)k.a * ri3, B = (F)k.b * ri3, C = (F)k.c * (ri3 * ri2) * rn;
#pragma unroll
    for (int q = 0; q < NR; q++) {
      const F* g = f + 3 * q;
      F df = dx * g[0] + dy * g[1] + dz * g[2], nf = n[0] * g[0] + n[1] * g[1] + n[2] * g[2];
      F ar = A * rn, an = A * df, bt = B * nf + C * df;
      u[q][0] += ar * g[0] + an * n[0] + bt * dx;
      u[q][1] += ar * g[1] + an * n[1] + bt * dy;
      u[q][2] += ar * g[2] + an * n[2] + bt * dz;
    }
  }
};

// ---------------------------------------------------------------------------
// permutations
// One caller-order point per thread: its densities are read across the
// right-hand sides (coalesced over the threads for each one) and written as
// the point's contiguous destination row (a full line for 16 RHS), instead of
// gathering 8 bytes from a different line per (point, RHS).
__global__ void k_gather(int n, int dof, int nr, const int* __restrict__ iperm, const double* __restrict__ src,
                         long src_rs, double* __restrict__ dst, float* __restrict__ dst32)
{
  const int per = nr * dof;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double* out = dst ? dst + (long)iperm[i] * per : nullptr;
    float* out32 = dst32 ? dst32 + (long)iperm[i] * per : nullptr;
    const double* in = src + i * dof;
    for (int e0 = 0; e0 < per; e0 += 8) {  // 8 loads in flight, then 8 stores
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int e = min(e0 + j, per - 1), r = e / dof;
        v[j] = in[r * src_rs + (e - r * dof)];
      }
      if (!out) {
      } else if ((per & 1) == 0 && e0 + 8 <= per) {  // 16-byte stores (rows of even length are 16-byte aligned)
#pragma unroll
        for (int j = 0; j < 8; j += 2) *(double2*)(out + e0 + j) = make_double2(v[j], v[j + 1]);
      } else {
#pragma unroll
        for (int j = 0; j < 8; j++)
          if (e0 + j < per) out[e0 + j] = v[j];
      }
      if (out32) {
        if ((per & 3) == 0 && e0 + 8 <= per) {  // rows of a multiple of 4 are 16-byte aligned
          *(float4*)(out32 + e0) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
          *(float4*)(out32 + e0 + 4) = make_float4((float)v[4], (float)v[5], (float)v[6], (float)v[7]);
        } else {
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (e0 + j < per) out32[e0 + j] = (float)v[j];
        }
      }
    }
  }
}

// Target values leave in two passes so that both sides of the permutation
// move whole lines: k_rows_to_slots transposes the leaf-order values
// [r][slot*dof + d] into slot-major rows tmp[slot][r*dof + d] through LDS
// (64 slots x 16 right-hand sides per block), then k_slots_to_caller takes one
// caller-order point per thread, reads its slot's row (128 bytes for 16 RHS)
// and writes the values across the right-hand sides, coalesced over the
// threads for each one.  The one-pass form writes 8 bytes to a different line
// per (point, RHS).
constexpr int ROWS_TS = 64;

__global__ __launch_bounds__(256) void k_rows_to_slots(int n, int dof, int nr, const double* __restrict__ src,
                                                       long src_rs, double* __restrict__ tmp)
{
  __shared__ double s_v[16][ROWS_TS * 3 + 1];  // dof <= 3
  const int k0 = blockIdx.x * ROWS_TS, r0 = blockIdx.y * 16;
  const int nk = min(ROWS_TS, n - k0), nrr = min(16, nr - r0);
  const int len = nk * dof;
  for (int e = threadIdx.x; e < nrr * len; e += 256) {
    const int rr = e / len, j = e - rr * len;
    s_v[rr][j] = src[(long)(r0 + rr) * src_rs + (long)k0 * dof + j];
  }
  __syncthreads();
  const int per = nrr * dof;
  for (int e = threadIdx.x; e < nk * per; e += 256) {
    const int kk = e / per, q = e - kk * per, rr = q / dof, d = q - rr * dof;
    tmp[((long)(k0 + kk) * nr + r0 + rr) * dof + d] = s_v[rr][kk * dof + d];
  }
}

__global__ void k_slots_to_caller(int n, int dof, int nr, const int* __restrict__ iperm, const double* __restrict__ tmp,
                                  double* __restrict__ dst, long dst_rs)
{
  const int per = nr * dof;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double* row = tmp + (long)iperm[i] * per;
    for (int e0 = 0; e0 < per; e0 += 8) {  // 8 loads in flight, then 8 stores
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) v[j] = row[min(e0 + j, per - 1)];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int e = e0 + j, r = e / dof;
        if (e < per) dst[r * dst_rs + i * dof + (e - r * dof)] = v[j];
      }
    }
  }
}

__global__ void k_scatter(int n, int dof, int nr, const int* __restrict__ perm, const double* __restrict__ src,
                          long src_rs, double* __restrict__ dst, long dst_rs)
{
  long tot = (long)n * dof * nr;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < tot; i += (long)gridDim.x * blockDim.x) {
    long per = (long)n * dof;
    int r = (int)(i / per);
    long rem = i - r * per;
    int k = (int)(rem / dof), d = (int)(rem - (long)k * dof);
    dst[r * dst_rs + (long)perm[k] * dof + d] = src[r * src_rs + rem];
  }
}

static int grid_for(long tot, int bs)
{
  long g = (tot + bs - 1) / bs;
  if (g > 65536L * 8) g = 65536L * 8;
  return (int)(g > 0 ? g : 1);
}

// ---------------------------------------------------------------------------
// Scratch array of the no-allocation evaluate.  In host code `scratch` names a
// host-side stand-in, not device memory: kernels and copies get the address
// that hipGetSymbolAddress returns for the current device (each GPU holds its
// own copy at its own address), kept per device.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_MAXDEV = 64;
std::mutex scratch_mu;
void* scratch_dev[SCRATCH_MAXDEV];
}  // namespace

hipError_t scratch_ptr(void** out)
{
  int d = 0;
  hipError_t e = hipGetDevice(&d);
  if (e != hipSuccess) return e;
  if (d < 0 || d >= SCRATCH_MAXDEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[d]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    scratch_dev[d] = p;
  }
  *out = scratch_dev[d];
  return hipSuccess;
}

// Zero fill by kernel (also inside captured graphs, where memset nodes are avoided)
__global__ void k_zero(long n, double* __restrict__ p)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

void zero(hipStream_t s, double* p, long n)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n, 256)), dim3(256), 0, s, n, p);
}

__global__ void k_zero_rows(long rowlen, const int* __restrict__ rows, int nrows, double* __restrict__ p)
{
  const long n = (long)nrows * rowlen;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long q = i / rowlen;
    p[(long)rows[q] * rowlen + (i - q * rowlen)] = 0.0;
  }
}

void zero_rows(hipStream_t s, double* p, const int* rows, int nrows, long rowlen)
{
  const long n = (long)nrows * rowlen;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_rows, dim3(grid_for(n, 256)), dim3(256), 0, s, rowlen, rows, nrows, p);
}

// Streaming copy of n doubles (reads n*8 bytes, writes n*8 bytes): the
// reference kernel for scaling memory-traffic counters.
__global__ void k_copy(long n, const double* __restrict__ src, double* __restrict__ dst)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) dst[i] = src[i];
}

void copy(hipStream_t s, const double* src, double* dst, long n)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256)), dim3(256), 0, s, n, src, dst);
}

void gather(hipStream_t s, int n, int dof, int nr, const int* iperm, const double* src, long src_rs, double* dst,
            float* dst32)
{
  if ((long)n * dof * nr == 0) return;
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256)), dim3(256), 0, s, n, dof, nr, iperm, src, src_rs, dst, dst32);
}

void scatter(hipStream_t s, int n, int dof, int nr, const int* perm, const double* src, long src_rs, double* dst,
             long dst_rs)
{
  long tot = (long)n * dof * nr;
  if (tot == 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(tot, 256)), dim3(256), 0, s, n, dof, nr, perm, src, src_rs, dst, dst_rs);
}

void scatter_rows(hipStream_t s, int n, int dof, int nr, const int* iperm, const double* src, long src_rs,
                  double* tmp, double* dst, long dst_rs)
{
  if ((long)n * dof * nr == 0 || dof > 3) return;  // dof <= 3: the caller checks
  hipLaunchKernelGGL(k_rows_to_slots, dim3((unsigned)((n + ROWS_TS - 1) / ROWS_TS), (unsigned)((nr + 15) / 16)),
                     dim3(256), 0, s, n, dof, nr, src, src_rs, tmp);
  hipLaunchKernelGGL(k_slots_to_caller, dim3(grid_for(n, 256)), dim3(256), 0, s, n, dof, nr, iperm, tmp, dst, dst_rs);
}

// ---------------------------------------------------------------------------
// near field (U list and direct W/X)
constexpr int P2P_BS = 128;
constexpr int TILE = 64;

// The generic near-field kernels (k_p2p, k_p2p32) give one lane to each
// target, except in a wave that holds at most 32 (16) of its chunk's
// targets: there each target gets 2 (4) lanes, which take every second
// (fourth) source, and sum() adds their partial sums into the first of them
// (lanes l + 32 and l + 16 into lane l, in that order).  Near field per
// evaluate: 1M Navier clusters 5.4 -> 4.5 ms (FP64), 2.8 -> 2.4 ms (FP32);
// 262k Stokes sphere 0.95 -> 0.72 ms.  (The L2T / M2T kernel k_s2p measured
// no gain from it.)  ON false (16-RHS blocks, whose register budget the
// shuffled sums would overrun): one lane per target throughout.
template <bool ON>
struct LaneSplit {
  int nsub, sub, ti;  // lanes per target, this lane's share, its target in the chunk
  __device__ LaneSplit(int t, int tn)
  {
    const int w = t >> 6, l = t & 63;
    const int mw = min(64, max(0, tn - 64 * w));  // the wave's targets (wave-uniform)
    nsub = !ON ? 1 : mw <= 16 ? 4 : mw <= 32 ? 2 : 1;
    const int lpt = 64 / nsub;
    ti = 64 * w + (l & (lpt - 1));
    sub = l / lpt;
  }
  template <int A, int B>
  __device__ void sum(double (&u)[A][B]) const
  {
    if (!ON || nsub == 1) return;  // wave-uniform
#pragma unroll
    for (int a = 0; a < A; a++)
#pragma unroll
      for (int b = 0; b < B; b++) {
        double v = u[a][b] + __shfl_down(u[a][b], 32);
        if (nsub == 4) v += __shfl_down(v, 16);
        u[a][b] = v;
      }
  }
};

// Blocks are dealt to the 8 XCDs round-robin by block index; handing each XCD
// a contiguous range of chunks instead keeps the source leaves that
// neighbouring chunks share in one L2.  Grid x is 8 * ceil(n / 8).
__device__ inline int xcd_item(int b, int n)
{
  const int per = (n + 7) / 8;
  return (b & 7) * per + (b >> 3);
}
static inline unsigned xcd_grid(int n) { return 8u * (unsigned)((n + 7) / 8); }

// Densities of sources j0 .. j0+cnt for NR right-hand sides into LDS,
// s_den[jj][q*S + d]; den points at right-hand side r0 of the block (FP64,
// or the FP32 copy).
template <int S, int NR, int BS, class E, class D>
__device__ inline void stage_den(E (*s_den)[NR * S], const D* den, const Points& src, int j0, int cnt, int t)
{
  constexpr int ND = NR * S;
  for (int i = t; i < cnt * ND; i += BS) {
    int jj = i / ND, rem = i - jj * ND;
    int q = rem / S, d = rem - q * S;
    s_den[jj][rem] = (E)den[(long)(j0 + jj) * src.den_js + (long)q * src.den_rs + d];
  }
}

template <int KT, int NR>
__global__ __launch_bounds__(P2P_BS) void k_p2p(const int* __restrict__ tbeg, const int* __restrict__ tnum,
                                                const int* __restrict__ eptr, const int2* __restrict__ ent, Points src,
                                                Targets trg, int roff, KC kc)
{
  using K = Knl<KT>;
  constexpr int NPOS = K::NRM ? 6 : 3;
  constexpr int ND = NR * K::S;
  __shared__ double s_pos[TILE][NPOS];
  __shared__ double s_den[TILE][ND];
  const int c = blockIdx.x, t = threadIdx.x;
  const int r0 = roff + blockIdx.y * NR;
  const int tb = tbeg[c], tn = tnum[c];
  const LaneSplit<(NR < 16)> ls(t, tn);
  const bool act = ls.ti < tn;
  double x = 0, y = 0, z = 0;
  if (act) {
    x = trg.x[tb + ls.ti];
    y = trg.y[tb + ls.ti];
    z = trg.z[tb + ls.ti];
  }
  double u[NR][K::T];
#pragma unroll
  for (int q = 0; q < NR; q++)
#pragma unroll
    for (int d = 0; d < K::T; d++) u[q][d] = 0.0;
  // FP64 densities, or (src.den null: the mixed form's pass over the entries
  // its FP32 kernels skip) the FP32 copy
  const double* den = src.den ? src.den + (long)r0 * src.den_rs : nullptr;
  const float* den32 = src.den ? nullptr : src.den32 + (long)r0 * src.den_rs;
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    const int2 se = ent[e];
    for (int b0 = 0; b0 < se.y; b0 += TILE) {
      const int cnt = min(TILE, se.y - b0);
      const int j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt; i += P2P_BS) {
        s_pos[i][0] = src.x[j0 + i];
        s_pos[i][1] = src.y[j0 + i];
        s_pos[i][2] = src.z[j0 + i];
        if (K::NRM) {
          s_pos[i][NPOS - 3] = src.nx[j0 + i];
          s_pos[i][NPOS - 2] = src.ny[j0 + i];
          s_pos[i][NPOS - 1] = src.nz[j0 + i];
        }
      }
      if (den)
        stage_den<K::S, NR, P2P_BS>(s_den, den, src, j0, cnt, t);
      else
        stage_den<K::S, NR, P2P_BS>(s_den, den32, src, j0, cnt, t);
      __syncthreads();
      if (act)
        for (int jj = ls.sub; jj < cnt; jj += ls.nsub)
          K::template add<NR>(x - s_pos[jj][0], y - s_pos[jj][1], z - s_pos[jj][2], &s_pos[jj][NPOS - 3], s_den[jj], u,
                              kc);
    }
  }
  ls.sum(u);
  if (act && ls.sub == 0) {
    double* val = trg.val + (long)r0 * trg.val_rs + (long)(tb + ls.ti) * K::T;
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int d = 0; d < K::T; d++) val[(long)q * trg.val_rs + d] += u[q][d];
  }
}

// Mixed-precision form of k_p2p for the kernels without a matrix-core form:
// coordinates relative to the chunk's first target computed in FP64 and
// rounded, normals and densities rounded, the kernel evaluated in FP32 and the
// sums of each staged tile (at most TILE sources) added in FP64.  Entries
// flagged 2 in echk (a pair may be closer than the relative FP32 coordinates
// resolve, so the r < 1e-8 rule needs the FP64 positions) are skipped: a
// separate FP64 pass over those entries adds them.
template <int KT, int NR>
__global__ __launch_bounds__(P2P_BS) void k_p2p32(const int* __restrict__ tbeg, const int* __restrict__ tnum,
                                                  const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                  const int* __restrict__ echk, Points src, Targets trg, int roff,
                                                  KC kc)
{
  using K = Knl<KT>;
  constexpr int NPOS = K::NRM ? 6 : 3;
  constexpr int ND = NR * K::S;
  __shared__ float s_pos[TILE][NPOS];
  __shared__ float s_den[TILE][ND];
  const int c = blockIdx.x, t = threadIdx.x;
  const int r0 = roff + blockIdx.y * NR;
  const int tb = tbeg[c], tn = tnum[c];
  const LaneSplit<(NR < 16)> ls(t, tn);
  const bool act = ls.ti < tn;
  const int tb0 = tn > 0 ? tb : 0;  // origin of the relative coordinates
  const double ox = trg.x[tb0], oy = trg.y[tb0], oz = trg.z[tb0];
  float x = 0, y = 0, z = 0;
  if (act) {
    x = (float)(trg.x[tb + ls.ti] - ox);
    y = (float)(trg.y[tb + ls.ti] - oy);
    z = (float)(trg.z[tb + ls.ti] - oz);
  }
  double u[NR][K::T];
#pragma unroll
  for (int q = 0; q < NR; q++)
#pragma unroll
    for (int d = 0; d < K::T; d++) u[q][d] = 0.0;
  const double* den = src.den + (long)r0 * src.den_rs;
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    if (echk && echk[e] == 2) continue;  // block-uniform: the FP64 pass's
    const int2 se = ent[e];
    for (int b0 = 0; b0 < se.y; b0 += TILE) {
      const int cnt = min(TILE, se.y - b0);
      const int j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt; i += P2P_BS) {
        s_pos[i][0] = (float)(src.x[j0 + i] - ox);
        s_pos[i][1] = (float)(src.y[j0 + i] - oy);
        s_pos[i][2] = (float)(src.z[j0 + i] - oz);
        if (K::NRM) {
          s_pos[i][NPOS - 3] = (float)src.nx[j0 + i];
          s_pos[i][NPOS - 2] = (float)src.ny[j0 + i];
          s_pos[i][NPOS - 1] = (float)src.nz[j0 + i];
        }
      }
      stage_den<K::S, NR, P2P_BS>(s_den, den, src, j0, cnt, t);
      __syncthreads();
      if (act) {
        float p[NR][K::T];
#pragma unroll
        for (int q = 0; q < NR; q++)
#pragma unroll
          for (int d = 0; d < K::T; d++) p[q][d] = 0.0f;
        for (int jj = ls.sub; jj < cnt; jj += ls.nsub)
          K::template add<NR>(x - s_pos[jj][0], y - s_pos[jj][1], z - s_pos[jj][2], &s_pos[jj][NPOS - 3], s_den[jj],
                              p, kc);
#pragma unroll
        for (int q = 0; q < NR; q++)
#pragma unroll
          for (int d = 0; d < K::T; d++) u[q][d] += (double)p[q][d];
      }
    }
  }
  ls.sum(u);
  if (act && ls.sub == 0) {
    double* val = trg.val + (long)r0 * trg.val_rs + (long)(tb + ls.ti) * K::T;
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int d = 0; d < K::T; d++) val[(long)q * trg.val_rs + d] += u[q][d];
  }
}

// Near field with the sources read through the scalar data cache: one wave
// per half chunk (64 targets), the wave walks the chunk's source leaves with
// wave-uniform addresses, so positions and the contiguous NR*S densities of a
// source arrive in scalar registers and feed the FMAs directly (no LDS, no
// barriers).  Requires src.den_rs == S.
template <int KT, int NR>
__global__ __launch_bounds__(64) void k_p2p_s(const int* __restrict__ tbeg, const int* __restrict__ tnum,
                                              const int* __restrict__ eptr, const int2* __restrict__ ent, Points src,
                                              Targets trg, int roff, KC kc)
{
  using K = Knl<KT>;
  const int c = blockIdx.x >> 1, h = (blockIdx.x & 1) * 64, t = threadIdx.x;
  const int r0 = roff + blockIdx.y * NR;
  const int tb = tbeg[c] + h, tn = tnum[c] - h;
  if (tn <= 0) return;
  const bool act = t < tn;
  const int it = act ? t : 0;
  const double x = trg.x[tb + it], y = trg.y[tb + it], z = trg.z[tb + it];
  double u[NR][K::T];
#pragma unroll
  for (int q = 0; q < NR; q++)
#pragma unroll
    for (int d = 0; d < K::T; d++) u[q][d] = 0.0;
  const double* __restrict__ den = src.den + (long)r0 * K::S;
  const long js = src.den_js;
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    const int2 se = ent[e];
#pragma unroll 2
    for (int j = se.x; j < se.x + se.y; j++) {
      double nrm[3];
      if (K::NRM) {
        nrm[0] = src.nx[j];
        nrm[1] = src.ny[j];
        nrm[2] = src.nz[j];
      }
      K::template add<NR>(x - src.x[j], y - src.y[j], z - src.z[j], nrm, den + j * js, u, kc);
    }
  }
  if (act) {
    double* val = trg.val + (long)r0 * trg.val_rs + (long)(tb + t) * K::T;
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int d = 0; d < K::T; d++) val[(long)q * trg.val_rs + d] += u[q][d];
  }
}

// Laplace near field for blocks of 16 right-hand sides on the FP64 matrix
// cores.  For a group of 4 sources the wave forms
//   U[q][t] += sum_k F[k][q] G[t][k]      (v_mfma_f64_16x16x4f64)
// with A = F^T (16 RHS x 4 sources, lane l holds F[k = l/16][q = l%16]) and
// B = G^T (4 sources x 16 targets, lane l holds G(target l%16, source l/16)),
// so every lane evaluates distinct pairs and the densities come from LDS as
// one contiguous 512-byte row per group instead of a broadcast per pair.  A
// wave covers 64 targets as four 16-target column blocks; the accumulator
// lane l, item v holds right-hand side l/16 + 4*v of target l%16 (the D layout
// of v_mfma_f64_16x16x4f64: row l/16 + 4v, column l%16).  Source
// groups are padded to a multiple of 4 with zero densities.  Summation order
// differs from the scalar kernels only by FP64 rounding.
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// One staged source tile against the wave's first NG 16-target column blocks
// (NG = the blocks that hold targets, so ragged leaves do not pay for empty
// blocks).
// GUARD: the r < 1e-8 exclusion; an entry whose closest pair of distinct
// points is farther apart than 1e-7 (flag 0 of the set-up's near-field
// pass: two different leaves, so no equal points either) goes without it.
template <int NG, bool GUARD = true>
__device__ __forceinline__ void p2p_tile64(int cnt4, int kk, int col, const double (*s_pos)[TILE],
                                           const double (*s_den)[16], const double* tx, const double* ty,
                                           const double* tz, v4d* acc)
{
  for (int g = 0; g < cnt4; g += 4) {
    const double sx = s_pos[0][g + kk], sy = s_pos[1][g + kk], sz = s_pos[2][g + kk];
    const double f = s_den[g + kk][col];
#pragma unroll
    for (int i = 0; i < NG; i++) {
      const double dx = tx[i] - sx, dy = ty[i] - sy, dz = tz[i] - sz;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double gv = GUARD ? rinv(r2) : rsqrt_nr(r2);
      acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(f, gv, acc[i], 0, 0, 0);
    }
  }
}

// sources per staged tile of the FP32 kernels: leaves of up to 160 points in
// one tile (4M points / 16 RHS, against 128: near field 7.69 -> 7.55 ms, S2M
// 0.92 -> 0.85 ms; 128 was 2 % faster than 64)
constexpr int TILE32 = 160;

// The reference's r < 1e-8 rule (kernel3d.cpp:21, :132) by entry, as the
// set-up flagged it (echk):
//   CHK 0: every pair of the entry is farther apart than its FP32 offsets
//          can lose and than 1e-8 (two different leaves without points next
//          to their shared boundary), so r^2 > 0: 1/r with no guard
//          (padded sources sit far away with zero density);
//   CHK 1: the target's own leaf with no two distinct points that close:
//          equal points (equal FP32 offsets) have r^2 = 0 and drop out.
// (Flag 2 -- a pair can be closer than the FP32 offsets resolve: leaves with
// points next to a shared boundary, or a leaf holding two distinct
// near-coincident points -- is left to an FP64 pass over those entries.)
template <int NG, int CHK, int TW = TILE32>
__device__ __forceinline__ void p2p_tile32(int cnt4, int kk, int col, const float (*s_pos)[TW],
                                           const float (*s_den)[16], const float* tx, const float* ty,
                                           const float* tz, double (*acc)[4])
{
  v4f part[NG];
#pragma unroll
  for (int i = 0; i < NG; i++) part[i] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
  for (int g = 0; g < cnt4; g += 4) {
    const float sx = s_pos[0][g + kk], sy = s_pos[1][g + kk], sz = s_pos[2][g + kk];
    const float f = s_den[g + kk][col];
#pragma unroll
    for (int i = 0; i < NG; i++) {
      const float dx = tx[i] - sx, dy = ty[i] - sy, dz = tz[i] - sz;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float gv = CHK == 1 ? (r2 < 1e-16f ? 0.0f : __builtin_amdgcn_rsqf(r2)) : __builtin_amdgcn_rsqf(r2);
      part[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(f, gv, part[i], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < NG; i++)
#pragma unroll
    for (int v = 0; v < 4; v++) acc[i][v] += (double)part[i][v];
}

__global__ __launch_bounds__(P2P_BS) void k_p2p_mfma111(int nchunk, const int* __restrict__ tbeg,
                                                        const int* __restrict__ tnum, const int* __restrict__ eptr,
                                                        const int2* __restrict__ ent, const int* __restrict__ echk,
                                                        Points src, Targets trg, int roff, KC kc)
{
  constexpr int NR = 16;
  __shared__ double s_pos[3][TILE];
  __shared__ double s_den[TILE][NR];
  const int c = xcd_item(blockIdx.x, nchunk), t = threadIdx.x, w = t >> 6, l = t & 63;
  if (c >= nchunk) return;
  const int col = l & 15, kk = l >> 4;
  const int r0 = roff + blockIdx.y * NR;
  const int tb = tbeg[c], tn = tnum[c];
  const bool wact = 64 * w < tn;  // wave-uniform
  const int ng = wact ? min(4, (tn - 64 * w + 15) >> 4) : 0;  // 16-target blocks with targets
  double tx[4], ty[4], tz[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int tt = 64 * w + 16 * i + col;
    tt = tt < tn ? tt : 0;
    tx[i] = trg.x[tb + tt];
    ty[i] = trg.y[tb + tt];
    tz[i] = trg.z[tb + tt];
  }
  v4d acc[4];
#pragma unroll
  for (int i = 0; i < 4; i++) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};
  const double* den = src.den + (long)r0 * src.den_rs;
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    const int2 se = ent[e];
    const bool guard = !echk || echk[e] != 0;  // block-uniform
    for (int b0 = 0; b0 < se.y; b0 += TILE) {
      const int cnt = min(TILE, se.y - b0), cnt4 = (cnt + 3) & ~3;
      const int j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt4; i += P2P_BS) {
        const bool in = i < cnt;
        s_pos[0][i] = in ? src.x[j0 + i] : 0.0;
        s_pos[1][i] = in ? src.y[j0 + i] : 0.0;
        s_pos[2][i] = in ? src.z[j0 + i] : 0.0;
      }
      stage_den<1, NR, P2P_BS>(s_den, den, src, j0, cnt, t);
      for (int i = cnt * NR + t; i < cnt4 * NR; i += P2P_BS) s_den[i / NR][i % NR] = 0.0;
      __syncthreads();
      if (!guard) {
        if (ng == 4)
          p2p_tile64<4, false>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 3)
          p2p_tile64<3, false>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 2)
          p2p_tile64<2, false>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 1)
          p2p_tile64<1, false>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      } else if (ng == 4)
        p2p_tile64<4>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 3)
        p2p_tile64<3>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 2)
        p2p_tile64<2>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 1)
        p2p_tile64<1>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
    }
  }
  if (!wact) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int tt = 64 * w + 16 * i + col;
    if (tt < tn) {
      double* val = trg.val + (long)(r0 + kk) * trg.val_rs + tb + tt;
#pragma unroll
      for (int v = 0; v < 4; v++) val[(long)(4 * v) * trg.val_rs] += kc.a * acc[i][v];
    }
  }
}

// Mixed-precision form of k_p2p_mfma111: the same wave layout on the FP32
// matrix cores (v_mfma_f32_16x16x4f32, accumulator item v of lane l holds
// right-hand side 4*(l/16) + v of target l%16).  Coordinates are taken
// relative to the chunk's first target in FP64 and then rounded, so FP32
// keeps the differences of nearby points to ~1e-7 of the leaf size; 1/r
// comes from v_rsq_f32.  The FP32 products of one staged source tile (at most
// 128 sources) are summed into FP64 accumulators.  echk[e] selects how the
// pairs of entry e take the r < 1e-8 exclusion (p2p_tile32); entries flagged
// 2 are skipped (an FP64 pass adds them).
__global__ __launch_bounds__(P2P_BS) void k_p2p_mfma111_f32(int nchunk, const int* __restrict__ tbeg,
                                                            const int* __restrict__ tnum, const int* __restrict__ eptr,
                                                            const int2* __restrict__ ent,
                                                            const int* __restrict__ echk, Points src, Targets trg,
                                                            int roff, KC kc, int nrv)
{
  constexpr int NR = 16;
  __shared__ float s_pos[3][TILE32];
  __shared__ __align__(16) float s_den[TILE32][NR];
  const int c = xcd_item(blockIdx.x, nchunk), t = threadIdx.x, w = t >> 6, l = t & 63;
  if (c >= nchunk) return;
  const int col = l & 15, kk = l >> 4;
  const int r0 = roff + blockIdx.y * NR;
  const int tb = tbeg[c], tn = tnum[c];
  const bool wact = 64 * w < tn;  // wave-uniform
  const int ng = wact ? min(4, (tn - 64 * w + 15) >> 4) : 0;  // 16-target blocks with targets
  const double ox = trg.x[tb], oy = trg.y[tb], oz = trg.z[tb];
  float tx[4], ty[4], tz[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int tt = 64 * w + 16 * i + col;
    tt = tt < tn ? tt : 0;
    tx[i] = (float)(trg.x[tb + tt] - ox);
    ty[i] = (float)(trg.y[tb + tt] - oy);
    tz[i] = (float)(trg.z[tb + tt] - oz);
  }
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int v = 0; v < 4; v++) acc[i][v] = 0.0;
  const double* den = src.den + (long)r0 * src.den_rs;
  const float* den32 = src.den32 ? src.den32 + (long)r0 * src.den_rs : nullptr;
  // all 16 right-hand sides real and every source's row of 16 FP32
  // densities 16-byte aligned: one thread stages a source's row as four
  // 16-byte loads
  const bool vec = den32 && nrv == NR && src.den_rs == 1 && src.den_js % 4 == 0 &&
                   ((unsigned long)den32 & 15) == 0;
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    const int chk = echk[e];  // block-uniform
    if (chk == 2) continue;   // the FP64 pass's
    const int2 se = ent[e];
    for (int b0 = 0; b0 < se.y; b0 += TILE32) {
      const int cnt = min(TILE32, se.y - b0), cnt4 = (cnt + 3) & ~3;
      const int j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt4; i += P2P_BS) {
        const bool in = i < cnt;
        s_pos[0][i] = in ? (float)(src.x[j0 + i] - ox) : 1e6f;  // padding: far away, zero density
        s_pos[1][i] = in ? (float)(src.y[j0 + i] - oy) : 0.0f;
        s_pos[2][i] = in ? (float)(src.z[j0 + i] - oz) : 0.0f;
      }
      if (vec) {
        for (int jj = t; jj < cnt4; jj += P2P_BS) {
          // named registers (an array here went to scratch memory)
          float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0, q3 = q0;
          if (jj < cnt) {
            const float4* row = reinterpret_cast<const float4*>(den32 + (long)(j0 + jj) * src.den_js);
            q0 = row[0];
            q1 = row[1];
            q2 = row[2];
            q3 = row[3];
          }
          float4* dst = reinterpret_cast<float4*>(s_den[jj]);
          dst[0] = q0;
          dst[1] = q1;
          dst[2] = q2;
          dst[3] = q3;
        }
      } else {
        for (int i = t; i < cnt4 * NR; i += P2P_BS) {
          const int jj = i / NR, q = i - jj * NR;
          const long at = (long)(j0 + jj) * src.den_js + (long)q * src.den_rs;
          s_den[jj][q] = jj < cnt && q < nrv ? (den32 ? den32[at] : (float)den[at]) : 0.0f;
        }
      }
      __syncthreads();
      if (chk == 1) {
        if (ng == 4)
          p2p_tile32<4, 1>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 3)
          p2p_tile32<3, 1>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 2)
          p2p_tile32<2, 1>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 1)
          p2p_tile32<1, 1>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      } else {
        if (ng == 4)
          p2p_tile32<4, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 3)
          p2p_tile32<3, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 2)
          p2p_tile32<2, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
        else if (ng == 1)
          p2p_tile32<1, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      }
    }
  }
  if (!wact) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int tt = 64 * w + 16 * i + col;
    if (tt < tn) {
      double* val = trg.val + (long)(r0 + 4 * kk) * trg.val_rs + tb + tt;
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (4 * kk + v < nrv) val[(long)v * trg.val_rs] += kc.a * acc[i][v];
    }
  }
}

// Launches of the FP32 matrix-core kernels: nr / 16 full blocks of 16
// right-hand sides, then a remainder of at least PAD16_MIN as one block of 16
// of which nrv are real (zero densities staged, nothing stored for the rest);
// fn(nb, roff, nrv).  Returns the right-hand sides left for the vector
// kernels (a remainder below PAD16_MIN: measured faster there at 1 RHS).
template <class F>
static int rhs_blocks16(int nr, F&& fn)
{
  if (nr >= 16) fn(nr / 16, 0, 16);
  const int rem = nr % 16;
  if (rem >= PAD16_MIN) fn(1, nr - rem, rem);
  return rem >= PAD16_MIN ? 0 : rem;
}

// Split nr right-hand sides into launches of NR in {16, 4, 1}; fn(NR, nblocks, roff).
template <class F>
static void rhs_blocks(int nr, F&& fn)
{
  int r = 0;
  while (r < nr) {
    int left = nr - r;
    int NR = left >= 16 ? 16 : left >= 4 ? 4 : 1;
    int nb = left / NR;
    fn(NR, nb, r);
    r += nb * NR;
  }
}

// ---------------------------------------------------------------------------
// Set-up pass of the mixed near field: for every near-field entry e of every
// chunk c (the chunk's targets against one source leaf), in FP64, the
// smallest distance between a target and a source that are not the same
// point (out[e].x, squared; DBL_MAX when there is none) and the largest
// distance of any of the entry's sources from the chunk's first target, the
// origin of the FP32 offsets (out[e].y, squared), and the same for the
// chunk's targets (tout[c], squared).  Block per chunk, one target per
// thread, sources staged in LDS, block-wide min / max.
constexpr int NMIN_TILE = 128;
__global__ __launch_bounds__(P2P_BS) void k_near_extent(const int* __restrict__ tbeg, const int* __restrict__ tnum,
                                                        const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                        Points src, Targets trg, double2* __restrict__ out,
                                                        double* __restrict__ tout)
{
  __shared__ double s_p[3][NMIN_TILE];
  __shared__ double s_red[2][P2P_BS / 64];
  const int c = blockIdx.x, t = threadIdx.x, w = t >> 6, l = t & 63;
  const int tb = tbeg[c], tn = tnum[c];
  const bool act = t < tn;
  const double ox = trg.x[tb], oy = trg.y[tb], oz = trg.z[tb];
  double x = ox, y = oy, z = oz;
  if (act) {
    x = trg.x[tb + t];
    y = trg.y[tb + t];
    z = trg.z[tb + t];
  }
  auto wave_min = [](double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
  };
  auto wave_max = [](double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
  };
  {
    const double dx = x - ox, dy = y - oy, dz = z - oz;
    const double m = wave_max(dx * dx + dy * dy + dz * dz);
    if (l == 0) s_red[1][w] = m;
    __syncthreads();
    if (t == 0) {
      double mm = 0.0;
      for (int q = 0; q < P2P_BS / 64; q++) mm = fmax(mm, s_red[1][q]);
      tout[c] = mm;
    }
  }
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    const int2 se = ent[e];
    double rmin = DBL_MAX, omax = 0.0;
    for (int b0 = 0; b0 < se.y; b0 += NMIN_TILE) {
      const int cnt = min(NMIN_TILE, se.y - b0), j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt; i += P2P_BS) {
        s_p[0][i] = src.x[j0 + i];
        s_p[1][i] = src.y[j0 + i];
        s_p[2][i] = src.z[j0 + i];
        const double dx = s_p[0][i] - ox, dy = s_p[1][i] - oy, dz = s_p[2][i] - oz;
        omax = fmax(omax, dx * dx + dy * dy + dz * dz);
      }
      __syncthreads();
      if (act)
        for (int jj = 0; jj < cnt; jj++) {
          const double dx = x - s_p[0][jj], dy = y - s_p[1][jj], dz = z - s_p[2][jj];
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 > 0.0) rmin = fmin(rmin, r2);
        }
    }
    rmin = wave_min(rmin);
    omax = wave_max(omax);
    __syncthreads();  // the previous entry's reduction has been read
    if (l == 0) {
      s_red[0][w] = rmin;
      s_red[1][w] = omax;
    }
    __syncthreads();
    if (t == 0) {
      double a = DBL_MAX, m = 0.0;
      for (int q = 0; q < P2P_BS / 64; q++) {
        a = fmin(a, s_red[0][q]);
        m = fmax(m, s_red[1][q]);
      }
      out[e] = make_double2(a, m);
    }
  }
}

void near_extent(hipStream_t s, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int2* ent,
                 Points src, Targets trg, double2* out, double* tout)
{
  if (nchunk <= 0) return;
  hipLaunchKernelGGL(k_near_extent, dim3(nchunk), dim3(P2P_BS), 0, s, tbeg, tnum, eptr, ent, src, trg, out, tout);
}

// ---------------------------------------------------------------------------
// FP64 direct sum at selected targets (FMM3d::check, fmm3d_check.cpp:48-54):
// block (chunk of 128 targets, source segment) writes a partial sum; the
// partials are reduced in segment order, so the result is deterministic.
template <int KT, int NR>
__global__ __launch_bounds__(P2P_BS) void k_direct_part(int nchk, const double* __restrict__ cx,
                                                        const double* __restrict__ cy, const double* __restrict__ cz,
                                                        int nsrc, int seglen, Points src, double* __restrict__ part,
                                                        int roff, KC kc)
{
  using K = Knl<KT>;
  constexpr int NPOS = K::NRM ? 6 : 3;
  constexpr int ND = NR * K::S;
  __shared__ double s_pos[TILE][NPOS];
  __shared__ double s_den[TILE][ND];
  const int t = threadIdx.x, k = blockIdx.x * P2P_BS + t, seg = blockIdx.y;
  const int r0 = roff + blockIdx.z * NR;
  const bool act = k < nchk;
  double x = act ? cx[k] : 0.0, y = act ? cy[k] : 0.0, z = act ? cz[k] : 0.0;
  double u[NR][K::T];
#pragma unroll
  for (int q = 0; q < NR; q++)
#pragma unroll
    for (int d = 0; d < K::T; d++) u[q][d] = 0.0;
  const double* den = src.den + (long)r0 * src.den_rs;
  const int sb = seg * seglen, se = min(nsrc, sb + seglen);
  for (int b0 = sb; b0 < se; b0 += TILE) {
    const int cnt = min(TILE, se - b0);
    __syncthreads();
    for (int i = t; i < cnt; i += P2P_BS) {
      s_pos[i][0] = src.x[b0 + i];
      s_pos[i][1] = src.y[b0 + i];
      s_pos[i][2] = src.z[b0 + i];
      if (K::NRM) {
        s_pos[i][NPOS - 3] = src.nx[b0 + i];
        s_pos[i][NPOS - 2] = src.ny[b0 + i];
        s_pos[i][NPOS - 1] = src.nz[b0 + i];
      }
    }
    stage_den<K::S, NR, P2P_BS>(s_den, den, src, b0, cnt, t);
    __syncthreads();
    if (act)
      for (int jj = 0; jj < cnt; jj++)
        K::template add<NR>(x - s_pos[jj][0], y - s_pos[jj][1], z - s_pos[jj][2], &s_pos[jj][NPOS - 3], s_den[jj], u,
                            kc);
  }
  if (act) {
    const int nseg = gridDim.y;
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int d = 0; d < K::T; d++) part[(((long)(r0 + q) * nseg + seg) * nchk + k) * K::T + d] = u[q][d];
  }
}

__global__ void k_direct_reduce(int n, int nseg, int nr, const double* __restrict__ part, double* __restrict__ out)
{
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;  // over nr * n values
  if (i >= (long)nr * n) return;
  int r = (int)(i / n);
  long k = i - (long)r * n;
  double s = 0.0;
  for (int g = 0; g < nseg; g++) s += part[((long)r * nseg + g) * n + k];
  out[i] = s;
}

template <int KT>
static void direct_kt(hipStream_t s, int nchk, const double* cx, const double* cy, const double* cz, int nsrc,
                      int nseg, int seglen, Points src, double* part, int nr, KC kc)
{
  rhs_blocks(nr, [&](int NR, int nb, int roff) {
    dim3 grid((nchk + P2P_BS - 1) / P2P_BS, nseg, nb);
    if (NR == 16)
      hipLaunchKernelGGL((k_direct_part<KT, 16>), grid, dim3(P2P_BS), 0, s, nchk, cx, cy, cz, nsrc, seglen, src,
                         part, roff, kc);
    else if (NR == 4)
      hipLaunchKernelGGL((k_direct_part<KT, 4>), grid, dim3(P2P_BS), 0, s, nchk, cx, cy, cz, nsrc, seglen, src, part,
                         roff, kc);
    else
      hipLaunchKernelGGL((k_direct_part<KT, 1>), grid, dim3(P2P_BS), 0, s, nchk, cx, cy, cz, nsrc, seglen, src, part,
                         roff, kc);
  });
}

void direct(hipStream_t s, int kt, int tdof, int nchk, const double* cx, const double* cy, const double* cz,
            int nsrc, int nseg, Points src, double* part, double* out, int nr, KC kc)
{
  if (nchk <= 0) return;
  const int seglen = (nsrc + nseg - 1) / nseg;
  switch (kt) {
#define KIFMM_DIRECT(KT) \
  case KT: direct_kt<KT>(s, nchk, cx, cy, cz, nsrc, nseg, seglen, src, part, nr, kc); break;
  KIFMM_DIRECT(111) KIFMM_DIRECT(121) KIFMM_DIRECT(311) KIFMM_DIRECT(312) KIFMM_DIRECT(321) KIFMM_DIRECT(322)
  KIFMM_DIRECT(511) KIFMM_DIRECT(521)
#undef KIFMM_DIRECT
  }
  long tot = (long)nr * nchk * tdof;
  hipLaunchKernelGGL(k_direct_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, nchk * tdof, nseg, nr,
                     part, out);
}

// ---------------------------------------------------------------------------
// points -> check surface (S2M, S2L)
constexpr int P2S_BS = 256;

template <int KT, int NR>
__global__ __launch_bounds__(P2S_BS) void k_p2s(const int* __restrict__ out_node, const double4* __restrict__ geom,
                                                const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                const double* __restrict__ unit, int npt, Points src,
                                                double* __restrict__ out, int Rc, int pln, int roff, KC kc, int accum)
{
  using K = Knl<KT>;
  constexpr int NPOS = K::NRM ? 6 : 3;
  constexpr int ND = NR * K::S;
  __shared__ double s_pos[TILE][NPOS];
  __shared__ double s_den[TILE][ND];
  const int it = blockIdx.x, t = threadIdx.x;
  const int p = blockIdx.y * P2S_BS + t;
  const int r0 = roff + blockIdx.z * NR;
  const bool act = p < npt;
  const double4 g = geom[it];
  double x = 0, y = 0, z = 0;
  if (act) {
    x = g.x + g.w * unit[3 * p];
    y = g.y + g.w * unit[3 * p + 1];
    z = g.z + g.w * unit[3 * p + 2];
  }
  double u[NR][K::T];
#pragma unroll
  for (int q = 0; q < NR; q++)
#pragma unroll
    for (int d = 0; d < K::T; d++) u[q][d] = 0.0;
  const double* den = src.den + (long)r0 * src.den_rs;
  for (int e = eptr[it]; e < eptr[it + 1]; e++) {
    const int2 se = ent[e];
    for (int b0 = 0; b0 < se.y; b0 += TILE) {
      const int cnt = min(TILE, se.y - b0);
      const int j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt; i += P2S_BS) {
        s_pos[i][0] = src.x[j0 + i];
        s_pos[i][1] = src.y[j0 + i];
        s_pos[i][2] = src.z[j0 + i];
        if (K::NRM) {
          s_pos[i][NPOS - 3] = src.nx[j0 + i];
          s_pos[i][NPOS - 2] = src.ny[j0 + i];
          s_pos[i][NPOS - 1] = src.nz[j0 + i];
        }
      }
      stage_den<K::S, NR, P2S_BS>(s_den, den, src, j0, cnt, t);
      __syncthreads();
      if (act)
        for (int jj = 0; jj < cnt; jj++)
          K::template add<NR>(x - s_pos[jj][0], y - s_pos[jj][1], z - s_pos[jj][2], &s_pos[jj][NPOS - 3], s_den[jj], u,
                              kc);
    }
  }
  if (act) {
    const long node = out_node[it];
#pragma unroll
    for (int q = 0; q < NR; q++) {
      double* o = out + (node * Rc + r0 + q) * pln + (long)p * K::T;
#pragma unroll
      for (int d = 0; d < K::T; d++) o[d] = u[q][d] + (accum ? o[d] : 0.0);
    }
  }
}

// ---------------------------------------------------------------------------
// equivalent surface -> targets (L2T, M2T), mm kernel
template <int KT, int NR>
__global__ __launch_bounds__(P2P_BS) void k_s2p(const int* __restrict__ tbeg, const int* __restrict__ tnum,
                                                const int* __restrict__ eptr, const int* __restrict__ enode,
                                                const int* __restrict__ ekind, const double4* __restrict__ egeom,
                                                const double* __restrict__ unit_de, const double* __restrict__ unit_ue,
                                                int nsp, const double* __restrict__ dneden,
                                                const double* __restrict__ upeden, int Rc, int pln, Targets trg,
                                                int roff, KC kc)
{
  using K = Knl<KT>;
  constexpr int ND = NR * K::S;
  __shared__ double s_pos[TILE][3];
  __shared__ double s_den[TILE][ND];
  const int c = blockIdx.x, t = threadIdx.x;
  const int r0 = roff + blockIdx.y * NR;
  const int tb = tbeg[c], tn = tnum[c];
  const bool act = t < tn;
  double x = 0, y = 0, z = 0;
  if (act) {
    x = trg.x[tb + t];
    y = trg.y[tb + t];
    z = trg.z[tb + t];
  }
  double u[NR][K::T];
#pragma unroll
  for (int q = 0; q < NR; q++)
#pragma unroll
    for (int d = 0; d < K::T; d++) u[q][d] = 0.0;
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    const int kind = ekind[e];
    const double* unit = kind == 0 ? unit_de : unit_ue;
    const double* buf = (kind == 0 ? dneden : upeden) + ((long)enode[e] * Rc + r0) * pln;
    const double4 g = egeom[e];
    for (int b0 = 0; b0 < nsp; b0 += TILE) {
      const int cnt = min(TILE, nsp - b0);
      __syncthreads();
      for (int i = t; i < cnt; i += P2P_BS) {
        s_pos[i][0] = g.x + g.w * unit[3 * (b0 + i)];
        s_pos[i][1] = g.y + g.w * unit[3 * (b0 + i) + 1];
        s_pos[i][2] = g.z + g.w * unit[3 * (b0 + i) + 2];
      }
      const int per = cnt * K::S;
      for (int i = t; i < NR * per; i += P2P_BS) {
        int q = i / per, rem = i - q * per;
        int jj = rem / K::S, sd = rem - jj * K::S;
        s_den[jj][q * K::S + sd] = buf[(long)q * pln + (long)b0 * K::S + rem];
      }
      __syncthreads();
      if (act)
        for (int jj = 0; jj < cnt; jj++)
          K::template add<NR>(x - s_pos[jj][0], y - s_pos[jj][1], z - s_pos[jj][2], nullptr, s_den[jj], u, kc);
    }
  }
  if (act) {
    double* val = trg.val + (long)r0 * trg.val_rs + (long)(tb + t) * K::T;
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int d = 0; d < K::T; d++) val[(long)q * trg.val_rs + d] += u[q][d];
  }
}

template <int KT, int NR>
static void p2p_launch(hipStream_t s, int variant, int nchunk, int nb, const int* tbeg, const int* tnum,
                       const int* eptr, const int2* ent, const int* echk, Points src, Targets trg, int roff, KC kc)
{
  if (variant >= 2 && KT == 111 && NR == 16 && src.den_rs == 1)
    hipLaunchKernelGGL(k_p2p_mfma111, dim3(xcd_grid(nchunk), nb), dim3(P2P_BS), 0, s, nchunk, tbeg, tnum, eptr, ent,
                       echk, src, trg, roff, kc);
  else if (variant == 1)
    hipLaunchKernelGGL((k_p2p_s<KT, NR>), dim3(2 * nchunk, nb), dim3(64), 0, s, tbeg, tnum, eptr, ent, src, trg, roff,
                       kc);
  else if (variant == 4)
    hipLaunchKernelGGL((k_p2p32<KT, NR>), dim3(nchunk, nb), dim3(P2P_BS), 0, s, tbeg, tnum, eptr, ent, echk, src, trg,
                       roff, kc);
  else
    hipLaunchKernelGGL((k_p2p<KT, NR>), dim3(nchunk, nb), dim3(P2P_BS), 0, s, tbeg, tnum, eptr, ent, src, trg, roff,
                       kc);
}

template <int KT>
static int p2p_kt(hipStream_t s, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int2* ent,
                  const int* echk, Points src, Targets trg, int nr, KC kc, int variant)
{
  int r0 = 0;
  if (KT == 111 && variant == 3 && echk) {
    const int left = rhs_blocks16(nr, [&](int nb, int roff, int nrv) {
      hipLaunchKernelGGL(k_p2p_mfma111_f32, dim3(xcd_grid(nchunk), nb), dim3(P2P_BS), 0, s, nchunk, tbeg, tnum, eptr,
                         ent, echk, src, trg, roff, kc, nrv);
    });
    r0 = nr - left;
  }
  rhs_blocks(nr - r0, [&](int NR, int nb, int roff) {
    if (NR == 16)
      p2p_launch<KT, 16>(s, variant, nchunk, nb, tbeg, tnum, eptr, ent, echk, src, trg, r0 + roff, kc);
    else if (NR == 4)
      p2p_launch<KT, 4>(s, variant, nchunk, nb, tbeg, tnum, eptr, ent, echk, src, trg, r0 + roff, kc);
    else
      p2p_launch<KT, 1>(s, variant, nchunk, nb, tbeg, tnum, eptr, ent, echk, src, trg, r0 + roff, kc);
  });
  // right-hand sides [0, n) went through a kernel that skips flagged entries
  return variant == 4 && echk ? nr : r0;
}

int p2p(hipStream_t s, int kt, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int2* ent,
        const int* echk, Points src, Targets trg, int nr, KC kc, int variant)
{
  if (nchunk <= 0) return 0;
  switch (kt) {
#define KIFMM_P2P(KT) \
  case KT: return p2p_kt<KT>(s, nchunk, tbeg, tnum, eptr, ent, echk, src, trg, nr, kc, variant);
  KIFMM_P2P(111) KIFMM_P2P(121) KIFMM_P2P(311) KIFMM_P2P(312) KIFMM_P2P(321) KIFMM_P2P(322) KIFMM_P2P(511)
  KIFMM_P2P(521)
#undef KIFMM_P2P
  }
  return 0;
}

// S2M / S2L for the Laplace kernel and 16-RHS blocks on the FP64 matrix
// cores: the wave layout of k_p2p_mfma111 with the check-surface points of an
// item as the targets (128 per block, positions centre + radius x unit
// surface) and the item's source points as the sources.  FP64 in both forms:
// the check values feed the pseudoinverses.
__global__ __launch_bounds__(P2P_BS) void k_p2s_mfma111(int nitem, int npc, const int* __restrict__ out_node,
                                                        const double4* __restrict__ geom,
                                                        const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                        const double* __restrict__ unit, int npt, Points src,
                                                        double* __restrict__ out, int Rc, int pln, int roff, KC kc,
                                                        int accum)
{
  constexpr int NR = 16;
  __shared__ double s_pos[3][TILE];
  __shared__ double s_den[TILE][NR];
  const int nb = nitem * npc, b = xcd_item(blockIdx.x, nb);
  if (b >= nb) return;  // block-uniform
  const int it = b / npc, pc = b - it * npc;
  const int t = threadIdx.x, w = t >> 6, l = t & 63, col = l & 15, kk = l >> 4;
  const int r0 = roff + blockIdx.y * NR;
  const int pb = pc * P2P_BS, tn = min(P2P_BS, npt - pb);  // surface points of this block
  const bool wact = 64 * w < tn;                           // wave-uniform
  const int ng = wact ? min(4, (tn - 64 * w + 15) >> 4) : 0;
  const double4 g = geom[it];
  double tx[4], ty[4], tz[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int tt = 64 * w + 16 * i + col;
    tt = pb + (tt < tn ? tt : 0);
    tx[i] = g.x + g.w * unit[3 * tt];
    ty[i] = g.y + g.w * unit[3 * tt + 1];
    tz[i] = g.z + g.w * unit[3 * tt + 2];
  }
  v4d acc[4];
#pragma unroll
  for (int i = 0; i < 4; i++) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};
  const double* den = src.den + (long)r0 * src.den_rs;
  for (int e = eptr[it]; e < eptr[it + 1]; e++) {
    const int2 se = ent[e];
    for (int b0 = 0; b0 < se.y; b0 += TILE) {
      const int cnt = min(TILE, se.y - b0), cnt4 = (cnt + 3) & ~3;
      const int j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt4; i += P2P_BS) {
        const bool in = i < cnt;
        s_pos[0][i] = in ? src.x[j0 + i] : 0.0;
        s_pos[1][i] = in ? src.y[j0 + i] : 0.0;
        s_pos[2][i] = in ? src.z[j0 + i] : 0.0;
      }
      stage_den<1, NR, P2P_BS>(s_den, den, src, j0, cnt, t);
      for (int i = cnt * NR + t; i < cnt4 * NR; i += P2P_BS) s_den[i / NR][i % NR] = 0.0;
      __syncthreads();
      if (ng == 4)
        p2p_tile64<4>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 3)
        p2p_tile64<3>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 2)
        p2p_tile64<2>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 1)
        p2p_tile64<1>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
    }
  }
  if (!wact) return;
  const long node = out_node[it];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int tt = 64 * w + 16 * i + col;
    if (tt >= tn) continue;
    double* o = out + (node * Rc + r0 + kk) * pln + pb + tt;  // item v: right-hand side kk + 4v
#pragma unroll
    for (int v = 0; v < 4; v++) o[(long)(4 * v) * pln] = kc.a * acc[i][v] + (accum ? o[(long)(4 * v) * pln] : 0.0);
  }
}

// Mixed-precision S2M / S2L: the layout of k_p2s_mfma111 on the FP32 matrix
// cores, coordinates relative to the item's box centre rounded to FP32 (the
// surface points exactly, c + r u - c = r u), FP32 products of one source
// tile summed in FP64.  Check surfaces never touch the item's sources, so
// 1/r is taken of max(r^2, 1e-16) without the exclusion test.  A block of NW
// waves takes 64 NW surface points of an item (all of them for np <= 8 when
// NW = ceil(npt / 64)), so each source tile is staged once per item; FP32
// density rows of 16 right-hand sides are staged as four 16-byte loads.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_p2s_mfma111_f32(int nitem, int npc, const int* __restrict__ out_node,
                                                             const double4* __restrict__ geom,
                                                             const int* __restrict__ eptr,
                                                             const int2* __restrict__ ent,
                                                             const double* __restrict__ unit, int npt, Points src,
                                                             double* __restrict__ out, int Rc, int pln, int roff,
                                                             KC kc, int accum, int nrv)
{
  constexpr int NR = 16, BS = 64 * NW;
  __shared__ float s_pos[3][TILE32];
  __shared__ __align__(16) float s_den[TILE32][NR];
  const int nb = nitem * npc, b = xcd_item(blockIdx.x, nb);
  if (b >= nb) return;  // block-uniform
  const int it = b / npc, pc = b - it * npc;
  const int t = threadIdx.x, w = t >> 6, l = t & 63, col = l & 15, kk = l >> 4;
  const int r0 = roff + blockIdx.y * NR;
  const int pb = pc * BS, tn = min(BS, npt - pb);  // surface points of this block
  const bool wact = 64 * w < tn;                    // wave-uniform
  const int ng = wact ? min(4, (tn - 64 * w + 15) >> 4) : 0;
  const double4 g = geom[it];
  float tx[4], ty[4], tz[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int tt = 64 * w + 16 * i + col;
    tt = pb + (tt < tn ? tt : 0);
    tx[i] = (float)(g.w * unit[3 * tt]);
    ty[i] = (float)(g.w * unit[3 * tt + 1]);
    tz[i] = (float)(g.w * unit[3 * tt + 2]);
  }
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int v = 0; v < 4; v++) acc[i][v] = 0.0;
  const double* den = src.den + (long)r0 * src.den_rs;
  const float* den32 = src.den32 ? src.den32 + (long)r0 * src.den_rs : nullptr;
  const bool vec = den32 && nrv == NR && src.den_rs == 1 && src.den_js % 4 == 0 &&
                   ((unsigned long)den32 & 15) == 0;
  for (int e = eptr[it]; e < eptr[it + 1]; e++) {
    const int2 se = ent[e];
    for (int b0 = 0; b0 < se.y; b0 += TILE32) {
      const int cnt = min(TILE32, se.y - b0), cnt4 = (cnt + 3) & ~3;
      const int j0 = se.x + b0;
      __syncthreads();
      for (int i = t; i < cnt4; i += BS) {
        const bool in = i < cnt;
        s_pos[0][i] = in ? (float)(src.x[j0 + i] - g.x) : 0.0f;
        s_pos[1][i] = in ? (float)(src.y[j0 + i] - g.y) : 0.0f;
        s_pos[2][i] = in ? (float)(src.z[j0 + i] - g.z) : 0.0f;
      }
      if (vec) {
        for (int jj = t; jj < cnt4; jj += BS) {
          // named registers (an array here goes to scratch memory)
          float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0, q3 = q0;
          if (jj < cnt) {
            const float4* row = reinterpret_cast<const float4*>(den32 + (long)(j0 + jj) * src.den_js);
            q0 = row[0];
            q1 = row[1];
            q2 = row[2];
            q3 = row[3];
          }
          float4* dst = reinterpret_cast<float4*>(s_den[jj]);
          dst[0] = q0;
          dst[1] = q1;
          dst[2] = q2;
          dst[3] = q3;
        }
      } else {
        for (int i = t; i < cnt4 * NR; i += BS) {
          const int jj = i / NR, q = i - jj * NR;
          const long at = (long)(j0 + jj) * src.den_js + (long)q * src.den_rs;
          s_den[jj][q] = jj < cnt && q < nrv ? (den32 ? den32[at] : (float)den[at]) : 0.0f;
        }
      }
      __syncthreads();
      if (ng == 4)
        p2p_tile32<4, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 3)
        p2p_tile32<3, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 2)
        p2p_tile32<2, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
      else if (ng == 1)
        p2p_tile32<1, 0>(cnt4, kk, col, s_pos, s_den, tx, ty, tz, acc);
    }
  }
  if (!wact) return;
  const long node = out_node[it];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int tt = 64 * w + 16 * i + col;
    if (tt >= tn) continue;
    double* o = out + (node * Rc + r0 + 4 * kk) * pln + pb + tt;  // item v: right-hand side 4 kk + v
#pragma unroll
    for (int v = 0; v < 4; v++)
      if (4 * kk + v < nrv) o[(long)v * pln] = kc.a * acc[i][v] + (accum ? o[(long)v * pln] : 0.0);
  }
}

// Launch of k_p2s_mfma111_f32 with NW = ceil(npt / 64) waves per block (at
// most 8; more points than 512 take several blocks per item).
static void p2s_f32_launch(hipStream_t s, int nb, int nitem, const int* out_node, const double4* geom,
                           const int* eptr, const int2* ent, const double* unit, int npt, Points src, double* out,
                           int Rc, int pln, int roff, KC kc, int accum, int nrv)
{
  const int nw = std::min(8, std::max(2, (npt + 63) / 64));
#define KIFMM_P2S32(NW)                                                                                         \
  do {                                                                                                          \
    const int npc = (npt + 64 * NW - 1) / (64 * NW);                                                            \
    hipLaunchKernelGGL(k_p2s_mfma111_f32<NW>, dim3(xcd_grid(nitem * npc), nb), dim3(64 * NW), 0, s, nitem, npc, \
                       out_node, geom, eptr, ent, unit, npt, src, out, Rc, pln, roff, kc, accum, nrv);          \
  } while (0)
  switch (nw) {
  case 2: KIFMM_P2S32(2); break;
  case 3: KIFMM_P2S32(3); break;
  case 4: KIFMM_P2S32(4); break;
  case 5: KIFMM_P2S32(5); break;
  case 6: KIFMM_P2S32(6); break;
  case 7: KIFMM_P2S32(7); break;
  default: KIFMM_P2S32(8); break;
  }
#undef KIFMM_P2S32
}

template <int KT>
static void p2s_kt(hipStream_t s, int nitem, const int* out_node, const double4* geom, const int* eptr,
                   const int2* ent, const double* unit, int npt, Points src, double* out, int Rc, int pln, int nr,
                   KC kc, int variant, int accum)
{
  int r0 = 0;
  if (KT == 111 && variant == 3) {
    r0 = nr - rhs_blocks16(nr, [&](int nb, int roff, int nrv) {
      p2s_f32_launch(s, nb, nitem, out_node, geom, eptr, ent, unit, npt, src, out, Rc, pln, roff, kc, accum, nrv);
    });
  }
  rhs_blocks(nr - r0, [&](int NR, int nb, int roff0) {
    const int roff = r0 + roff0;
    dim3 grid(nitem, (npt + P2S_BS - 1) / P2S_BS, nb);
    if (KT == 111 && NR == 16 && variant >= 2 && src.den_rs == 1) {
      const int npc = (npt + P2P_BS - 1) / P2P_BS;
      hipLaunchKernelGGL(k_p2s_mfma111, dim3(xcd_grid(nitem * npc), nb), dim3(P2P_BS), 0, s, nitem, npc, out_node,
                         geom, eptr, ent, unit, npt, src, out, Rc, pln, roff, kc, accum);
    } else if (NR == 16)
      hipLaunchKernelGGL((k_p2s<KT, 16>), grid, dim3(P2S_BS), 0, s, out_node, geom, eptr, ent, unit, npt, src, out,
                         Rc, pln, roff, kc, accum);
    else if (NR == 4)
      hipLaunchKernelGGL((k_p2s<KT, 4>), grid, dim3(P2S_BS), 0, s, out_node, geom, eptr, ent, unit, npt, src, out,
                         Rc, pln, roff, kc, accum);
    else
      hipLaunchKernelGGL((k_p2s<KT, 1>), grid, dim3(P2S_BS), 0, s, out_node, geom, eptr, ent, unit, npt, src, out,
                         Rc, pln, roff, kc, accum);
  });
}

void p2s(hipStream_t s, int kt, int nitem, const int* out_node, const double4* geom, const int* eptr, const int2* ent,
         const double* unit, int npt, Points src, double* out, int Rc, int pln, int nr, KC kc, int variant, int accum)
{
  if (nitem <= 0) return;
  switch (kt) {
#define KIFMM_P2S(KT) \
  case KT: p2s_kt<KT>(s, nitem, out_node, geom, eptr, ent, unit, npt, src, out, Rc, pln, nr, kc, variant, accum); break;
  KIFMM_P2S(111) KIFMM_P2S(121) KIFMM_P2S(311) KIFMM_P2S(312) KIFMM_P2S(321) KIFMM_P2S(322) KIFMM_P2S(511)
  KIFMM_P2S(521)
#undef KIFMM_P2S
  }
}

constexpr int S2P_TILE32 = 160;

// L2T + M2T for the Laplace mm kernel and 16-RHS blocks on the matrix cores:
// the wave layout of k_p2p_mfma111 (FP64) and k_p2p_mfma111_f32 (FP32, mixed
// form) with the surface points of each entry as the sources -- positions
// centre + radius x unit surface, densities from the node's [rhs][pln] row
// of the downward (kind 0) or upward (kind 1) equivalent densities.
template <bool F32>
__global__ __launch_bounds__(P2P_BS) void k_s2p_mfma111(int nchunk, const int* __restrict__ tbeg,
                                                        const int* __restrict__ tnum, const int* __restrict__ eptr,
                                                        const int* __restrict__ enode, const int* __restrict__ ekind,
                                                        const double4* __restrict__ egeom,
                                                        const double* __restrict__ unit_de,
                                                        const double* __restrict__ unit_ue, int nsp,
                                                        const double* __restrict__ dneden,
                                                        const double* __restrict__ upeden, int Rc, int pln,
                                                        Targets trg, int roff, KC kc, int nrv)
{
  typedef typename std::conditional<F32, float, double>::type R;
  // FP32: 160 surface points per staged tile, all of an np 6 surface (152) in one
  constexpr int NR = 16, TL = F32 ? S2P_TILE32 : TILE;
  __shared__ R s_pos[3][TL];
  __shared__ R s_den[TL][NR];
  const int c = xcd_item(blockIdx.x, nchunk), t = threadIdx.x, w = t >> 6, l = t & 63;
  if (c >= nchunk) return;  // block-uniform
  const int col = l & 15, kk = l >> 4;
  const int r0 = roff + blockIdx.y * NR;
  const int tb = tbeg[c], tn = tnum[c];
  const bool wact = 64 * w < tn;  // wave-uniform
  const int ng = wact ? min(4, (tn - 64 * w + 15) >> 4) : 0;
  // FP32: coordinates relative to the chunk's first target; FP64: absolute
  const double ox = F32 ? trg.x[tb] : 0.0, oy = F32 ? trg.y[tb] : 0.0, oz = F32 ? trg.z[tb] : 0.0;
  R tx[4], ty[4], tz[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int tt = 64 * w + 16 * i + col;
    tt = tt < tn ? tt : 0;
    tx[i] = (R)(trg.x[tb + tt] - ox);
    ty[i] = (R)(trg.y[tb + tt] - oy);
    tz[i] = (R)(trg.z[tb + tt] - oz);
  }
  double acc64[4][4];
  v4d accm[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    accm[i] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; v++) acc64[i][v] = 0.0;
  }
  for (int e = eptr[c]; e < eptr[c + 1]; e++) {
    const int kind = ekind[e];
    const double* unit = kind == 0 ? unit_de : unit_ue;
    const double* buf = (kind == 0 ? dneden : upeden) + ((long)enode[e] * Rc + r0) * pln;
    const double4 g = egeom[e];
    for (int b0 = 0; b0 < nsp; b0 += TL) {
      const int cnt = min(TL, nsp - b0), cnt4 = (cnt + 3) & ~3;
      __syncthreads();
      for (int i = t; i < cnt4; i += P2P_BS) {
        const bool in = i < cnt;
        const int pi = in ? b0 + i : 0;
        s_pos[0][i] = in ? (R)(g.x + g.w * unit[3 * pi] - ox) : (R)0;
        s_pos[1][i] = in ? (R)(g.y + g.w * unit[3 * pi + 1] - oy) : (R)0;
        s_pos[2][i] = in ? (R)(g.z + g.w * unit[3 * pi + 2] - oz) : (R)0;
      }
      for (int i = t; i < cnt4 * NR; i += P2P_BS) {
        const int q = i / cnt4, jj = i - q * cnt4;
        s_den[jj][q] = jj < cnt && q < nrv ? (R)buf[(long)q * pln + b0 + jj] : (R)0;
      }
      __syncthreads();
      if constexpr (F32) {
        if (ng == 4)
          p2p_tile32<4, 1, TL>(cnt4, kk, col, (const float(*)[TL])s_pos, (const float(*)[16])s_den, (const float*)tx,
                        (const float*)ty, (const float*)tz, acc64);
        else if (ng == 3)
          p2p_tile32<3, 1, TL>(cnt4, kk, col, (const float(*)[TL])s_pos, (const float(*)[16])s_den, (const float*)tx,
                        (const float*)ty, (const float*)tz, acc64);
        else if (ng == 2)
          p2p_tile32<2, 1, TL>(cnt4, kk, col, (const float(*)[TL])s_pos, (const float(*)[16])s_den, (const float*)tx,
                        (const float*)ty, (const float*)tz, acc64);
        else if (ng == 1)
          p2p_tile32<1, 1, TL>(cnt4, kk, col, (const float(*)[TL])s_pos, (const float(*)[16])s_den, (const float*)tx,
                        (const float*)ty, (const float*)tz, acc64);
      } else {
        if (ng == 4)
          p2p_tile64<4>(cnt4, kk, col, (const double(*)[TILE])s_pos, (const double(*)[16])s_den, (const double*)tx,
                        (const double*)ty, (const double*)tz, accm);
        else if (ng == 3)
          p2p_tile64<3>(cnt4, kk, col, (const double(*)[TILE])s_pos, (const double(*)[16])s_den, (const double*)tx,
                        (const double*)ty, (const double*)tz, accm);
        else if (ng == 2)
          p2p_tile64<2>(cnt4, kk, col, (const double(*)[TILE])s_pos, (const double(*)[16])s_den, (const double*)tx,
                        (const double*)ty, (const double*)tz, accm);
        else if (ng == 1)
          p2p_tile64<1>(cnt4, kk, col, (const double(*)[TILE])s_pos, (const double(*)[16])s_den, (const double*)tx,
                        (const double*)ty, (const double*)tz, accm);
      }
    }
  }
  if (!wact) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int tt = 64 * w + 16 * i + col;
    if (tt >= tn) continue;
    if constexpr (F32) {  // FP32 item v of lane l: right-hand side 4*(l/16) + v
      double* val = trg.val + (long)(r0 + 4 * kk) * trg.val_rs + tb + tt;
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (4 * kk + v < nrv) val[(long)v * trg.val_rs] += kc.a * acc64[i][v];
    } else {  // FP64 item v of lane l: right-hand side l/16 + 4v
      double* val = trg.val + (long)(r0 + kk) * trg.val_rs + tb + tt;
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (kk + 4 * v < nrv) val[(long)(4 * v) * trg.val_rs] += kc.a * accm[i][v];
    }
  }
}

template <int KT>
static void s2p_kt(hipStream_t s, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int* enode,
                   const int* ekind, const double4* egeom, const double* unit_de, const double* unit_ue, int nsp,
                   const double* dneden, const double* upeden, int Rc, int pln, Targets trg, int nr, KC kc,
                   int variant)
{
  int r0 = 0;
  if (KT == 111 && variant == 3) {
    r0 = nr - rhs_blocks16(nr, [&](int nb, int roff, int nrv) {
      hipLaunchKernelGGL((k_s2p_mfma111<true>), dim3(xcd_grid(nchunk), nb), dim3(P2P_BS), 0, s, nchunk, tbeg, tnum,
                         eptr, enode, ekind, egeom, unit_de, unit_ue, nsp, dneden, upeden, Rc, pln, trg, roff, kc, nrv);
    });
  }
  rhs_blocks(nr - r0, [&](int NR, int nb, int roff0) {
    const int roff = r0 + roff0;
    dim3 grid(nchunk, nb);
    if (KT == 111 && NR == 16 && variant == 2)
      hipLaunchKernelGGL((k_s2p_mfma111<false>), dim3(xcd_grid(nchunk), nb), dim3(P2P_BS), 0, s, nchunk, tbeg, tnum,
                         eptr, enode, ekind, egeom, unit_de, unit_ue, nsp, dneden, upeden, Rc, pln, trg, roff, kc, 16);
    else if (NR == 16)
      hipLaunchKernelGGL((k_s2p<KT, 16>), grid, dim3(P2P_BS), 0, s, tbeg, tnum, eptr, enode, ekind, egeom, unit_de,
                         unit_ue, nsp, dneden, upeden, Rc, pln, trg, roff, kc);
    else if (NR == 4)
      hipLaunchKernelGGL((k_s2p<KT, 4>), grid, dim3(P2P_BS), 0, s, tbeg, tnum, eptr, enode, ekind, egeom, unit_de,
                         unit_ue, nsp, dneden, upeden, Rc, pln, trg, roff, kc);
    else
      hipLaunchKernelGGL((k_s2p<KT, 1>), grid, dim3(P2P_BS), 0, s, tbeg, tnum, eptr, enode, ekind, egeom, unit_de,
                         unit_ue, nsp, dneden, upeden, Rc, pln, trg, roff, kc);
  });
}

void s2p(hipStream_t s, int kt, int nchunk, const int* tbeg, const int* tnum, const int* eptr, const int* enode,
         const int* ekind, const double4* egeom, const double* unit_de, const double* unit_ue, int nsp,
         const double* dneden, const double* upeden, int Rc, int pln, Targets trg, int nr, KC kc, int variant)
{
  if (nchunk <= 0) return;
  switch (kt) {
  case 111:
    s2p_kt<111>(s, nchunk, tbeg, tnum, eptr, enode, ekind, egeom, unit_de, unit_ue, nsp, dneden, upeden, Rc, pln, trg,
                nr, kc, variant);
    break;
  case 301:
    s2p_kt<301>(s, nchunk, tbeg, tnum, eptr, enode, ekind, egeom, unit_de, unit_ue, nsp, dneden, upeden, Rc, pln, trg,
                nr, kc, variant);
    break;
  case 511:
    s2p_kt<511>(s, nchunk, tbeg, tnum, eptr, enode, ekind, egeom, unit_de, unit_ue, nsp, dneden, upeden, Rc, pln, trg,
                nr, kc, variant);
    break;
  }
}

// ---------------------------------------------------------------------------
// batched indexed FP64 GEMM: 64 x 64 tile of C per block, K in steps of 16.
//
// These products feed or apply the truncated pseudoinverses, whose singular
// values run down to 0.1^(np+1) of the largest: y = P x cancels massively, so
// the rounding of a plain FP64 dot product depends on the summation order at
// the 1e-11 (np 6) to 1e-9 (np 8) level of the final potential.  Each output
// is therefore accumulated with the error-free transformations TwoProd (FMA)
// and TwoSum (Knuth), i.e. the compensated dot product "Dot2" of Ogita, Rump
// and Oishi (SIAM J. Sci. Comput. 26, 2005), and rounded once at the end.
__device__ inline void dot2_acc(double& hi, double& lo, double a, double b)
{
#pragma clang fp contract(off)
  double p = a * b;
  double e = fma(a, b, -p);
  double s = hi + p;
  double bb = s - hi;
  double t = (hi - (s - bb)) + (p - bb);
  hi = s;
  lo += t + e;
}

// (o + hi + lo) rounded once (up to the last TwoSum)
__device__ inline double dot2_finish(double o, double hi, double lo)
{
#pragma clang fp contract(off)
  double s = o + hi;
  double bb = s - o;
  double t = (o - (s - bb)) + (hi - bb);
  return s + (t + lo);
}

template <bool DOT2>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs p)
{
  const int grp = blockIdx.z;
  const int n0 = p.gbeg[grp], n1 = p.gbeg[grp + 1];
  const long cb = (long)n0 * p.Rc + (long)blockIdx.x * 64, ce = (long)n1 * p.Rc;
  if (cb >= ce) return;
  const int rb = blockIdx.y * 64;
  const double* A = p.A + grp * p.a_gstride;
  const int K = p.nseg * p.Kseg;
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  double hi[4][4], lo[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) hi[i][j] = lo[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int e = t + 256 * i, row = e & 63, kk = e >> 6;
      int gr = rb + row, gk = k0 + kk;
      As[kk][row] = (gr < p.M && gk < K) ? A[(long)gk * p.M + gr] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int e = t + 256 * i, kk = e & 15, cl = e >> 4;
      int gk = k0 + kk;
      long j = cb + cl;
      double v = 0.0;
      if (j < ce && gk < K) {
        int n = (int)(j / p.Rc), r = (int)(j - (long)n * p.Rc);
        int seg = gk / p.Kseg, kin = gk - seg * p.Kseg;
        int bn = p.bnode[(long)seg * p.ncol + n];
        if (bn >= 0) {
          v = p.B[((long)bn * p.Rc + r) * p.Kseg + kin];
          if (p.bscale) v *= p.bscale[kin];  // power of two: exact
        }
      }
      Bs[kk][cl] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (DOT2)
            dot2_acc(hi[i][j], lo[i][j], a[i], b[j]);
          else
            hi[i][j] = fma(a[i], b[j], hi[i][j]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    long jj = cb + ty + 16 * j;
    if (jj >= ce) continue;
    int n = (int)(jj / p.Rc), r = (int)(jj - (long)n * p.Rc);
    double* out = p.C + ((long)p.cnode[n] * p.Rc + r) * p.M;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int row = rb + tx + 16 * i;
      if (row >= p.M) continue;
      double h = hi[i][j], l = lo[i][j];
      if (p.cscale) {  // power of two: exact
        h *= p.cscale[row];
        l *= p.cscale[row];
      }
      if (DOT2)
        out[row] = dot2_finish(p.accumulate ? out[row] : 0.0, h, l);
      else
        out[row] = p.accumulate ? out[row] + h : h;
    }
  }
}

// Compensated (Dot2) product, staged like k_gemm_mx: 64 x 64 tile of C per
// block, each thread a 4 x 4 block of (hi, lo) accumulators; K in steps of 8
// (segments are multiples of 8), the A and B tiles double-buffered in LDS with
// the next step's values loaded into registers while the current step is
// accumulated; each thread's staging addresses are fixed per segment (no
// per-element index arithmetic).  One barrier per step.
__global__ __launch_bounds__(256) void k_gemm_dot2(GemmArgs p)
{
  const int grp = blockIdx.z;
  const int n0 = p.gbeg[grp], n1 = p.gbeg[grp + 1];
  const long cb = (long)n0 * p.Rc + (long)blockIdx.x * 64, ce = (long)n1 * p.Rc;
  if (cb >= ce) return;  // block-uniform
  const int rb = blockIdx.y * 64;
  const double* A = p.A + grp * p.a_gstride;
  const int nstep = p.nseg * p.Kseg / 8, sps = p.Kseg / 8;
  __shared__ double As[2][8][64];
  __shared__ double Bs[2][8][65];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  // staging: A elements (kk = e >> 6, row = e & 63), B elements (kk = e & 7,
  // column e >> 3) for e = t and t + 256
  int akk[2], arow[2], bkk[2], bcl[2], bn_n[2], bn_r[2];
  bool aok[2], bok[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const int e = t + 256 * i;
    akk[i] = e >> 6;
    arow[i] = e & 63;
    aok[i] = rb + arow[i] < p.M;
    bkk[i] = e & 7;
    bcl[i] = e >> 3;
    const long j = cb + bcl[i];
    bok[i] = j < ce;
    bn_n[i] = bok[i] ? (int)(j / p.Rc) : 0;
    bn_r[i] = bok[i] ? (int)(j - (long)bn_n[i] * p.Rc) : 0;
  }
  const double* bcol[2];
  double bmul[2];
  auto seg_ptrs = [&](int seg) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int bn = bok[i] ? p.bnode[(long)seg * p.ncol + bn_n[i]] : -1;
      bcol[i] = p.B + (bn >= 0 ? ((long)bn * p.Rc + bn_r[i]) * p.Kseg : 0) + bkk[i];
      bmul[i] = bn >= 0 ? 1.0 : 0.0;
    }
  };
  double ra[2], rbv[2];
  auto load = [&](int st) {
    const int k0 = st * 8, kin = (st - (st / sps) * sps) * 8;
#pragma unroll
    for (int i = 0; i < 2; i++) {
      ra[i] = aok[i] ? A[(long)(k0 + akk[i]) * p.M + rb + arow[i]] : 0.0;
      double v = bcol[i][kin] * bmul[i];
      if (p.bscale) v *= p.bscale[kin + bkk[i]];  // power of two: exact
      rbv[i] = v;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      As[buf][akk[i]][arow[i]] = ra[i];
      Bs[buf][bkk[i]][bcl[i]] = rbv[i];
    }
  };
  double hi[4][4], lo[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) hi[i][j] = lo[i][j] = 0.0;
  seg_ptrs(0);
  load(0);
  store(0);
  __syncthreads();
  for (int st = 0; st < nstep; st++) {
    const int buf = st & 1;
    const bool more = st + 1 < nstep;
    if (more) {
      if ((st + 1) % sps == 0) seg_ptrs((st + 1) / sps);
      load(st + 1);
    }
#pragma unroll 2
    for (int kk = 0; kk < 8; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[buf][kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[buf][kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) dot2_acc(hi[i][j], lo[i][j], a[i], b[j]);
    }
    if (more) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const long jj = cb + ty + 16 * j;
    if (jj >= ce) continue;
    const int n = (int)(jj / p.Rc), r = (int)(jj - (long)n * p.Rc);
    double* out = p.C + ((long)p.cnode[n] * p.Rc + r) * p.M;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = rb + tx + 16 * i;
      if (row >= p.M) continue;
      double h = hi[i][j], l = lo[i][j];
      if (p.cscale) {  // power of two: exact
        h *= p.cscale[row];
        l *= p.cscale[row];
      }
      out[row] = dot2_finish(p.accumulate ? out[row] : 0.0, h, l);
    }
  }
}

// Plain (uncompensated) form on the FP64 matrix cores: the same 64 x 64 tile
// and operand staging; wave w owns the 32 x 32 quarter (rows 32*(w/2), columns
// 32*(w%2)) as 2 x 2 tiles of v_mfma_f64_16x16x4f64 (A lane l: row l%16, k
// l/16; B lane l: k l/16, column l%16; D item v of lane l: row l/16 + 4v).
__global__ __launch_bounds__(256) void k_gemm_mfma(GemmArgs p)
{
  const int grp = blockIdx.z;
  const int n0 = p.gbeg[grp], n1 = p.gbeg[grp + 1];
  const long cb = (long)n0 * p.Rc + (long)blockIdx.x * 64, ce = (long)n1 * p.Rc;
  if (cb >= ce) return;
  const int rb = blockIdx.y * 64;
  const double* A = p.A + grp * p.a_gstride;
  const int K = p.nseg * p.Kseg;
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int e = t + 256 * i, row = e & 63, kk = e >> 6;
      int gr = rb + row, gk = k0 + kk;
      As[kk][row] = (gr < p.M && gk < K) ? A[(long)gk * p.M + gr] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int e = t + 256 * i, kk = e & 15, cl = e >> 4;
      int gk = k0 + kk;
      long j = cb + cl;
      double v = 0.0;
      if (j < ce && gk < K) {
        int n = (int)(j / p.Rc), r = (int)(j - (long)n * p.Rc);
        int seg = gk / p.Kseg, kin = gk - seg * p.Kseg;
        int bn = p.bnode[(long)seg * p.ncol + n];
        if (bn >= 0) {
          v = p.B[((long)bn * p.Rc + r) * p.Kseg + kin];
          if (p.bscale) v *= p.bscale[kin];  // power of two: exact
        }
      }
      Bs[kk][cl] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const int kk = ks * 4 + lk;
      const double a0 = As[kk][wr + li], a1 = As[kk][wr + 16 + li];
      const double b0 = Bs[kk][wc + li], b1 = Bs[kk][wc + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    long jj = cb + wc + 16 * j + li;
    if (jj >= ce) continue;
    int n = (int)(jj / p.Rc), r = (int)(jj - (long)n * p.Rc);
    double* out = p.C + ((long)p.cnode[n] * p.Rc + r) * p.M;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row = rb + wr + 16 * i + lk + 4 * v;
        if (row >= p.M) continue;
        double h = acc[i][j][v];
        if (p.cscale) h *= p.cscale[row];  // power of two: exact
        out[row] = p.accumulate ? out[row] + h : h;
      }
  }
}

// Plain product on the FP64 matrix cores, wide form.  A block covers RT x 16
// rows (every row of an operator of up to 160 rows) and 64 G columns; wave w
// owns columns [16 G w, 16 G (w + 1)) of the block as RT x G tiles of
// v_mfma_f64_16x16x4f64.  K runs in steps of 8 -- every surface length
// nS(n) x dof is a multiple of 8, so a step never crosses a segment.  The A
// tile of a step (RT*16 x 8, with the power-of-two B scale of its K rows
// folded in) is double-buffered in LDS; each lane loads the two B values it
// feeds straight from its column (contiguous in K) PB steps ahead.  Lane
// (k, col) feeds B[k0 + 2k + s][col] to MFMA slice s, so slice s reads A
// rows at K index k0 + 2k + s.  One barrier per step.
//
// With ks > 1 the K steps are split over ks blocks (blockIdx.z = group x ks +
// split), each writing its raw partial product to work[split][column][row];
// k_gemm_reduce then sums the partials in split order, so the result does not
// depend on scheduling.
template <int RT, int G, int PB = 1>
__global__ __launch_bounds__(256) void k_gemm_mx(GemmArgs p, int ks)
{
  constexpr int MR = RT * 16, NAT = (8 * MR + 255) / 256;
  const int grp = blockIdx.z / ks, sp = blockIdx.z - grp * ks;
  const int n0 = p.gbeg[grp], n1 = p.gbeg[grp + 1];
  const long cb = (long)n0 * p.Rc + (long)blockIdx.x * 64 * G, ce = (long)n1 * p.Rc;
  if (cb >= ce) return;  // block-uniform
  const int rb = blockIdx.y * MR;
  const double* A = p.A + grp * p.a_gstride;
  const int nstep = p.nseg * p.Kseg / 8, sps = p.Kseg / 8;  // steps, steps per segment
  const int st0 = (int)((long)nstep * sp / ks), st1 = (int)((long)nstep * (sp + 1) / ks);
  __shared__ double sA[2][8][MR];
  const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
  int cn[G], cr[G];
  bool cok[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const long j = cb + (long)(w * G + g) * 16 + li;
    cok[g] = j < ce;
    cn[g] = cok[g] ? (int)(j / p.Rc) : 0;
    cr[g] = cok[g] ? (int)(j - (long)cn[g] * p.Rc) : 0;
  }
  // B column pointers of segment seg (clamped to column 0 of B, masked by bm)
  const double* bp[G];
  double bm[G];
  auto seg_ptrs = [&](int seg) {
    const double* Bb = (seg > 0 && p.B2) ? p.B2 : p.B;
#pragma unroll
    for (int g = 0; g < G; g++) {
      const int bn = cok[g] ? p.bnode[(long)seg * p.ncol + cn[g]] : -1;
      bp[g] = Bb + (bn >= 0 ? ((long)bn * p.Rc + cr[g]) * p.Kseg : 0) + 2 * lk;
      bm[g] = bn >= 0 ? 1.0 : 0.0;
    }
  };
  double ar[NAT];
  auto load_a = [&](int st) {
    const int k0 = st * 8;
#pragma unroll
    for (int i = 0; i < NAT; i++) {
      const int e = t + 256 * i, kk = e / MR, row = e - kk * MR, gk = k0 + kk;
      double v = 0.0;
      if (e < 8 * MR && rb + row < p.M) {
        v = A[(long)gk * p.M + rb + row];
        if (p.bscale && gk >= p.bscale_seg * p.Kseg) v *= p.bscale[gk - (gk / p.Kseg) * p.Kseg];  // power of two: exact
      }
      ar[i] = v;
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NAT; i++) {
      const int e = t + 256 * i, kk = e / MR, row = e - kk * MR;
      if (e < 8 * MR) sA[buf][kk][row] = ar[i];
    }
  };
  // B values of steps st .. st + PB - 1 in flight (bq[0] is step st's);
  // the column pointers follow the step being loaded
  double2 bq[PB][G];
  auto load_b = [&](int st, double2* dst) {
    const int kin = (st - (st / sps) * sps) * 8;
#pragma unroll
    for (int g = 0; g < G; g++) {
      const double2 v = *(const double2*)(bp[g] + kin);
      dst[g] = make_double2(v.x * bm[g], v.y * bm[g]);
    }
  };
  v4d acc[RT][G];
#pragma unroll
  for (int rt = 0; rt < RT; rt++)
#pragma unroll
    for (int g = 0; g < G; g++) acc[rt][g] = v4d{0.0, 0.0, 0.0, 0.0};
  seg_ptrs(st0 / sps);
  load_a(st0);
#pragma unroll
  for (int d = 0; d < PB; d++) {
    const int ls = st0 + d;
    if (ls < st1) {
      if (d > 0 && ls % sps == 0) seg_ptrs(ls / sps);
      load_b(ls, bq[d]);
    }
  }
  store_a(0);
  __syncthreads();
  for (int st = st0; st < st1; st++) {
    const int buf = (st - st0) & 1;
    const bool more = st + 1 < st1;
    if (more) load_a(st + 1);
    double2 bn[G];
    const int ls = st + PB;  // the step whose B values are loaded now
    if (ls < st1) {
      if (ls % sps == 0) seg_ptrs(ls / sps);
      load_b(ls, bn);
    }
#pragma unroll
    for (int sl = 0; sl < 2; sl++)
#pragma unroll
      for (int rt = 0; rt < RT; rt++) {
        const double a = sA[buf][2 * lk + sl][rt * 16 + li];
#pragma unroll
        for (int g = 0; g < G; g++)
          acc[rt][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sl ? bq[0][g].y : bq[0][g].x, acc[rt][g], 0, 0, 0);
      }
    if (more) store_a(buf ^ 1);
#pragma unroll
    for (int d = 0; d + 1 < PB; d++)
#pragma unroll
      for (int g = 0; g < G; g++) bq[d][g] = bq[d + 1][g];
    if (ls < st1)
#pragma unroll
      for (int g = 0; g < G; g++) bq[PB - 1][g] = bn[g];
    __syncthreads();
  }
#pragma unroll
  for (int g = 0; g < G; g++) {
    if (!cok[g]) continue;
    const long jj = cb + (long)(w * G + g) * 16 + li;
    double* out = ks > 1 ? p.work + ((long)sp * p.ctot + jj) * p.M
                         : p.C + ((long)p.cnode[cn[g]] * p.Rc + cr[g]) * p.M;
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = rb + rt * 16 + lk + 4 * v;
        if (row >= p.M) continue;
        double h = acc[rt][g][v];
        if (ks > 1) {
          out[row] = h;
          continue;
        }
        if (p.cscale) h *= p.cscale[row];  // power of two: exact
        out[row] = p.accumulate ? out[row] + h : h;
      }
  }
}

__global__ __launch_bounds__(256) void k_gemm_reduce(GemmArgs p, int ks)
{
  const long n = p.ctot * p.M;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const long jj = i / p.M;
    const int row = (int)(i - jj * p.M);
    double h = 0.0;
    for (int sp = 0; sp < ks; sp++) h += p.work[(long)sp * n + i];
    if (p.cscale) h *= p.cscale[row];  // power of two: exact
    const int cn = (int)(jj / p.Rc), cr = (int)(jj - (long)cn * p.Rc);
    double* out = p.C + ((long)p.cnode[cn] * p.Rc + cr) * p.M + row;
    *out = p.accumulate ? *out + h : h;
  }
}

// wide tile: GMX_RT 16-row tiles x GMX_G 16-column tiles per wave (measured
// best of 10x2, 10x1 and 5x4 on the 4M-point Laplace and 1M-point Navier runs)
constexpr int GMX_RT = 10, GMX_G = 2;

// Splits of K that bring a wide-tile product to about a.split_blocks blocks
// (default 1024, 4 per CU; 1 = no split; each split keeps at least 4 steps of 8).
static int gemm_split_of(const GemmArgs& a, long nb)
{
  const long target = a.split_blocks > 0 ? a.split_blocks : 1024;
  if (nb >= target / 2) return 1;
  const int nstep = a.nseg * a.Kseg / 8;
  long ks = (target + nb - 1) / nb;
  ks = std::min<long>(ks, std::min(16, nstep / 4));
  return (int)std::max<long>(ks, 1);
}

int gemm_split(const GemmArgs& a)
{
  if (a.compensated || a.mfma != 2 || a.Kseg % 8 != 0 || a.ngroup <= 0 || a.maxcols <= 0) return 1;
  const long cols = (long)a.maxcols * a.Rc;
  const long nb = a.M <= 64 ? ((cols + 255) / 256) * a.ngroup
                            : ((cols + 64 * GMX_G - 1) / (64 * GMX_G)) * ((a.M + 16 * GMX_RT - 1) / (16 * GMX_RT)) * a.ngroup;
  return gemm_split_of(a, nb);
}

void gemm(hipStream_t s, const GemmArgs& a)
{
  if (a.ngroup <= 0 || a.maxcols <= 0) return;
  long cols = (long)a.maxcols * a.Rc;
  // wide tiles, with K split over blocks (partials in a.work) when the
  // columns alone give fewer than 2 blocks per CU; 64 x 64 tiles without a
  // workspace
  if (!a.compensated && a.mfma == 2 && a.Kseg % 8 == 0) {
    int ks = gemm_split(a);
    if (a.B2 && !(a.work && (long)ks * a.ctot * a.M <= a.work_n)) ks = 1;  // two B buffers: wide tiles only
    if (ks == 1 || (a.work && (long)ks * a.ctot * a.M <= a.work_n)) {
      if (a.M <= 64) {
        dim3 grid((unsigned)((cols + 255) / 256), 1, a.ngroup * ks);
        hipLaunchKernelGGL((k_gemm_mx<4, 4>), grid, dim3(256), 0, s, a, ks);
      } else {
        dim3 grid((unsigned)((cols + 64 * GMX_G - 1) / (64 * GMX_G)), (a.M + 16 * GMX_RT - 1) / (16 * GMX_RT), a.ngroup * ks);
        // B values 2 steps ahead (4M points / 16 RHS, mixed: up 2.35 -> 2.2 ms,
        // down 1.89 -> 1.76 ms; 3 ahead: 3.3 and 2.6 ms); KIFMM3D_GEMM_PB = 1, 2, 3
        static const int pb = [] {
          const char* v = std::getenv("KIFMM3D_GEMM_PB");
          return v ? std::atoi(v) : 2;
        }();
        if (pb == 2)
          hipLaunchKernelGGL((k_gemm_mx<GMX_RT, GMX_G, 2>), grid, dim3(256), 0, s, a, ks);
        else if (pb == 3)
          hipLaunchKernelGGL((k_gemm_mx<GMX_RT, GMX_G, 3>), grid, dim3(256), 0, s, a, ks);
        else
          hipLaunchKernelGGL((k_gemm_mx<GMX_RT, GMX_G>), grid, dim3(256), 0, s, a, ks);
      }
      if (ks > 1) {
        const long n = a.ctot * a.M;
        hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, s, a, ks);
      }
      return;
    }
  }
  dim3 grid((unsigned)((cols + 63) / 64), (a.M + 63) / 64, a.ngroup);
  if (a.compensated && a.Kseg % 8 == 0)
    hipLaunchKernelGGL(k_gemm_dot2, grid, dim3(256), 0, s, a);
  else if (a.compensated)
    hipLaunchKernelGGL(k_gemm<true>, grid, dim3(256), 0, s, a);
  else if (a.mfma)
    hipLaunchKernelGGL(k_gemm_mfma, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_gemm<false>, grid, dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// FFTs on the (2np)^3 grid in FFTW 2 rfftwnd layout ([z][y][x], x halved).
// Only the np^3 cube at offset np/2 holds nonzero data (matmgnt3d.cpp:256,
// :286), so the forward pass sums over np points per dimension and the
// inverse evaluates only the surface points.
//
// Spectra are stored per slot either right-hand-side major ([slot][r][f][C],
// rhs_inner = 0) or right-hand-side innermost ([slot][f][C][Rc], rhs_inner =
// 1, the layout of the tiled matrix-core M2L); CT is double2 or float2.

// Arithmetic of the transforms: FP64 for FP64 spectra, FP32 for FP32 spectra
// (the spectra are rounded to FP32 for the M2L products either way).
template <class CT> struct FftMath { using A = double; using A2 = double2; };
template <> struct FftMath<float2> { using A = float; using A2 = float2; };
template <class T2> __device__ inline T2 cz()
{
  T2 v;
  v.x = 0;
  v.y = 0;
  return v;
}
template <class T2> __device__ inline T2 cmul2(T2 a, T2 b)
{
  T2 v;
  v.x = a.x * b.x - a.y * b.y;
  v.y = a.x * b.y + a.y * b.x;
  return v;
}
template <class T2> __device__ inline T2 cmulc(T2 a, T2 b)  // a * conj(b)
{
  T2 v;
  v.x = a.x * b.x + a.y * b.y;
  v.y = a.y * b.x - a.x * b.y;
  return v;
}
template <class D, class S> __device__ inline D cget(const S& s)
{
  D d;
  d.x = s.x;
  d.y = s.y;
  return d;
}
template <class D, class S> __device__ inline void cset(D& d, S s)
{
  d.x = s.x;
  d.y = s.y;
}
template <int NP, class A2>
__device__ inline void load_twiddles_t(A2 (*s_wt)[2 * NP], const double2* tw, int t, int nt)
{
  constexpr int N = 2 * NP;
  for (int e = t; e < NP * N; e += nt) {
    const int i = e / N, m = e - i * N;
    s_wt[i][m] = cget<A2>(tw[((i + NP / 2) * m) % N]);
  }
}


// Half-length DFT stages.  The twiddle of input i (placed at i + NP/2 on the
// grid of N = 2 NP points) and output m + NP is the one of output m times
// (-1)^(i + NP/2), since w(t + N/2) = -w(t).  So the N outputs of NP inputs are
// P + Q at m and P - Q at m + NP, with P (Q) the sum over the inputs of even
// (odd) i + NP/2 of the twiddled values at m: NP^2 complex products instead of
// 2 NP^2.  The adjoint stage folds its N inputs the same way first.
template <class A2>
__device__ __forceinline__ void fft_half_mac(A2 x, A2 w, bool odd, A2& sp, A2& sn)
{
  const A2 v = cmul2(x, w);
  if (odd) {
    sn.x += v.x;
    sn.y += v.y;
  } else {
    sp.x += v.x;
    sp.y += v.y;
  }
}
template <class A2, class LT>
__device__ __forceinline__ void fft_half_out(A2 sp, A2 sn, LT& lo, LT& hi)
{
  A2 a, b;
  a.x = sp.x + sn.x;
  a.y = sp.y + sn.y;
  b.x = sp.x - sn.x;
  b.y = sp.y - sn.y;
  cset(lo, a);
  cset(hi, b);
}
template <class A2>
__device__ __forceinline__ void fft_half_in(A2 lo, A2 hi, A2& s, A2& d)
{
  s.x = lo.x + hi.x;
  s.y = lo.y + hi.y;
  d.x = lo.x - hi.x;
  d.y = lo.y - hi.y;
}

template <int NP, class CT>
__global__ __launch_bounds__(256) void k_fft_fwd(const int* __restrict__ slot_node, int Rc,
                                                 const double* __restrict__ upeden, int pln, int sm,
                                                 const double* __restrict__ dscale, const short* __restrict__ c2s,
                                                 CT* __restrict__ effden, const double2* __restrict__ tw, int rhs_inner)
{
  constexpr int N = 2 * NP, H = NP + 1, C = N * N * H, NC = NP * NP * NP;
  constexpr int NSP = NP * NP * NP - (NP - 2) * (NP - 2) * (NP - 2);
  using A = typename FftMath<CT>::A;
  using A2 = typename FftMath<CT>::A2;
  __shared__ A s_den[NSP * 4];
  __shared__ short s_c2s[NC];
  __shared__ A2 s_w[N];
  __shared__ A2 X1[NP * NP * H];
  __shared__ A2 X2[NP * N * H];
  const int slot = blockIdx.x, r = blockIdx.y, t = threadIdx.x;
  const double* den = upeden + ((long)slot_node[slot] * Rc + r) * pln;
  for (int i = t; i < pln; i += 256) s_den[i] = (A)(den[i] * dscale[i % sm]);
  for (int i = t; i < NC; i += 256) s_c2s[i] = c2s[i];
  if (t < N) s_w[t] = cget<A2>(tw[t]);
  __syncthreads();
  const long os = rhs_inner ? Rc : 1;
  CT* out = effden + (long)slot * Rc * sm * C + (rhs_inner ? (long)r : (long)r * sm * C);
  for (int f = 0; f < sm; f++) {
    for (int o = t; o < NP * NP * H; o += 256) {  // along x: X1[k][j][kx]
      int kx = o % H, j = (o / H) % NP, k = o / (H * NP);
      A2 acc = cz<A2>();
#pragma unroll
      for (int i = 0; i < NP; i++) {
        int si = s_c2s[(i * NP + j) * NP + k];
        if (si >= 0) {
          const A v = s_den[si * sm + f];
          const A2 w = s_w[((i + NP / 2) * kx) % N];
          acc.x += v * w.x;
          acc.y += v * w.y;
        }
      }
      X1[o] = acc;
    }
    __syncthreads();
    for (int o = t; o < NP * N * H; o += 256) {  // along y: X2[k][ky][kx]
      int kx = o % H, ky = (o / H) % N, k = o / (H * N);
      A2 acc = cz<A2>();
#pragma unroll
      for (int j = 0; j < NP; j++) {
        const A2 v = cmul2(X1[(k * NP + j) * H + kx], s_w[((j + NP / 2) * ky) % N]);
        acc.x += v.x;
        acc.y += v.y;
      }
      X2[o] = acc;
    }
    __syncthreads();
    for (int o = t; o < C; o += 256) {  // along z
      int kx = o % H, ky = (o / H) % N, kz = o / (H * N);
      A2 acc = cz<A2>();
#pragma unroll
      for (int k = 0; k < NP; k++) {
        const A2 v = cmul2(X2[(k * N + ky) * H + kx], s_w[((k + NP / 2) * kz) % N]);
        acc.x += v.x;
        acc.y += v.y;
      }
      cset(out[((long)f * C + o) * os], acc);
    }
    __syncthreads();
  }
}

template <int NP, class CT>
__global__ __launch_bounds__(256) void k_fft_inv(const int* __restrict__ slot_node, int Rc,
                                                 const CT* __restrict__ effval, int tm,
                                                 const int* __restrict__ s2c, int nsp, double* __restrict__ dncval,
                                                 int pln, const double2* __restrict__ tw, int rhs_inner)
{
  constexpr int N = 2 * NP, H = NP + 1, C = N * N * H;
  using A = typename FftMath<CT>::A;
  using A2 = typename FftMath<CT>::A2;
  __shared__ A2 s_x[C];           // spectrum, later Y2[k][j][kx]
  __shared__ A2 Y1[NP * N * H];   // [k][ky][kx]
  __shared__ A2 s_w[N];
  const int slot = blockIdx.x, r = blockIdx.y, t = threadIdx.x;
  if (t < N) s_w[t] = cget<A2>(tw[t]);
  double* out = dncval + ((long)slot_node[slot] * Rc + r) * pln;
  const long os = rhs_inner ? Rc : 1;
  const CT* in = effval + (long)slot * Rc * tm * C + (rhs_inner ? (long)r : (long)r * tm * C);
  for (int f = 0; f < tm; f++) {
    __syncthreads();
    for (int o = t; o < C; o += 256) s_x[o] = cget<A2>(in[((long)f * C + o) * os]);
    __syncthreads();
    for (int o = t; o < NP * N * H; o += 256) {  // along z, only the cube's planes
      int kx = o % H, ky = (o / H) % N, k = o / (H * N);
      A2 acc = cz<A2>();
#pragma unroll
      for (int kz = 0; kz < N; kz++) {
        const A2 v = cmulc(s_x[(kz * N + ky) * H + kx], s_w[((k + NP / 2) * kz) % N]);
        acc.x += v.x;
        acc.y += v.y;
      }
      Y1[o] = acc;
    }
    __syncthreads();
    for (int o = t; o < NP * NP * H; o += 256) {  // along y, only the cube's rows
      int kx = o % H, j = (o / H) % NP, k = o / (H * NP);
      A2 acc = cz<A2>();
#pragma unroll
      for (int ky = 0; ky < N; ky++) {
        const A2 v = cmulc(Y1[(k * N + ky) * H + kx], s_w[((j + NP / 2) * ky) % N]);
        acc.x += v.x;
        acc.y += v.y;
      }
      s_x[o] = acc;
    }
    __syncthreads();
    for (int p = t; p < nsp; p += 256) {  // along x (c2r) at the surface points
      int cc = s2c[p];
      int i = cc / (NP * NP), j = (cc / NP) % NP, k = cc % NP;
      const A2* y = &s_x[(k * NP + j) * H];
      int xi = i + NP / 2;
      A v = y[0].x + ((xi & 1) ? -y[NP].x : y[NP].x);
#pragma unroll
      for (int kx = 1; kx < NP; kx++) {
        const A2 w = s_w[(xi * kx) % N];
        v += (A)2 * (y[kx].x * w.x + y[kx].y * w.y);
      }
      out[p * tm + f] = v;
    }
  }
}

// Right-hand-side-innermost forms for scalar kernels (sm = tm = 1).  A block
// of 64 R threads transforms R right-hand sides of one slot.  Every stage of
// the separable DFT gives a thread one (rhs, line) pair: it loads the line's
// inputs into registers once and produces all of the line's outputs, with the
// twiddles w((i + np/2) m) read from a small table as LDS broadcasts, so LDS
// traffic is one read and one write per intermediate value.  Global accesses
// put the R right-hand sides of a frequency next to each other.  Intermediate
// spectra in LDS are FP64 for FP64 spectra, FP32 for FP32 spectra; per-RHS rows
// are padded by one element so the R rows of a wave's accesses fall in
// different banks.

// Blocks (slot, group of R right-hand sides): the groups of one slot share
// the 128-byte lines of its spectra, so they are dealt to one XCD side by side.
__device__ inline bool fft_block(int nslot, int nrg, int R, int& slot, int& r0)
{
  const int nitem = nslot * nrg, item = xcd_item(blockIdx.x, nitem);
  slot = item / nrg;
  r0 = (item - slot * nrg) * R;
  return item < nitem;
}

template <int NP, class CT, int R>
__global__ __launch_bounds__(64 * R) void k_fft_fwd_ri(int nslot, const int* __restrict__ slot_node, int Rc,
                                                       const double* __restrict__ upeden, int pln,
                                                       const double* __restrict__ dscale, const short* __restrict__ c2s,
                                                       CT* __restrict__ effden, const double2* __restrict__ tw)
{
  constexpr int N = 2 * NP, H = NP + 1, C = N * N * H, NC = NP * NP * NP, BS = 64 * R;
  constexpr int NSP = NP * NP * NP - (NP - 2) * (NP - 2) * (NP - 2);
  typedef CT LT;
  using A = typename FftMath<CT>::A;
  using A2 = typename FftMath<CT>::A2;
  __shared__ A s_den[R][NSP];
  __shared__ short s_c2s[NC];
  __shared__ A2 s_wt[NP][N];
  __shared__ LT X1[R][NP * NP * H + 1];  // [k][j][kx]
  __shared__ LT X2[R][NP * N * H + 1];   // [k][ky][kx]
  int slot, r0;
  if (!fft_block(nslot, Rc / R, R, slot, r0)) return;  // block-uniform
  const int t = threadIdx.x;
  for (int i = t; i < NC; i += BS) s_c2s[i] = c2s[i];
  load_twiddles_t<NP>(s_wt, tw, t, BS);
  const double ds = dscale[0];
  const double* den = upeden + ((long)slot_node[slot] * Rc + r0) * pln;
  for (int e = t; e < R * NSP; e += BS) {
    const int rr = e / NSP, i = e - rr * NSP;
    s_den[rr][i] = (A)(den[(long)rr * pln + i] * ds);
  }
  __syncthreads();
  for (int e = t; e < R * NP * NP; e += BS) {  // along x: line (k, j)
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e % R, col = e / R, k = col / NP, j = col - k * NP;
    A2 acc[H];
#pragma unroll
    for (int kx = 0; kx < H; kx++) acc[kx] = cz<A2>();
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const int si = s_c2s[(i * NP + j) * NP + k];
      if (si >= 0) {
        const A v = s_den[rr][si];
#pragma unroll
        for (int kx = 0; kx < H; kx++) {
          acc[kx].x += v * s_wt[i][kx].x;
          acc[kx].y += v * s_wt[i][kx].y;
        }
      }
    }
#pragma unroll
    for (int kx = 0; kx < H; kx++) cset(X1[rr][(k * NP + j) * H + kx], acc[kx]);
  }
  __syncthreads();
  for (int e = t; e < R * NP * H; e += BS) {  // along y: line (k, kx)
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e % R, col = e / R, k = col / H, kx = col - k * H;
    A2 x[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) x[j] = cget<A2>(X1[rr][(k * NP + j) * H + kx]);
#pragma unroll
    for (int ky = 0; ky < NP; ky++) {  // outputs ky and ky + NP together
      A2 sp = cz<A2>(), sn = cz<A2>();
#pragma unroll
      for (int j = 0; j < NP; j++) fft_half_mac(x[j], s_wt[j][ky], ((j + NP / 2) & 1) != 0, sp, sn);
      fft_half_out(sp, sn, X2[rr][(k * N + ky) * H + kx], X2[rr][(k * N + ky + NP) * H + kx]);
    }
  }
  __syncthreads();
  CT* out = effden + (long)slot * C * Rc + r0;
  for (int e = t; e < R * N * H; e += BS) {  // along z: line (ky, kx), right-hand sides innermost
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e % R, col = e / R;
    A2 x[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) x[k] = cget<A2>(X2[rr][k * N * H + col]);
#pragma unroll
    for (int kz = 0; kz < NP; kz++) {  // outputs kz and kz + NP together
      A2 sp = cz<A2>(), sn = cz<A2>();
#pragma unroll
      for (int k = 0; k < NP; k++) fft_half_mac(x[k], s_wt[k][kz], ((k + NP / 2) & 1) != 0, sp, sn);
      fft_half_out(sp, sn, out[(long)(kz * N * H + col) * Rc + rr], out[(long)((kz + NP) * N * H + col) * Rc + rr]);
    }
  }
}

// The forward transform with the x and y stages fused: a thread takes the
// line (k, kx) of one right-hand side, sums the x direction of the np rows j
// of plane k straight into registers (cube points off the surface read a zero
// kept at index NSP of the density row) and transforms them along y at once,
// so the x-stage values never pass through LDS: one barrier fewer, and at np 6
// 38 KB of LDS per 8 FP32 right-hand sides instead of 54 KB.
template <int NP, class CT, int R>
__global__ __launch_bounds__(64 * R) void k_fft_fwd_xy(int nslot, const int* __restrict__ slot_node, int Rc,
                                                             const double* __restrict__ upeden, int pln,
                                                             const double* __restrict__ dscale,
                                                             const short* __restrict__ c2s, CT* __restrict__ effden,
                                                             const double2* __restrict__ tw)
{
  constexpr int N = 2 * NP, H = NP + 1, C = N * N * H, NC = NP * NP * NP, BS = 64 * R;
  constexpr int NSP = NP * NP * NP - (NP - 2) * (NP - 2) * (NP - 2);
  typedef CT LT;
  using A = typename FftMath<CT>::A;
  using A2 = typename FftMath<CT>::A2;
  __shared__ A s_den[R][NSP + 1];
  __shared__ short s_c2s[NC];
  __shared__ A2 s_wt[NP][N];
  __shared__ LT X2[R][NP * N * H + 1];  // [k][ky][kx]
  int slot, r0;
  if (!fft_block(nslot, Rc / R, R, slot, r0)) return;  // block-uniform
  const int t = threadIdx.x;
  for (int i = t; i < NC; i += BS) {
    const short v = c2s[i];
    s_c2s[i] = v >= 0 ? v : (short)NSP;
  }
  load_twiddles_t<NP>(s_wt, tw, t, BS);
  const double ds = dscale[0];
  const double* den = upeden + ((long)slot_node[slot] * Rc + r0) * pln;
  for (int e = t; e < R * NSP; e += BS) {
    const int rr = e / NSP, i = e - rr * NSP;
    s_den[rr][i] = (A)(den[(long)rr * pln + i] * ds);
  }
  if (t < R) s_den[t][NSP] = 0;
  __syncthreads();
  for (int e = t; e < R * NP * H; e += BS) {  // along x and y: line (k, kx)
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e % R, col = e / R, k = col / H, kx = col - k * H;
    A2 x[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) {
      A2 a = cz<A2>();
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const A v = s_den[rr][s_c2s[(i * NP + j) * NP + k]];
        a.x += v * s_wt[i][kx].x;
        a.y += v * s_wt[i][kx].y;
      }
      x[j] = a;
    }
#pragma unroll
    for (int ky = 0; ky < NP; ky++) {  // outputs ky and ky + NP together
      A2 sp = cz<A2>(), sn = cz<A2>();
#pragma unroll
      for (int j = 0; j < NP; j++) fft_half_mac(x[j], s_wt[j][ky], ((j + NP / 2) & 1) != 0, sp, sn);
      fft_half_out(sp, sn, X2[rr][(k * N + ky) * H + kx], X2[rr][(k * N + ky + NP) * H + kx]);
    }
  }
  __syncthreads();
  CT* out = effden + (long)slot * C * Rc + r0;
  for (int e = t; e < R * N * H; e += BS) {  // along z: line (ky, kx), right-hand sides innermost
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e % R, col = e / R;
    A2 x[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) x[k] = cget<A2>(X2[rr][k * N * H + col]);
#pragma unroll
    for (int kz = 0; kz < NP; kz++) {  // outputs kz and kz + NP together
      A2 sp = cz<A2>(), sn = cz<A2>();
#pragma unroll
      for (int k = 0; k < NP; k++) fft_half_mac(x[k], s_wt[k][kz], ((k + NP / 2) & 1) != 0, sp, sn);
      fft_half_out(sp, sn, out[(long)(kz * N * H + col) * Rc + rr], out[(long)((kz + NP) * N * H + col) * Rc + rr]);
    }
  }
}

// BS threads per block (64 R by default).
template <int NP, class CT, int R, int BS = 64 * R>
__global__ __launch_bounds__(BS) void k_fft_inv_ri(int nslot, const int* __restrict__ slot_node, int Rc,
                                                       const CT* __restrict__ effval, const int* __restrict__ s2c,
                                                       int nsp, double* __restrict__ dncval, int pln,
                                                       const double2* __restrict__ tw)
{
  constexpr int N = 2 * NP, H = NP + 1, C = N * N * H;
  typedef CT LT;
  using A = typename FftMath<CT>::A;
  using A2 = typename FftMath<CT>::A2;
  __shared__ A2 s_wt[NP][N];
  __shared__ LT Y1[R][NP * N * H + 1];   // [k][ky][kx], only the cube's planes
  __shared__ LT Y2[R][NP * NP * H + 1];  // [k][j][kx], only the cube's rows
  int slot, r0;
  if (!fft_block(nslot, Rc / R, R, slot, r0)) return;  // block-uniform
  const int t = threadIdx.x;
  load_twiddles_t<NP>(s_wt, tw, t, BS);
  __syncthreads();
  const CT* in = effval + (long)slot * C * Rc + r0;
  for (int e = t; e < R * N * H; e += BS) {  // along z: line (ky, kx), spectra read once from HBM
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e % R, col = e / R;
    A2 xs[NP], xd[NP];  // x[kz] + x[kz + NP], x[kz] - x[kz + NP]
#pragma unroll
    for (int kz = 0; kz < NP; kz++)
      fft_half_in(cget<A2>(in[(long)(kz * N * H + col) * Rc + rr]),
                  cget<A2>(in[(long)((kz + NP) * N * H + col) * Rc + rr]), xs[kz], xd[kz]);
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const A2* x = ((k + NP / 2) & 1) ? xd : xs;
      A2 acc = cz<A2>();
#pragma unroll
      for (int kz = 0; kz < NP; kz++) {
        const A2 v = cmulc(x[kz], s_wt[k][kz]);
        acc.x += v.x;
        acc.y += v.y;
      }
      cset(Y1[rr][k * N * H + col], acc);
    }
  }
  __syncthreads();
  for (int e = t; e < R * NP * H; e += BS) {  // along y: line (k, kx)
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e % R, col = e / R, k = col / H, kx = col - k * H;
    A2 ys[NP], yd[NP];  // y[ky] + y[ky + NP], y[ky] - y[ky + NP]
#pragma unroll
    for (int ky = 0; ky < NP; ky++)
      fft_half_in(cget<A2>(Y1[rr][(k * N + ky) * H + kx]), cget<A2>(Y1[rr][(k * N + ky + NP) * H + kx]), ys[ky],
                  yd[ky]);
#pragma unroll
    for (int j = 0; j < NP; j++) {
      const A2* y = ((j + NP / 2) & 1) ? yd : ys;
      A2 acc = cz<A2>();
#pragma unroll
      for (int ky = 0; ky < NP; ky++) {
        const A2 v = cmulc(y[ky], s_wt[j][ky]);
        acc.x += v.x;
        acc.y += v.y;
      }
      cset(Y2[rr][(k * NP + j) * H + kx], acc);
    }
  }
  __syncthreads();
  double* out = dncval + ((long)slot_node[slot] * Rc + r0) * pln;
  for (int e = t; e < R * nsp; e += BS) {  // along x (c2r) at the surface points
    asm volatile("" ::: "memory");  // keep the twiddle loads inside the loop (registers)
    const int rr = e / nsp, p = e - rr * nsp;
    const int cc = s2c[p];
    const int i = cc / (NP * NP), j = (cc / NP) % NP, k = cc % NP;
    const LT* y = &Y2[rr][(k * NP + j) * H];
    const int xi = i + NP / 2;
    A v = cget<A2>(y[0]).x + ((xi & 1) ? -cget<A2>(y[NP]).x : cget<A2>(y[NP]).x);
#pragma unroll
    for (int kx = 1; kx < NP; kx++) {
      const A2 w = s_wt[i][kx], yk = cget<A2>(y[kx]);
      v += (A)2 * (yk.x * w.x + yk.y * w.y);
    }
    out[(long)rr * pln + p] = v;
  }
}

// Right-hand sides per block of the FP32-spectrum transforms (FP64: 4).
// Measured at 4M points / 16 RHS with FP32 arithmetic: forward 8 -> 1.74 ms,
// 4 -> 1.92 ms; inverse 8 -> 1.35 ms, 16 -> 1.63 ms.
constexpr int FFT_RW32_FWD = 8, FFT_RW32_INV = 8;

template <class CT>
static void fft_fwd_t(hipStream_t s, int np, int nslot, const int* slot_node, int Rc, const double* upeden, int pln,
                      int sm, const double* dscale, const short* cube2surf, CT* effden, const double2* tw, int ri)
{
  if (nslot <= 0) return;
  if (ri && sm == 1) {
    constexpr int RW = sizeof(CT) == 8 ? FFT_RW32_FWD : 4;
    // np 6: the x and y stages fused, a block writing whole 128-byte lines
    // of spectra: 16 FP32 right-hand sides (1024 threads) or 8 FP64 ones
    // (512).  4M points / 16 RHS, a forward pass: FP32 1.71 -> 1.58 ms fused
    // with 8 right-hand sides a block, 1.54 ms with 16; FP64 3.20 ms (three
    // stages, 4 a block) -> 3.75 ms fused with 4, 2.44 ms fused with 8; 384
    // or 256 threads per block of 8 FP32 right-hand sides 1.95 and 1.74 ms
    constexpr int RX = sizeof(CT) == 8 ? 16 : 8;
    if (np == 6 && Rc % RX == 0) {
      hipLaunchKernelGGL((k_fft_fwd_xy<6, CT, RX>), dim3(xcd_grid(nslot * (Rc / RX))), dim3(64 * RX), 0, s, nslot,
                         slot_node, Rc, upeden, pln, dscale, cube2surf, effden, tw);
      return;
    }
    switch (np) {
    case 4: hipLaunchKernelGGL((k_fft_fwd_ri<4, CT, RW>), dim3(xcd_grid(nslot * (Rc / RW))), dim3(64 * RW), 0, s, nslot, slot_node, Rc, upeden, pln, dscale, cube2surf, effden, tw); break;
    case 6: hipLaunchKernelGGL((k_fft_fwd_ri<6, CT, RW>), dim3(xcd_grid(nslot * (Rc / RW))), dim3(64 * RW), 0, s, nslot, slot_node, Rc, upeden, pln, dscale, cube2surf, effden, tw); break;
    case 8: hipLaunchKernelGGL((k_fft_fwd_ri<8, CT, RW / 2>), dim3(xcd_grid(nslot * (Rc / (RW / 2)))), dim3(32 * RW), 0, s, nslot, slot_node, Rc, upeden, pln, dscale, cube2surf, effden, tw); break;
    }
    return;
  }
  dim3 grid(nslot, Rc);
  switch (np) {
  case 4: hipLaunchKernelGGL((k_fft_fwd<4, CT>), grid, dim3(256), 0, s, slot_node, Rc, upeden, pln, sm, dscale, cube2surf, effden, tw, ri); break;
  case 6: hipLaunchKernelGGL((k_fft_fwd<6, CT>), grid, dim3(256), 0, s, slot_node, Rc, upeden, pln, sm, dscale, cube2surf, effden, tw, ri); break;
  case 8: hipLaunchKernelGGL((k_fft_fwd<8, CT>), grid, dim3(256), 0, s, slot_node, Rc, upeden, pln, sm, dscale, cube2surf, effden, tw, ri); break;
  }
}

template <class CT>
static void fft_inv_t(hipStream_t s, int np, int nslot, const int* slot_node, int Rc, const CT* effval, int tm,
                      const int* surf2cube, int nsp, double* dncval, int pln, const double2* tw, int ri)
{
  if (nslot <= 0) return;
  if (ri && tm == 1) {
    constexpr int RW = sizeof(CT) == 8 ? FFT_RW32_INV : 4;
    // np 6, FP64: 6 waves for the 4 right-hand sides of a block, so the z
    // stage's 336 lines take one pass and 3 blocks fit a CU (4M points /
    // 16 RHS: 2.94 -> 2.53 ms an inverse pass; for 8 FP32 right-hand sides
    // 384 and 256 threads measured 1.54 and 1.35 ms against 1.39 at 512)
    if constexpr (sizeof(CT) == 16)
      if (np == 6) {
        hipLaunchKernelGGL((k_fft_inv_ri<6, CT, RW, 384>), dim3(xcd_grid(nslot * (Rc / RW))), dim3(384), 0, s, nslot,
                           slot_node, Rc, effval, surf2cube, nsp, dncval, pln, tw);
        return;
      }
    switch (np) {
    case 4: hipLaunchKernelGGL((k_fft_inv_ri<4, CT, RW>), dim3(xcd_grid(nslot * (Rc / RW))), dim3(64 * RW), 0, s, nslot, slot_node, Rc, effval, surf2cube, nsp, dncval, pln, tw); break;
    case 6: hipLaunchKernelGGL((k_fft_inv_ri<6, CT, RW>), dim3(xcd_grid(nslot * (Rc / RW))), dim3(64 * RW), 0, s, nslot, slot_node, Rc, effval, surf2cube, nsp, dncval, pln, tw); break;
    case 8: hipLaunchKernelGGL((k_fft_inv_ri<8, CT, RW / 2>), dim3(xcd_grid(nslot * (Rc / (RW / 2)))), dim3(32 * RW), 0, s, nslot, slot_node, Rc, effval, surf2cube, nsp, dncval, pln, tw); break;
    }
    return;
  }
  dim3 grid(nslot, Rc);
  switch (np) {
  case 4: hipLaunchKernelGGL((k_fft_inv<4, CT>), grid, dim3(256), 0, s, slot_node, Rc, effval, tm, surf2cube, nsp, dncval, pln, tw, ri); break;
  case 6: hipLaunchKernelGGL((k_fft_inv<6, CT>), grid, dim3(256), 0, s, slot_node, Rc, effval, tm, surf2cube, nsp, dncval, pln, tw, ri); break;
  case 8: hipLaunchKernelGGL((k_fft_inv<8, CT>), grid, dim3(256), 0, s, slot_node, Rc, effval, tm, surf2cube, nsp, dncval, pln, tw, ri); break;
  }
}

void fft_fwd(hipStream_t s, int np, int nslot, const int* slot_node, int Rc, const double* upeden, int pln, int sm,
             const double* dscale, const short* cube2surf, void* effden, int single, int rhs_inner, const double2* tw)
{
  if (single)
    fft_fwd_t(s, np, nslot, slot_node, Rc, upeden, pln, sm, dscale, cube2surf, (float2*)effden, tw, rhs_inner);
  else
    fft_fwd_t(s, np, nslot, slot_node, Rc, upeden, pln, sm, dscale, cube2surf, (double2*)effden, tw, rhs_inner);
}

void fft_inv(hipStream_t s, int np, int nslot, const int* slot_node, int Rc, const void* effval, int single,
             int rhs_inner, int tm, const int* surf2cube, int nsp, double* dncval, int pln, const double2* tw)
{
  if (single)
    fft_inv_t(s, np, nslot, slot_node, Rc, (const float2*)effval, tm, surf2cube, nsp, dncval, pln, tw, rhs_inner);
  else
    fft_inv_t(s, np, nslot, slot_node, Rc, (const double2*)effval, tm, surf2cube, nsp, dncval, pln, tw, rhs_inner);
}

// ---------------------------------------------------------------------------
// M2L products in the frequency domain, one frequency per lane:
// effval[t][r][i][q] = sum_{V pairs} sum_j T_op[j][i][q] effden[s][r][j][q]
template <int SM, int TM, int RB>
__global__ __launch_bounds__(256) void k_m2l(int C, int Rc, const int* __restrict__ vptr, const int* __restrict__ vsrc,
                                             const int* __restrict__ vop, const double2* __restrict__ ops,
                                             const double2* __restrict__ effden, double2* __restrict__ effval, int roff)
{
  const int q = blockIdx.y * 256 + threadIdx.x;
  const int tr = blockIdx.x;
  const int r0 = roff + blockIdx.z * RB;
  if (q >= C) return;
  double2 acc[RB][TM];
#pragma unroll
  for (int rr = 0; rr < RB; rr++)
#pragma unroll
    for (int i = 0; i < TM; i++) acc[rr][i] = make_double2(0, 0);
  for (int e = vptr[tr]; e < vptr[tr + 1]; e++) {
    const double2* T = ops + (long)vop[e] * SM * TM * C + q;
    double2 tv[SM][TM];
#pragma unroll
    for (int j = 0; j < SM; j++)
#pragma unroll
      for (int i = 0; i < TM; i++) tv[j][i] = T[(long)(j * TM + i) * C];
    const double2* D = effden + ((long)vsrc[e] * Rc + r0) * SM * C + q;
#pragma unroll
    for (int rr = 0; rr < RB; rr++) {
#pragma unroll
      for (int j = 0; j < SM; j++) {
        double2 d = D[((long)rr * SM + j) * C];
#pragma unroll
        for (int i = 0; i < TM; i++) {
          acc[rr][i].x += tv[j][i].x * d.x - tv[j][i].y * d.y;
          acc[rr][i].y += tv[j][i].x * d.y + tv[j][i].y * d.x;
        }
      }
    }
  }
  double2* out = effval + ((long)tr * Rc + r0) * TM * C + q;
#pragma unroll
  for (int rr = 0; rr < RB; rr++)
#pragma unroll
    for (int i = 0; i < TM; i++) out[((long)rr * TM + i) * C] = acc[rr][i];
}

// M2L grouped by target parent.  One lane per frequency; the lane keeps the
// spectra of all 8 target children for RB right-hand sides in registers and
// streams the children of the 26 neighbour parents, so every source spectrum
// is read once per target parent instead of once per V pair, and each
// operator value serves RB right-hand sides.  Blocks are ordered parent-
// fastest within a frequency block and parents are dealt to the 8 XCDs in
// contiguous ranges, so neighbouring parents share an L2.  A block covers BS
// frequencies: the spectra of one (frequency block, right-hand-side block)
// over a whole level are the working set of the parent sweep, so BS sets how
// much of it stays in the L2 and Infinity Cache.
template <int SM, int TM, int RB, int BS>
__global__ __launch_bounds__(BS) void k_m2l_grp(int C, int Rc, int npar, const int* __restrict__ tslot,
                                                const int* __restrict__ sslot, const int* __restrict__ optab,
                                                const double2* __restrict__ ops, const double2* __restrict__ effden,
                                                double2* __restrict__ effval, int roff)
{
  const int per = (npar + 7) / 8;
  const int b = blockIdx.x;
  const int pidx = (b & 7) * per + (b >> 3);
  const int q = blockIdx.y * BS + threadIdx.x;
  const int r0 = roff + blockIdx.z * RB;
  if (pidx >= npar || q >= C) return;
  double2 acc[8][RB][TM];
#pragma unroll
  for (int c = 0; c < 8; c++)
#pragma unroll
    for (int rr = 0; rr < RB; rr++)
#pragma unroll
      for (int i = 0; i < TM; i++) acc[c][rr][i] = make_double2(0, 0);
  const int* ss = sslot + (long)pidx * 26 * 8;
  for (int n = 0; n < 26; n++) {
    for (int sc = 0; sc < 8; sc++) {
      const int sl = ss[n * 8 + sc];
      if (sl < 0) continue;
      double2 x[RB][SM];
      const double2* D = effden + ((long)sl * Rc + r0) * SM * C + q;
#pragma unroll
      for (int rr = 0; rr < RB; rr++)
#pragma unroll
        for (int j = 0; j < SM; j++) x[rr][j] = D[((long)rr * SM + j) * C];
      const int* ot = optab + (n * 8 + sc) * 8;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int o = ot[c];
        if (o < 0) continue;
        const double2* Tp = ops + (long)o * SM * TM * C + q;
#pragma unroll
        for (int j = 0; j < SM; j++)
#pragma unroll
          for (int i = 0; i < TM; i++) {
            const double2 t = Tp[(long)(j * TM + i) * C];
#pragma unroll
            for (int rr = 0; rr < RB; rr++) {
              acc[c][rr][i].x += t.x * x[rr][j].x - t.y * x[rr][j].y;
              acc[c][rr][i].y += t.x * x[rr][j].y + t.y * x[rr][j].x;
            }
          }
      }
    }
  }
  const int* ts = tslot + (long)pidx * 8;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int tl = ts[c];
    if (tl < 0) continue;
    double2* out = effval + ((long)tl * Rc + r0) * TM * C + q;
#pragma unroll
    for (int rr = 0; rr < RB; rr++)
#pragma unroll
      for (int i = 0; i < TM; i++) out[((long)rr * TM + i) * C] = acc[c][rr][i];
  }
}

template <int SM, int TM, int RB, int BS>
static void m2l_grp_launch(hipStream_t s, int C, int npar, int Rc, int nb, int roff, const int* tslot,
                           const int* sslot, const int* optab, const double2* ops, const double2* effden,
                           double2* effval)
{
  const int per = (npar + 7) / 8;
  dim3 grid(8 * per, (C + BS - 1) / BS, nb);
  hipLaunchKernelGGL((k_m2l_grp<SM, TM, RB, BS>), grid, dim3(BS), 0, s, C, Rc, npar, tslot, sslot, optab, ops,
                     effden, effval, roff);
}

// Right-hand sides are covered by launches of RBMAX per lane and a remainder
// of single right-hand sides.  (Lanes sharing one frequency across right-hand
// side groups, so that operator loads coincide, measured slower on MI355X: the
// texture path costs per lane, not per distinct address.)
template <int SM, int TM, int RBMAX, int BS>
static void m2l_grp_st(hipStream_t s, int C, int npar, int Rc, const int* tslot, const int* sslot, const int* optab,
                       const double2* ops, const double2* effden, double2* effval)
{
  int r = 0;
  while (r < Rc) {
    const int left = Rc - r;
    if (left >= RBMAX) {
      const int nb = left / RBMAX;
      m2l_grp_launch<SM, TM, RBMAX, BS>(s, C, npar, Rc, nb, r, tslot, sslot, optab, ops, effden, effval);
      r += nb * RBMAX;
    } else {
      m2l_grp_launch<SM, TM, 1, BS>(s, C, npar, Rc, left, r, tslot, sslot, optab, ops, effden, effval);
      r += left;
    }
  }
}

void m2l_grouped(hipStream_t s, int sm, int tm, int C, int npar, int Rc, int rb, int bs, const int* tslot,
                 const int* sslot, const int* optab, const double2* ops, const double2* effden, double2* effval)
{
  if (npar <= 0) return;
  const bool wide = bs == 256;
#define KIFMM_M2L_ST(SMv, TMv, RBv)                                                          \
  (wide ? m2l_grp_st<SMv, TMv, RBv, 256>(s, C, npar, Rc, tslot, sslot, optab, ops, effden, effval) \
        : m2l_grp_st<SMv, TMv, RBv, 64>(s, C, npar, Rc, tslot, sslot, optab, ops, effden, effval))
  if (sm == 1 && tm == 1) {
    if (rb == 8)
      KIFMM_M2L_ST(1, 1, 8);
    else
      KIFMM_M2L_ST(1, 1, 4);
  } else if (sm == 4 && tm == 3) {
    if (rb == 2)
      KIFMM_M2L_ST(4, 3, 2);
    else
      KIFMM_M2L_ST(4, 3, 1);
  } else if (sm == 3 && tm == 3) {
    if (rb == 2)
      KIFMM_M2L_ST(3, 3, 2);
    else
      KIFMM_M2L_ST(3, 3, 1);
  }
#undef KIFMM_M2L_ST
}

template <int SM, int TM>
static void m2l_st(hipStream_t s, int C, int ntrg, int Rc, const int* vptr, const int* vsrc, const int* vop,
                   const double2* ops, const double2* effden, double2* effval)
{
  int r = 0;
  while (r < Rc) {
    int left = Rc - r;
    int RB = left >= 4 ? 4 : 1;
    int nb = left / RB;
    dim3 grid(ntrg, (C + 255) / 256, nb);
    if (RB == 4)
      hipLaunchKernelGGL((k_m2l<SM, TM, 4>), grid, dim3(256), 0, s, C, Rc, vptr, vsrc, vop, ops, effden, effval, r);
    else
      hipLaunchKernelGGL((k_m2l<SM, TM, 1>), grid, dim3(256), 0, s, C, Rc, vptr, vsrc, vop, ops, effden, effval, r);
    r += nb * RB;
  }
}

void m2l(hipStream_t s, int sm, int tm, int C, int ntrg, int Rc, const int* vptr, const int* vsrc, const int* vop,
         const double2* ops, const double2* effden, double2* effval)
{
  if (ntrg <= 0) return;
  if (sm == 1 && tm == 1) m2l_st<1, 1>(s, C, ntrg, Rc, vptr, vsrc, vop, ops, effden, effval);
  else if (sm == 4 && tm == 3) m2l_st<4, 3>(s, C, ntrg, Rc, vptr, vsrc, vop, ops, effden, effval);
  else if (sm == 3 && tm == 3) m2l_st<3, 3>(s, C, ntrg, Rc, vptr, vsrc, vop, ops, effden, effval);
}

// ---------------------------------------------------------------------------
// Tiled M2L on the matrix cores (scalar kernels, 16 right-hand sides).
//
// A tile is the 2x2x2 block of target parents under one grandparent; their 26
// neighbours lie in the 4x4x4 block of parents around it (the halo).  Block
// (tile, frequency f, group of 16 right-hand sides) stages the spectra of the
// halo's children at f in LDS once (each is then used by up to 8 target
// parents instead of being re-read per parent) and wave p of the block forms,
// for target parent p and each neighbour direction n,
//   Y[c][r] += sum_s M_n(f)[c][s] X[s][r]          (c, s children; r RHS)
// with M_n(f)[c][s] = T_{optab[n][s][c]}(f), as the real 16x16x16 product
//   [Yr; Yi] += [[Mr, -Mi], [Mi, Mr]] [Xr; Xi]
// in four 16x16x4 MFMAs.  Lane l holds A[i = l%16][k = l/16] and
// B[k][j = l%16]; the K index of slice ks is source child k (ks even) or 4+k
// (ks odd), real part for ks < 2 and imaginary part otherwise.  Rows of A are
// ordered so that lane (k, r) ends up with Y of children 2k and 2k+1 as
// (re, im) pairs: accumulator item v holds child 2k + v/2, part v%2.  The
// position of item v in the 16x16 result is row k + 4v for the FP64 MFMA and
// row 4k + v for the FP32 one.

template <class R>
struct Mt;
template <>
struct Mt<double> {
  typedef double2 C2;
  typedef v4d V4;
  static constexpr bool KV = true;  // D row of lane k, item v is k + 4v
  __device__ static inline V4 mfma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct Mt<float> {
  typedef float2 C2;
  typedef v4f V4;
  static constexpr bool KV = false;  // D row of lane k, item v is 4k + v
  __device__ static inline V4 mfma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

constexpr int M2LT_BS = 512;  // 8 waves: one per target parent of the tile
constexpr int M2LT_FPB = 16;  // frequencies per block, visited in a software-pipelined loop (8: +1 %, 32: same, 63: +5 %)
constexpr int NAENT = 26 * 32;  // A-operand table entries: direction x (source k, target child c)
// FP32 form: directions summed on the matrix cores in FP32 before their sum
// is added to the FP64 accumulators (0: all 26 in FP32)
#ifndef M2L_F32_FLUSH
#define M2L_F32_FLUSH 0
#endif

// The A operands depend only on (direction, frequency, lane), not on the
// target parent, so the block stages them once per frequency: entry
// n*32 + k*8 + c holds (T_{optab[n][k][c]}, T_{optab[n][4+k][c]}) (zero where
// the child pair is not in a V list), read contiguously from a table expanded
// at setup ([C][NAENT][2]), and every wave reads its lane's entry with one LDS
// load.  (FP32: two blocks per CU fit the LDS, so registers are
// capped to match: the second launch bound is waves per SIMD.)
template <class R>
__global__ __launch_bounds__(M2LT_BS, sizeof(R) == 4 ? 4 : 2) void k_m2l_tile(
    int C, int Rc, int ntile, const int* __restrict__ ttile, const int* __restrict__ htile,
    const typename Mt<R>::C2* __restrict__ atab, const typename Mt<R>::C2* __restrict__ effden,
    typename Mt<R>::C2* __restrict__ effval, int roff, int fpart)
{
  using C2 = typename Mt<R>::C2;
  using V4 = typename Mt<R>::V4;
  __shared__ C2 xs[64 * 8 * 16];  // [halo parent][source child][rhs] at the current frequency
  __shared__ C2 s_A[NAENT][2];    // A-operand table at the current frequency
  __shared__ int s_hm[64];        // halo parent has a child in some V list
  // tiles are dealt to the 8 XCDs in contiguous ranges so that neighbouring
  // tiles (whose halos overlap) share an L2
  // fpart 0: tiles in 8 contiguous ranges, one per XCD, frequency blocks in
  // grid y; fpart 1: frequency blocks dealt to the XCDs (b % 8), each XCD
  // sweeping every tile of a frequency block before the next
  const int b = blockIdx.x;
  int tile, fb;
  if (fpart) {
    const int j = b >> 3;
    tile = j % ntile;
    fb = (b & 7) + 8 * (j / ntile);
  } else {
    const int per = (ntile + 7) / 8;
    tile = (b & 7) * per + (b >> 3);
    fb = blockIdx.y;
  }
  if (tile >= ntile || fb * M2LT_FPB >= C) return;
  const int f0 = fb * M2LT_FPB, f1 = min(C, f0 + M2LT_FPB), r0 = roff + blockIdx.z * 16;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int* hs = htile + (long)tile * 512;
  if (t < 64) {
    int m = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) m |= hs[t * 8 + c] >= 0;
    s_hm[t] = m;
  }
  // staging: thread t moves entries e = it*512 + t, i.e. halo child row
  // it*32 + t/16 and right-hand side t%16 (FP64), or the 16-byte pair of
  // right-hand sides 2(t%8), 2(t%8)+1 of halo child row it*64 + t/8 (FP32:
  // half the loads, stores and slot registers), and A-table entries t and t + 512
  constexpr bool P2 = sizeof(R) == 4;
  constexpr int NIT = P2 ? 8 : 16;
  int hsl[NIT];
#pragma unroll
  for (int it = 0; it < NIT; it++) hsl[it] = P2 ? hs[it * 64 + (t >> 3)] : hs[it * 32 + (t >> 4)];
  const C2* eb = effden + r0 + (P2 ? 2 * (t & 7) : (t & 15));
  C2 pre[16];  // FP32: pre[2 it], pre[2 it + 1] hold the pair of iteration it
  C2 pa00, pa01, pa10, pa11;  // A-table entries t and t + 512 (named: no local array)
  auto fetch = [&](int f) {
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      if constexpr (P2) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (hsl[it] >= 0) v = *(const float4*)(eb + ((long)hsl[it] * C + f) * Rc);
        pre[2 * it].x = v.x;
        pre[2 * it].y = v.y;
        pre[2 * it + 1].x = v.z;
        pre[2 * it + 1].y = v.w;
      } else {
        pre[it].x = 0;
        pre[it].y = 0;
        if (hsl[it] >= 0) pre[it] = eb[((long)hsl[it] * C + f) * Rc];
      }
    }
    // A table of frequency f: a contiguous [NAENT][2] block of the expanded table
    const C2* at = atab + (long)f * NAENT * 2;
    pa00 = at[2 * t];
    pa01 = at[2 * t + 1];
    if (t + M2LT_BS < NAENT) {  // (measured 1.5 % faster than a clamped unconditional load)
      pa10 = at[2 * (t + M2LT_BS)];
      pa11 = at[2 * (t + M2LT_BS) + 1];
    }
  };
  fetch(f0);
  const int* ts = ttile + ((long)tile * 8 + w) * 8;
  int live = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) live |= ts[c] >= 0;
  const int px = w >> 2, py = (w >> 1) & 1, pz = w & 1;
  const int hc = ((px + 1) * 4 + (py + 1)) * 4 + (pz + 1);
  const int i = l & 15, k = l >> 4;
  // output (child, part) of row i of A
  int kr, vr;
  if (Mt<R>::KV) {
    kr = i & 3;
    vr = i >> 2;
  } else {
    kr = i >> 2;
    vr = i & 3;
  }
  const int ca = 2 * kr + (vr >> 1), pim = vr & 1;
  const int aent = k * 8 + ca;
  __syncthreads();
  // directions n (in (dx, dy, dz) order, dz fastest, centre skipped) whose
  // halo parent has children: a wave-uniform bit mask
  unsigned dmask = 0;
  for (int n = 0; n < 26; n++) {
    const int m = n + (n >= 13);
    const int h = hc + (m / 9 - 1) * 16 + ((m / 3) % 3 - 1) * 4 + (m % 3 - 1);
    if (s_hm[h]) dmask |= 1u << n;
  }
  dmask = __builtin_amdgcn_readfirstlane(dmask);  // uniform per wave (one parent): scalar branches
  for (int f = f0; f < f1; f++) {
    __syncthreads();  // the previous frequency's products are done with xs and s_A
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      if constexpr (P2)
        *(float4*)&xs[it * 1024 + 2 * t] = make_float4(pre[2 * it].x, pre[2 * it].y, pre[2 * it + 1].x, pre[2 * it + 1].y);
      else
        xs[it * M2LT_BS + t] = pre[it];
    }
    s_A[t][0] = pa00;  // t < 512 < NAENT
    s_A[t][1] = pa01;
    if (t + M2LT_BS < NAENT) {
      s_A[t + M2LT_BS][0] = pa10;
      s_A[t + M2LT_BS][1] = pa11;
    }
    __syncthreads();
    if (f + 1 < f1) fetch(f + 1);  // in flight while the matrix cores work on f
    if (!live) continue;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    V4 d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};  // two chains for matrix-core overlap
    int nacc = 0;
    auto direction = [&](int n) {
      const int m = n + (n >= 13);
      const int h = hc + (m / 9 - 1) * 16 + ((m / 3) % 3 - 1) * 4 + (m % 3 - 1);
      const C2 m0 = s_A[n * 32 + aent][0], m1 = s_A[n * 32 + aent][1];
      const R a0 = pim ? m0.y : m0.x, a1 = pim ? m1.y : m1.x;
      const R a2 = pim ? m0.x : -m0.y, a3 = pim ? m1.x : -m1.y;
      const C2 x0 = xs[(h * 8 + k) * 16 + i], x1 = xs[(h * 8 + 4 + k) * 16 + i];
      d0 = Mt<R>::mfma(a0, x0.x, d0);
      d1 = Mt<R>::mfma(a2, x0.y, d1);
      d0 = Mt<R>::mfma(a1, x1.x, d0);
      d1 = Mt<R>::mfma(a3, x1.y, d1);
      if (!Mt<R>::KV && M2L_F32_FLUSH && ++nacc == M2L_F32_FLUSH) {
        // single precision: partial sums of M2L_F32_FLUSH directions go to FP64
#pragma unroll
        for (int v = 0; v < 4; v++) {
          acc[v] += (double)d0[v] + (double)d1[v];
          d0[v] = 0;
          d1[v] = 0;
        }
        nacc = 0;
      }
    };
    // fully unrolled (measured fastest in both precisions: 26 > 8 > 4 > 2).
    // A parent with all 26 neighbour directions takes a branch-free sequence
    // whose LDS operand loads run one direction ahead of the matrix cores.
    if (dmask == 0x3ffffffu) {
      auto hof = [&](int n) {
        const int m = n + (n >= 13);
        return hc + (m / 9 - 1) * 16 + ((m / 3) % 3 - 1) * 4 + (m % 3 - 1);
      };
      C2 p0 = s_A[aent][0], p1 = s_A[aent][1];
      C2 q0 = xs[(hof(0) * 8 + k) * 16 + i], q1 = xs[(hof(0) * 8 + 4 + k) * 16 + i];
#pragma unroll
      for (int n = 0; n < 26; n++) {
        const C2 m0 = p0, m1 = p1, x0 = q0, x1 = q1;
        if (n + 1 < 26) {
          const int h = hof(n + 1);
          p0 = s_A[(n + 1) * 32 + aent][0];
          p1 = s_A[(n + 1) * 32 + aent][1];
          q0 = xs[(h * 8 + k) * 16 + i];
          q1 = xs[(h * 8 + 4 + k) * 16 + i];
        }
        const R a0 = pim ? m0.y : m0.x, a1 = pim ? m1.y : m1.x;
        const R a2 = pim ? m0.x : -m0.y, a3 = pim ? m1.x : -m1.y;
        d0 = Mt<R>::mfma(a0, x0.x, d0);
        d1 = Mt<R>::mfma(a2, x0.y, d1);
        d0 = Mt<R>::mfma(a1, x1.x, d0);
        d1 = Mt<R>::mfma(a3, x1.y, d1);
        asm volatile("" ::: "memory");  // keep the loads one direction ahead, no further
      }
    } else {
#pragma unroll
      for (int n = 0; n < 26; n++)
        if (dmask >> n & 1u) direction(n);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) acc[v] += (double)d0[v] + (double)d1[v];
#pragma unroll
    for (int h2 = 0; h2 < 2; h2++) {
      const int sl = ts[2 * k + h2];
      if (sl >= 0) {
        C2 v;
        v.x = (R)acc[2 * h2];
        v.y = (R)acc[2 * h2 + 1];
        effval[((long)sl * C + f) * Rc + r0 + i] = v;
      }
    }
  }
}

void m2l_tiled(hipStream_t s, int C, int ntile, int Rc, const int* ttile, const int* htile, const void* atab,
               const void* effden, void* effval, int single, int fpart)
{
  if (ntile <= 0 || Rc < 16) return;
  const int per = (ntile + 7) / 8, nfb = (C + M2LT_FPB - 1) / M2LT_FPB;
  const dim3 grid = fpart ? dim3((unsigned)(8 * ntile * ((nfb + 7) / 8)), 1, Rc / 16) : dim3(8 * per, nfb, Rc / 16);
  if (single)
    hipLaunchKernelGGL(k_m2l_tile<float>, grid, dim3(M2LT_BS), 0, s, C, Rc, ntile, ttile, htile,
                       (const float2*)atab, (const float2*)effden, (float2*)effval, 0, fpart);
  else
    hipLaunchKernelGGL(k_m2l_tile<double>, grid, dim3(M2LT_BS), 0, s, C, Rc, ntile, ttile, htile,
                       (const double2*)atab, (const double2*)effden, (double2*)effval, 0, fpart);
}

// ---------------------------------------------------------------------------
// Tiled FP32 M2L, one wave per (tile, frequency): the products of
// k_m2l_tile<float> (same A-operand table, lane layout and output rows) with
// the work split the other way round.  A wave takes whole tiles at one
// frequency f.  The block stages the A operands of all 26 directions at f
// in LDS once, as the four values each lane feeds; a wave holds the
// accumulators of all 8 target parents of its tile (4 per parent and lane)
// and walks the 64 halo parents: each halo parent's children spectra are
// loaded from memory once, straight into the B operands (D halo parents
// ahead of the matrix cores), and applied to every target parent of the tile
// that has the halo parent as a neighbour -- 1 to 8 of them, known at compile
// time from the halo position -- with one 16-byte LDS read of A each.  One
// block barrier, after the staging.  Tile slots travel in registers (lane
// l holds entries q * 64 + l of the tile's halo table and entry l of its
// target table) and are fetched with cross-lane reads.  Halo parents without
// any source child are skipped; target parents without target children are
// computed and not stored.  FP32 accumulation over the 26 directions, as in
// k_m2l_tile<float>.
//
// Blocks (frequency f, group of 4 TPW tiles, TPW = tiles per wave): the frequencies are dealt to
// the 8 XCDs (block b runs on XCD b % 8) and each XCD sweeps all tile groups
// of a frequency before the next, so every tile that reads a halo spectrum
// line (slot, f) runs on the one L2 that holds it, at about the same time.
constexpr int M2LW_BS = 256;

// Direction index (0 .. 25, dz fastest, centre skipped) in which halo parent
// h (position (h / 16, h / 4 % 4, h % 4)) neighbours target parent p of the
// tile (position (p / 4, p / 2 % 2, p % 2) + 1 in the halo), or -1.
__host__ __device__ constexpr int m2lw_dir(int h, int p)
{
  const int dx = (h >> 4) - (p >> 2) - 1, dy = ((h >> 2) & 3) - ((p >> 1) & 1) - 1, dz = (h & 3) - (p & 1) - 1;
  if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1 || (dx == 0 && dy == 0 && dz == 0)) return -1;
  const int m = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
  return m - (m > 13);
}

// f(std::integral_constant<int, H>) for H = 0 .. 63 in order: every halo
// position is a compile-time constant in the generated code
template <class F, int... H>
__device__ __forceinline__ void m2lw_each(std::integer_sequence<int, H...>, F&& f)
{
  (f(std::integral_constant<int, H>{}), ...);
}

// R: float (the mixed form; FP32 accumulation over the 26 directions) or
// double (the plain form, FP64 matrix cores).  OFF32: the level's spectra and
// its zero slot span less than 4 GiB, so a spectrum address is a 32-bit byte
// offset from the frequency's row of slot 0 (one VALU instruction per load
// and the SGPR-based load form); otherwise a 64-bit shift-add.
template <class R> struct Mw;
template <> struct Mw<float> {
  typedef float4 A4;
  __device__ static inline A4 make(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
};
template <> struct Mw<double> {
  typedef double4 A4;
  __device__ static inline A4 make(double a, double b, double c, double d) { return make_double4(a, b, c, d); }
};

template <class R, int D, int TPW, int OCC, int ORD, bool OFF32>
__global__ __launch_bounds__(M2LW_BS, OCC) void k_m2l_wave(int C, int Rc, int ntile, int ngrp, int zslot,
                                                        const int* __restrict__ ttile, const int* __restrict__ htile,
                                                        const typename Mt<R>::C2* __restrict__ atab,
                                                        const typename Mt<R>::C2* __restrict__ effden,
                                                        typename Mt<R>::C2* __restrict__ effval)
{
  using C2 = typename Mt<R>::C2;
  using V4 = typename Mt<R>::V4;
  using A4 = typename Mw<R>::A4;
  // the A operands of frequency f, [direction][lane] as the four values each
  // lane feeds: one LDS read per (halo parent, target parent)
  __shared__ A4 s_A4[26][64];
  // XCD b % 8 takes the frequencies f = b % 8 (mod 8) one after another and,
  // for each, every tile group of the level: the blocks resident on one XCD
  // work on one or two frequencies, whose spectra (every slot of the level at
  // f) its L2 holds while the 8 tiles sharing each line read it
  const int b = blockIdx.x, j = b >> 3;
  const int f = (b & 7) + 8 * (j / ngrp), g = j % ngrp;
  if (f >= C) return;
  const int r0 = blockIdx.y * 16;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int i = l & 15, k = l >> 4;
  // Lane (i, k) feeds row i of A, which holds output (child 2 kr + vr / 2,
  // part vr % 2) (accumulator item v of lane (i, k) is row 4k + v for the FP32
  // MFMA, k + 4v for the FP64 one), in K slices 0 .. 3: with (m.x, m.y) the
  // operator of source child k and (m.z, m.w) that of source child 4 + k,
  // slice 0 meets the real part of child k's spectrum, 1 its imaginary part,
  // 2 and 3 those of child 4 + k.
  for (int e = t; e < 26 * 64; e += M2LW_BS) {
    const int n = e >> 6, ln = e & 63, ii = ln & 15, kk = ln >> 4;
    const int kr = Mt<R>::KV ? ii & 3 : ii >> 2, vr = Mt<R>::KV ? ii >> 2 : ii & 3;
    const int ca = 2 * kr + (vr >> 1), pim = vr & 1;
    const A4 m = *reinterpret_cast<const A4*>(atab + ((long)f * NAENT + n * 32 + kk * 8 + ca) * 2);
    s_A4[n][ln] = pim ? Mw<R>::make(m.y, m.x, m.w, m.z) : Mw<R>::make(m.x, -m.y, m.z, -m.w);
  }
  __syncthreads();
  // the spectrum of slot s at f for right-hand side r0 + i is eb0[16 o + i]
  // with o = s * C * Rc / 16 (< 2^31, checked by the launcher); absent halo
  // children read slot zslot, which holds zeros, so every load is
  // unconditional and the loads in flight are counted exactly
  const C2* eb0 = effden + (long)f * Rc + r0;
  constexpr int SH = sizeof(C2) == 8 ? 7 : 8;  // bytes of 16 spectra, log2
  const unsigned ib = (unsigned)sizeof(C2) * i;
  const int cr16 = C * (Rc >> 4);
  for (int tt = 0; tt < TPW; tt++) {
    const int tile = (g * 4 + w) * TPW + tt;
    if (tile >= ntile) break;  // wave-uniform
    // hv: the offsets of the halo children; present: bit h set if halo
    // parent h has a source child (the entries 8h .. 8h+7 of halo parent h
    // sit in lanes 8 (h % 8) .. 8 (h % 8) + 7 of hv[h / 8])
    int hv[8];
    unsigned long long present = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int sl = htile[(long)tile * 512 + q * 64 + l];
      hv[q] = (sl >= 0 ? sl : zslot) * cr16;
      const unsigned long long bal = __ballot(sl >= 0);
#pragma unroll
      for (int bb = 0; bb < 8; bb++)
        if ((bal >> (8 * bb)) & 0xffull) present |= 1ull << (8 * q + bb);
    }
    const int tv = ttile[(long)tile * 64 + l];
    // offsets of halo parent h, children k and 4 + k (cross-lane reads)
    auto offs = [&](int h, int* o) {
      o[0] = __shfl(hv[h >> 3], 8 * (h & 7) + k);
      o[1] = __shfl(hv[h >> 3], 8 * (h & 7) + 4 + k);
    };
    // their spectra for the lane's right-hand side
    auto load = [&](const int* o, C2* x) {
      if constexpr (OFF32) {
        const char* e8 = reinterpret_cast<const char*>(eb0);
        x[0] = *reinterpret_cast<const C2*>(e8 + (((unsigned)o[0] << SH) + ib));
        x[1] = *reinterpret_cast<const C2*>(e8 + (((unsigned)o[1] << SH) + ib));
      } else {
        x[0] = eb0[((size_t)(unsigned)o[0] << 4) + i];
        x[1] = eb0[((size_t)(unsigned)o[1] << 4) + i];
      }
    };
    V4 acc[8];
#pragma unroll
    for (int p = 0; p < 8; p++) acc[p] = V4{0, 0, 0, 0};
    C2 ring[D][2];  // the spectra of the next D halo parents, in flight
    int on[2];          // the offsets of the halo parent after those
#pragma unroll
    for (int d = 0; d < D; d++) {
      offs(d, on);
      load(on, ring[d]);
    }
    offs(D, on);
    m2lw_each(std::make_integer_sequence<int, 64>{}, [&](auto hc) {
      constexpr int h = decltype(hc)::value;
      const C2 x0 = ring[h % D][0], x1 = ring[h % D][1];
      if constexpr (h + D < 64) {
        load(on, ring[h % D]);
        if constexpr (h + D + 1 < 64) offs(h + D + 1, on);
      }
      if ((present >> h) & 1ull) {
        // the target parents that have halo parent h as a neighbour; K
        // slices in turn over them (independent accumulators between
        // dependent products)
        A4 av[8];
#pragma unroll
        for (int p = 0; p < 8; p++) {
          const int n = m2lw_dir(h, p);
          if (n >= 0) av[p] = s_A4[n][l];
        }
        auto mf = [&](int ks, int p) {
          const R xb = ks == 0 ? x0.x : ks == 1 ? x0.y : ks == 2 ? x1.x : x1.y;
          const R ab = ks == 0 ? av[p].x : ks == 1 ? av[p].y : ks == 2 ? av[p].z : av[p].w;
          acc[p] = Mt<R>::mfma(ab, xb, acc[p]);
        };
        if constexpr (ORD == 0) {
#pragma unroll
          for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int p = 0; p < 8; p++)
              if (m2lw_dir(h, p) >= 0) mf(ks, p);
        } else {  // one target parent after the other: its A operand is needed first
#pragma unroll
          for (int p = 0; p < 8; p++)
#pragma unroll
            for (int ks = 0; ks < 4; ks++)
              if (m2lw_dir(h, p) >= 0) mf(ks, p);
        }
      }
      asm volatile("" ::: "memory");  // spectra loads D halo parents ahead, no further
    });
#pragma unroll
    for (int p = 0; p < 8; p++)
#pragma unroll
      for (int h2 = 0; h2 < 2; h2++) {
        const int sl = __shfl(tv, p * 8 + 2 * k + h2);
        if (sl >= 0) {
          C2 v;
          v.x = acc[p][2 * h2];
          v.y = acc[p][2 * h2 + 1];
          effval[((long)sl * C + f) * Rc + r0 + i] = v;
        }
      }
  }
}

void m2l_wave(hipStream_t s, int C, int ntile, int Rc, int nslot, const int* ttile, const int* htile,
              const void* atab, const void* effden, void* effval, int depth, int single)
{
  if (ntile <= 0 || Rc < 16) return;
  if ((long)(nslot + 1) * C * (Rc / 16) > INT_MAX) {  // halo offsets beyond 32 bits: the block-per-tile kernel
    m2l_tiled(s, C, ntile, Rc, ttile, htile, atab, effden, effval, single);
    return;
  }
  // Measured at 4M points / 16 RHS, levels 4-5 (KIFMM3D_M2L_WAVE = the four
  // template arguments as digits, for launch-shape sweeps):
  //  * one tile per wave: the resident blocks of an XCD hold one frequency at
  //    a time, whose spectra fit its L2: 6.7 GB read per evaluate, 1.4x the
  //    unique spectra; two tiles per wave (4241) amortise the A-table staging
  //    over 8 tiles a block but keep two frequencies resident: 18 GB read and
  //    slower;
  //  * the products of one target parent after the other (ORD 1) rather than
  //    K slice by K slice over all of them: 8.7 -> 8.1 ms for the M2L (the A
  //    operands are read from LDS two parents at a time, 98 instead of 116
  //    VGPRs); 2 halo parents ahead 8.7 ms, 4, 6 or 8 ahead 8.1-8.3 ms, 5
  //    waves per SIMD no faster.
  static const int var = [] {
    const char* v = std::getenv("KIFMM3D_M2L_WAVE");
    return v ? std::atoi(v) : 4141;
  }();
  const int tpw = (var / 100) % 10 == 2 ? 2 : 1;
  const int ngrp = (ntile + 4 * tpw - 1) / (4 * tpw);
  dim3 grid((unsigned)(8 * ngrp * ((C + 7) / 8)), Rc / 16);
  // slot nslot holds zeros (the caller's); byte offsets below 4 GiB: 32-bit
  const size_t esz = single ? sizeof(float2) : sizeof(double2);
  const bool off32 = (unsigned long long)(nslot + 1) * C * Rc * esz <= 0xffffffffull;
#define KIFMM_M2LW_R(R, C2, DD, TT, OO, RR)                                                                     \
  do {                                                                                                          \
    if (off32)                                                                                                  \
      hipLaunchKernelGGL((k_m2l_wave<R, DD, TT, OO, RR, true>), grid, dim3(M2LW_BS), 0, s, C, Rc, ntile, ngrp,  \
                         nslot, ttile, htile, (const C2*)atab, (const C2*)effden, (C2*)effval);                 \
    else                                                                                                        \
      hipLaunchKernelGGL((k_m2l_wave<R, DD, TT, OO, RR, false>), grid, dim3(M2LW_BS), 0, s, C, Rc, ntile, ngrp, \
                         nslot, ttile, htile, (const C2*)atab, (const C2*)effden, (C2*)effval);                 \
  } while (0)
#define KIFMM_M2LW(DD, TT, OO, RR) KIFMM_M2LW_R(float, float2, DD, TT, OO, RR)
  // FP64 (the plain form): 53 KB of A operands per block, 3 blocks (waves per
  // SIMD) per CU.  4M points / 16 RHS, levels 4-5: M2L 21.1 -> 13.9 ms against
  // the block-per-tile kernel; 2 blocks per CU (KIFMM3D_M2L_WAVE64=2) 14.5 ms,
  // 3 halo parents ahead (=3) the same as 4.
  if (!single) {
    static const int v64 = [] {
      const char* v = std::getenv("KIFMM3D_M2L_WAVE64");
      return v ? std::atoi(v) : 0;
    }();
    if (depth <= 1)
      KIFMM_M2LW_R(double, double2, 1, 1, 3, 1);
    else if (v64 == 2)
      KIFMM_M2LW_R(double, double2, 4, 1, 2, 1);
    else if (v64 == 3)
      KIFMM_M2LW_R(double, double2, 3, 1, 3, 1);
    else
      KIFMM_M2LW_R(double, double2, 4, 1, 3, 1);
    return;
  }
  if (depth <= 1)
    KIFMM_M2LW(1, 1, 4, 1);
  else if (var == 4140)
    KIFMM_M2LW(4, 1, 4, 0);
  else if (var == 4241)
    KIFMM_M2LW(4, 2, 4, 1);
  else
    KIFMM_M2LW(4, 1, 4, 1);
#undef KIFMM_M2LW
#undef KIFMM_M2LW_R
}

// ---------------------------------------------------------------------------
// M2L on the matrix cores for any kernel and any number of right-hand sides
// (spectra in the [slot][rhs][dof][C] layout of the grouped plan).
//
// For one frequency f and neighbour direction n, the M2L of a target parent
// T is the complex product Y_T[(c,t)] += sum_{(s,j)} M_n[(c,t)][(s,j)]
// X_{T+n}[(s,j)] over the 8 TM outputs (child c, dof t) and the 8 SM inputs
// (source child s, dof j), M_n[(c,t)][(s,j)] = T_{optab[n][s][c]}[j][t](f).
// M_n does not depend on T, so the product is batched over columns =
// (target parent, right-hand side) pairs, 16 per matrix-core tile, in real form
//   [Yr; Yi] += [[Mr, -Mi], [Mi, Mr]] [Xr; Xi]
// with v_mfma_f64_16x16x4f64.  Row tile rt < TM holds 16 real rows: row r16
// is output rt*8 + r16%8 (flattened (c, t)), real part for r16 < 8.  The 16 SM
// real inputs are cut into 4 SM slices of 4; input kappa is part kappa/(8 SM)
// of flattened (s, j) = kappa % (8 SM).  With the output layout of the
// instruction (item v of lane l is row l/16 + 4v) lane (k, col) ends with
// (re, im) of outputs rt*8 + k and rt*8 + k + 4.
//
// A block covers FB consecutive frequencies and 4 tiles of 16 columns (one
// per wave).  For each direction the block builds [[Mr,-Mi],[Mi,Mr]] of each
// frequency in LDS from the frequency-major operator copy opsT [C][343][SM][TM]
// through a frequency-independent index table, and every wave applies them to
// its tile; a lane loads the FB frequencies of each source spectrum as one
// contiguous run.  pmask[q] has bit n set if
// parent q has a source child in direction n; a tile whose parents have none
// skips the direction.  The blocks of one column block and 8 neighbouring
// frequencies are dealt to one XCD side by side, so each 128-byte line of
// spectra (8 frequencies) is fetched into its L2 once for all 8.
constexpr int M2LV_BS = 256;

// Raw buffer loads (gfx9 resource word 3; offsets past the resource's size
// read zeros).
constexpr int BUF_DW3 = 0x00020000;
template <class R> __device__ inline R buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off, int soff);
template <> __device__ inline float buf_ld<float>(__amdgpu_buffer_rsrc_t r, unsigned off, int soff)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}
template <> __device__ inline double buf_ld<double>(__amdgpu_buffer_rsrc_t r, unsigned off, int soff)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}

// Precision traits of the parent-grouped M2L.  FP64: row r16 of a row tile is
// output r16 % 8, part r16 / 8, so accumulator items (h, h + 2) of lane k are
// (re, im) of output k + 4h.  FP32 (output item v of lane l at row
// 4*(l/16) + v): rows interleave (re, im), row r16 is output r16 / 2, part
// r16 % 2, so items (2h, 2h + 1) of lane k are (re, im) of output 2k + h.
template <class R>
struct Mv;
template <>
struct Mv<double> {
  typedef double2 C2;
  typedef v4d V4;
  __device__ static inline V4 mfma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static inline int row_out(int r16) { return r16 & 7; }
  __device__ static inline int row_im(int r16) { return r16 >> 3; }
  __device__ static inline int item_out(int k, int h) { return k + 4 * h; }
  __device__ static inline int item_re(int h) { return h; }
  __device__ static inline int item_im(int h) { return h + 2; }
};
template <>
struct Mv<float> {
  typedef float2 C2;
  typedef v4f V4;
  __device__ static inline V4 mfma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static inline int row_out(int r16) { return r16 >> 1; }
  __device__ static inline int row_im(int r16) { return r16 & 1; }
  __device__ static inline int item_out(int k, int h) { return 2 * k + h; }
  __device__ static inline int item_re(int h) { return 2 * h; }
  __device__ static inline int item_im(int h) { return 2 * h + 1; }
};

// The loads of a direction are issued before its barriers: the operator-
// table entries and source slots of direction n + 1 are loaded while
// direction n is applied, and the operator values (through a buffer resource:
// one 32-bit offset per entry for both frequencies) and the source spectra of
// a direction are fetched together, so a direction waits on one memory round
// trip instead of four.  Measured against the direction-by-direction order
// (1M Navier clusters, 262k Stokes sphere, M2L per evaluate): FP64 3x3 29.0 ->
// 26.7 ms, FP64 4x3 5.4 -> 4.7 ms, FP32 4x3 3.5 -> 2.9 ms, FP32 3x3 19.8 ->
// 19.9 ms.
//
// XA: the A operands come from the table expanded at set-up (atabv, [C][26]
// [NA] in the block's entry order, signs and zeros applied), one coalesced
// run per (frequency, direction), instead of gathers through aidx from the
// operators (1M Navier clusters, FP32: M2L 19.9 -> 16.8 ms).
template <class R, int SM, int TM, int FB, bool XA>
__global__ __launch_bounds__(M2LV_BS) void k_m2l_vec(int C, int Rc, long ncol, int ncb, const int* __restrict__ tslot,
                                                     const int* __restrict__ sslot, const int* __restrict__ pmask,
                                                     const int* __restrict__ aidx,
                                                     const typename Mv<R>::C2* __restrict__ opsT,
                                                     const R* __restrict__ atabv,
                                                     const typename Mv<R>::C2* __restrict__ effden,
                                                     typename Mv<R>::C2* __restrict__ effval)
{
  typedef typename Mv<R>::C2 C2;
  typedef typename Mv<R>::V4 V4;
  constexpr int NKS = 4 * SM, NH = NKS / 2, ST = SM * TM, NA = TM * NKS * 64;
  __shared__ R s_A[FB][NA];
  // items (group of 8 frequencies, column block, FB-frequency run in the
  // group): the runs of one 128-byte line of spectra run side by side on one XCD
  constexpr int RUNS = 8 / FB;
  const int ngrp = (C + 7) / 8, nitem = ngrp * ncb * RUNS;
  const int item = xcd_item(blockIdx.x, nitem);
  const int fg = item / (RUNS * ncb), rem = item - fg * RUNS * ncb, cb = rem / RUNS;
  const int f0 = fg * 8 + (rem - cb * RUNS) * FB;
  if (item >= nitem || f0 >= C) return;  // block-uniform
  const int nf = min(FB, C - f0);
  const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
  const long col = ((long)cb * 4 + w) * 16 + li;
  const int q = col < ncol ? (int)(col / Rc) : -1;
  const int r = col < ncol ? (int)(col - (long)(col / Rc) * Rc) : 0;
  const int msk = q >= 0 ? pmask[q] : 0;
  const int* ss = sslot + (long)(q < 0 ? 0 : q) * 26 * 8;
  V4 acc[FB][TM];
#pragma unroll
  for (int fb = 0; fb < FB; fb++)
#pragma unroll
    for (int rt = 0; rt < TM; rt++) acc[fb][rt] = V4{0, 0, 0, 0};
  // operator-table entries and source slots of the next direction, loaded one
  // direction ahead; the operator values and the source spectra of a
  // direction are then fetched together, before its barriers
  int ai[ST], sl[NH];
  auto fetch_idx = [&](int n) {
    if constexpr (!XA)
#pragma unroll
      for (int i = 0; i < ST; i++) ai[i] = aidx[(long)n * NA + t + M2LV_BS * i];
#pragma unroll
    for (int kh = 0; kh < NH; kh++) sl[kh] = q >= 0 ? ss[n * 8 + (kh * 4 + lk) / SM] : -1;
  };
  fetch_idx(0);
  // the operators of the block's frequencies through a buffer resource based
  // at frequency f0: one 32-bit byte offset per entry serves all FB
  // frequencies, which lie a scalar FSTR bytes apart
  constexpr unsigned FSTR = 343u * ST * sizeof(C2);
  const __amdgpu_buffer_rsrc_t orsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)(opsT + (long)f0 * 343 * ST), (short)0, (int)(FSTR * nf), BUF_DW3);
  for (int n = 0; n < 26; n++) {
    // the direction-n blocks [[Mr,-Mi],[Mi,Mr]] of the FB frequencies: entry
    // e is aidx[n][e] = 2 x (index of the real value among one frequency's
    // 343 sm tm complex operators) + 1 if negated, or -1 (adjacent pair)
    R av[FB][ST];
    if constexpr (XA) {
#pragma unroll
      for (int fb = 0; fb < FB; fb++) {
        const R* ap = atabv + ((long)(f0 + (fb < nf ? fb : 0)) * 26 + n) * NA + t;
#pragma unroll
        for (int i = 0; i < ST; i++) av[fb][i] = ap[M2LV_BS * i];
      }
    } else {
      unsigned ao[ST];
#pragma unroll
      for (int i = 0; i < ST; i++) ao[i] = (unsigned)(ai[i] < 0 ? 0 : ai[i] >> 1) * (unsigned)sizeof(R);
#pragma unroll
      for (int fb = 0; fb < FB; fb++)
#pragma unroll
        for (int i = 0; i < ST; i++) av[fb][i] = buf_ld<R>(orsrc, ao[i], fb < nf ? fb * (int)FSTR : 0);
    }
    // this lane's NH source spectra for the FB frequencies (contiguous in the
    // spectra layout), clamped to slot 0 / the run and zeroed below
    const bool use = __any((msk >> n) & 1);  // wave-uniform: a parent of the tile has direction n
    C2 x[NH][FB];
    bool xz[NH];
    if (use) {
#pragma unroll
      for (int kh = 0; kh < NH; kh++) {
        const int j = (kh * 4 + lk) % SM;
        xz[kh] = sl[kh] < 0;
        const C2* xp = effden + (((long)(xz[kh] ? 0 : sl[kh]) * Rc + r) * SM + j) * C + f0;
#pragma unroll
        for (int fb = 0; fb < FB; fb++) x[kh][fb] = xp[fb < nf ? fb : 0];
      }
    }
    // the entries' signs (waits for the operator values only: the spectra
    // loads stay in flight), then the next direction's indices
    if constexpr (!XA)
#pragma unroll
      for (int fb = 0; fb < FB; fb++)
#pragma unroll
        for (int i = 0; i < ST; i++) av[fb][i] = ai[i] < 0 ? (R)0 : ((ai[i] & 1) ? -av[fb][i] : av[fb][i]);
    if (n + 1 < 26) fetch_idx(n + 1);
    __syncthreads();  // the previous direction is done with s_A
#pragma unroll
    for (int fb = 0; fb < FB; fb++)
#pragma unroll
      for (int i = 0; i < ST; i++) s_A[fb][t + M2LV_BS * i] = av[fb][i];
    __syncthreads();
    if (!use) continue;
#pragma unroll
    for (int kh = 0; kh < NH; kh++)
#pragma unroll
      for (int fb = 0; fb < FB; fb++) {
        const R xr = xz[kh] ? (R)0 : x[kh][fb].x, xi = xz[kh] ? (R)0 : x[kh][fb].y;
#pragma unroll
        for (int rt = 0; rt < TM; rt++) {
          acc[fb][rt] = Mv<R>::mfma(s_A[fb][(rt * NKS + kh) * 64 + l], xr, acc[fb][rt]);
          acc[fb][rt] = Mv<R>::mfma(s_A[fb][(rt * NKS + kh + NH) * 64 + l], xi, acc[fb][rt]);
        }
      }
  }
  if (q < 0) return;
#pragma unroll
  for (int rt = 0; rt < TM; rt++)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int qo = rt * 8 + Mv<R>::item_out(lk, h), c = qo / TM, ti = qo - c * TM;
      const int tl = tslot[(long)q * 8 + c];
      if (tl < 0) continue;
      C2* out = effval + (((long)tl * Rc + r) * TM + ti) * C + f0;
#pragma unroll
      for (int fb = 0; fb < FB; fb++)
        if (fb < nf) {
          C2 v;
          v.x = acc[fb][rt][Mv<R>::item_re(h)];
          v.y = acc[fb][rt][Mv<R>::item_im(h)];
          out[fb] = v;
        }
    }
}

template <class R, int SM, int TM, int FB>
static void m2l_vec_launch(hipStream_t s, int C, int npar, int Rc, const int* tslot, const int* sslot,
                           const int* pmask, const int* aidx, const void* opsT, const void* atabv, const void* effden,
                           void* effval)
{
  typedef typename Mv<R>::C2 C2;
  const long ncol = (long)npar * Rc;
  const int ncb = (int)((ncol + 63) / 64);
  const dim3 grid(xcd_grid((C + 7) / 8 * ncb * (8 / FB)));
  if constexpr (sizeof(R) == 4)
    if (atabv) {
      hipLaunchKernelGGL((k_m2l_vec<R, SM, TM, FB, true>), grid, dim3(M2LV_BS), 0, s, C, Rc, ncol, ncb, tslot, sslot,
                         pmask, aidx, (const C2*)opsT, (const R*)atabv, (const C2*)effden, (C2*)effval);
      return;
    }
  hipLaunchKernelGGL((k_m2l_vec<R, SM, TM, FB, false>), grid, dim3(M2LV_BS), 0, s, C, Rc, ncol, ncb, tslot, sslot,
                     pmask, aidx, (const C2*)opsT, (const R*)nullptr, (const C2*)effden, (C2*)effval);
}

// atabv[(f * 26 + n) * na + e]: entry e of aidx[n] resolved against the
// operators of frequency f (FP32, signs applied, zeros for -1)
__global__ void k_expand_m2l_vec(int C, int na, int st, const int* __restrict__ aidx, const float* __restrict__ opsT,
                                 float* __restrict__ atabv)
{
  const long n = (long)C * 26 * na;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long fn = i / na;
    const int e = (int)(i - fn * na), f = (int)(fn / 26), d = (int)(fn - (long)f * 26);
    const int ai = aidx[(long)d * na + e];
    float v = 0.0f;
    if (ai >= 0) {
      v = opsT[(long)f * 343 * st * 2 + (ai >> 1)];
      if (ai & 1) v = -v;
    }
    atabv[i] = v;
  }
}

void expand_m2l_vec(hipStream_t s, int C, int sm, int tm, const int* aidx32, const float* opsT32, float* atabv)
{
  const int na = tm * 4 * sm * 64;
  const long n = (long)C * 26 * na;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_expand_m2l_vec, dim3(grid_for(n, 256)), dim3(256), 0, s, C, na, sm * tm, aidx32, opsT32, atabv);
}

template <class R>
static void m2l_matrix_t(hipStream_t s, int sm, int tm, int C, int npar, int Rc, const int* tslot, const int* sslot,
                         const int* pmask, const int* aidx, const void* opsT, const void* atabv, const void* effden,
                         void* effval)
{
  // frequencies per block: 2 measured best in both precisions (of 1, 2, 4, 8)
  constexpr int FB = 2;
  if (sm == 1 && tm == 1)
    m2l_vec_launch<R, 1, 1, FB>(s, C, npar, Rc, tslot, sslot, pmask, aidx, opsT, atabv, effden, effval);
  else if (sm == 4 && tm == 3)
    m2l_vec_launch<R, 4, 3, FB>(s, C, npar, Rc, tslot, sslot, pmask, aidx, opsT, atabv, effden, effval);
  else if (sm == 3 && tm == 3)
    m2l_vec_launch<R, 3, 3, FB>(s, C, npar, Rc, tslot, sslot, pmask, aidx, opsT, atabv, effden, effval);
}

void m2l_matrix(hipStream_t s, int sm, int tm, int C, int npar, int Rc, const int* tslot, const int* sslot,
                const int* pmask, const int* aidx, const void* opsT, const void* atabv, const void* effden,
                void* effval, int single)
{
  if (npar <= 0) return;
  if (single)
    m2l_matrix_t<float>(s, sm, tm, C, npar, Rc, tslot, sslot, pmask, aidx, opsT, atabv, effden, effval);
  else
    m2l_matrix_t<double>(s, sm, tm, C, npar, Rc, tslot, sslot, pmask, aidx, opsT, nullptr, effden, effval);
}

}  // namespace dev
}  // namespace kifmm
