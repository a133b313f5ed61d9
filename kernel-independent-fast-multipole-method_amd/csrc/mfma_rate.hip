// mfma_rate.hip -- sustained matrix-core rate of the MI355X for the MFMA
// shapes the evaluator uses (16x16x4 FP32 and FP64) and the 16x16x32 BF16 /
// FP16 ones: every wave runs ITER rounds of 8 independent accumulation
// chains from registers, no memory traffic inside the loop.  The rates are
// the practical ceilings quoted beside the published peaks in the roofline
// notes (the published peaks assume the 2.4 GHz peak clock).
//
// Second table: whether vector (VALU) instructions overlap FP32 MFMAs on a
// SIMD.  Each round of 8 chains is interleaved with NV independent FP32 FMAs
// per MFMA (one of them a v_rsq_f32 in the last rows), and the same FMAs are
// timed alone; the time of one MFMA slot is printed as SIMD cycles at the
// 2.4 GHz clock.  If the units overlapped, MFMA + FMAs would cost the larger
// of the two, not their sum.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

constexpr int BS = 256, NCHAIN = 8;

__global__ __launch_bounds__(BS) void k_f32(int iter, float* out)
{
  v4f acc[NCHAIN];
  for (int i = 0; i < NCHAIN; i++) acc[i] = v4f{0, 0, 0, 0};
  const float a = threadIdx.x * 1e-3f, b = blockIdx.x * 1e-3f;
  for (int it = 0; it < iter; it++)
#pragma unroll
    for (int i = 0; i < NCHAIN; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
  float s = 0;
  for (int i = 0; i < NCHAIN; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * BS + threadIdx.x] = s;
}

__global__ __launch_bounds__(BS) void k_f64(int iter, float* out)
{
  v4d acc[NCHAIN];
  for (int i = 0; i < NCHAIN; i++) acc[i] = v4d{0, 0, 0, 0};
  const double a = threadIdx.x * 1e-3, b = blockIdx.x * 1e-3;
  for (int it = 0; it < iter; it++)
#pragma unroll
    for (int i = 0; i < NCHAIN; i++) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  double s = 0;
  for (int i = 0; i < NCHAIN; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * BS + threadIdx.x] = (float)s;
}

__global__ __launch_bounds__(BS) void k_bf16(int iter, float* out)
{
  v4f acc[NCHAIN];
  for (int i = 0; i < NCHAIN; i++) acc[i] = v4f{0, 0, 0, 0};
  v8s a, b;
  for (int i = 0; i < 8; i++) {
    a[i] = (short)(threadIdx.x + i);
    b[i] = (short)(blockIdx.x + i);
  }
  for (int it = 0; it < iter; it++)
#pragma unroll
    for (int i = 0; i < NCHAIN; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
  float s = 0;
  for (int i = 0; i < NCHAIN; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * BS + threadIdx.x] = s;
}

__global__ __launch_bounds__(BS) void k_f16(int iter, float* out)
{
  v4f acc[NCHAIN];
  for (int i = 0; i < NCHAIN; i++) acc[i] = v4f{0, 0, 0, 0};
  v8h a, b;
  for (int i = 0; i < 8; i++) {
    a[i] = (_Float16)(threadIdx.x * 1e-3f);
    b[i] = (_Float16)(blockIdx.x * 1e-3f);
  }
  for (int it = 0; it < iter; it++)
#pragma unroll
    for (int i = 0; i < NCHAIN; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[i], 0, 0, 0);
  float s = 0;
  for (int i = 0; i < NCHAIN; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * BS + threadIdx.x] = s;
}

template <int NV, int MF, int RSQ>
__global__ __launch_bounds__(BS) void k_mix(int iter, float* out)
{
  v4f acc[NCHAIN];
  for (int i = 0; i < NCHAIN; i++) acc[i] = v4f{0, 0, 0, 0};
  float f[8];
  for (int j = 0; j < 8; j++) f[j] = threadIdx.x * 1e-3f + j;
  const float a = threadIdx.x * 1e-3f, b = blockIdx.x * 1e-3f;
  for (int it = 0; it < iter; it++) {
#pragma unroll
    for (int i = 0; i < NCHAIN; i++) {
      if (MF) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const int j = (i * NV + v) & 7;
        if (RSQ && v == 0)
          f[j] = __builtin_amdgcn_rsqf(f[j] + 1.0f);
        else
          f[j] = __builtin_fmaf(f[j], 0.999f, 0.5f);
      }
    }
  }
  float s = 0;
  for (int i = 0; i < NCHAIN; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  for (int j = 0; j < 8; j++) s += f[j];
  out[blockIdx.x * BS + threadIdx.x] = s;
}

bool ok(hipError_t e, const char* what)
{
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "mfma_rate: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main()
{
  const int nb = 256 * 8, iter = 4000;  // 8 blocks of 4 waves per CU
  float* out = nullptr;
  if (!ok(hipMalloc((void**)&out, sizeof(float) * nb * BS), "hipMalloc")) return 1;
  hipEvent_t a, b;
  if (!ok(hipEventCreate(&a), "hipEventCreate") || !ok(hipEventCreate(&b), "hipEventCreate")) return 1;
  struct Shape {
    const char* name;
    double flops;  // 2 M N K per instruction
    double peak;   // published dense peak, TFLOP/s
  } sh[4] = {{"FP32 16x16x4", 2.0 * 16 * 16 * 4, 157.3},
             {"FP64 16x16x4", 2.0 * 16 * 16 * 4, 78.6},
             {"BF16 16x16x32", 2.0 * 16 * 16 * 32, 2500.0},
             {"FP16 16x16x32", 2.0 * 16 * 16 * 32, 2500.0}};
  for (int rep = 0; rep < 2; rep++)  // the first round warms the clocks up
    for (int k = 0; k < 4; k++) {
      if (!ok(hipEventRecord(a), "hipEventRecord")) return 1;
      if (k == 0) hipLaunchKernelGGL(k_f32, dim3(nb), dim3(BS), 0, 0, iter, out);
      if (k == 1) hipLaunchKernelGGL(k_f64, dim3(nb), dim3(BS), 0, 0, iter, out);
      if (k == 2) hipLaunchKernelGGL(k_bf16, dim3(nb), dim3(BS), 0, 0, iter, out);
      if (k == 3) hipLaunchKernelGGL(k_f16, dim3(nb), dim3(BS), 0, 0, iter, out);
      if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(b), "hipEventRecord") ||
          !ok(hipEventSynchronize(b), "hipEventSynchronize"))
        return 1;
      float ms = 0;
      if (!ok(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime")) return 1;
      const double tf = (double)nb * (BS / 64) * iter * NCHAIN * sh[k].flops / (ms * 1e-3) / 1e12;
      if (rep == 1)
        std::printf("%-14s %8.3f ms  %7.1f TFLOP/s  (%.0f %% of the published %.1f)\n", sh[k].name, ms, tf,
                    100.0 * tf / sh[k].peak, sh[k].peak);
    }
  struct Mix {
    const char* name;
    void (*k)(int, float*);
  } mix[] = {{"MFMA", k_mix<0, 1, 0>},
             {"MFMA + 4 FMA", k_mix<4, 1, 0>},
             {"MFMA + 8 FMA", k_mix<8, 1, 0>},
             {"MFMA + 7 FMA + 1 rsq", k_mix<8, 1, 1>},
             {"4 FMA", k_mix<4, 0, 0>},
             {"8 FMA", k_mix<8, 0, 0>},
             {"7 FMA + 1 rsq", k_mix<8, 0, 1>}};
  std::printf("\nFP32 16x16x4 MFMA beside independent FP32 vector instructions (cycles per MFMA slot per SIMD, 2.4 GHz):\n");
  for (const Mix& m : mix) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
      if (!ok(hipEventRecord(a), "hipEventRecord")) return 1;
      hipLaunchKernelGGL(m.k, dim3(nb), dim3(BS), 0, 0, iter / 2, out);
      if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(b), "hipEventRecord") ||
          !ok(hipEventSynchronize(b), "hipEventSynchronize"))
        return 1;
      float ms = 0;
      if (!ok(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime")) return 1;
      best = ms < best ? ms : best;
    }
    const double slots = (double)nb * (BS / 64) * (iter / 2) * NCHAIN / 1024.0;  // per SIMD
    std::printf("%-22s %8.3f ms  %6.1f cycles\n", m.name, best, best * 1e-3 * 2.4e9 / slots);
  }
  (void)hipFree(out);
  return 0;
}
