// issue_rates.hip -- register-only probes of the instruction mixes that bound
// the mixed-precision Biot-Savart pair kernel (k_bve_mfma) on gfx950, for the
// per-kernel floor table of the measurement notes.  Each probe runs a fixed
// instruction mix `iters` times on registers only (no memory traffic in the
// loop), every wave of the grid doing the same; the host times the launch and
// also reads the shader clock (s_memtime) around the loop in every wave.
//
// Mixes (one iteration each):
//   0  16 v_fma_f32, independent chains                    (full-rate reference)
//   1  16 v_rcp_f32, independent chains (x -> 1/x)         (reciprocal rate)
//   2   4 v_mfma_f32_16x16x4_f32, independent accumulators
//   3   4 v_mfma_f32_4x4x1_16b_f32, independent accumulators
//   4  the kernel's 256-pair step: one 16x16x4 product, 4 adds and 4
//      reciprocals on its output, four 4x4x1 accumulations (as k_bve_mfma)
//   5  as 4 with the four reciprocals batched: 3 products, one reciprocal,
//      6 products (y_q = 1/x_q from r = 1/(x0 x1 x2 x3))
//   6  half the waves run mix 2, the other half mix 0 (vector/matrix co-issue)
//   7   4 v_mfma_f32_16x16x32_bf16, independent accumulators
//   8  the 256-pair step with the first product as one 16x16x32 bf16 product
//      (operands split into three bf16 parts), alpha as its accumulator input
//      (no adds) and the four reciprocals batched
//   9  as 8 with four reciprocals
// C ABI: sbslpm_probe_issue(mix, blocks, waves_per_block, iters, out_cycles,
// out_ms) launches and synchronises; out_cycles receives the mean shader clock
// count per iteration per wave, out_ms the launch's event time.  The loops
// carry a dependence from one iteration to the next so that nothing is hoisted.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int MIX>
__global__ __launch_bounds__(1024) void k_probe(int iters, float seed, unsigned long long* cycles, float* sink) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  float r[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) r[i] = seed + 0.001f * (lane + i);
  v4f acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = v4f{seed, seed * 0.5f, 0.25f, 0.125f};
  v8bf abf, bbf;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    abf[k] = (__bf16)(r[k] * 0.5f);
    bbf[k] = (__bf16)(r[k + 8] * 0.25f);
  }
  const unsigned long long t0 = __builtin_amdgcn_s_memtime();
  for (int it = 0; it < iters; ++it) {
    if (MIX == 0 || (MIX == 6 && (wave & 1))) {
#pragma unroll
      for (int i = 0; i < 16; ++i) r[i] = fmaf(r[i], 0.999f, 0.0001f);
    } else if (MIX == 1) {
#pragma unroll
      for (int i = 0; i < 16; ++i) r[i] = __builtin_amdgcn_rcpf(r[i]);
    } else if (MIX == 2 || MIX == 6) {
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(r[i], r[i + 4], acc[i], 0, 0, 0);
    } else if (MIX == 3) {
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = __builtin_amdgcn_mfma_f32_4x4x1f32(r[i], r[i + 4], acc[i], 0, 0, 0);
    } else if (MIX == 7) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v8bf a, b;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          a[k] = (__bf16)r[(i + k) & 15];
          b[k] = (__bf16)r[(i + k + 4) & 15];
        }
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
      }
    } else if (MIX == 4 || MIX == 5 || MIX == 8 || MIX == 9) {
      v4f d;
      float x[4];
      if (MIX == 4 || MIX == 5) {
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(r[0], r[1], v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = d[q] + r[2 + q];
      } else {
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abf, bbf, v4f{r[2], r[3], r[4], r[5]}, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = d[q];
      }
      float y[4];
      if (MIX == 4 || MIX == 9) {
#pragma unroll
        for (int q = 0; q < 4; ++q) y[q] = __builtin_amdgcn_rcpf(x[q]);
      } else {
        const float p01 = x[0] * x[1], p23 = x[2] * x[3];
        const float rr = __builtin_amdgcn_rcpf(p01 * p23);
        const float r01 = rr * p23, r23 = rr * p01;
        y[0] = r01 * x[1];
        y[1] = r01 * x[0];
        y[2] = r23 * x[3];
        y[3] = r23 * x[2];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(r[8 + q], y[q], acc[q], 0, 0, 0);
      // carry a dependence into the next iteration's operands
      r[0] = acc[0][0] * 1e-30f + r[0];
      r[1] = acc[1][1] * 1e-30f + r[1];
      if (MIX >= 8) {
        r[2] = acc[2][2] * 1e-30f + r[2];
        abf[0] = (__bf16)r[2];
      }
    }
  }
  const unsigned long long t1 = __builtin_amdgcn_s_memtime();
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 16; ++i) s += r[i];
#pragma unroll
  for (int i = 0; i < 4; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  sink[blockIdx.x * blockDim.x + threadIdx.x] = s;
  if (lane == 0) atomicAdd(cycles, t1 - t0);
}

}  // namespace

extern "C" int sbslpm_probe_issue(int mix, int blocks, int waves_per_block, int iters, double* out_cycles,
                                  double* out_ms) {
  if (mix < 0 || mix > 9 || blocks < 1 || waves_per_block < 1 || waves_per_block > 16 || iters < 1) return 1;
  unsigned long long* cyc = nullptr;
  float* sink = nullptr;
  const size_t nthreads = (size_t)blocks * waves_per_block * 64;
  if (hipMalloc(&cyc, sizeof(unsigned long long)) != hipSuccess) return 2;
  if (hipMalloc(&sink, nthreads * sizeof(float)) != hipSuccess) {
    (void)hipFree(cyc);
    return 2;
  }
  int rc = 0;
  hipEvent_t e0, e1;
  if (hipMemset(cyc, 0, sizeof(unsigned long long)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = 3;
  if (!rc) {
    const dim3 grid(blocks), block(64 * waves_per_block);
    (void)hipEventRecord(e0, nullptr);
    switch (mix) {
      case 0: hipLaunchKernelGGL(k_probe<0>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 1: hipLaunchKernelGGL(k_probe<1>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 2: hipLaunchKernelGGL(k_probe<2>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 3: hipLaunchKernelGGL(k_probe<3>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 4: hipLaunchKernelGGL(k_probe<4>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 5: hipLaunchKernelGGL(k_probe<5>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 6: hipLaunchKernelGGL(k_probe<6>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 7: hipLaunchKernelGGL(k_probe<7>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      case 8: hipLaunchKernelGGL(k_probe<8>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
      default: hipLaunchKernelGGL(k_probe<9>, grid, block, 0, nullptr, iters, 1.0f, cyc, sink); break;
    }
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess) rc = 4;
  }
  if (!rc) {
    float ms = 0.f;
    unsigned long long total = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (hipMemcpy(&total, cyc, sizeof(total), hipMemcpyDeviceToHost) != hipSuccess) rc = 5;
    *out_cycles = (double)total / ((double)blocks * waves_per_block * iters);
    *out_ms = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(cyc);
  (void)hipFree(sink);
  return rc;
}
