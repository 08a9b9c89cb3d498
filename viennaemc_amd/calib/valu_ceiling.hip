// valu_ceiling.hip -- issue-rate calibration of the vector ALU on one GPU.
//
// Streams of one instruction class each (eight independent chains per lane,
// no memory traffic inside the loop, so only issue limits the rate), timed
// with HIP events on the null stream.  Each kernel runs M iterations; the
// instructions per iteration of each stream are fixed by its loop body:
//   fma   8 x v_fma_f32
//   int   8 x v_add_u32 + 8 x v_xor_b32
//   mad64 8 x v_mad_u64_u32 + 8 x v_xor_b32 (the Philox multiply)
//   trans 8 x v_exp_f32 + 8 x v_mul_f32
//   mix   4 x v_fma_f32 + 4 x v_add_u32 + 4 x v_xor_b32
//   sel   8 x v_add_u32 + 8 x v_cndmask_b32
// Output: one JSON line with the launch shape, the device clock attribute
// and the best time of three launches per stream.  rocprofv3 --pmc on this
// program gives each stream's instruction-class counts, so the counters'
// meaning and the per-class issue cost can be read off directly.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kChains = 8;

__global__ __launch_bounds__(256) void s_fma(float *out, int M, float b, float c) {
  float a[kChains];
#pragma unroll
  for (int i = 0; i < kChains; i++)
    a[i] = threadIdx.x * 0.001f + i;
  for (int m = 0; m < M; m++) {
#pragma unroll
    for (int i = 0; i < kChains; i++)
      a[i] = __builtin_fmaf(a[i], b, c);
  }
  float s = 0;
#pragma unroll
  for (int i = 0; i < kChains; i++)
    s += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void s_int(uint32_t *out, int M, uint32_t k) {
  uint32_t x[kChains];
#pragma unroll
  for (int i = 0; i < kChains; i++)
    x[i] = threadIdx.x * 2654435761u + i * k;
  for (int m = 0; m < M; m++) {
#pragma unroll
    for (int i = 0; i < kChains; i++)
      x[i] ^= x[(i + 1) & (kChains - 1)] + k;
  }
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kChains; i++)
    s ^= x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void s_mad64(uint32_t *out, int M, uint32_t K) {
  uint32_t x[kChains];
#pragma unroll
  for (int i = 0; i < kChains; i++)
    x[i] = threadIdx.x * 2654435761u + i;
  for (int m = 0; m < M; m++) {
#pragma unroll
    for (int i = 0; i < kChains; i++) {
      const uint64_t p = (uint64_t)x[i] * K;
      x[i] = (uint32_t)(p >> 32) ^ (uint32_t)p;
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kChains; i++)
    s ^= x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void s_trans(float *out, int M, float c) {
  float a[kChains];
#pragma unroll
  for (int i = 0; i < kChains; i++)
    a[i] = threadIdx.x * 1e-6f + i * 1e-3f;
  for (int m = 0; m < M; m++) {
#pragma unroll
    for (int i = 0; i < kChains; i++)
      a[i] = __builtin_amdgcn_exp2f(a[i]) * c;
  }
  float s = 0;
#pragma unroll
  for (int i = 0; i < kChains; i++)
    s += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void s_mix(float *out, int M, float b, float c, uint32_t k) {
  float a[4];
  uint32_t x[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    a[i] = threadIdx.x * 0.001f + i;
    x[i] = threadIdx.x * 2654435761u + i * k;
  }
  for (int m = 0; m < M; m++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      a[i] = __builtin_fmaf(a[i], b, c);
      x[i] ^= x[(i + 1) & 3] + k;
    }
  }
  float s = 0;
#pragma unroll
  for (int i = 0; i < 4; i++)
    s += a[i] + (float)(x[i] & 7u);
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void s_sel(uint32_t *out, int M, uint32_t k) {
  uint32_t x[kChains];
#pragma unroll
  for (int i = 0; i < kChains; i++)
    x[i] = threadIdx.x * 2654435761u + i * k;
  const bool p = (threadIdx.x & 1) != 0;
  for (int m = 0; m < M; m++) {
#pragma unroll
    for (int i = 0; i < kChains; i++)
      x[i] = p ? x[(i + 1) & (kChains - 1)] : x[(i + 3) & (kChains - 1)] + k;
  }
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kChains; i++)
    s ^= x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

bool ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return false;
  }
  return true;
}

}  // namespace

int main() {
  const int blocks = 8192, threads = 256, M = 2048;
  void *out = nullptr;
  if (!ok(hipMalloc(&out, (size_t)blocks * threads * 4), "hipMalloc"))
    return 1;
  int clk = 0, cus = 0;
  if (!ok(hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, 0), "clock") ||
      !ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "cus"))
    return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event"))
    return 1;
  const char *names[6] = {"fma", "int", "mad64", "trans", "mix", "sel"};
  const int per_iter[6] = {8, 16, 16, 16, 12, 16};
  std::printf("{\"blocks\": %d, \"threads\": %d, \"iterations\": %d, \"clock_khz\": %d, "
              "\"cus\": %d, \"streams\": {",
              blocks, threads, M, clk, cus);
  for (int kind = 0; kind < 6; kind++) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
      if (!ok(hipEventRecord(e0, 0), "record"))
        return 1;
      if (kind == 0)
        s_fma<<<blocks, threads>>>((float *)out, M, 0.999f, 0.001f);
      if (kind == 1)
        s_int<<<blocks, threads>>>((uint32_t *)out, M, 0x9E3779B9u);
      if (kind == 2)
        s_mad64<<<blocks, threads>>>((uint32_t *)out, M, 0xD2511F53u);
      if (kind == 3)
        s_trans<<<blocks, threads>>>((float *)out, M, 0.5f);
      if (kind == 4)
        s_mix<<<blocks, threads>>>((float *)out, M, 0.999f, 0.001f, 0x9E3779B9u);
      if (kind == 5)
        s_sel<<<blocks, threads>>>((uint32_t *)out, M, 0x9E3779B9u);
      if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1, 0), "record") ||
          !ok(hipEventSynchronize(e1), "sync"))
        return 1;
      float ms = 0;
      if (!ok(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
        return 1;
      best = ms < best ? ms : best;
    }
    // wave64 instructions of the loop body per wave, over the whole grid
    const double waves = (double)blocks * threads / 64.0;
    const double instr = waves * (double)M * per_iter[kind];
    std::printf("%s\"%s\": {\"ms\": %.4f, \"loop_wave_instr\": %.6g}", kind ? ", " : "",
                names[kind], best, instr);
  }
  std::printf("}}\n");
  (void)hipFree(out);
  return 0;
}
