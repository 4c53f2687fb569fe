// Measurement probes of the device's own ceilings, for the roofline figures:
// the HBM bandwidth a plain copy reaches, the LDS bandwidth of 128-bit reads
// (ds_read_b128), the shader clock during that LDS probe (cycle counter
// over the kernel's own time), and the FP32 FMA issue rate of the VALU.  No reference counterpart: the reference has
// no timing at all.
#include <functional>

#include "nts_internal.h"

// Four 16-byte loads in flight per lane before the stores.
__global__ void __launch_bounds__(256) k_probe_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    uint64_t n4) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n4; i += stride) dst[i] = src[i];
}

// Every lane issues eight 128-bit LDS reads (ds_read_b128) per iteration,
// all in one assembly block that also waits for them, so no read is still in
// flight when the block ends.  The 64 lanes of a wave read 1 KB of
// consecutive words per read (no bank conflicts).  One word of each read is
// folded into a result that only a result no launch can produce would store.
#define PROBE_LDS_WORDS 2048
__global__ void __launch_bounds__(256) k_probe_lds(uint32_t iters, uint32_t* __restrict__ sink,
                                                   uint64_t* __restrict__ cycles) {
  __shared__ uint4 s[PROBE_LDS_WORDS];
  for (uint32_t i = threadIdx.x; i < PROBE_LDS_WORDS; i += blockDim.x) s[i] = make_uint4(i, i * 3u, i * 5u, i * 7u);
  __syncthreads();
  // LDS byte addresses of the eight words this lane reads (the low half of
  // the flat address of a workgroup-local variable is its LDS offset)
  uint32_t ad[8];
#pragma unroll
  for (int k = 0; k < 8; ++k)
    ad[k] = (uint32_t)(uintptr_t)&s[(threadIdx.x + 256u * (uint32_t)k) & (PROBE_LDS_WORDS - 1u)];
  const uint64_t c0 = __builtin_readcyclecounter();
  const uint64_t w0 = wall_clock64();
  uint32_t acc = 0;
  for (uint32_t it = 0; it < iters; it += 8) {
    uint4 v0, v1, v2, v3, v4, v5, v6, v7;
    __asm__ volatile(
        "ds_read_b128 %0, %8\n\t"
        "ds_read_b128 %1, %9\n\t"
        "ds_read_b128 %2, %10\n\t"
        "ds_read_b128 %3, %11\n\t"
        "ds_read_b128 %4, %12\n\t"
        "ds_read_b128 %5, %13\n\t"
        "ds_read_b128 %6, %14\n\t"
        "ds_read_b128 %7, %15\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=v"(v0), "=v"(v1), "=v"(v2), "=v"(v3), "=v"(v4), "=v"(v5), "=v"(v6), "=v"(v7)
        : "v"(ad[0]), "v"(ad[1]), "v"(ad[2]), "v"(ad[3]), "v"(ad[4]), "v"(ad[5]), "v"(ad[6]), "v"(ad[7])
        : "memory");
    acc ^= v0.x ^ v1.y ^ v2.z ^ v3.w ^ v4.x ^ v5.y ^ v6.z ^ v7.w;
  }
  const uint64_t c1 = __builtin_readcyclecounter();
  const uint64_t w1 = wall_clock64();
  // the words are below 2^14, so is their XOR: this store never happens
  if (acc == 0xdeadbeefu) sink[threadIdx.x & 255u] = acc;
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    cycles[0] = c1 - c0;
    cycles[1] = w1 - w0;
  }
}

// Eight independent FP32 FMA chains per lane (v_fma_f32 on wave64), the
// loop counter and branch scalar: the VALU issue ceiling for FP32 work.  The
// operands come from the runtime value s (1.0) so nothing folds; a result no
// launch produces would be stored.  Block 0 times its loop with the cycle
// counter and the wall clock (for the clock of this run).
#define PROBE_FMA_CHAINS 8
__global__ void __launch_bounds__(256) k_probe_fma(uint32_t iters, float s, float* __restrict__ sink,
                                                   uint64_t* __restrict__ cycles) {
  const float lane = (float)(threadIdx.x & 63u);
  float a0 = lane * 1e-3f * s, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
  float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
  const float b = 0.999f * s, c = 1e-3f * s;
  const uint64_t c0 = __builtin_readcyclecounter();
  const uint64_t w0 = wall_clock64();
  for (uint32_t it = 0; it < iters; ++it) {
    a0 = __builtin_fmaf(a0, b, c);
    a1 = __builtin_fmaf(a1, b, c);
    a2 = __builtin_fmaf(a2, b, c);
    a3 = __builtin_fmaf(a3, b, c);
    a4 = __builtin_fmaf(a4, b, c);
    a5 = __builtin_fmaf(a5, b, c);
    a6 = __builtin_fmaf(a6, b, c);
    a7 = __builtin_fmaf(a7, b, c);
  }
  const uint64_t c1 = __builtin_readcyclecounter();
  const uint64_t w1 = wall_clock64();
  const float r = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
  if (r == -1.0f) sink[threadIdx.x] = r;  // the chains stay positive
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    cycles[0] = c1 - c0;
    cycles[1] = w1 - w0;
  }
}

static int time_launches(nts_ctx* ctx, int reps, double* best_ms, const std::function<int()>& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  NTS_HIP_TRY(hipEventCreate(&a));
  NTS_HIP_TRY(hipEventCreate(&b));
  *best_ms = 1e30;
  int rc = NTS_OK;
  for (int r = 0; r < reps && rc == NTS_OK; ++r) {
    (void)hipEventRecord(a, ctx->stream);
    rc = launch();
    (void)hipEventRecord(b, ctx->stream);
    if (hipEventSynchronize(b) != hipSuccess) rc = nts_fail(NTS_E_HIP, "probe: event wait failed");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    if (r > 0 && ms < *best_ms) *best_ms = ms;  // the first launch warms up
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

extern "C" int nts_probe_peaks(nts_ctx* ctx, uint64_t copy_bytes, double* hbm_copy_gbs, double* lds_read_tbs,
                               double* clock_ghz) {
  NTS_CHECK_ARG(ctx && hbm_copy_gbs && lds_read_tbs && clock_ghz && copy_bytes >= (1u << 20));
  const uint64_t n4 = copy_bytes / 16;
  void* buf = nullptr;
  int rc = nts_dev_alloc(ctx, &buf, 2 * n4 * 16 + 4096);
  if (rc) return rc;
  float4* src = (float4*)buf;
  float4* dst = src + n4;
  uint64_t* cycles = (uint64_t*)(dst + n4);
  uint32_t* sink = (uint32_t*)(cycles + 8);
  rc = nts_fill_u32(ctx, buf, 0x3f800000u, 2 * n4 * 4);
  double copy_ms = 0.0, lds_ms = 0.0;
  if (rc == NTS_OK)
    rc = time_launches(ctx, 8, &copy_ms, [&]() {
      NTS_LAUNCH(ctx, "probe_copy", k_probe_copy, dim3(4096), dim3(256), 0, (const float4*)src, dst, n4);
      return NTS_OK;
    });
  // 4096 workgroups of 256 lanes, 16 B per lane per iteration
  const uint32_t grid = 4096, iters = 8192;
  if (rc == NTS_OK)
    rc = time_launches(ctx, 5, &lds_ms, [&]() {
      NTS_LAUNCH(ctx, "probe_lds", k_probe_lds, dim3(grid), dim3(256), 0, iters, sink, cycles);
      return NTS_OK;
    });
  uint64_t cyc[2] = {0, 0};
  if (rc == NTS_OK) {
    if (hipMemcpy(cyc, cycles, 16, hipMemcpyDeviceToHost) != hipSuccess) rc = nts_fail(NTS_E_HIP, "probe readback");
  }
  (void)hipStreamSynchronize(ctx->stream);
  nts_dev_free(ctx, buf, 2 * n4 * 16 + 4096);
  if (rc) return rc;
  *hbm_copy_gbs = 2.0 * (double)(n4 * 16) / (copy_ms * 1e-3) / 1e9;
  const double lds_bytes = (double)grid * 256.0 * (double)iters * 16.0;
  *lds_read_tbs = lds_bytes / (lds_ms * 1e-3) / 1e12;
  // workgroup 0's read loop: shader cycles over its wall-clock time
  *clock_ghz = 0.0;
  int dev = 0, wall_khz = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && wall_khz > 0 &&
      cyc[1] > 0)
    *clock_ghz = (double)cyc[0] / ((double)cyc[1] / ((double)wall_khz * 1e3)) / 1e9;
  (void)hipGetLastError();
  return NTS_OK;
}

extern "C" int nts_probe_valu(nts_ctx* ctx, double* fma_tflops, double* fma_per_simd_cycle) {
  NTS_CHECK_ARG(ctx && fma_tflops && fma_per_simd_cycle);
  int dev = 0, cus = 0, wall_khz = 0;
  NTS_HIP_TRY(hipGetDevice(&dev));
  NTS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  NTS_HIP_TRY(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev));
  if (cus < 1) return nts_fail(NTS_E_HIP, "no compute units");
  void* buf = nullptr;
  int rc = nts_dev_alloc(ctx, &buf, 4096);
  if (rc) return rc;
  uint64_t* cycles = (uint64_t*)buf;
  float* sink = (float*)((char*)buf + 256);
  // eight 256-lane workgroups per CU: eight waves per SIMD
  const uint32_t grid = (uint32_t)cus * 8u, iters = 16384;
  double ms = 0.0;
  rc = time_launches(ctx, 5, &ms, [&]() {
    NTS_LAUNCH(ctx, "probe_fma", k_probe_fma, dim3(grid), dim3(256), 0, iters, 1.0f, sink, cycles);
    return NTS_OK;
  });
  uint64_t cyc[2] = {0, 0};
  if (rc == NTS_OK && hipMemcpy(cyc, cycles, 16, hipMemcpyDeviceToHost) != hipSuccess)
    rc = nts_fail(NTS_E_HIP, "probe readback");
  (void)hipStreamSynchronize(ctx->stream);
  nts_dev_free(ctx, buf, 4096);
  if (rc) return rc;
  const double lane_fmas = (double)grid * 256.0 * (double)iters * PROBE_FMA_CHAINS;
  *fma_tflops = 2.0 * lane_fmas / (ms * 1e-3) / 1e12;
  // wave instructions per SIMD per shader cycle, at the clock of this run
  double clk = 0.0;
  if (wall_khz > 0 && cyc[1] > 0) clk = (double)cyc[0] / ((double)cyc[1] / ((double)wall_khz * 1e3));
  *fma_per_simd_cycle = clk > 0.0 ? (lane_fmas / 64.0) / (4.0 * cus) / (ms * 1e-3 * clk) : 0.0;
  return NTS_OK;
}
