// valu_rates.hip -- issue-rate probe of gfx950's integer VALU operations and
// LDS read bandwidth, the two roofs of the step loop.
//
// VALU: every variant runs ITER x 8 copies of one instruction per wave with
// 4 waves per SIMD (1024 workgroups of 256 threads on 256 CUs), the operands
// independent of the previous copy, and prints cycles per wave-instruction per
// SIMD at the stated clock.  LDS: every lane reads its own slots of a 16 KB
// workgroup array in a loop, four reads in flight per wave (ds_read_b32 / _b64
// / _b128), and the probe prints bytes per clock per CU.  Nothing is read or written outside the
// kernel's own output array and LDS array.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#define ITER 100000
#define CLOCK_HZ 2.4e9
#define VK(name, body)                                                                                              \
  __global__ void __launch_bounds__(256) name(int *out, int x0) {                                                   \
    uint32_t a = threadIdx.x + x0, b = a * 3u, c = a ^ 5u, d = a + 7u;                                             \
    uint64_t p = ((uint64_t)a << 32) | b, q = ((uint64_t)c << 32) | d;                                            \
    for (int i = 0; i < ITER; ++i) { body }                                                                         \
    out[blockIdx.x * blockDim.x + threadIdx.x] = (int)(a + b + c + d + (uint32_t)p + (uint32_t)q);              \
  }
#define R8(s) s s s s s s s s
VK(k_add, asm volatile(R8("v_add_u32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_sub, asm volatile(R8("v_sub_u32_e32 %0, 60, %1\n") : "+v"(a) : "v"(b));)
VK(k_and, asm volatile(R8("v_and_b32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_or, asm volatile(R8("v_or_b32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_xor, asm volatile(R8("v_xor_b32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_lshr, asm volatile(R8("v_lshrrev_b32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_ashr, asm volatile(R8("v_ashrrev_i32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_lshl, asm volatile(R8("v_lshlrev_b32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_add3, asm volatile(R8("v_add3_u32 %0, %1, %2, -1\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_and_or, asm volatile(R8("v_and_or_b32 %0, %1, %2, %3\n") : "+v"(a) : "v"(b), "v"(c), "v"(d));)
VK(k_lshl_add, asm volatile(R8("v_lshl_add_u32 %0, %1, 3, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_lshl_or, asm volatile(R8("v_lshl_or_b32 %0, %1, 16, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_bfe, asm volatile(R8("v_bfe_u32 %0, %1, %2, 1\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_bfe_i32, asm volatile(R8("v_bfe_i32 %0, %1, 2, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_min, asm volatile(R8("v_min_u32_e32 %0, %1, %2\n") : "+v"(a) : "v"(b), "v"(c));)
VK(k_cndmask, asm volatile(R8("v_cndmask_b32_e64 %0, %1, %2, s[20:21]\n") : "+v"(a) : "v"(b), "v"(c) : "s20", "s21");)
VK(k_cmp32, asm volatile(R8("v_cmp_eq_u32_e64 s[20:21], %0, %1\n") :: "v"(a), "v"(b) : "s20", "s21");)
VK(k_cmp64, asm volatile(R8("v_cmp_eq_u64_e64 s[20:21], %0, %1\n") :: "v"(p), "v"(q) : "s20", "s21");)
VK(k_add_bfe, asm volatile(R8("v_add_u32_e32 %0, %2, %3\n v_bfe_u32 %1, %2, %3, 1\n") : "+v"(a), "+v"(d) : "v"(b), "v"(c));)

#define LDS_WORDS 4096
// Four reads in flight per wave between waits (the wait sits in the same asm
// block as the reads, so no read is outstanding when the block ends); the
// lane's slot is (tid * 16) mod 4 KB plus offsets 0, 4, 8 and 12 KB, inside
// the 16 KB array.
template <int W>
__global__ void __launch_bounds__(256) k_lds(uint32_t *out, int x0) {
  __shared__ __align__(16) uint32_t s[LDS_WORDS];
  for (int i = threadIdx.x; i < LDS_WORDS; i += blockDim.x) s[i] = (uint32_t)(i * 2654435761u) ^ (uint32_t)x0;
  __syncthreads();
  const uint32_t addr =
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) uint32_t *)s + ((threadIdx.x * 16u) & 4095u);
  uint32_t acc = 0;
  for (int i = 0; i < ITER / 10; ++i) {
    if (W == 1) {
      uint32_t v0, v1, v2, v3;
      asm volatile("ds_read_b32 %0, %4\n ds_read_b32 %1, %4 offset:4096\n ds_read_b32 %2, %4 offset:8192\n"
                   " ds_read_b32 %3, %4 offset:12288\n s_waitcnt lgkmcnt(0)"
                   : "=v"(v0), "=v"(v1), "=v"(v2), "=v"(v3) : "v"(addr));
      acc ^= v0 ^ v1 ^ v2 ^ v3;
    } else if (W == 2) {
      uint2 v0, v1, v2, v3;
      asm volatile("ds_read_b64 %0, %4\n ds_read_b64 %1, %4 offset:4096\n ds_read_b64 %2, %4 offset:8192\n"
                   " ds_read_b64 %3, %4 offset:12288\n s_waitcnt lgkmcnt(0)"
                   : "=v"(v0), "=v"(v1), "=v"(v2), "=v"(v3) : "v"(addr));
      acc ^= v0.x ^ v0.y ^ v1.x ^ v1.y ^ v2.x ^ v2.y ^ v3.x ^ v3.y;
    } else {
      uint4 v0, v1, v2, v3;
      asm volatile("ds_read_b128 %0, %4\n ds_read_b128 %1, %4 offset:4096\n ds_read_b128 %2, %4 offset:8192\n"
                   " ds_read_b128 %3, %4 offset:12288\n s_waitcnt lgkmcnt(0)"
                   : "=v"(v0), "=v"(v1), "=v"(v2), "=v"(v3) : "v"(addr));
      acc ^= v0.x ^ v0.y ^ v0.z ^ v0.w ^ v1.x ^ v1.y ^ v1.z ^ v1.w ^ v2.x ^ v2.y ^ v2.z ^ v2.w ^ v3.x ^ v3.y ^ v3.z ^
             v3.w;
    }
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

int main() {
  const int grid = 1024, wg = 256;
  int *out = nullptr;
  if (hipMalloc(&out, (size_t)grid * wg * sizeof(int)) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  typedef void (*F)(int *, int);
  const F fs[] = {k_add, k_sub, k_and, k_or, k_xor, k_lshr, k_ashr, k_lshl, k_add3, k_and_or, k_lshl_add,
                  k_lshl_or, k_bfe, k_bfe_i32, k_min, k_cndmask, k_cmp32, k_cmp64, k_add_bfe};
  const char *nm[] = {"v_add_u32", "v_sub_u32", "v_and_b32", "v_or_b32", "v_xor_b32", "v_lshrrev_b32",
                      "v_ashrrev_i32", "v_lshlrev_b32", "v_add3_u32", "v_and_or_b32", "v_lshl_add_u32",
                      "v_lshl_or_b32", "v_bfe_u32", "v_bfe_i32", "v_min_u32", "v_cndmask_b32", "v_cmp_eq_u32",
                      "v_cmp_eq_u64", "v_add_u32 + v_bfe_u32"};
  const int per[] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2};
  printf("# VALU: cycles per wave-instruction per SIMD at %.1f GHz, 4 waves per SIMD\n", CLOCK_HZ / 1e9);
  for (int k = 0; k < (int)(sizeof fs / sizeof fs[0]); ++k) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      hipLaunchKernelGGL(fs[k], dim3(grid), dim3(wg), 0, 0, out, rep);
      if (hipEventRecord(e0, 0) != hipSuccess) return 1;
      hipLaunchKernelGGL(fs[k], dim3(grid), dim3(wg), 0, 0, out, rep);
      if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
      float ms = 0;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
      best = ms < best ? ms : best;
    }
    const double cyc = best * 1e-3 * CLOCK_HZ, ninst = 4.0 * ITER * 8 * per[k];
    printf("%-24s %8.3f ms  %5.2f cycles\n", nm[k], best, cyc / ninst);
  }
  printf("# LDS reads: bytes per clock per CU at %.1f GHz (every lane its own slot, 16 waves per CU)\n", CLOCK_HZ / 1e9);
  typedef void (*G)(uint32_t *, int);
  const G gs[] = {k_lds<1>, k_lds<2>, k_lds<4>};
  const char *gn[] = {"ds_read_b32", "ds_read_b64", "ds_read_b128"};
  const int bytes[] = {4, 8, 16};
  for (int k = 0; k < 3; ++k) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      hipLaunchKernelGGL(gs[k], dim3(grid), dim3(wg), 0, 0, (uint32_t *)out, rep);
      if (hipEventRecord(e0, 0) != hipSuccess) return 1;
      hipLaunchKernelGGL(gs[k], dim3(grid), dim3(wg), 0, 0, (uint32_t *)out, rep);
      if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
      float ms = 0;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
      best = ms < best ? ms : best;
    }
    const double b = (double)grid * wg * (ITER / 10) * 4.0 * bytes[k];
    printf("%-24s %8.3f ms  %6.1f B/clk/CU  (%.2f TB/s)\n", gn[k], best, b / (best * 1e-3 * CLOCK_HZ) / 256.0,
           b / (best * 1e-3) / 1e12);
  }
  return 0;
}
