// valu_rates.hip -- issue rate of the VALU instructions the walk kernel uses,
// on gfx950: eight independent dependency chains per lane, 4 or 8 waves per
// SIMD, two loop lengths differenced to cancel launch costs.  Prints lane-ops
// per second and per CU per clock (at 2.4 GHz), and the shader clock
// (s_memtime against the 100 MHz s_memrealtime).  The roofline of bench.py
// takes its MUL32/ALU32 issue ratio from here.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>
#define B8(S) S S S S S S S S
#define CH(INS, A) asm volatile(INS : "+v"(A) : "v"(b));
#define BODY(INS) CH(INS,a0) CH(INS,a1) CH(INS,a2) CH(INS,a3) CH(INS,a4) CH(INS,a5) CH(INS,a6) CH(INS,a7)
#define K32(NAME, INS) \
__global__ void NAME(uint32_t *out, int iters, uint64_t *clk) { \
  uint32_t a0=threadIdx.x,a1=a0+1,a2=a0+2,a3=a0+3,a4=a0+4,a5=a0+5,a6=a0+6,a7=a0+7, b=blockIdx.x|1; \
  uint64_t t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime(); \
  for (int i = 0; i < iters; ++i) { BODY(INS) BODY(INS) BODY(INS) BODY(INS) } \
  uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime(); \
  out[blockIdx.x*blockDim.x+threadIdx.x] = a0^a1^a2^a3^a4^a5^a6^a7; \
  if (threadIdx.x==0 && blockIdx.x==0) { clk[0] = t1 - t0; clk[1] = r1 - r0; } }
K32(k_add, "v_add_u32 %0, %0, %1")
K32(k_xor, "v_xor_b32 %0, %0, %1")
K32(k_mov, "v_mov_b32 %0, %1")
K32(k_mullo, "v_mul_lo_u32 %0, %0, %1")
K32(k_mulhi, "v_mul_hi_u32 %0, %0, %1")
K32(k_ffbl, "v_ffbl_b32 %0, %0")
K32(k_addmul, "v_add_u32 %0, %0, %1\n\tv_mul_lo_u32 %0, %0, %1")
#define K64(NAME, INS, CLOB, BT) \
__global__ void NAME(uint32_t *out, int iters, uint64_t *clk) { \
  uint64_t a0=threadIdx.x,a1=a0+1,a2=a0+2,a3=a0+3,a4=a0+4,a5=a0+5,a6=a0+6,a7=a0+7; BT b=blockIdx.x|1; \
  uint64_t t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime(); \
  for (int i = 0; i < iters; ++i) { \
    asm volatile(INS : "+v"(a0) : "v"(b) : CLOB); asm volatile(INS : "+v"(a1) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a2) : "v"(b) : CLOB); asm volatile(INS : "+v"(a3) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a4) : "v"(b) : CLOB); asm volatile(INS : "+v"(a5) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a6) : "v"(b) : CLOB); asm volatile(INS : "+v"(a7) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a0) : "v"(b) : CLOB); asm volatile(INS : "+v"(a1) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a2) : "v"(b) : CLOB); asm volatile(INS : "+v"(a3) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a4) : "v"(b) : CLOB); asm volatile(INS : "+v"(a5) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a6) : "v"(b) : CLOB); asm volatile(INS : "+v"(a7) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a0) : "v"(b) : CLOB); asm volatile(INS : "+v"(a1) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a2) : "v"(b) : CLOB); asm volatile(INS : "+v"(a3) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a4) : "v"(b) : CLOB); asm volatile(INS : "+v"(a5) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a6) : "v"(b) : CLOB); asm volatile(INS : "+v"(a7) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a0) : "v"(b) : CLOB); asm volatile(INS : "+v"(a1) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a2) : "v"(b) : CLOB); asm volatile(INS : "+v"(a3) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a4) : "v"(b) : CLOB); asm volatile(INS : "+v"(a5) : "v"(b) : CLOB); \
    asm volatile(INS : "+v"(a6) : "v"(b) : CLOB); asm volatile(INS : "+v"(a7) : "v"(b) : CLOB); } \
  uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime(); \
  out[blockIdx.x*blockDim.x+threadIdx.x] = (uint32_t)(a0^a1^a2^a3^a4^a5^a6^a7); \
  if (threadIdx.x==0 && blockIdx.x==0) { clk[0] = t1 - t0; clk[1] = r1 - r0; } }
K64(k_lshr64, "v_lshrrev_b64 %0, 3, %0", "vcc", uint64_t)
K64(k_mad64, "v_mad_u64_u32 %0, s[40:41], %1, %1, %0", "s40", uint32_t)
K64(k_lshladd64, "v_lshl_add_u64 %0, %0, 0, %1", "vcc", uint64_t)
K64(k_cmp64, "v_cmp_le_u64 vcc, %0, %1", "vcc", uint64_t)
#define CHC(INS, A) asm volatile(INS : "+v"(A) : "v"(b) : "vcc");
#define BODYC(INS) CHC(INS,a0) CHC(INS,a1) CHC(INS,a2) CHC(INS,a3) CHC(INS,a4) CHC(INS,a5) CHC(INS,a6) CHC(INS,a7)
__global__ void k_subco(uint32_t *out, int iters, uint64_t *clk) {
  uint32_t a0=threadIdx.x,a1=a0+1,a2=a0+2,a3=a0+3,a4=a0+4,a5=a0+5,a6=a0+6,a7=a0+7, b=blockIdx.x|1;
  uint64_t t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
  for (int i = 0; i < iters; ++i) { BODYC("v_sub_co_u32 %0, vcc, %0, %1") BODYC("v_sub_co_u32 %0, vcc, %0, %1") BODYC("v_sub_co_u32 %0, vcc, %0, %1") BODYC("v_sub_co_u32 %0, vcc, %0, %1") }
  uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  out[blockIdx.x*blockDim.x+threadIdx.x] = a0^a1^a2^a3^a4^a5^a6^a7;
  if (threadIdx.x==0 && blockIdx.x==0) { clk[0] = t1 - t0; clk[1] = r1 - r0; } }

// more 32-bit forms: compares into VCC or an SGPR pair, selects, logic,
// funnel shift, three-operand adds
#define CHK(INS, A) asm volatile(INS : "+v"(A) : "v"(b) : "vcc", "s40", "s41");
#define BODYK(INS) CHK(INS,a0) CHK(INS,a1) CHK(INS,a2) CHK(INS,a3) CHK(INS,a4) CHK(INS,a5) CHK(INS,a6) CHK(INS,a7)
#define K32C(NAME, INS) \
__global__ void NAME(uint32_t *out, int iters, uint64_t *clk) { \
  uint32_t a0=threadIdx.x,a1=a0+1,a2=a0+2,a3=a0+3,a4=a0+4,a5=a0+5,a6=a0+6,a7=a0+7, b=blockIdx.x|1; \
  uint64_t t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime(); \
  for (int i = 0; i < iters; ++i) { BODYK(INS) BODYK(INS) BODYK(INS) BODYK(INS) } \
  uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime(); \
  out[blockIdx.x*blockDim.x+threadIdx.x] = a0^a1^a2^a3^a4^a5^a6^a7; \
  if (threadIdx.x==0 && blockIdx.x==0) { clk[0] = t1 - t0; clk[1] = r1 - r0; } }
K32C(k_cmp32vcc, "v_cmp_le_u32 vcc, %0, %1")
K32C(k_cmp32s, "v_cmp_le_u32_e64 s[40:41], %0, %1")
K32C(k_cndmask, "v_cndmask_b32 %0, %0, %1, vcc")
K32C(k_and, "v_and_b32 %0, %0, %1")
K32C(k_sub, "v_sub_u32 %0, %0, %1")
K32C(k_xad, "v_xad_u32 %0, %0, %1, %0")
K32C(k_add3, "v_add3_u32 %0, %0, %1, %0")
K32C(k_alignbit, "v_alignbit_b32 %0, %0, %1, 5")
K32C(k_lshr32, "v_lshrrev_b32 %0, 3, %0")
K32C(k_addco, "v_add_co_u32 %0, vcc, %0, %1")
K32C(k_addc, "v_addc_co_u32 %0, vcc, %0, %1, vcc")
K32C(k_min, "v_min_u32 %0, %0, %1")
K32C(k_lshlor, "v_lshl_or_b32 %0, %0, 1, %1")
K32C(k_bfe, "v_bfe_u32 %0, %0, %1, 7")

typedef void (*kfn)(uint32_t*, int, uint64_t*);
int main() {
  hipDeviceProp_t prop; (void)hipGetDeviceProperties(&prop, 0);
  const int cus = prop.multiProcessorCount;
  printf("device %s CUs %d clockRate %d kHz\n", prop.gcnArchName, cus, prop.clockRate);
  const int block = 256;
  uint32_t *out; uint64_t *clk; (void)hipMalloc(&out, (size_t)cus*16*block*4); (void)hipMalloc(&clk, 16);
  struct { const char *n; kfn f; int per; } ks[] = {
    {"v_add_u32",k_add,32},{"v_xor_b32",k_xor,32},{"v_mov_b32",k_mov,32},{"v_mul_lo_u32",k_mullo,32},
    {"v_mul_hi_u32",k_mulhi,32},{"v_ffbl_b32",k_ffbl,32},{"add+mul_lo",k_addmul,64},
    {"v_lshrrev_b64",k_lshr64,32},{"v_mad_u64_u32",k_mad64,32},{"v_lshl_add_u64",k_lshladd64,32},
    {"v_cmp_le_u64",k_cmp64,32},{"v_sub_co_u32",k_subco,32},
    {"v_cmp_le_u32 vcc",k_cmp32vcc,32},{"v_cmp_le_u32 sgpr",k_cmp32s,32},{"v_cndmask_b32",k_cndmask,32},
    {"v_and_b32",k_and,32},{"v_sub_u32",k_sub,32},{"v_xad_u32",k_xad,32},{"v_add3_u32",k_add3,32},
    {"v_alignbit_b32",k_alignbit,32},{"v_lshrrev_b32",k_lshr32,32},{"v_add_co_u32",k_addco,32},
    {"v_addc_co_u32",k_addc,32},{"v_min_u32",k_min,32},{"v_lshl_or_b32",k_lshlor,32},{"v_bfe_u32",k_bfe,32}};
  hipEvent_t e0, e1; (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  for (int wpsimd : {4, 8}) {
    const int grid = cus * wpsimd;  // 4 waves per workgroup: wpsimd waves per SIMD
    printf("-- %d waves per SIMD\n", wpsimd);
    for (auto &k : ks) {
      float ms[2]; uint64_t c[2];
      int its[2] = {4000, 12000};
      hipLaunchKernelGGL(k.f, dim3(grid), dim3(block), 0, 0, out, 10, clk);
      (void)hipDeviceSynchronize();
      for (int j = 0; j < 2; ++j) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(k.f, dim3(grid), dim3(block), 0, 0, out, its[j], clk);
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms[j], e0, e1);
        (void)hipMemcpy(c, clk, 16, hipMemcpyDeviceToHost);
      }
      const double dinstr = (double)grid * (block / 64) * (its[1] - its[0]) * k.per;  // wave-instructions
      const double dt = (ms[1] - ms[0]) * 1e-3;
      const double f = c[1] ? (double)c[0] / c[1] * 100e6 : 0;  // s_memtime vs 100 MHz s_memrealtime
      printf("%-16s %7.2f ms  %.3e lane-ops/s  %6.2f lane-ops/CU/clk@2.4GHz  memtime/realtime*100MHz=%.3f GHz\n",
             k.n, ms[1], dinstr * 64 / dt, dinstr * 64 / dt / cus / 2.4e9, f / 1e9);
    }
  }
  return 0;
}
