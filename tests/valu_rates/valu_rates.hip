// Issue cost of single VALU instruction kinds on gfx950 (MI355X), in cycles
// per wave64 instruction per SIMD.
//
// Each kernel runs, in every lane, ITER iterations of NCH independent chains
// of one instruction on registers (no memory access inside the loop), then
// writes one word per lane.  The grid has W one-wave workgroups per SIMD
// (W = 4 and 8), so the SIMDs stay busy; cycles = kernel time x 2.4 GHz
// (the engine clock, which GRBM_GUI_ACTIVE confirms) / instructions per SIMD.
// The single-chain rows measure a dependent chain (latency), the rest
// throughput.  Used to read the BC7 encoder's counters: which of its
// instruction kinds issue every ~2.4 cycles and which every ~4.2.
//
//     tests/valu_rates/_build/valu_rates
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define ITER 4096

#define DEFK(NAME, NCH, BODY)                                      \
  __global__ void NAME(uint32_t* out, uint32_t seed) {             \
    uint32_t a[8], b = seed ^ threadIdx.x, c = seed * 3u + 1u;     \
    for (int i = 0; i < 8; i++) a[i] = seed + i * 77u + threadIdx.x; \
    for (int it = 0; it < ITER; it++) {                            \
      _Pragma("unroll") for (int i = 0; i < NCH; i++) { BODY; }    \
    }                                                              \
    uint32_t r = 0;                                                \
    for (int i = 0; i < 8; i++) r ^= a[i];                         \
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;                \
  }

// a[i] = op(a[i], b) / op(a[i], b, c) on VGPRs; S: the second source in an SGPR.
#define R2(op) asm volatile(op " %0, %0, %1" : "+v"(a[i]) : "v"(b))
#define R3(op) asm volatile(op " %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "v"(c))
#define R3D(op) asm volatile(op " %0, %1, %2, %0" : "+v"(a[i]) : "v"(b), "v"(c))
#define R1(op) asm volatile(op " %0, %0" : "+v"(a[i]))

DEFK(v_add_f32, 8, R2("v_add_f32"))
DEFK(v_mul_f32, 8, R2("v_mul_f32"))
DEFK(v_fma_f32, 8, R3("v_fma_f32"))
DEFK(v_add_u32, 8, R2("v_add_u32"))
DEFK(v_sub_u32, 8, R2("v_sub_u32"))
DEFK(v_and_b32, 8, R2("v_and_b32"))
DEFK(v_xor_b32, 8, R2("v_xor_b32"))
DEFK(v_lshrrev_b32, 8, R2("v_lshrrev_b32"))
DEFK(v_lshlrev_b32, 8, R2("v_lshlrev_b32"))
DEFK(v_bitop3_b32, 8, asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(a[i]) : "v"(b), "v"(c)))
DEFK(v_mov_b32, 8, asm volatile("v_mov_b32 %0, %1" : "=v"(a[i]) : "v"(a[(i + 1) & 7])))
DEFK(v_add_f32_sgpr, 8, asm volatile("v_add_f32 %0, %1, %0" : "+v"(a[i]) : "s"(c)))
DEFK(v_min_i32, 8, R2("v_min_i32"))
DEFK(v_min_f32, 8, R2("v_min_f32"))
DEFK(v_cmp_lt_i32, 8, asm volatile("v_cmp_lt_i32 vcc, %0, %1" : : "v"(a[i]), "v"(b) : "vcc"))
DEFK(v_cmp_cndmask, 8, asm volatile("v_cmp_lt_i32 vcc, %0, %1\n\tv_cndmask_b32 %0, %0, %2, vcc" : "+v"(a[i]) : "v"(b), "v"(c) : "vcc"))
DEFK(v_cvt_f32_i32, 8, R1("v_cvt_f32_i32"))
DEFK(v_cvt_i32_f32, 8, R1("v_cvt_i32_f32"))
DEFK(v_cvt_f32_ubyte0, 8, R1("v_cvt_f32_ubyte0"))
DEFK(v_bfe_u32, 8, R3("v_bfe_u32"))
DEFK(v_perm_b32, 8, R3("v_perm_b32"))
DEFK(v_lshl_or_b32, 8, R3("v_lshl_or_b32"))
DEFK(v_or3_b32, 8, R3("v_or3_b32"))
DEFK(v_med3_f32, 8, R3("v_med3_f32"))
DEFK(v_mul_i32_i24, 8, R2("v_mul_i32_i24"))
DEFK(v_mad_i32_i24, 8, R3("v_mad_i32_i24"))
DEFK(v_mul_lo_u32, 8, R2("v_mul_lo_u32"))
DEFK(v_pk_mad_u16, 8, R3("v_pk_mad_u16"))
DEFK(v_pk_ashrrev_i16, 8, asm volatile("v_pk_ashrrev_i16 %0, 6, %0" : "+v"(a[i])))
DEFK(v_dot2_i32_i16, 8, R3D("v_dot2_i32_i16"))
DEFK(v_dot4_u32_u8, 8, R3D("v_dot4_u32_u8"))
DEFK(v_rcp_f32, 8, R1("v_rcp_f32"))
DEFK(v_sqrt_f32, 8, R1("v_sqrt_f32"))
DEFK(v_mov_b32_dpp, 8, asm volatile("v_mov_b32_dpp %0, %1 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "=v"(a[i]) : "v"(a[(i + 1) & 7])))
// Operand kinds: a non-inline literal constant, an inline constant, an SGPR.
DEFK(v_mul_f32_lit, 8, asm volatile("v_mul_f32 %0, 0x37800000, %0" : "+v"(a[i])))
DEFK(v_mul_f32_inl, 8, asm volatile("v_mul_f32 %0, 0.5, %0" : "+v"(a[i])))
DEFK(v_mul_f32_e64, 8, asm volatile("v_mul_f32_e64 %0, %0, %1" : "+v"(a[i]) : "v"(b)))
DEFK(v_and_b32_lit, 8, asm volatile("v_and_b32 %0, 0xff00ff, %0" : "+v"(a[i])))
DEFK(v_and_b32_inl, 8, asm volatile("v_and_b32 %0, 15, %0" : "+v"(a[i])))
DEFK(v_and_b32_sgpr, 8, asm volatile("v_and_b32 %0, %1, %0" : "+v"(a[i]) : "s"(c)))
DEFK(v_fmac_f32, 8, asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(a[i]) : "v"(b), "v"(c)))
DEFK(v_fmac_f32_lit, 8, asm volatile("v_fmac_f32 %0, 0x37800000, %1" : "+v"(a[i]) : "v"(b)))
DEFK(v_fmamk_f32, 8, asm volatile("v_fmamk_f32 %0, %0, 0x37800000, %1" : "+v"(a[i]) : "v"(b)))
DEFK(v_fma_f32_sgpr, 8, asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[i]) : "s"(c), "v"(b)))
DEFK(v_mov_b32_lit, 8, asm volatile("v_mov_b32 %0, 0x12345" : "=v"(a[i])))
DEFK(v_or_b32, 8, R2("v_or_b32"))
DEFK(v_bitop3_b32_sgpr, 8, asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(a[i]) : "s"(c), "v"(b)))
DEFK(v_perm_b32_sgpr, 8, asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "s"(c)))
DEFK(v_sub_f32, 8, R2("v_sub_f32"))
DEFK(v_max_f32, 8, R2("v_max_f32"))
DEFK(v_lshlrev_b32_inl, 8, asm volatile("v_lshlrev_b32 %0, 5, %0" : "+v"(a[i])))
DEFK(v_ashrrev_i32, 8, R2("v_ashrrev_i32"))
DEFK(v_lshl_add_u32, 8, R3("v_lshl_add_u32"))
DEFK(v_add3_u32, 8, R3("v_add3_u32"))
DEFK(v_alignbit_b32, 8, R3("v_alignbit_b32"))
DEFK(v_max3_f32, 8, R3("v_max3_f32"))
DEFK(v_rndne_f32, 8, R1("v_rndne_f32"))
DEFK(v_mul_u32_u24, 8, R2("v_mul_u32_u24"))
DEFK(v_cvt_pkrtz_f16_f32, 8, R2("v_cvt_pkrtz_f16_f32"))
DEFK(v_pk_add_u16, 8, R2("v_pk_add_u16"))
DEFK(v_pk_min_u16, 8, R2("v_pk_min_u16"))
DEFK(v_add_f16, 8, R2("v_add_f16"))
DEFK(v_dot2_f32_f16, 8, R3D("v_dot2_f32_f16"))
DEFK(v_dot2c_f32_f16, 8, asm volatile("v_dot2c_f32_f16 %0, %1, %2" : "+v"(a[i]) : "v"(b), "v"(c)))
// A select whose lane mask was written by a v_cmp in the same iteration
// (one compare for eight selects), and one whose mask is an SGPR pair that no
// VALU instruction writes in the loop.
DEFK(v_cndmask_after_cmp, 8,
     if (i == 0) asm volatile("v_cmp_lt_i32 vcc, %0, %1" : : "v"(b), "v"(c) : "vcc");
     asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(a[i]) : "v"(b) : "vcc"))
DEFK(v_cndmask_e64_sreg, 8,
     asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "s"((uint64_t)c * 0x9E3779B97F4A7C15ull)))
DEFK(v_cndmask_pair_cmp, 8,
     if ((i & 1) == 0) asm volatile("v_cmp_lt_i32 vcc, %0, %1" : : "v"(a[i]), "v"(c) : "vcc");
     asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(a[i]) : "v"(b) : "vcc"))
DEFK(v_cndmask_e64_vcc, 8,
     if (i == 0) asm volatile("v_cmp_lt_i32 vcc, %0, %1" : : "v"(b), "v"(c) : "vcc");
     asm volatile("v_cndmask_b32_e64 %0, %0, %1, vcc" : "+v"(a[i]) : "v"(b) : "vcc"))
DEFK(v_cndmask_e64_cmp_sreg, 8,
     uint64_t msk;
     if (i == 0) asm volatile("v_cmp_lt_i32_e64 %0, %1, %2" : "=s"(msk) : "v"(b), "v"(c));
     asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "s"(msk)))
DEFK(v_cndmask_add_mix, 8,
     if (i == 0) asm volatile("v_cmp_lt_i32 vcc, %0, %1" : : "v"(b), "v"(c) : "vcc");
     if (i & 1) asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(a[i]) : "v"(b) : "vcc");
     else asm volatile("v_add_f32 %0, %0, %1" : "+v"(a[i]) : "v"(b)))
DEFK(chain_v_add_f32, 1, R2("v_add_f32"))
DEFK(chain_v_pk_mad_u16, 1, R3("v_pk_mad_u16"))

typedef void (*kfn)(uint32_t*, uint32_t);
struct K {
  const char* name;
  kfn f;
  int nch;
};

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
      return 1;                                                              \
    }                                                                        \
  } while (0)

int main() {
#define E(n, c) {#n, n, c},
  K ks[] = {E(v_add_f32, 8) E(v_mul_f32, 8) E(v_fma_f32, 8) E(v_add_u32, 8) E(v_sub_u32, 8)
            E(v_and_b32, 8) E(v_xor_b32, 8) E(v_lshrrev_b32, 8) E(v_lshlrev_b32, 8)
            E(v_bitop3_b32, 8) E(v_mov_b32, 8) E(v_add_f32_sgpr, 8) E(v_min_i32, 8)
            E(v_min_f32, 8) E(v_cmp_lt_i32, 8) E(v_cmp_cndmask, 8) E(v_cvt_f32_i32, 8)
            E(v_cvt_i32_f32, 8) E(v_cvt_f32_ubyte0, 8) E(v_bfe_u32, 8) E(v_perm_b32, 8)
            E(v_lshl_or_b32, 8) E(v_or3_b32, 8) E(v_med3_f32, 8) E(v_mul_i32_i24, 8)
            E(v_mad_i32_i24, 8) E(v_mul_lo_u32, 8) E(v_pk_mad_u16, 8) E(v_pk_ashrrev_i16, 8)
            E(v_dot2_i32_i16, 8) E(v_dot4_u32_u8, 8) E(v_rcp_f32, 8) E(v_sqrt_f32, 8)
            E(v_mov_b32_dpp, 8) E(v_mul_f32_lit, 8) E(v_mul_f32_inl, 8) E(v_mul_f32_e64, 8)
            E(v_and_b32_lit, 8) E(v_and_b32_inl, 8) E(v_and_b32_sgpr, 8) E(v_fmac_f32, 8)
            E(v_fmac_f32_lit, 8) E(v_fmamk_f32, 8) E(v_fma_f32_sgpr, 8) E(v_mov_b32_lit, 8)
            E(v_or_b32, 8) E(v_bitop3_b32_sgpr, 8) E(v_perm_b32_sgpr, 8)
            E(v_sub_f32, 8) E(v_max_f32, 8) E(v_lshlrev_b32_inl, 8) E(v_ashrrev_i32, 8)
            E(v_lshl_add_u32, 8) E(v_add3_u32, 8) E(v_alignbit_b32, 8) E(v_max3_f32, 8)
            E(v_rndne_f32, 8) E(v_mul_u32_u24, 8) E(v_cvt_pkrtz_f16_f32, 8) E(v_pk_add_u16, 8)
            E(v_pk_min_u16, 8) E(v_add_f16, 8) E(v_dot2_f32_f16, 8) E(v_dot2c_f32_f16, 8) E(v_cndmask_after_cmp, 8) E(v_cndmask_e64_sreg, 8) E(v_cndmask_pair_cmp, 8) E(v_cndmask_e64_vcc, 8) E(v_cndmask_e64_cmp_sreg, 8) E(v_cndmask_add_mix, 8)
            E(chain_v_add_f32, 1) E(chain_v_pk_mad_u16, 1)};
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const double clk = 2.4e9;
  const int waves[2] = {4, 8};
  uint32_t* out;
  CHECK(hipMalloc(&out, (size_t)cus * 4 * 8 * 64 * sizeof(uint32_t)));  // largest grid
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  printf("%d CUs; cycles per wave64 instruction per SIMD at 4 and 8 waves per SIMD\n", cus);
  for (const K& k : ks) {
    printf("%-22s", k.name);
    for (int w : waves) {
      const int blocks = cus * 4 * w;  // one 64-lane wave per workgroup
      hipLaunchKernelGGL(k.f, dim3(blocks), dim3(64), 0, 0, out, 1u);
      CHECK(hipEventRecord(e0));
      for (int r = 0; r < 5; r++) hipLaunchKernelGGL(k.f, dim3(blocks), dim3(64), 0, 0, out, 2u + r);
      CHECK(hipEventRecord(e1));
      CHECK(hipEventSynchronize(e1));
      float ms;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      const double per_simd = (double)w * ITER * k.nch;
      printf(" %7.2f", ms / 5 * 1e-3 * clk / per_simd);
    }
    printf("\n");
  }
  CHECK(hipFree(out));
  return 0;
}
