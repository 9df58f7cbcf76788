// mfma_round: the rounding of v_mfma_f32_16x16x4f32 and v_mfma_f32_16x16x32_bf16 on small known cases.
// One wave; every row of A and every column of B are equal, so every output entry is
// C + sum_k a_k b_k.  Lane l holds K index l / 16 (FP32) or 8 (l / 16) + e, e = 0..7 (BF16).  Prints one
// line per case: "<case> <f32 result> <bf16 result>" as hex floats, after checking that all 256
// outputs of each instruction agree (else "<case> DIFFER").  Test infrastructure: run by
// tests/test_gpu_mfma_rounding.py, which holds the expected values and the model they follow.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

struct Case {
  float a[32], b[32], c;
};
constexpr int kCases = 12;

__global__ void probe(const Case* cs, float* out) {
  const int lane = threadIdx.x;  // one wave of 64
  for (int t = 0; t < kCases; t++) {
    const Case& c = cs[t];
    const f4 acc = {c.c, c.c, c.c, c.c};
    const f4 d = __builtin_amdgcn_mfma_f32_16x16x4f32(c.a[lane / 16], c.b[lane / 16], acc, 0, 0, 0);
    bf8 A, B;
    for (int e = 0; e < 8; e++) {
      A[e] = (__bf16)c.a[8 * (lane / 16) + e];
      B[e] = (__bf16)c.b[8 * (lane / 16) + e];
    }
    const f4 d2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
    for (int r = 0; r < 4; r++) {
      out[(2 * t) * 256 + lane * 4 + r] = d[r];
      out[(2 * t + 1) * 256 + lane * 4 + r] = d2[r];
    }
  }
}

int main() {
  static Case h[kCases];
  std::memset(h, 0, sizeof h);
  const float u = 0x1p-23f;  // ulp of 1
  // 0: 1 + 0.75 ulp (both products at K 0, 1)
  h[0].a[0] = 1; h[0].b[0] = 1; h[0].a[1] = 0.75f * u; h[0].b[1] = 1;
  // 1: C = 1 plus a product of 0.75 ulp
  h[1].c = 1; h[1].a[0] = 0.75f * u; h[1].b[0] = 1;
  // 2: 1 + 3 x 0.375 ulp (K 0..3)
  h[2].a[0] = 1; h[2].b[0] = 1;
  for (int k = 1; k < 4; k++) { h[2].a[k] = 0.375f * u; h[2].b[k] = 1; }
  // 3: (1 + u)^2 - (1 + 2u): exact 2^-46 (FP32 operands; BF16 rounds the operands)
  h[3].a[0] = 1 + u; h[3].b[0] = 1 + u; h[3].a[1] = -(1 + 2 * u); h[3].b[1] = 1;
  // 4: 2^10 - 2^10 + 2^-20 (K 0..2)
  h[4].a[0] = 1024; h[4].b[0] = 1; h[4].a[1] = -1024; h[4].b[1] = 1; h[4].a[2] = 0x1p-20f; h[4].b[2] = 1;
  // 5: 1 - 0.75 ulp
  h[5].a[0] = 1; h[5].b[0] = 1; h[5].a[1] = -0.75f * u; h[5].b[1] = 1;
  // 6: C = -1, then 1 + 16 x 2^-25 (K 0 and 1..16)
  h[6].c = -1; h[6].a[0] = 1; h[6].b[0] = 1;
  for (int k = 1; k < 17; k++) { h[6].a[k] = 0x1p-25f; h[6].b[k] = 1; }
  // 7: C = 0.75 x 2^-24, products 1 (K 0) and -1 (K 8)
  h[7].c = 0.75f * 0x1p-24f; h[7].a[0] = 1; h[7].b[0] = 1; h[7].a[8] = -1; h[7].b[8] = 1;
  // 8: C = -0.5, 1 x 0.5 + 1.5 x 2^-24 (K 0, 1)
  h[8].c = -0.5f; h[8].a[0] = 1; h[8].b[0] = 0.5f; h[8].a[1] = 1.5f * 0x1p-24f; h[8].b[1] = 1;
  // 9: C = -2.25, 1.5 x 1.5 + 2^-24 (K 0, 1)
  h[9].c = -2.25f; h[9].a[0] = 1.5f; h[9].b[0] = 1.5f; h[9].a[1] = 0x1p-24f; h[9].b[1] = 1;
  // 10: C = -1, 1 (K 0) + 1.5 x 2^-24 (K 4): one group of 8
  h[10].c = -1; h[10].a[0] = 1; h[10].b[0] = 1; h[10].a[4] = 1.5f * 0x1p-24f; h[10].b[4] = 1;
  // 11: C = -1, 1 (K 0) + 1.5 x 2^-24 (K 8): two groups
  h[11].c = -1; h[11].a[0] = 1; h[11].b[0] = 1; h[11].a[8] = 1.5f * 0x1p-24f; h[11].b[8] = 1;

  Case* dc = nullptr;
  float* dout = nullptr;
  if (hipMalloc(&dc, sizeof h) != hipSuccess || hipMalloc(&dout, 2 * kCases * 256 * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 2;
  }
  if (hipMemcpy(dc, h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) return 2;
  hipLaunchKernelGGL(probe, dim3(1), dim3(64), 0, 0, dc, dout);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 2;
  static float o[2 * kCases * 256];
  if (hipMemcpy(o, dout, sizeof o, hipMemcpyDeviceToHost) != hipSuccess) return 2;
  for (int t = 0; t < kCases; t++) {
    bool same = true;
    for (int m = 0; m < 2; m++)
      for (int i = 1; i < 256; i++) same &= std::memcmp(&o[(2 * t + m) * 256 + i], &o[(2 * t + m) * 256], 4) == 0;
    if (same)
      std::printf("%d %a %a\n", t, (double)o[(2 * t) * 256], (double)o[(2 * t + 1) * 256]);
    else
      std::printf("%d DIFFER\n", t);
  }
  (void)hipFree(dc);
  (void)hipFree(dout);
  return 0;
}
