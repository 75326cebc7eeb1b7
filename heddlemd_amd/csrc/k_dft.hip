// SPME reciprocal DFT stages as split-precision FP16 MFMA contractions
// (gfx950), beside the FP32 radix FFT of k_spme.hip.
//
// A 1D DFT of length n over many lines is the product Y = W X with the
// n x n DFT matrix W.  Each operand is split into a high and a low FP16
// part (x = hi + lo, lo = fp16(x - hi)), and W X is accumulated in FP32
// as W_hi X_hi + W_hi X_lo + W_lo X_hi on the matrix cores
// (v_mfma_f32_32x32x8_f16), which keeps ~22 significant bits.  Operands
// are scaled by powers of two so that both parts stay in FP16's normal
// range.
#include "hd_device.h"
#include "hd_kernels.h"

typedef _Float16 hd_half4 __attribute__((ext_vector_type(4)));
typedef float hd_float16v __attribute__((ext_vector_type(16)));

// Lane layout of v_mfma_f32_32x32x8_f16 (one block):
//   A (32 x 8): lane l holds A[l % 32][4 (l / 32) + e], e = 0..3
//   B (8 x 32): lane l holds B[4 (l / 32) + e][l % 32]
//   D (32 x 32): lane l, element v holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32]
__device__ __forceinline__ unsigned mfma_d_row(unsigned lane, unsigned v) {
  return 8u * (v >> 2) + 4u * (lane >> 5) + (v & 3u);
}

// Self test of that layout: A[i][k] = (8 i + k) mod 61 - 30,
// B[k][j] = (32 k + j) mod 23 - 11, out[l * 16 + v] = D element v of lane l.
__global__ void k_selftest_mfma(float* out) {
  unsigned l = threadIdx.x;
  hd_half4 a, b;
  for (int e = 0; e < 4; ++e) {
    unsigned i = l % 32u, k = 4u * (l / 32u) + (unsigned)e;
    a[e] = (_Float16)(float)((int)((8u * i + k) % 61u) - 30);
    unsigned j = l % 32u;
    b[e] = (_Float16)(float)((int)((32u * k + j) % 23u) - 11);
  }
  hd_float16v c;
  for (int v = 0; v < 16; ++v) c[v] = 0.f;
  c = __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, c, 0, 0, 0);
  for (int v = 0; v < 16; ++v) out[l * 16u + (unsigned)v] = c[v];
}

void hd_launch_selftest_mfma(hipStream_t st, float* out) {
  hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, st, out);
}
