// SPME reciprocal DFT stages as split-precision FP16 MFMA contractions
// (gfx950), beside the FP32 radix FFT of k_spme.hip.
//
// A 1D DFT of length n over many lines is the product Y = W X with the
// n x n DFT matrix W.  Each operand is split into a high and a low FP16
// part (x = hi + lo, lo = fp16(x - hi)), and W X is accumulated in FP32
// as W_hi X_hi + W_hi X_lo + W_lo X_hi on the matrix cores
// (v_mfma_f32_32x32x16_f16, gfx950), which keeps ~22 significant bits.
// Operands are scaled by powers of two so that both parts stay in FP16's
// normal range.
#include "hd_device.h"
#include "hd_kernels.h"

typedef _Float16 hd_half4 __attribute__((ext_vector_type(4)));
typedef _Float16 hd_half8 __attribute__((ext_vector_type(8)));
typedef __bf16 hd_bf8 __attribute__((ext_vector_type(8)));
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

// The doubled-K gfx950 forms v_mfma_f32_32x32x16_{f16,bf16}:
//   A (32 x 16): lane l holds A[l % 32][8 (l / 32) + e], e = 0..7
//   B (16 x 32): lane l holds B[8 (l / 32) + e][l % 32]
//   D: the 32x32x8 layout above.
// Self test with A[i][k] = (16 i + k) mod 61 - 30, B[k][j] = (32 k + j)
// mod 23 - 11 (exact in both types).
template <bool BF16>
__global__ void k_selftest_mfma_k16(float* out) {
  unsigned l = threadIdx.x;
  hd_half8 a, b;
  hd_bf8 ab, bb;
  for (int e = 0; e < 8; ++e) {
    unsigned i = l % 32u, k = 8u * (l / 32u) + (unsigned)e, j = l % 32u;
    float av = (float)((int)((16u * i + k) % 61u) - 30);
    float bv = (float)((int)((32u * k + j) % 23u) - 11);
    a[e] = (_Float16)av; b[e] = (_Float16)bv;
    ab[e] = (__bf16)av; bb[e] = (__bf16)bv;
  }
  hd_float16v c;
  for (int v = 0; v < 16; ++v) c[v] = 0.f;
  if (BF16) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ab, bb, c, 0, 0, 0);
  else c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  for (int v = 0; v < 16; ++v) out[l * 16u + (unsigned)v] = c[v];
}

void hd_launch_selftest_mfma_k16(hipStream_t st, float* out, bool bf16) {
  if (bf16) hipLaunchKernelGGL(k_selftest_mfma_k16<true>, dim3(1), dim3(64), 0, st, out);
  else hipLaunchKernelGGL(k_selftest_mfma_k16<false>, dim3(1), dim3(64), 0, st, out);
}

// ---------------------------------------------------------------------
// DFT passes as GEMMs.  A workgroup takes DFT_NL lines (the GEMM's N),
// stages them in LDS as FP16 high/low parts of the line scaled by 2^-s
// (s chosen from the tile's largest magnitude), and its four waves each
// produce 32-row tiles of Y = W X (M rows; K = input length):
//   main = W_hi X_hi,  corr = W_hi X_lo' + W_lo' X_hi,
//   Y = (main + corr 2^-12) 2^s,
// where the low parts are stored scaled by 2^12 (X_lo' = 2^12 (x - x_hi),
// W_lo' = 2^12 (W - W_hi)) so that they stay in FP16's normal range.
//
// Passes over the (na, nb, nch = nc/2 + 1) complex grid, forward sign -1,
// inverse +1, unnormalised (the radix FFT's convention):
//   C_FWD: rows of the i64 spread grid (real, length nc) -> 2 nch outputs
//          (real parts, then imaginary parts);
//   CPLX:  complex lines along b or a, [re; im] stacked (K = M = 2n) with
//          the block matrix [[Cr, -Ci], [Ci, Cr]];
//   A_CONV: forward a-axis, influence multiply with the energy / virial
//          partials, inverse a-axis, in one workgroup pass;
//   C_INV: Hermitian half spectrum (K = 2 nch) -> nc real values.
// ---------------------------------------------------------------------
#define DFT_NL 32
#define DFT_THREADS 256
#define DFT_KMAX 256
#define DFT_KP (DFT_KMAX + 8)  // FP16 line pitch (halves): 16-byte rows, 132 dwords
#define DFT_XP (DFT_KMAX + 1)  // f32 line pitch: odd, so column walks hit distinct banks
#define DFT_LO_SCALE 4096.0f
#define DFT_LO_UNSCALE (1.0f / 4096.0f)

enum { DFT_C_FWD = 0, DFT_CPLX_B = 1, DFT_CPLX_A = 2, DFT_A_CONV = 3, DFT_C_INV = 4 };

// A DFT matrix W (M x K, M a multiple of 32, K of 16) in fragment order:
// the A operand of one 32x32x16 step (row tile mt, K step ks) is 64 lanes x
// 8 halves stored contiguously, element e of lane l at
// ((mt (K/16) + ks) 64 + l) 8 + e = W[32 mt + l % 32][16 ks + 8 (l / 32) + e],
// so a wave's fragment load is one contiguous 1 KB read.
struct DftMat {
  const _Float16 *hi, *lo;
  unsigned M, K;
};

// Global index (complex units) of element e of line `line` for the passes
// along b (lines = (a, c')) and along a (lines = (b, c')).
template <bool ALONG_B>
__device__ __forceinline__ size_t dft_cidx(const HdDftArgs& a, unsigned line, unsigned e) {
  unsigned o = line / a.nch, cc = line - o * a.nch;
  return ALONG_B ? ((size_t)o * a.nb + e) * a.nch + cc : ((size_t)e * a.nb + o) * a.nch + cc;
}

// Stage the workgroup's lines into s_x[line][k] (f32, zero beyond the
// valid range), with coalesced global reads: rows for the c-axis passes,
// consecutive c' for the b / a passes (real and imaginary parts of one
// complex element land in rows k and n + k).
// The global loads of DFT_SB elements per thread are issued before any of
// them is stored, so their latencies overlap.
#define DFT_SB 8
template <int MODE>
__device__ __forceinline__ void dft_stage(const HdDftArgs& a, unsigned line0, unsigned K,
                                          float* s_x) {
  unsigned t = threadIdx.x;
  if (MODE == DFT_C_FWD) {
    unsigned nrow = a.na * a.nb, total = DFT_NL * K;
    for (unsigned e0 = t; e0 < total; e0 += DFT_SB * DFT_THREADS) {
      float v[DFT_SB];
#pragma unroll
      for (int u = 0; u < DFT_SB; ++u) {
        unsigned e = e0 + u * DFT_THREADS;
        unsigned ln = e / K, k = e - ln * K;
        v[u] = 0.f;
        if (e < total && line0 + ln < nrow && k < a.nc)
          v[u] = (float)a.rho[(size_t)(line0 + ln) * a.nc + k] * (1.0f / 4294967296.0f);
      }
#pragma unroll
      for (int u = 0; u < DFT_SB; ++u) {
        unsigned e = e0 + u * DFT_THREADS;
        unsigned ln = e / K, k = e - ln * K;
        if (e < total) s_x[ln * DFT_XP + k] = v[u];
      }
    }
  } else if (MODE == DFT_C_INV) {
    unsigned nrow = a.na * a.nb, total = DFT_NL * a.nch;
    for (unsigned e0 = t; e0 < total; e0 += DFT_SB * DFT_THREADS) {
      float2 v[DFT_SB];
#pragma unroll
      for (int u = 0; u < DFT_SB; ++u) {
        unsigned e = e0 + u * DFT_THREADS;
        unsigned ln = e / a.nch, k = e - ln * a.nch;
        v[u] = make_float2(0.f, 0.f);
        if (e < total && line0 + ln < nrow) v[u] = a.cgrid[(size_t)(line0 + ln) * a.nch + k];
      }
#pragma unroll
      for (int u = 0; u < DFT_SB; ++u) {
        unsigned e = e0 + u * DFT_THREADS;
        unsigned ln = e / a.nch, k = e - ln * a.nch;
        if (e < total) {
          s_x[ln * DFT_XP + k] = v[u].x;
          s_x[ln * DFT_XP + a.nch + k] = v[u].y;
        }
      }
    }
    unsigned pad = K - 2 * a.nch;
    for (unsigned e = t; e < DFT_NL * pad; e += DFT_THREADS) {
      unsigned ln = e / pad, k = 2 * a.nch + (e - ln * pad);
      s_x[ln * DFT_XP + k] = 0.f;
    }
  } else {
    constexpr bool along_b = (MODE == DFT_CPLX_B);
    unsigned n = along_b ? a.nb : a.na;
    unsigned nl = (along_b ? a.na : a.nb) * a.nch, total = DFT_NL * n;
    for (unsigned e0 = t; e0 < total; e0 += DFT_SB * DFT_THREADS) {
      float2 v[DFT_SB];
#pragma unroll
      for (int u = 0; u < DFT_SB; ++u) {
        unsigned e = e0 + u * DFT_THREADS;
        unsigned ln = e % DFT_NL, kk = e / DFT_NL;
        v[u] = make_float2(0.f, 0.f);
        if (e < total && line0 + ln < nl) v[u] = a.cgrid[dft_cidx<along_b>(a, line0 + ln, kk)];
      }
#pragma unroll
      for (int u = 0; u < DFT_SB; ++u) {
        unsigned e = e0 + u * DFT_THREADS;
        unsigned ln = e % DFT_NL, kk = e / DFT_NL;
        if (e < total) {
          s_x[ln * DFT_XP + kk] = v[u].x;
          s_x[ln * DFT_XP + n + kk] = v[u].y;
        }
      }
    }
  }
}

// stage s_x[line][k] (f32, already in LDS) as FP16 high/low parts with a
// common power-of-two scale; returns the exponent s
__device__ __forceinline__ int dft_split(const float* s_x, _Float16* xh, _Float16* xl,
                                         unsigned K, unsigned* s_max) {
  unsigned t = threadIdx.x;
  if (t == 0) *s_max = 0u;
  __syncthreads();
  unsigned mx = 0u;
  for (unsigned e = t; e < DFT_NL * K; e += DFT_THREADS) {
    unsigned line = e / K, k = e - line * K;
    mx = max(mx, __float_as_uint(fabsf(s_x[line * DFT_XP + k])));
  }
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, (unsigned)__shfl_xor((int)mx, off));
  if ((t & 63u) == 0) atomicMax(s_max, mx);
  __syncthreads();
  float vmax = __uint_as_float(*s_max);
  int ex = 0;
  if (vmax > 0.f) (void)frexpf(vmax, &ex);
  int s = ex - 14;  // |x| 2^-s < 2^14
  for (unsigned e = t; e < DFT_NL * K; e += DFT_THREADS) {
    unsigned line = e / K, k = e - line * K;
    float x = ldexpf(s_x[line * DFT_XP + k], -s);
    _Float16 h = (_Float16)x;
    _Float16 l = (_Float16)((x - (float)h) * DFT_LO_SCALE);
    xh[line * DFT_KP + k] = h;
    xl[line * DFT_KP + k] = l;
  }
  __syncthreads();
  return s;
}

// one 32-row tile of Y = W X for the staged lines; acc in D layout.  The
// operand fragments of DFT_KB consecutive K steps are loaded before their
// MFMAs are issued, so the W reads (L2, coalesced) overlap.
#define DFT_KB 4
__device__ __forceinline__ void dft_tile(const DftMat& w, unsigned mt, const _Float16* xh,
                                         const _Float16* xl, hd_float16v& main,
                                         hd_float16v& corr) {
  unsigned lane = threadIdx.x & 63u;
  unsigned nks = w.K / 16u;
  for (int v = 0; v < 16; ++v) { main[v] = 0.f; corr[v] = 0.f; }
  const hd_half8* wh = reinterpret_cast<const hd_half8*>(w.hi) + (size_t)mt * nks * 64u + lane;
  const hd_half8* wl = reinterpret_cast<const hd_half8*>(w.lo) + (size_t)mt * nks * 64u + lane;
  unsigned xo = (lane & 31u) * DFT_KP + 8u * (lane >> 5);
  const hd_half8* bxh = reinterpret_cast<const hd_half8*>(xh + xo);
  const hd_half8* bxl = reinterpret_cast<const hd_half8*>(xl + xo);
  unsigned ks = 0;
  for (; ks + DFT_KB <= nks; ks += DFT_KB) {
    hd_half8 ah[DFT_KB], al[DFT_KB], bh[DFT_KB], bl[DFT_KB];
#pragma unroll
    for (int u = 0; u < DFT_KB; ++u) {
      ah[u] = wh[(size_t)(ks + u) * 64u];
      al[u] = wl[(size_t)(ks + u) * 64u];
    }
#pragma unroll
    for (int u = 0; u < DFT_KB; ++u) {
      bh[u] = bxh[2u * (ks + u)];
      bl[u] = bxl[2u * (ks + u)];
    }
#pragma unroll
    for (int u = 0; u < DFT_KB; ++u) {
      main = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[u], bh[u], main, 0, 0, 0);
      corr = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[u], bl[u], corr, 0, 0, 0);
      corr = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[u], bh[u], corr, 0, 0, 0);
    }
  }
  for (; ks < nks; ++ks) {
    hd_half8 ah = wh[(size_t)ks * 64u], al = wl[(size_t)ks * 64u];
    hd_half8 bh = bxh[2u * ks], bl = bxl[2u * ks];
    main = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, main, 0, 0, 0);
    corr = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, corr, 0, 0, 0);
    corr = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, corr, 0, 0, 0);
  }
}

// complex passes: wave w owns the real-part tile w and the imaginary-part
// tile w + n/32 of the same frequencies, so each lane holds whole complex
// outputs; with TO_LDS they go to s_x (rows k and n + k), else to the grid
template <bool ALONG_B, bool TO_LDS>
__device__ __forceinline__ void dft_cplx_gemm(const HdDftArgs& a, const DftMat& w,
                                              unsigned line0, unsigned n, const _Float16* xh,
                                              const _Float16* xl, float sc, float* s_x) {
  unsigned wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  unsigned half = n / 32u;
  unsigned nl = (ALONG_B ? a.na : a.nb) * a.nch;
  for (unsigned mt = wave; mt < half; mt += 4u) {
    hd_float16v mr, cr, mi, ci;
    dft_tile(w, mt, xh, xl, mr, cr);
    dft_tile(w, mt + half, xh, xl, mi, ci);
    unsigned ln = lane & 31u;
    for (unsigned v = 0; v < 16; ++v) {
      unsigned m = mt * 32u + mfma_d_row(lane, v);
      float yr = (mr[v] + cr[v] * DFT_LO_UNSCALE) * sc;
      float yi = (mi[v] + ci[v] * DFT_LO_UNSCALE) * sc;
      if (TO_LDS) {
        s_x[ln * DFT_XP + m] = yr;
        s_x[ln * DFT_XP + n + m] = yi;
      } else if (line0 + ln < nl) {
        a.cgrid[dft_cidx<ALONG_B>(a, line0 + ln, m)] = make_float2(yr, yi);
      }
    }
  }
}

template <int MODE>
__global__ void __launch_bounds__(DFT_THREADS) k_dft(HdDftArgs a, DftMat w, DftMat w2) {
  __shared__ float s_x[DFT_NL * DFT_XP];
  __shared__ __attribute__((aligned(16))) _Float16 xh[DFT_NL * DFT_KP];
  __shared__ __attribute__((aligned(16))) _Float16 xl[DFT_NL * DFT_KP];
  __shared__ unsigned s_max;
  __shared__ double red_e[DFT_THREADS], red_w[DFT_THREADS];
  unsigned t = threadIdx.x, wave = t >> 6, lane = t & 63u;
  unsigned line0 = blockIdx.x * DFT_NL;
  dft_stage<MODE>(a, line0, w.K, s_x);
  __syncthreads();
  int s = dft_split(s_x, xh, xl, w.K, &s_max);
  float sc = ldexpf(1.0f, s);
  if (MODE == DFT_CPLX_B) {
    dft_cplx_gemm<true, false>(a, w, line0, a.nb, xh, xl, sc, s_x);
    return;
  }
  if (MODE == DFT_CPLX_A) {
    dft_cplx_gemm<false, false>(a, w, line0, a.na, xh, xl, sc, s_x);
    return;
  }
  if (MODE == DFT_C_FWD || MODE == DFT_C_INV) {
    // result tiles through LDS, then row-contiguous stores
    unsigned ntiles = w.M / 32u;
    for (unsigned mt = wave; mt < ntiles; mt += 4u) {
      hd_float16v main, corr;
      dft_tile(w, mt, xh, xl, main, corr);
      for (unsigned v = 0; v < 16; ++v) {
        unsigned m = mt * 32u + mfma_d_row(lane, v);
        s_x[(lane & 31u) * DFT_XP + m] = (main[v] + corr[v] * DFT_LO_UNSCALE) * sc;
      }
    }
    __syncthreads();
    unsigned nrow = a.na * a.nb;
    if (MODE == DFT_C_FWD) {
      for (unsigned e = t; e < DFT_NL * a.nch; e += DFT_THREADS) {
        unsigned ln = e / a.nch, k = e - ln * a.nch;
        if (line0 + ln < nrow)
          a.cgrid[(size_t)(line0 + ln) * a.nch + k] =
              make_float2(s_x[ln * DFT_XP + k], s_x[ln * DFT_XP + a.nch + k]);
      }
    } else {
      for (unsigned e = t; e < DFT_NL * a.nc; e += DFT_THREADS) {
        unsigned ln = e / a.nc, c = e - ln * a.nc;
        if (line0 + ln < nrow) a.vgrid[(size_t)(line0 + ln) * a.nc + c] = s_x[ln * DFT_XP + c];
      }
    }
    return;
  }
  // DFT_A_CONV: forward a-axis into LDS, influence, inverse a-axis
  dft_cplx_gemm<false, true>(a, w, line0, a.na, xh, xl, sc, s_x);
  __syncthreads();
  // influence multiply on the forward a-spectrum of the staged lines
  // (spme_recip.cu:34-115, 296-342), energy/virial partials in FP64
  {
    HdLattice L = hd_load_lattice(a.lattice);
    float v_box = L.lx * L.ly * L.lz;
    float four_alpha2 = 4.0f * a.alpha * a.alpha;
    const float pi_r = 3.14159265358979f;
    float four_pi = 4.0f * pi_r, two_pi = 2.0f * pi_r;
    float pref = a.kc / v_box;
    float ilx = 1.0f / L.lx, ily = 1.0f / L.ly, ilz = 1.0f / L.lz;
    float ra_x = ilx, ra_y = -L.xy * ilx * ily, ra_z = (L.xy * L.yz - L.xz * L.ly) * ilx * ily * ilz;
    float rb_y = ily, rb_z = -L.yz * ily * ilz, rc_z = ilz;
    unsigned na = a.na, nl = a.nb * a.nch;
    double se = 0.0, sw = 0.0;
    for (unsigned e = t; e < DFT_NL * na; e += DFT_THREADS) {
      unsigned ln = e / na, ka = e - ln * na;
      unsigned line = line0 + ln;
      if (line >= nl) continue;
      unsigned b = line / a.nch, kcc = line - b * a.nch;
      float ma = (ka <= na / 2) ? (float)ka : (float)ka - (float)na;
      float mb = (b <= a.nb / 2) ? (float)b : (float)b - (float)a.nb;
      float mc = (kcc <= a.nc / 2) ? (float)kcc : (float)kcc - (float)a.nc;
      float kx = two_pi * (ma * ra_x);
      float ky = two_pi * (ma * ra_y + mb * rb_y);
      float kz = two_pi * (ma * ra_z + mb * rb_z + mc * rc_z);
      float k2 = kx * kx + ky * ky + kz * kz;
      float g = 0.f, vf = 0.f;
      if (k2 != 0.f) {
        g = pref * (four_pi / k2) * expf(-k2 / four_alpha2) * a.bfa[ka] * a.bfb[b] * a.bfc[kcc];
        vf = g * (1.0f - k2 / (2.0f * a.alpha * a.alpha));
      }
      float yr = s_x[ln * DFT_XP + ka], yi = s_x[ln * DFT_XP + na + ka];
      float rsq = yr * yr + yi * yi;
      double hw = (kcc == 0 || 2 * kcc == a.nc) ? 1.0 : 2.0;
      se += hw * (double)g * (double)rsq;
      sw += hw * (double)vf * (double)rsq;
      s_x[ln * DFT_XP + ka] = g * yr;
      s_x[ln * DFT_XP + na + ka] = g * yi;
    }
    red_e[t] = se;
    red_w[t] = sw;
    __syncthreads();
    for (unsigned st = 1; st < DFT_THREADS; st <<= 1) {
      if ((t % (2 * st)) == 0) {
        red_e[t] += red_e[t + st];
        red_w[t] += red_w[t + st];
      }
      __syncthreads();
    }
    if (t == 0) a.partials[blockIdx.x] = make_double2(red_e[0], red_w[0]);
  }
  s = dft_split(s_x, xh, xl, w2.K, &s_max);
  sc = ldexpf(1.0f, s);
  dft_cplx_gemm<false, false>(a, w2, line0, a.na, xh, xl, sc, s_x);
}

unsigned hd_dft_partials_count(unsigned nb, unsigned nc) {
  return (nb * (nc / 2 + 1) + DFT_NL - 1) / DFT_NL;
}

void hd_launch_dft(hipStream_t st, const HdDftArgs& a, const HdDftMats& m) {
  unsigned rows = a.na * a.nb;
  unsigned lb = a.na * a.nch, la = a.nb * a.nch;
  DftMat cf = {m.cf_hi, m.cf_lo, m.cf_m, m.cf_k}, ci = {m.ci_hi, m.ci_lo, m.ci_m, m.ci_k};
  DftMat bf = {m.bf_hi, m.bf_lo, 2 * a.nb, 2 * a.nb}, bi = {m.bi_hi, m.bi_lo, 2 * a.nb, 2 * a.nb};
  DftMat af = {m.af_hi, m.af_lo, 2 * a.na, 2 * a.na}, ai = {m.ai_hi, m.ai_lo, 2 * a.na, 2 * a.na};
  hipLaunchKernelGGL(k_dft<DFT_C_FWD>, dim3((rows + DFT_NL - 1) / DFT_NL), dim3(DFT_THREADS), 0,
                     st, a, cf, cf);
  hipLaunchKernelGGL(k_dft<DFT_CPLX_B>, dim3((lb + DFT_NL - 1) / DFT_NL), dim3(DFT_THREADS), 0,
                     st, a, bf, bf);
  hipLaunchKernelGGL(k_dft<DFT_A_CONV>, dim3((la + DFT_NL - 1) / DFT_NL), dim3(DFT_THREADS), 0,
                     st, a, af, ai);
  hipLaunchKernelGGL(k_dft<DFT_CPLX_B>, dim3((lb + DFT_NL - 1) / DFT_NL), dim3(DFT_THREADS), 0,
                     st, a, bi, bi);
  hipLaunchKernelGGL(k_dft<DFT_C_INV>, dim3((rows + DFT_NL - 1) / DFT_NL), dim3(DFT_THREADS), 0,
                     st, a, ci, ci);
}
