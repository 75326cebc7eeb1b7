// SPME reciprocal space on gfx950: charge spread, an own 3D real FFT in
// LDS, the influence-function multiply with its virial/energy reduction,
// and the force gather.
//
// Reference pipeline (src/forces/spme.rs:403-502):
//   influence (kernels/spme_recip.cu:34-115) -> spread, i64 fixed point at
//   2^32 with round-to-nearest (kernels/spme_spread_gather.cuh:151-221) ->
//   f32 conversion (spme_recip.cu:237-253) -> cuFFT R2C -> influence
//   multiply + per-block virial tree (spme_recip.cu:296-342) -> cuFFT C2R
//   -> reduce (spme_recip.cu:376-403) -> gather (spme_spread_gather.cuh:
//   223-334).
// MI355X design:
//   * no cuFFT/rocFFT: R2C along c (packed half-length complex FFT + post
//     twiddle), then b, then a, all radix-2 in LDS with per-axis twiddle
//     tables;
//   * the a-axis pass fuses forward FFT, influence multiply (G(k) and the
//     virial factor are evaluated in registers from the live lattice, so no
//     G/vf arrays are stored or re-read), the energy/virial partial sums in
//     FP64 and the inverse a-axis FFT: one read and one write of the
//     complex grid instead of five;
//   * the c-axis forward pass reads the i64 spread grid directly (the
//     fixed-point -> f32 conversion is fused);
//   * the gather is fused into the force finalize kernel (k_integrate.hip).
#include "hd_device.h"
#include "hd_kernels.h"

#define FFT_THREADS 256
#define FFT_TILE 2048  // complex elements staged per workgroup

// Cardinal B-spline by Cox-de Boor at order P (spme_spread_gather.cuh:36-54);
// the spread and the gather are built for P = 4 and 5.
template <int P>
__device__ __forceinline__ float bspline_w(float x) {
  float v[P];
#pragma unroll
  for (int i = 0; i < P; ++i) {
    float xi = x - (float)i;
    v[i] = (xi >= 0.f && xi < 1.f) ? 1.f : 0.f;
  }
#pragma unroll
  for (int order = 2; order <= P; ++order) {
    float inv = 1.0f / (float)(order - 1);
#pragma unroll
    for (int i = 0; i < P - order + 1; ++i) {
      float xi = x - (float)i;
      v[i] = xi * inv * v[i] + ((float)order - xi) * inv * v[i + 1];
    }
  }
  return v[0];
}

// Spline records, once per step in slot order: the FP64 spline origin
// (hd_spline_origin) of every charged atom, as (t_a, t_b, t_c, q) plus
// the packed origin cell g_a | g_b << 10 | g_c << 20 (SPME_NO_CELL for
// uncharged atoms).  Read by the spread and by the force gather.
#define SPME_NO_CELL HD_SPME_NO_CELL

__global__ void k_spline_setup(const float4* __restrict__ tile_posq,
                               const float* __restrict__ lattice, unsigned na, unsigned nb,
                               unsigned nc, float4* __restrict__ spl, unsigned* __restrict__ sg,
                               unsigned n) {
  unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  hd_spline_record(tile_posq[s], hd_load_lattice(lattice), na, nb, nc, spl, sg, s);
}

// Spread: the reference accumulates each atom's p^3 contributions with
// i64 atomics at 2^32, rounded to nearest (spme_spread_gather.cuh:151-221).
// Here every workgroup owns one brick of grid cells and computes it
// completely: it gathers the atoms of the neighbour list's voxels that can
// reach the brick, adds their in-brick contributions into LDS with 64-bit
// LDS atomics, and stores the brick with plain coalesced stores.  No global
// atomics (device-scope atomics on MI355X are resolved past the per-XCD L2)
// and no clearing pass.  The per-cell values are the same integers in any
// order, so the grid equals per-atom global accumulation bit for bit.
//
// Which atoms can reach a brick: cells [c0, c1) receive contributions from
// atoms whose origin cell lies in [c0, c1 + p - 1), i.e. whose fractional
// coordinate (+1/2) lies in [c0/n, (c1 + p - 1)/n).  Atoms sit in the voxel
// they were binned into at the last rebuild, and have moved less than half
// the skin since (the rebuild criterion), so the window is widened by mv
// voxels on each side (mv >= skin/2 in voxel units, plus one).  The
// window's atoms are flattened (per-voxel slot ranges + an LDS scan) so
// that the threads share them evenly.
#define SPREAD_THREADS 256
#define SPREAD_BRICK_MAX 2048  // cells per brick (16 KiB of i64)
#define SPREAD_WIN_MAX 3072    // voxels per window listed through LDS
#define SPREAD_PER_MAX (SPREAD_WIN_MAX / SPREAD_THREADS)
#define SPREAD_ATOM_MAX 4096   // window atoms listed in LDS

// round(x) as i64 for an integer-valued or rounded f32 x, the value
// __float2ll_rn gives: one conversion through i32 when |x| < 2^31 (every
// contribution below 0.5 e at the 2^32 scale), the general conversion
// otherwise.
__device__ __forceinline__ long long spread_fixed(float x) {
  float r = rintf(x);
  if (fabsf(r) < 2147483648.0f) return (long long)(int)r;
  return __float2ll_rn(r);
}

template <int SPME_P>
__device__ __forceinline__ void spread_atom(const float4* __restrict__ spl,
                                            const unsigned* __restrict__ sg, unsigned s,
                                            unsigned na, unsigned nb, unsigned nc, int c0a,
                                            int c0b, int c0c, unsigned ba, unsigned bb,
                                            unsigned bc, unsigned long long* acc) {
  unsigned g = sg[s];
  float4 r = spl[s];  // loaded with the origin, not after its test
  if (g == SPME_NO_CELL) return;
  int ga = (int)(g & 1023u), gb = (int)((g >> 10) & 1023u), gc = (int)(g >> 20);
  // brick-relative index of the stencil's first cell per axis; cell
  // (g - d) mod n lies in the brick when 0 <= rel - d < extent (g and the
  // brick origin both lie in [0, n), so one conditional add is mod n)
  int rla = ga - c0a, rlb = gb - c0b, rlc = gc - c0c;
  rla += rla < 0 ? (int)na : 0;
  rlb += rlb < 0 ? (int)nb : 0;
  rlc += rlc < 0 ? (int)nc : 0;
  if (rla >= (int)ba + SPME_P - 1 || rlb >= (int)bb + SPME_P - 1 ||
      rlc >= (int)bc + SPME_P - 1)
    return;
  float wa[SPME_P], wb[SPME_P], wc[SPME_P];
#pragma unroll
  for (int d = 0; d < SPME_P; ++d) {
    wa[d] = bspline_w<SPME_P>((float)d + r.x);
    wb[d] = bspline_w<SPME_P>((float)d + r.y);
    wc[d] = bspline_w<SPME_P>((float)d + r.z);
  }
#pragma unroll
  for (int iz = 0; iz < SPME_P; ++iz) {
    int oc = rlc - iz;
    if ((unsigned)oc >= bc) continue;
    // the 2^32 scale is a power of two: applying it to q w_c first gives
    // the same products as applying it last
    float dz = (r.w * wc[iz]) * 4294967296.0f;
#pragma unroll
    for (int da = 0; da < SPME_P; ++da) {
      int oa = rla - da;
      if ((unsigned)oa >= ba) continue;
      float dzda = dz * wa[da];
#pragma unroll
      for (int db = 0; db < SPME_P; ++db) {
        int ob = rlb - db;
        if ((unsigned)ob >= bb) continue;
        long long vf = spread_fixed(dzda * wb[db]);
        if (vf != 0)
          atomicAdd(&acc[((unsigned)oa * bb + (unsigned)ob) * bc + (unsigned)oc],
                    (unsigned long long)vf);
      }
    }
  }
}

template <int SPME_P>
__global__ void __launch_bounds__(SPREAD_THREADS) k_spread(
    const float4* __restrict__ spl, const unsigned* __restrict__ sg, unsigned na, unsigned nb,
    unsigned nc, unsigned ba, unsigned bb, unsigned bc, const unsigned* __restrict__ voxel_rank,
    const unsigned* __restrict__ voxel_off, unsigned nva, unsigned nvb, unsigned nvc, int mva,
    int mvb, int mvc, unsigned long long* __restrict__ rho) {
  __shared__ unsigned long long acc[SPREAD_BRICK_MAX];
  __shared__ unsigned wslot[SPREAD_ATOM_MAX];
  __shared__ unsigned part[SPREAD_THREADS];
  unsigned t = threadIdx.x;
  unsigned nbb = nb / bb, nbc = nc / bc;
  unsigned bi = blockIdx.x;
  unsigned bic = bi % nbc, bib = (bi / nbc) % nbb, bia = bi / (nbc * nbb);
  int c0a = (int)(bia * ba), c0b = (int)(bib * bb), c0c = (int)(bic * bc);
  unsigned vol = ba * bb * bc;
  for (unsigned k = t; k < vol; k += SPREAD_THREADS) acc[k] = 0ull;
  // voxel window per axis (periodic), clamped to one full period
  int vla = (int)(((long long)c0a * nva) / na) - mva;
  int vha = (int)((((long long)c0a + ba + SPME_P - 1) * nva - 1) / na) + mva;
  int vlb = (int)(((long long)c0b * nvb) / nb) - mvb;
  int vhb = (int)((((long long)c0b + bb + SPME_P - 1) * nvb - 1) / nb) + mvb;
  int vlc = (int)(((long long)c0c * nvc) / nc) - mvc;
  int vhc = (int)((((long long)c0c + bc + SPME_P - 1) * nvc - 1) / nc) + mvc;
  unsigned wa_n = (unsigned)min(vha - vla + 1, (int)nva);
  unsigned wb_n = (unsigned)min(vhb - vlb + 1, (int)nvb);
  unsigned wc_n = (unsigned)min(vhc - vlc + 1, (int)nvc);
  unsigned nwin = wa_n * wb_n * wc_n;
  bool flat = nwin <= SPREAD_WIN_MAX;  // uniform over the workgroup
  __syncthreads();  // acc cleared before any accumulation
  unsigned per = (nwin + SPREAD_THREADS - 1) / SPREAD_THREADS;
  // thread t owns the contiguous window range [t*per, t*per + per): the
  // slot range of each of its voxels, kept in registers
  unsigned vs0[SPREAD_PER_MAX], vcnt[SPREAD_PER_MAX];
  unsigned my = 0;
#pragma unroll
  for (unsigned j = 0; j < SPREAD_PER_MAX; ++j) {
    unsigned k = t * per + j;
    vs0[j] = 0u;
    vcnt[j] = 0u;
    if (flat && j < per && k < nwin) {
      unsigned kc = k % wc_n, kab = k / wc_n;
      unsigned kb = kab % wb_n, ka = kab / wb_n;
      // periodic wrap; the voxel grid sides are powers of two
      unsigned va = (unsigned)((vla + (int)ka) & ((int)nva - 1));
      unsigned vb = (unsigned)((vlb + (int)kb) & ((int)nvb - 1));
      unsigned vc = (unsigned)((vlc + (int)kc) & ((int)nvc - 1));
      unsigned r = voxel_rank[(va * nvb + vb) * nvc + vc];
      vs0[j] = voxel_off[r];
      vcnt[j] = voxel_off[r + 1] - vs0[j];
      my += vcnt[j];
    }
  }
  // exclusive scan of the per-thread totals
  part[t] = my;
  __syncthreads();
  for (unsigned off = 1; off < SPREAD_THREADS; off <<= 1) {
    unsigned u = (t >= off) ? part[t - off] : 0u;
    __syncthreads();
    part[t] += u;
    __syncthreads();
  }
  unsigned total = part[SPREAD_THREADS - 1];
  bool listed = flat && total <= SPREAD_ATOM_MAX;  // uniform over the workgroup
  if (listed) {
    // the window's atoms in window order, then shared evenly
    unsigned run = part[t] - my;
#pragma unroll
    for (unsigned j = 0; j < SPREAD_PER_MAX; ++j)
      for (unsigned i = 0; i < vcnt[j]; ++i) wslot[run++] = vs0[j] + i;
    __syncthreads();
    for (unsigned f = t; f < total; f += SPREAD_THREADS)
      spread_atom<SPME_P>(spl, sg, wslot[f], na, nb, nc, c0a, c0b, c0c, ba, bb, bc, acc);
  } else {
    // a window too wide to list: every thread walks its own voxels
    for (unsigned j = 0; j < per; ++j) {
      unsigned k = t * per + j;
      if (k >= nwin) break;
      unsigned kc = k % wc_n, kab = k / wc_n;
      unsigned kb = kab % wb_n, ka = kab / wb_n;
      unsigned va = (unsigned)((vla + (int)ka) & ((int)nva - 1));
      unsigned vb = (unsigned)((vlb + (int)kb) & ((int)nvb - 1));
      unsigned vc = (unsigned)((vlc + (int)kc) & ((int)nvc - 1));
      unsigned r = voxel_rank[(va * nvb + vb) * nvc + vc];
      unsigned s0 = voxel_off[r], s1 = voxel_off[r + 1];
      for (unsigned sl = s0; sl < s1; ++sl)
        spread_atom<SPME_P>(spl, sg, sl, na, nb, nc, c0a, c0b, c0c, ba, bb, bc, acc);
    }
  }
  __syncthreads();
  for (unsigned k = t; k < vol; k += SPREAD_THREADS) {
    unsigned kc = k % bc, kab = k / bc;
    unsigned kb = kab % bb, ka = kab / bb;
    rho[((size_t)(c0a + ka) * nb + (c0b + kb)) * nc + (c0c + kc)] = acc[k];
  }
}

// the largest divisor of n that is at most cap
static unsigned divisor_at_most(unsigned n, unsigned cap) {
  for (unsigned d = cap < n ? cap : n; d > 1; --d)
    if (n % d == 0) return d;
  return 1;
}

unsigned hd_spread_bricks(unsigned na, unsigned nb, unsigned nc, unsigned brick[3]) {
  // up to 8 x 16 x 16 cells, each side a divisor of the grid's (8 and 16
  // for the power-of-two grids; e.g. 5 x 10 x 16 for 20 x 20 x 80)
  brick[0] = divisor_at_most(na, 8);
  brick[1] = divisor_at_most(nb, 16);
  brick[2] = divisor_at_most(nc, 16);
  return (na / brick[0]) * (nb / brick[1]) * (nc / brick[2]);
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = 0ull;
}

// ---------------------------------------------------------------------
// Radix-2 complex FFT over `nlines` lines of length n held contiguously in
// LDS, input already in bit-reversed order.  tw[k] = exp(-2 pi i k / twn),
// k < twn/2, staged in LDS by the caller (fft_stage_twiddles) so that no
// butterfly waits on a global load; sign = -1 forward, +1 inverse
// (unnormalised).
// ---------------------------------------------------------------------
#define FFT_TW_MAX 256  // twn / 2 for the largest axis (nc <= 512)

__device__ __forceinline__ void fft_stage_twiddles(float2* stw, const float2* __restrict__ tw,
                                                   unsigned twn) {
  for (unsigned i = threadIdx.x; i < twn / 2; i += blockDim.x) stw[i] = tw[i];
}
__device__ __forceinline__ float2 cmul(float2 b, float2 w) {
  return make_float2(b.x * w.x - b.y * w.y, b.x * w.y + b.y * w.x);
}

// Two radix-2 stages per LDS pass (a radix-4 unit of 4 elements per thread:
// the same butterflies in the same order, half the passes and barriers),
// a last radix-2 pass when log2n is odd.
__device__ __forceinline__ void lds_fft(float2* d, unsigned n, unsigned log2n, unsigned nlines,
                                        unsigned lstride, int sign, const float2* tw,
                                        unsigned twn) {
  unsigned s = 1;
  for (; s + 1 <= log2n; s += 2) {
    unsigned h = 1u << (s - 1);         // half of stage s
    unsigned t1 = twn >> s, t2 = twn >> (s + 1);
    unsigned qn = n >> 2;                // radix-4 units per line
    unsigned nu = nlines * qn;
    for (unsigned idx = threadIdx.x; idx < nu; idx += blockDim.x) {
      unsigned line = idx >> (log2n - 2), u = idx & (qn - 1);
      unsigned j = u & (h - 1);
      unsigned i0 = ((u >> (s - 1)) << (s + 1)) + j;
      float2 w1 = tw[j * t1], w2 = tw[j * t2], w3 = tw[(j + h) * t2];
      if (sign > 0) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
      float2* row = d + line * lstride;
      float2 a0 = row[i0], a1 = row[i0 + h], a2 = row[i0 + 2 * h], a3 = row[i0 + 3 * h];
      float2 b1 = cmul(a1, w1), b3 = cmul(a3, w1);
      float2 y0 = make_float2(a0.x + b1.x, a0.y + b1.y), y1 = make_float2(a0.x - b1.x, a0.y - b1.y);
      float2 y2 = make_float2(a2.x + b3.x, a2.y + b3.y), y3 = make_float2(a2.x - b3.x, a2.y - b3.y);
      float2 c2 = cmul(y2, w2), c3 = cmul(y3, w3);
      row[i0] = make_float2(y0.x + c2.x, y0.y + c2.y);
      row[i0 + 2 * h] = make_float2(y0.x - c2.x, y0.y - c2.y);
      row[i0 + h] = make_float2(y1.x + c3.x, y1.y + c3.y);
      row[i0 + 3 * h] = make_float2(y1.x - c3.x, y1.y - c3.y);
    }
    __syncthreads();
  }
  if (s == log2n) {
    unsigned half = 1u << (s - 1);
    unsigned tstr = twn >> s;
    unsigned hn = n >> 1;
    unsigned nbf = nlines * hn;
    for (unsigned idx = threadIdx.x; idx < nbf; idx += blockDim.x) {
      unsigned line = idx >> (log2n - 1);
      unsigned k = idx & (hn - 1);
      unsigned j = k & (half - 1);
      unsigned i0 = ((k >> (s - 1)) << s) + j;
      unsigned i1 = i0 + half;
      float2 w = tw[j * tstr];
      if (sign > 0) w.y = -w.y;
      float2* row = d + line * lstride;
      float2 a = row[i0], bw = cmul(row[i1], w);
      row[i0] = make_float2(a.x + bw.x, a.y + bw.y);
      row[i1] = make_float2(a.x - bw.x, a.y - bw.y);
    }
    __syncthreads();
  }
}

// Decimation-in-frequency twin: natural-order input, bit-reversed output;
// stages log2n .. 1, two per pass (the radix-2 DIF butterflies in order).
__device__ __forceinline__ void lds_fft_dif(float2* d, unsigned n, unsigned log2n,
                                            unsigned nlines, unsigned lstride, int sign,
                                            const float2* tw, unsigned twn) {
  unsigned s = log2n;
  for (; s >= 2; s -= 2) {
    unsigned H = 1u << (s - 2);          // half of stage s - 1
    unsigned t1 = twn >> s, t2 = twn >> (s - 1);
    unsigned qn = n >> 2;
    unsigned nu = nlines * qn;
    for (unsigned idx = threadIdx.x; idx < nu; idx += blockDim.x) {
      unsigned line = idx >> (log2n - 2), u = idx & (qn - 1);
      unsigned j = u & (H - 1);
      unsigned i0 = ((u >> (s - 2)) << s) + j;
      float2 w1 = tw[j * t1], w1b = tw[(j + H) * t1], w2 = tw[j * t2];
      if (sign > 0) { w1.y = -w1.y; w1b.y = -w1b.y; w2.y = -w2.y; }
      float2* row = d + line * lstride;
      float2 x0 = row[i0], x1 = row[i0 + H], x2 = row[i0 + 2 * H], x3 = row[i0 + 3 * H];
      float2 y0 = make_float2(x0.x + x2.x, x0.y + x2.y);
      float2 y2 = cmul(make_float2(x0.x - x2.x, x0.y - x2.y), w1);
      float2 y1 = make_float2(x1.x + x3.x, x1.y + x3.y);
      float2 y3 = cmul(make_float2(x1.x - x3.x, x1.y - x3.y), w1b);
      row[i0] = make_float2(y0.x + y1.x, y0.y + y1.y);
      row[i0 + H] = cmul(make_float2(y0.x - y1.x, y0.y - y1.y), w2);
      row[i0 + 2 * H] = make_float2(y2.x + y3.x, y2.y + y3.y);
      row[i0 + 3 * H] = cmul(make_float2(y2.x - y3.x, y2.y - y3.y), w2);
    }
    __syncthreads();
  }
  if (s == 1) {
    unsigned hn = n >> 1;
    unsigned nbf = nlines * hn;
    for (unsigned idx = threadIdx.x; idx < nbf; idx += blockDim.x) {
      unsigned line = idx >> (log2n - 1);
      unsigned k = idx & (hn - 1);
      unsigned i0 = k << 1, i1 = i0 + 1;
      float2 w = tw[0];
      if (sign > 0) w.y = -w.y;
      float2* row = d + line * lstride;
      float2 a = row[i0], b = row[i1];
      float2 dlt = make_float2(a.x - b.x, a.y - b.y);
      row[i0] = make_float2(a.x + b.x, a.y + b.y);
      row[i1] = cmul(dlt, w);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------
// Mixed-radix line FFT for lengths n = 2^a 3^b 5^c that are not powers of
// two (the reference sizes its grids for cuFFT's small primes, e.g.
// 48 x 48 x 96 and 20 x 20 x 80).  Self-sorting (Stockham) stages of radix
// 4, then 2, 3, 5: natural order in and out, so no bit reversal.  A stage
// of radix R with span ns (the product of the earlier radices) takes, for
// every j < n/R, the R values x[j + r n/R], multiplies value r by
// w^(r k) with w = exp(sign 2 pi i / (ns R)) and k = j mod ns, applies the
// R-point DFT and stores output r at (j div ns) ns R + k + r ns.  Each
// thread holds its butterflies' outputs in registers across a barrier, so
// the stage runs in place in LDS.
// ---------------------------------------------------------------------
#define FFT_MR_PER_THREAD 8  // butterflies per thread: nlines * n / R <= 2048 at 256 threads

// exp(sign * 2 pi i t / twn) for 0 <= t < twn from the half-circle table
// stw[t] = exp(-2 pi i t / twn), t < twn / 2
__device__ __forceinline__ float2 tw_full(const float2* stw, unsigned t, unsigned twn, int sign) {
  unsigned half = twn >> 1;
  float2 w = (t < half) ? stw[t] : make_float2(-stw[t - half].x, -stw[t - half].y);
  if (sign > 0) w.y = -w.y;
  return w;
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float f) { return make_float2(a.x * f, a.y * f); }
// a * (i f)
__device__ __forceinline__ float2 cmul_if(float2 a, float f) { return make_float2(-a.y * f, a.x * f); }

// R-point DFT with exp(sign 2 pi i jk / R), in place
template <int R>
__device__ __forceinline__ void small_dft(float2 (&a)[R], int sign) {
  const float sg = (float)sign;
  if constexpr (R == 2) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  } else if constexpr (R == 4) {
    float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]), t3 = cmul_if(csub(a[1], a[3]), sg);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  } else if constexpr (R == 3) {
    const float h3 = 0.86602540378443865f;  // sin(2 pi / 3)
    float2 t = cadd(a[1], a[2]);
    float2 u = cmul_if(csub(a[1], a[2]), sg * h3);
    float2 m = csub(a[0], cscale(t, 0.5f));
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, u);
    a[2] = csub(m, u);
  } else {  // R == 5
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;  // cos 2pi/5, cos 4pi/5
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;   // sin 2pi/5, sin 4pi/5
    float2 b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
    float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
    float2 e1 = cadd(a[0], cadd(cscale(b1, c1), cscale(b2, c2)));
    float2 e2 = cadd(a[0], cadd(cscale(b1, c2), cscale(b2, c1)));
    float2 f1 = cmul_if(cadd(cscale(d1, s1), cscale(d2, s2)), sg);
    float2 f2 = cmul_if(csub(cscale(d1, s2), cscale(d2, s1)), sg);
    a[0] = cadd(a[0], cadd(b1, b2));
    a[1] = cadd(e1, f1);
    a[4] = csub(e1, f1);
    a[2] = cadd(e2, f2);
    a[3] = csub(e2, f2);
  }
}

template <int R>
__device__ __forceinline__ void mr_stage(float2* d, unsigned n, unsigned nlines, unsigned lstride,
                                         unsigned ns, int sign, const float2* stw, unsigned twn) {
  const unsigned m = n / R, total = nlines * m;
  const unsigned tstep = twn / (ns * R);
  float2 v[FFT_MR_PER_THREAD][R];
  unsigned dst[FFT_MR_PER_THREAD];
#pragma unroll
  for (int b = 0; b < FFT_MR_PER_THREAD; ++b) {
    unsigned idx = threadIdx.x + (unsigned)b * blockDim.x;
    dst[b] = 0xFFFFFFFFu;
    if (idx < total) {
      unsigned line = idx / m, j = idx - line * m;
      unsigned k = j % ns;
      const float2* row = d + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = row[j + (unsigned)r * m];
        v[b][r] = r ? cmul(x, tw_full(stw, (unsigned)r * k * tstep, twn, sign)) : x;
      }
      small_dft<R>(v[b], sign);
      dst[b] = line * lstride + (j / ns) * ns * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < FFT_MR_PER_THREAD; ++b)
    if (dst[b] != 0xFFFFFFFFu) {
#pragma unroll
      for (int r = 0; r < R; ++r) d[dst[b] + (unsigned)r * ns] = v[b][r];
    }
  __syncthreads();
}

// Complex FFT of `nlines` LDS lines of length n (2^a 3^b 5^c), natural
// order in and out; stw: exp(-2 pi i t / twn), t < twn / 2, with n | twn.
__device__ __forceinline__ void lds_fft_mixed(float2* d, unsigned n, unsigned nlines,
                                              unsigned lstride, int sign, const float2* stw,
                                              unsigned twn) {
  unsigned ns = 1, rem = n;
  while (rem > 1) {
    if (rem % 4 == 0) {
      mr_stage<4>(d, n, nlines, lstride, ns, sign, stw, twn);
      ns *= 4; rem /= 4;
    } else if (rem % 2 == 0) {
      mr_stage<2>(d, n, nlines, lstride, ns, sign, stw, twn);
      ns *= 2; rem /= 2;
    } else if (rem % 3 == 0) {
      mr_stage<3>(d, n, nlines, lstride, ns, sign, stw, twn);
      ns *= 3; rem /= 3;
    } else {
      mr_stage<5>(d, n, nlines, lstride, ns, sign, stw, twn);
      ns *= 5; rem /= 5;
    }
  }
}

__device__ __forceinline__ unsigned bitrev(unsigned k, unsigned log2n) {
  return log2n ? (__brev(k) >> (32 - log2n)) : 0u;
}

// c-axis forward R2C: rows of nc reals read from the i64 spread grid
// (converted to f32 exactly as spme_recip.cu:237-253), packed as nc/2
// complex values, FFT'd, then untangled into nc/2+1 spectrum values.
// twc: twiddles for length nc (exp(-2 pi i k/nc), k < nc/2).
template <bool MIXED>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_c_fwd(
    const long long* __restrict__ rho, float2* __restrict__ cgrid, unsigned nrows, unsigned nc,
    unsigned log2h, const float2* __restrict__ twc) {
  __shared__ float2 d[FFT_TILE];
  __shared__ float2 stw[FFT_TW_MAX];
  unsigned h = nc >> 1, nch = h + 1;
  unsigned rows = FFT_TILE / h;
  unsigned r0 = blockIdx.x * rows;
  unsigned nr = min(rows, nrows - r0);
  const float inv = 1.0f / 4294967296.0f;
  fft_stage_twiddles(stw, twc, nc);
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, m = e - r * h;
    longlong2 v = reinterpret_cast<const longlong2*>(rho + (size_t)(r0 + r) * nc)[m];
    d[r * h + (MIXED ? m : bitrev(m, log2h))] = make_float2((float)v.x * inv, (float)v.y * inv);
  }
  __syncthreads();
  if (MIXED)
    lds_fft_mixed(d, h, nr, h, -1, stw, nc);
  else
    lds_fft(d, h, log2h, nr, h, -1, stw, nc);
  for (unsigned e = threadIdx.x; e < nr * nch; e += blockDim.x) {
    unsigned r = e / nch, k = e - r * nch;
    float2 zk = d[r * h + (k % h)];
    float2 zc = d[r * h + ((h - k) % h)];
    zc.y = -zc.y;  // conj(Z[h-k])
    float2 E = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    // O = (Z[k] - conj(Z[h-k])) / (2i)
    float2 O = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
    float2 w = (k < h) ? stw[k] : make_float2(-1.f, 0.f);  // exp(-2 pi i k / nc)
    float2 wO = make_float2(w.x * O.x - w.y * O.y, w.x * O.y + w.y * O.x);
    cgrid[(size_t)(r0 + r) * nch + k] = make_float2(E.x + wO.x, E.y + wO.y);
  }
}

// c-axis inverse C2R: nc/2+1 spectrum values per row -> nc reals,
// unnormalised (round trip = nc * x).
// 256 threads; se, sw: 256 doubles of LDS each
static_assert(FFT_THREADS == 256, "the inverse c pass runs the 256-thread reduction");
__device__ __forceinline__ void recip_reduce_block(const double2* __restrict__ partials,
                                                   unsigned np, double* __restrict__ recip,
                                                   float* __restrict__ w_share, unsigned n,
                                                   double* se, double* sw) {
  double e = 0.0, w = 0.0;
  for (unsigned i = threadIdx.x; i < np; i += 256) {
    double2 p = partials[i];
    e += p.x;
    w += p.y;
  }
  se[threadIdx.x] = e;
  sw[threadIdx.x] = w;
  __syncthreads();
  for (unsigned s = 1; s < 256; s <<= 1) {
    if ((threadIdx.x % (2 * s)) == 0) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sw[threadIdx.x] += sw[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    recip[0] = 0.5 * se[0];
    recip[1] = 0.5 * sw[0];
    w_share[0] = n ? (float)(0.5 * sw[0] / (double)n) : 0.f;
  }
}

template <bool MIXED>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_c_inv(
    const float2* __restrict__ cgrid, float* __restrict__ vgrid, unsigned nrows, unsigned nc,
    unsigned log2h, const float2* __restrict__ twc, const double2* __restrict__ partials,
    unsigned np, double* __restrict__ recip, float* __restrict__ w_share, unsigned n_atoms) {
  __shared__ float2 d[FFT_TILE];
  __shared__ float2 stw[FFT_TW_MAX];
  if (partials && blockIdx.x == 0) {
    // the convolution's energy and virial partials are complete (written
    // by the a-axis pass before this launch): block 0 reduces them first,
    // in k_recip_reduce's order, with the tile's LDS as its scratch
    double* se = reinterpret_cast<double*>(d);
    recip_reduce_block(partials, np, recip, w_share, n_atoms, se, se + 256);
    __syncthreads();
  }
  unsigned h = nc >> 1, nch = h + 1;
  unsigned rows = FFT_TILE / h;
  unsigned r0 = blockIdx.x * rows;
  unsigned nr = min(rows, nrows - r0);
  fft_stage_twiddles(stw, twc, nc);
  __syncthreads();
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, k = e - r * h;
    const float2* row = cgrid + (size_t)(r0 + r) * nch;
    float2 xk = row[k], xc = row[h - k];
    xc.y = -xc.y;  // conj(X[h-k])
    float2 A = make_float2(xk.x + xc.x, xk.y + xc.y);
    float2 B = make_float2(xk.x - xc.x, xk.y - xc.y);
    float2 w = stw[k];
    w.y = -w.y;  // exp(+2 pi i k / nc)
    float2 wB = make_float2(w.x * B.x - w.y * B.y, w.x * B.y + w.y * B.x);
    // Z' = A + i * wB
    d[r * h + (MIXED ? k : bitrev(k, log2h))] = make_float2(A.x - wB.y, A.y + wB.x);
  }
  __syncthreads();
  if (MIXED)
    lds_fft_mixed(d, h, nr, h, +1, stw, nc);
  else
    lds_fft(d, h, log2h, nr, h, +1, stw, nc);
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, m = e - r * h;
    reinterpret_cast<float2*>(vgrid + (size_t)(r0 + r) * nc)[m] = d[r * h + m];
  }
}

// Strided axis pass over the complex grid: lines along an axis of length
// n with element stride `estride` (complex units); `nouter` groups of
// lines start at outer*ostride; within a group, CW consecutive columns.
#define FFT_CW 16
// LDS lines are padded to n+1 complex values to spread the column-wise
// accesses over the banks.  `rev_in` stores in bit-reversed order (for a
// DIT pass); `rev_out` reads a bit-reversed line (after a DIF pass).
__device__ __forceinline__ void strided_load(float2* d, const float2* __restrict__ g,
                                             size_t base, unsigned n, unsigned log2n,
                                             size_t estride, unsigned ncol, bool rev_in) {
  for (unsigned e = threadIdx.x; e < n * FFT_CW; e += blockDim.x) {
    unsigned k = e / FFT_CW, c = e - k * FFT_CW;
    float2 v = make_float2(0.f, 0.f);
    if (c < ncol) v = g[base + (size_t)k * estride + c];
    d[c * (n + 1) + (rev_in ? bitrev(k, log2n) : k)] = v;
  }
}
__device__ __forceinline__ void strided_store(const float2* d, float2* __restrict__ g,
                                              size_t base, unsigned n, unsigned log2n,
                                              size_t estride, unsigned ncol, bool rev_out) {
  for (unsigned e = threadIdx.x; e < n * FFT_CW; e += blockDim.x) {
    unsigned k = e / FFT_CW, c = e - k * FFT_CW;
    if (c < ncol) g[base + (size_t)k * estride + c] = d[c * (n + 1) + (rev_out ? bitrev(k, log2n) : k)];
  }
}

// b-axis pass (forward or inverse), in place.  grid: (na * nchunks).
template <bool MIXED>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_b(float2* __restrict__ cgrid, unsigned na,
                                                       unsigned nb, unsigned nch,
                                                       unsigned log2b, int sign,
                                                       const float2* __restrict__ twb) {
  __shared__ float2 d[FFT_CW * 257];
  __shared__ float2 stw[FFT_TW_MAX];
  unsigned nchunk = (nch + FFT_CW - 1) / FFT_CW;
  unsigned a = blockIdx.x / nchunk, c0 = (blockIdx.x - a * nchunk) * FFT_CW;
  unsigned ncol = min(FFT_CW, nch - c0);
  size_t base = (size_t)a * nb * nch + c0;
  fft_stage_twiddles(stw, twb, nb);
  strided_load(d, cgrid, base, nb, log2b, nch, ncol, !MIXED);
  __syncthreads();
  if (MIXED)
    lds_fft_mixed(d, nb, FFT_CW, nb + 1, sign, stw, nb);
  else
    lds_fft(d, nb, log2b, FFT_CW, nb + 1, sign, stw, nb);
  strided_store(d, cgrid, base, nb, log2b, nch, ncol, false);
}

// Grids whose a-plane fits in LDS (nb * nc/2 <= FFT_TILE complex values
// and the plane's nc/2 + 1 columns of nb, padded, within FFT_PLANE_MAX):
// one workgroup per a-plane runs the c-axis and b-axis passes of the
// forward transform back to back (and of the inverse, b then c), with the
// plane staged in LDS between them instead of a round trip through the
// complex grid and a second launch.  Each line is transformed by the same
// LDS routines, on the same values, as in k_fft_c_fwd / k_fft_b /
// k_fft_c_inv; the compiler may round a few multiply-adds differently in
// the two kernels, so the results agree to f32 rounding.
#define FFT_PLANE_MAX 2176  // 33 columns x 65 (c <= 64, b <= 64)

__host__ __device__ inline bool fft_plane_fits(unsigned nb, unsigned nc) {
  return nb * (nc / 2) <= FFT_TILE && (nc / 2 + 1) * (nb + 1) <= FFT_PLANE_MAX;
}

template <bool MC, bool MB>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_cb_fwd(
    const long long* __restrict__ rho, float2* __restrict__ cgrid, unsigned nb, unsigned nc,
    unsigned log2h, unsigned log2b, const float2* __restrict__ twc,
    const float2* __restrict__ twb) {
  __shared__ float2 d[FFT_TILE];
  __shared__ float2 col[FFT_PLANE_MAX];
  __shared__ float2 stw[FFT_TW_MAX], stwb[FFT_TW_MAX];
  const unsigned h = nc >> 1, nch = h + 1, nr = nb, ls = nb + 1;
  const unsigned r0 = blockIdx.x * nb;
  const float inv = 1.0f / 4294967296.0f;
  fft_stage_twiddles(stw, twc, nc);
  fft_stage_twiddles(stwb, twb, nb);
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, m = e - r * h;
    longlong2 v = reinterpret_cast<const longlong2*>(rho + (size_t)(r0 + r) * nc)[m];
    d[r * h + (MC ? m : bitrev(m, log2h))] = make_float2((float)v.x * inv, (float)v.y * inv);
  }
  __syncthreads();
  if (MC)
    lds_fft_mixed(d, h, nr, h, -1, stw, nc);
  else
    lds_fft(d, h, log2h, nr, h, -1, stw, nc);
  // k_fft_c_fwd's untangling, into the b-axis columns
  for (unsigned e = threadIdx.x; e < nr * nch; e += blockDim.x) {
    unsigned r = e / nch, k = e - r * nch;
    float2 zk = d[r * h + (k % h)];
    float2 zc = d[r * h + ((h - k) % h)];
    zc.y = -zc.y;
    float2 E = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    float2 O = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
    float2 w = (k < h) ? stw[k] : make_float2(-1.f, 0.f);
    float2 wO = make_float2(w.x * O.x - w.y * O.y, w.x * O.y + w.y * O.x);
    col[k * ls + (MB ? r : bitrev(r, log2b))] = make_float2(E.x + wO.x, E.y + wO.y);
  }
  __syncthreads();
  if (MB)
    lds_fft_mixed(col, nb, nch, ls, -1, stwb, nb);
  else
    lds_fft(col, nb, log2b, nch, ls, -1, stwb, nb);
  for (unsigned e = threadIdx.x; e < nr * nch; e += blockDim.x) {
    unsigned r = e / nch, k = e - r * nch;
    cgrid[(size_t)(r0 + r) * nch + k] = col[k * ls + r];
  }
}

template <bool MB, bool MC>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_bc_inv(
    const float2* __restrict__ cgrid, float* __restrict__ vgrid, unsigned nb, unsigned nc,
    unsigned log2h, unsigned log2b, const float2* __restrict__ twc,
    const float2* __restrict__ twb, const double2* __restrict__ partials, unsigned np,
    double* __restrict__ recip, float* __restrict__ w_share, unsigned n_atoms) {
  __shared__ float2 d[FFT_TILE];
  __shared__ float2 col[FFT_PLANE_MAX];
  __shared__ float2 stw[FFT_TW_MAX], stwb[FFT_TW_MAX];
  const unsigned h = nc >> 1, nch = h + 1, nr = nb, ls = nb + 1;
  const unsigned r0 = blockIdx.x * nb;
  if (partials && blockIdx.x == 0) {
    // as in k_fft_c_inv: block 0 reduces the energy and virial partials
    double* se = reinterpret_cast<double*>(d);
    recip_reduce_block(partials, np, recip, w_share, n_atoms, se, se + 256);
    __syncthreads();
  }
  fft_stage_twiddles(stw, twc, nc);
  fft_stage_twiddles(stwb, twb, nb);
  for (unsigned e = threadIdx.x; e < nr * nch; e += blockDim.x) {
    unsigned r = e / nch, k = e - r * nch;
    col[k * ls + (MB ? r : bitrev(r, log2b))] = cgrid[(size_t)(r0 + r) * nch + k];
  }
  __syncthreads();
  if (MB)
    lds_fft_mixed(col, nb, nch, ls, +1, stwb, nb);
  else
    lds_fft(col, nb, log2b, nch, ls, +1, stwb, nb);
  // k_fft_c_inv's packing of the row spectrum, from the b-axis columns
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, k = e - r * h;
    float2 xk = col[k * ls + r], xc = col[(h - k) * ls + r];
    xc.y = -xc.y;
    float2 A = make_float2(xk.x + xc.x, xk.y + xc.y);
    float2 B = make_float2(xk.x - xc.x, xk.y - xc.y);
    float2 w = stw[k];
    w.y = -w.y;
    float2 wB = make_float2(w.x * B.x - w.y * B.y, w.x * B.y + w.y * B.x);
    d[r * h + (MC ? k : bitrev(k, log2h))] = make_float2(A.x - wB.y, A.y + wB.x);
  }
  __syncthreads();
  if (MC)
    lds_fft_mixed(d, h, nr, h, +1, stw, nc);
  else
    lds_fft(d, h, log2h, nr, h, +1, stw, nc);
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, m = e - r * h;
    reinterpret_cast<float2*>(vgrid + (size_t)(r0 + r) * nc)[m] = d[r * h + m];
  }
}

// a-axis pass fused with the influence multiply:
//   forward FFT over a -> for every k: G(k), virial factor vf(k)
//   (spme_recip.cu:34-115 in f32 from the live lattice), energy and
//   virial partials hw*G*|rho^|^2 and hw*vf*|rho^|^2 (spme_recip.cu:
//   296-342) summed in FP64 by a fixed tree -> V^ = G rho^ -> inverse FFT
//   over a.  grid: (nb * nchunks); partials[blockIdx] = (E, W).
template <bool MIXED>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_a_conv(
    float2* __restrict__ cgrid, unsigned na, unsigned nb, unsigned nc, unsigned nch,
    unsigned log2a, const float2* __restrict__ twa, const float* __restrict__ bfa,
    const float* __restrict__ bfb, const float* __restrict__ bfc,
    const float* __restrict__ lattice, float kc, float alpha, double2* __restrict__ partials) {
  __shared__ float2 d[FFT_CW * 257];
  __shared__ float2 stw[FFT_TW_MAX];
  __shared__ double red_e[FFT_THREADS], red_w[FFT_THREADS];
  unsigned nchunk = (nch + FFT_CW - 1) / FFT_CW;
  unsigned b = blockIdx.x / nchunk, c0 = (blockIdx.x - b * nchunk) * FFT_CW;
  unsigned ncol = min(FFT_CW, nch - c0);
  size_t base = (size_t)b * nch + c0;
  size_t estride = (size_t)nb * nch;
  unsigned ls = na + 1;
  fft_stage_twiddles(stw, twa, na);
  strided_load(d, cgrid, base, na, log2a, estride, ncol, !MIXED);
  __syncthreads();
  if (MIXED)
    lds_fft_mixed(d, na, FFT_CW, ls, -1, stw, na);
  else
    lds_fft(d, na, log2a, FFT_CW, ls, -1, stw, na);
  HdLattice L = hd_load_lattice(lattice);
  float v_box = L.lx * L.ly * L.lz;
  float four_alpha2 = 4.0f * alpha * alpha;
  const float pi_r = 3.14159265358979f;
  float four_pi = 4.0f * pi_r, two_pi = 2.0f * pi_r;
  float pref = kc / v_box;
  float ilx = 1.0f / L.lx, ily = 1.0f / L.ly, ilz = 1.0f / L.lz;
  float ra_x = ilx, ra_y = -L.xy * ilx * ily, ra_z = (L.xy * L.yz - L.xz * L.ly) * ilx * ily * ilz;
  float rb_y = ily, rb_z = -L.yz * ily * ilz, rc_z = ilz;
  float mb = (b <= nb / 2) ? (float)b : (float)b - (float)nb;
  double se = 0.0, sw = 0.0;
  for (unsigned e = threadIdx.x; e < na * FFT_CW; e += blockDim.x) {
    unsigned c = e / na, a = e - c * na;
    if (c >= ncol) continue;
    unsigned kcc = c0 + c;
    float ma = (a <= na / 2) ? (float)a : (float)a - (float)na;
    float mc = (kcc <= nc / 2) ? (float)kcc : (float)kcc - (float)nc;
    float kx = two_pi * (ma * ra_x);
    float ky = two_pi * (ma * ra_y + mb * rb_y);
    float kz = two_pi * (ma * ra_z + mb * rb_z + mc * rc_z);
    float k2 = kx * kx + ky * ky + kz * kz;
    float g = 0.f, vf = 0.f;
    if (k2 != 0.f) {
      g = pref * (four_pi / k2) * expf(-k2 / four_alpha2) * bfa[a] * bfb[b] * bfc[kcc];
      vf = g * (1.0f - k2 / (2.0f * alpha * alpha));
    }
    float2 v = d[c * ls + a];
    float rsq = v.x * v.x + v.y * v.y;
    double hw = (kcc == 0 || 2 * kcc == nc) ? 1.0 : 2.0;
    se += hw * (double)g * (double)rsq;
    sw += hw * (double)vf * (double)rsq;
    d[c * ls + a] = make_float2(g * v.x, g * v.y);
  }
  red_e[threadIdx.x] = se;
  red_w[threadIdx.x] = sw;
  __syncthreads();
  for (unsigned s = 1; s < FFT_THREADS; s <<= 1) {
    if ((threadIdx.x % (2 * s)) == 0) {
      red_e[threadIdx.x] += red_e[threadIdx.x + s];
      red_w[threadIdx.x] += red_w[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = make_double2(red_e[0], red_w[0]);
  if (MIXED) {
    lds_fft_mixed(d, na, FFT_CW, ls, +1, stw, na);
    strided_store(d, cgrid, base, na, log2a, estride, ncol, false);
  } else {
    lds_fft_dif(d, na, log2a, FFT_CW, ls, +1, stw, na);
    strided_store(d, cgrid, base, na, log2a, estride, ncol, true);
  }
}

// Fixed-order FP64 reduction of the (E, W) partials; writes
//   recip[0] = E_recip (Ewald sum without self term), recip[1] = W_recip,
//   w_share[0] = (f32) W_recip / N  (the uniform per-atom share,
//   spme_recip.cu:376-403).
__global__ void __launch_bounds__(256) k_recip_reduce(const double2* __restrict__ partials,
                                                      unsigned np, double* __restrict__ recip,
                                                      float* __restrict__ w_share, unsigned n) {
  __shared__ double se[256], sw[256];
  recip_reduce_block(partials, np, recip, w_share, n, se, sw);
}

static unsigned ilog2u(unsigned v) {
  unsigned r = 0;
  while ((1u << r) < v) ++r;
  return r;
}

void hd_launch_spme_spread(hipStream_t st, const HdSpmeArgs& a) {
  unsigned brick[3];
  unsigned nbricks = hd_spread_bricks(a.na, a.nb, a.nc, brick);
  if (a.n > 0 && !a.splines_done)
    hipLaunchKernelGGL(k_spline_setup, dim3((a.n + 255) / 256), dim3(256), 0, st, a.tile_posq,
                       a.lattice, a.na, a.nb, a.nc, a.spl, a.sg, a.n);
  if (a.order == 5)
    hipLaunchKernelGGL(k_spread<5>, dim3(nbricks), dim3(SPREAD_THREADS), 0, st, a.spl, a.sg,
                       a.na, a.nb, a.nc, brick[0], brick[1], brick[2], a.voxel_rank, a.voxel_off,
                       a.nv[0], a.nv[1], a.nv[2], a.mv[0], a.mv[1], a.mv[2],
                       (unsigned long long*)a.rho_fixed);
  else
    hipLaunchKernelGGL(k_spread<4>, dim3(nbricks), dim3(SPREAD_THREADS), 0, st, a.spl, a.sg,
                       a.na, a.nb, a.nc, brick[0], brick[1], brick[2], a.voxel_rank, a.voxel_off,
                       a.nv[0], a.nv[1], a.nv[2], a.mv[0], a.mv[1], a.mv[2],
                       (unsigned long long*)a.rho_fixed);
}

void hd_launch_recip_reduce(hipStream_t st, const HdSpmeArgs& a, unsigned n_partials) {
  hipLaunchKernelGGL(k_recip_reduce, dim3(1), dim3(256), 0, st, a.partials, n_partials, a.recip,
                     a.w_share, a.n);
}

static bool pow2u(unsigned v) { return v && !(v & (v - 1)); }

template <bool MC, bool MB>
static void launch_plane_passes(hipStream_t st, const HdSpmeArgs& a, bool fwd, unsigned log2h,
                                unsigned log2b, unsigned np) {
  if (fwd)
    hipLaunchKernelGGL((k_fft_cb_fwd<MC, MB>), dim3(a.na), dim3(FFT_THREADS), 0, st,
                       a.rho_fixed, a.cgrid, a.nb, a.nc, log2h, log2b, a.tw_c, a.tw_b);
  else
    hipLaunchKernelGGL((k_fft_bc_inv<MB, MC>), dim3(a.na), dim3(FFT_THREADS), 0, st, a.cgrid,
                       a.vgrid, a.nb, a.nc, log2h, log2b, a.tw_c, a.tw_b, a.partials, np,
                       a.recip, a.w_share, a.n);
}

static void plane_passes(hipStream_t st, const HdSpmeArgs& a, bool fwd, bool mc, bool mb,
                         unsigned log2h, unsigned log2b, unsigned np) {
  if (mc && mb) launch_plane_passes<true, true>(st, a, fwd, log2h, log2b, np);
  else if (mc) launch_plane_passes<true, false>(st, a, fwd, log2h, log2b, np);
  else if (mb) launch_plane_passes<false, true>(st, a, fwd, log2h, log2b, np);
  else launch_plane_passes<false, false>(st, a, fwd, log2h, log2b, np);
}

void hd_launch_spme(hipStream_t st, const HdSpmeArgs& a) {
  unsigned nch = a.nc / 2 + 1;
  hd_launch_spme_spread(st, a);
  unsigned nrows = a.na * a.nb;
  unsigned h = a.nc / 2;
  unsigned rows = FFT_TILE / h;
  dim3 gc((nrows + rows - 1) / rows), tb(FFT_THREADS);
  bool mc = !pow2u(h), mb = !pow2u(a.nb), ma = !pow2u(a.na);
  if (!a.no_plane && fft_plane_fits(a.nb, a.nc)) {
    // c and b passes per a-plane in one launch each way
    unsigned nchunk = (nch + FFT_CW - 1) / FFT_CW;
    unsigned lh = mc ? 0u : ilog2u(h), lb = mb ? 0u : ilog2u(a.nb);
    plane_passes(st, a, true, mc, mb, lh, lb, 0);
    dim3 ga(a.nb * nchunk);
    if (ma)
      hipLaunchKernelGGL(k_fft_a_conv<true>, ga, tb, 0, st, a.cgrid, a.na, a.nb, a.nc, nch, 0u,
                         a.tw_a, a.bf_a, a.bf_b, a.bf_c, a.lattice, (float)a.kc, (float)a.alpha,
                         a.partials);
    else
      hipLaunchKernelGGL(k_fft_a_conv<false>, ga, tb, 0, st, a.cgrid, a.na, a.nb, a.nc, nch,
                         ilog2u(a.na), a.tw_a, a.bf_a, a.bf_b, a.bf_c, a.lattice, (float)a.kc,
                         (float)a.alpha, a.partials);
    plane_passes(st, a, false, mc, mb, lh, lb, a.nb * nchunk);
    return;
  }
  if (mc)
    hipLaunchKernelGGL(k_fft_c_fwd<true>, gc, tb, 0, st, a.rho_fixed, a.cgrid, nrows, a.nc, 0u,
                       a.tw_c);
  else
    hipLaunchKernelGGL(k_fft_c_fwd<false>, gc, tb, 0, st, a.rho_fixed, a.cgrid, nrows, a.nc,
                       ilog2u(h), a.tw_c);
  unsigned nchunk = (nch + FFT_CW - 1) / FFT_CW;
  dim3 gb(a.na * nchunk), ga(a.nb * nchunk);
  for (int sign = -1; sign <= 1; sign += 2) {
    if (sign > 0) {
      if (ma)
        hipLaunchKernelGGL(k_fft_a_conv<true>, ga, tb, 0, st, a.cgrid, a.na, a.nb, a.nc, nch, 0u,
                           a.tw_a, a.bf_a, a.bf_b, a.bf_c, a.lattice, (float)a.kc,
                           (float)a.alpha, a.partials);
      else
        hipLaunchKernelGGL(k_fft_a_conv<false>, ga, tb, 0, st, a.cgrid, a.na, a.nb, a.nc, nch,
                           ilog2u(a.na), a.tw_a, a.bf_a, a.bf_b, a.bf_c, a.lattice, (float)a.kc,
                           (float)a.alpha, a.partials);
    }
    if (mb)
      hipLaunchKernelGGL(k_fft_b<true>, gb, tb, 0, st, a.cgrid, a.na, a.nb, nch, 0u, sign,
                         a.tw_b);
    else
      hipLaunchKernelGGL(k_fft_b<false>, gb, tb, 0, st, a.cgrid, a.na, a.nb, nch, ilog2u(a.nb),
                         sign, a.tw_b);
  }
  // the inverse c pass also reduces the energy and virial partials
  if (mc)
    hipLaunchKernelGGL(k_fft_c_inv<true>, gc, tb, 0, st, a.cgrid, a.vgrid, nrows, a.nc, 0u,
                       a.tw_c, a.partials, a.nb * nchunk, a.recip, a.w_share, a.n);
  else
    hipLaunchKernelGGL(k_fft_c_inv<false>, gc, tb, 0, st, a.cgrid, a.vgrid, nrows, a.nc,
                       ilog2u(h), a.tw_c, a.partials, a.nb * nchunk, a.recip, a.w_share, a.n);
}

unsigned hd_spme_partials_count(unsigned nb, unsigned nc) {
  unsigned nch = nc / 2 + 1;
  return nb * ((nch + FFT_CW - 1) / FFT_CW);
}
