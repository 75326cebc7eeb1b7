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

#define SPME_P 4
#define FFT_THREADS 256
#define FFT_TILE 2048  // complex elements staged per workgroup

// Cardinal B-spline by Cox-de Boor at order SPME_P (spme_spread_gather.cuh:36-54).
__device__ __forceinline__ float bspline_w(float x) {
  float v[SPME_P];
#pragma unroll
  for (int i = 0; i < SPME_P; ++i) {
    float xi = x - (float)i;
    v[i] = (xi >= 0.f && xi < 1.f) ? 1.f : 0.f;
  }
#pragma unroll
  for (int order = 2; order <= SPME_P; ++order) {
    float inv = 1.0f / (float)(order - 1);
#pragma unroll
    for (int i = 0; i < SPME_P - order + 1; ++i) {
      float xi = x - (float)i;
      v[i] = xi * inv * v[i] + ((float)order - xi) * inv * v[i + 1];
    }
  }
  return v[0];
}

// Spline setup (spme_spread_gather.cuh:90-126): wrap, fractional + 1/2,
// u = s*n, primary bin floor(u), fractional offset t.
__device__ __forceinline__ void spline_setup(float px, float py, float pz, const HdLattice& L,
                                             unsigned na, unsigned nb, unsigned nc, float& ta,
                                             float& tb, float& tc, int& ga, int& gb, int& gc) {
  int ia, ib, ic;
  hd_wrap_with_image(px, py, pz, ia, ib, ic, L);
  float sa, sb, sc;
  hd_cart_to_frac(px, py, pz, L, sa, sb, sc);
  float ua = (sa + 0.5f) * (float)na, ub = (sb + 0.5f) * (float)nb, uc = (sc + 0.5f) * (float)nc;
  float fa = floorf(ua), fb = floorf(ub), fc = floorf(uc);
  ta = ua - fa; tb = ub - fb; tc = uc - fc;
  int la = (int)fa, lb = (int)fb, lc = (int)fc;
  ga = ((la % (int)na) + (int)na) % (int)na;
  gb = ((lb % (int)nb) + (int)nb) % (int)nb;
  gc = ((lc % (int)nc) + (int)nc) % (int)nc;
}

// Spread: SPME_P threads per atom (slot order, so neighbouring threads
// touch neighbouring cells); thread iz owns one c-slice and adds its
// P*P cells with i64 atomics.  Integer addition is associative, so the
// grid is bit-identical run to run.
__global__ void k_spread(const float4* __restrict__ tile_posq, const float* __restrict__ lattice,
                         unsigned na, unsigned nb, unsigned nc,
                         unsigned long long* __restrict__ rho, unsigned n) {
  unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * SPME_P) return;
  unsigned s = gid / SPME_P, iz = gid - s * SPME_P;
  float4 pq = tile_posq[s];
  float qi = pq.w;
  if (qi == 0.f) return;
  HdLattice L = hd_load_lattice(lattice);
  float ta, tb, tc;
  int ga, gb, gc;
  hd_spline_origin(pq.x, pq.y, pq.z, L, na, nb, nc, ta, tb, tc, ga, gb, gc);
  float wa[SPME_P], wb[SPME_P];
#pragma unroll
  for (int d = 0; d < SPME_P; ++d) {
    wa[d] = bspline_w((float)d + ta);
    wb[d] = bspline_w((float)d + tb);
  }
  float wc = bspline_w((float)iz + tc);
  unsigned gcz = ((unsigned)gc + nc - iz) % nc;
  float dz = qi * wc;
#pragma unroll
  for (int da = 0; da < SPME_P; ++da) {
    unsigned gaa = ((unsigned)ga + na - da) % na;
    float dzda = dz * wa[da];
#pragma unroll
    for (int db = 0; db < SPME_P; ++db) {
      float v = dzda * wb[db];
      long long vf = __float2ll_rn(v * 4294967296.0f);
      if (vf != 0) {
        unsigned gbb = ((unsigned)gb + nb - db) % nb;
        atomicAdd(rho + ((size_t)gaa * nb + gbb) * nc + gcz, (unsigned long long)vf);
      }
    }
  }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = 0ull;
}

// ---------------------------------------------------------------------
// Radix-2 complex FFT over `nlines` lines of length n held contiguously in
// LDS, input already in bit-reversed order.  tw[k] = exp(-2 pi i k / twn),
// k < twn/2; sign = -1 forward, +1 inverse (unnormalised).
// ---------------------------------------------------------------------
__device__ __forceinline__ void lds_fft(float2* d, unsigned n, unsigned log2n, unsigned nlines,
                                        unsigned lstride, int sign,
                                        const float2* __restrict__ tw, unsigned twn) {
  unsigned hn = n >> 1;
  unsigned nbf = nlines * hn;
  for (unsigned s = 1; s <= log2n; ++s) {
    unsigned half = 1u << (s - 1);
    unsigned tstr = twn >> s;
    for (unsigned idx = threadIdx.x; idx < nbf; idx += blockDim.x) {
      unsigned line = idx >> (log2n - 1);
      unsigned k = idx & (hn - 1);
      unsigned j = k & (half - 1);
      unsigned i0 = ((k >> (s - 1)) << s) + j;
      unsigned i1 = i0 + half;
      float2 w = tw[j * tstr];
      if (sign > 0) w.y = -w.y;
      float2* row = d + line * lstride;
      float2 a = row[i0], b = row[i1];
      float2 bw = make_float2(b.x * w.x - b.y * w.y, b.x * w.y + b.y * w.x);
      row[i0] = make_float2(a.x + bw.x, a.y + bw.y);
      row[i1] = make_float2(a.x - bw.x, a.y - bw.y);
    }
    __syncthreads();
  }
}

// Decimation-in-frequency twin: natural-order input, bit-reversed output.
__device__ __forceinline__ void lds_fft_dif(float2* d, unsigned n, unsigned log2n,
                                            unsigned nlines, unsigned lstride, int sign,
                                            const float2* __restrict__ tw, unsigned twn) {
  unsigned hn = n >> 1;
  unsigned nbf = nlines * hn;
  for (unsigned s = log2n; s >= 1; --s) {
    unsigned half = 1u << (s - 1);
    unsigned tstr = twn >> s;
    for (unsigned idx = threadIdx.x; idx < nbf; idx += blockDim.x) {
      unsigned line = idx >> (log2n - 1);
      unsigned k = idx & (hn - 1);
      unsigned j = k & (half - 1);
      unsigned i0 = ((k >> (s - 1)) << s) + j;
      unsigned i1 = i0 + half;
      float2 w = tw[j * tstr];
      if (sign > 0) w.y = -w.y;
      float2* row = d + line * lstride;
      float2 a = row[i0], b = row[i1];
      float2 dlt = make_float2(a.x - b.x, a.y - b.y);
      row[i0] = make_float2(a.x + b.x, a.y + b.y);
      row[i1] = make_float2(dlt.x * w.x - dlt.y * w.y, dlt.x * w.y + dlt.y * w.x);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ unsigned bitrev(unsigned k, unsigned log2n) {
  return log2n ? (__brev(k) >> (32 - log2n)) : 0u;
}

// c-axis forward R2C: rows of nc reals read from the i64 spread grid
// (converted to f32 exactly as spme_recip.cu:237-253), packed as nc/2
// complex values, FFT'd, then untangled into nc/2+1 spectrum values.
// twc: twiddles for length nc (exp(-2 pi i k/nc), k < nc/2).
__global__ void __launch_bounds__(FFT_THREADS) k_fft_c_fwd(
    const long long* __restrict__ rho, float2* __restrict__ cgrid, unsigned nrows, unsigned nc,
    unsigned log2h, const float2* __restrict__ twc) {
  __shared__ float2 d[FFT_TILE];
  unsigned h = nc >> 1, nch = h + 1;
  unsigned rows = FFT_TILE / h;
  unsigned r0 = blockIdx.x * rows;
  unsigned nr = min(rows, nrows - r0);
  const float inv = 1.0f / 4294967296.0f;
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, m = e - r * h;
    longlong2 v = reinterpret_cast<const longlong2*>(rho + (size_t)(r0 + r) * nc)[m];
    d[r * h + bitrev(m, log2h)] = make_float2((float)v.x * inv, (float)v.y * inv);
  }
  __syncthreads();
  lds_fft(d, h, log2h, nr, h, -1, twc, nc);
  for (unsigned e = threadIdx.x; e < nr * nch; e += blockDim.x) {
    unsigned r = e / nch, k = e - r * nch;
    float2 zk = d[r * h + (k % h)];
    float2 zc = d[r * h + ((h - k) % h)];
    zc.y = -zc.y;  // conj(Z[h-k])
    float2 E = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    // O = (Z[k] - conj(Z[h-k])) / (2i)
    float2 O = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
    float2 w = (k < h) ? twc[k] : make_float2(-1.f, 0.f);  // exp(-2 pi i k / nc)
    float2 wO = make_float2(w.x * O.x - w.y * O.y, w.x * O.y + w.y * O.x);
    cgrid[(size_t)(r0 + r) * nch + k] = make_float2(E.x + wO.x, E.y + wO.y);
  }
}

// c-axis inverse C2R: nc/2+1 spectrum values per row -> nc reals,
// unnormalised (round trip = nc * x).
__global__ void __launch_bounds__(FFT_THREADS) k_fft_c_inv(
    const float2* __restrict__ cgrid, float* __restrict__ vgrid, unsigned nrows, unsigned nc,
    unsigned log2h, const float2* __restrict__ twc) {
  __shared__ float2 d[FFT_TILE];
  unsigned h = nc >> 1, nch = h + 1;
  unsigned rows = FFT_TILE / h;
  unsigned r0 = blockIdx.x * rows;
  unsigned nr = min(rows, nrows - r0);
  for (unsigned e = threadIdx.x; e < nr * h; e += blockDim.x) {
    unsigned r = e / h, k = e - r * h;
    const float2* row = cgrid + (size_t)(r0 + r) * nch;
    float2 xk = row[k], xc = row[h - k];
    xc.y = -xc.y;  // conj(X[h-k])
    float2 A = make_float2(xk.x + xc.x, xk.y + xc.y);
    float2 B = make_float2(xk.x - xc.x, xk.y - xc.y);
    float2 w = twc[k];
    w.y = -w.y;  // exp(+2 pi i k / nc)
    float2 wB = make_float2(w.x * B.x - w.y * B.y, w.x * B.y + w.y * B.x);
    // Z' = A + i * wB
    d[r * h + bitrev(k, log2h)] = make_float2(A.x - wB.y, A.y + wB.x);
  }
  __syncthreads();
  lds_fft(d, h, log2h, nr, h, +1, twc, nc);
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
__global__ void __launch_bounds__(FFT_THREADS) k_fft_b(float2* __restrict__ cgrid, unsigned na,
                                                       unsigned nb, unsigned nch,
                                                       unsigned log2b, int sign,
                                                       const float2* __restrict__ twb) {
  __shared__ float2 d[FFT_CW * 257];
  unsigned nchunk = (nch + FFT_CW - 1) / FFT_CW;
  unsigned a = blockIdx.x / nchunk, c0 = (blockIdx.x - a * nchunk) * FFT_CW;
  unsigned ncol = min(FFT_CW, nch - c0);
  size_t base = (size_t)a * nb * nch + c0;
  strided_load(d, cgrid, base, nb, log2b, nch, ncol, true);
  __syncthreads();
  lds_fft(d, nb, log2b, FFT_CW, nb + 1, sign, twb, nb);
  strided_store(d, cgrid, base, nb, log2b, nch, ncol, false);
}

// a-axis pass fused with the influence multiply:
//   forward FFT over a -> for every k: G(k), virial factor vf(k)
//   (spme_recip.cu:34-115 in f32 from the live lattice), energy and
//   virial partials hw*G*|rho^|^2 and hw*vf*|rho^|^2 (spme_recip.cu:
//   296-342) summed in FP64 by a fixed tree -> V^ = G rho^ -> inverse FFT
//   over a.  grid: (nb * nchunks); partials[blockIdx] = (E, W).
__global__ void __launch_bounds__(FFT_THREADS) k_fft_a_conv(
    float2* __restrict__ cgrid, unsigned na, unsigned nb, unsigned nc, unsigned nch,
    unsigned log2a, const float2* __restrict__ twa, const float* __restrict__ bfa,
    const float* __restrict__ bfb, const float* __restrict__ bfc,
    const float* __restrict__ lattice, float kc, float alpha, double2* __restrict__ partials) {
  __shared__ float2 d[FFT_CW * 257];
  __shared__ double red_e[FFT_THREADS], red_w[FFT_THREADS];
  unsigned nchunk = (nch + FFT_CW - 1) / FFT_CW;
  unsigned b = blockIdx.x / nchunk, c0 = (blockIdx.x - b * nchunk) * FFT_CW;
  unsigned ncol = min(FFT_CW, nch - c0);
  size_t base = (size_t)b * nch + c0;
  size_t estride = (size_t)nb * nch;
  unsigned ls = na + 1;
  strided_load(d, cgrid, base, na, log2a, estride, ncol, true);
  __syncthreads();
  lds_fft(d, na, log2a, FFT_CW, ls, -1, twa, na);
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
  lds_fft_dif(d, na, log2a, FFT_CW, ls, +1, twa, na);
  strided_store(d, cgrid, base, na, log2a, estride, ncol, true);
}

// Fixed-order FP64 reduction of the (E, W) partials; writes
//   recip[0] = E_recip (Ewald sum without self term), recip[1] = W_recip,
//   w_share[0] = (f32) W_recip / N  (the uniform per-atom share,
//   spme_recip.cu:376-403).
__global__ void __launch_bounds__(256) k_recip_reduce(const double2* __restrict__ partials,
                                                      unsigned np, double* __restrict__ recip,
                                                      float* __restrict__ w_share, unsigned n) {
  __shared__ double se[256], sw[256];
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

static unsigned ilog2u(unsigned v) {
  unsigned r = 0;
  while ((1u << r) < v) ++r;
  return r;
}

void hd_launch_spme(hipStream_t st, const HdSpmeArgs& a) {
  size_t M = (size_t)a.na * a.nb * a.nc;
  unsigned nch = a.nc / 2 + 1;
  hipLaunchKernelGGL(k_zero_u64, dim3(1024), dim3(256), 0, st, (unsigned long long*)a.rho_fixed, M);
  if (a.n > 0)
    hipLaunchKernelGGL(k_spread, dim3((a.n * SPME_P + 255) / 256), dim3(256), 0, st, a.tile_posq,
                       a.lattice, a.na, a.nb, a.nc, (unsigned long long*)a.rho_fixed, a.n);
  unsigned nrows = a.na * a.nb;
  unsigned h = a.nc / 2;
  unsigned rows = FFT_TILE / h;
  hipLaunchKernelGGL(k_fft_c_fwd, dim3((nrows + rows - 1) / rows), dim3(FFT_THREADS), 0, st,
                     a.rho_fixed, a.cgrid, nrows, a.nc, ilog2u(h), a.tw_c);
  unsigned nchunk = (nch + FFT_CW - 1) / FFT_CW;
  hipLaunchKernelGGL(k_fft_b, dim3(a.na * nchunk), dim3(FFT_THREADS), 0, st, a.cgrid, a.na, a.nb,
                     nch, ilog2u(a.nb), -1, a.tw_b);
  hipLaunchKernelGGL(k_fft_a_conv, dim3(a.nb * nchunk), dim3(FFT_THREADS), 0, st, a.cgrid, a.na,
                     a.nb, a.nc, nch, ilog2u(a.na), a.tw_a, a.bf_a, a.bf_b, a.bf_c, a.lattice,
                     (float)a.kc, (float)a.alpha, a.partials);
  hipLaunchKernelGGL(k_fft_b, dim3(a.na * nchunk), dim3(FFT_THREADS), 0, st, a.cgrid, a.na, a.nb,
                     nch, ilog2u(a.nb), +1, a.tw_b);
  hipLaunchKernelGGL(k_fft_c_inv, dim3((nrows + rows - 1) / rows), dim3(FFT_THREADS), 0, st,
                     a.cgrid, a.vgrid, nrows, a.nc, ilog2u(h), a.tw_c);
  hipLaunchKernelGGL(k_recip_reduce, dim3(1), dim3(256), 0, st, a.partials, a.nb * nchunk,
                     a.recip, a.w_share, a.n);
}

unsigned hd_spme_partials_count(unsigned nb, unsigned nc) {
  unsigned nch = nc / 2 + 1;
  return nb * ((nch + FFT_CW - 1) / FFT_CW);
}
