// Fused Lennard-Jones + real-space Ewald pair forces over the packed
// 64-atom cluster list (gfx950, wave64).
//
// Reference: the JIT-composed packed pair kernel
// (src/forces/jit_composed.rs:1490-1838) with the LJ functor
// (src/forces/lj.rs:225-335) and the SPME real-space functor
// (src/forces/spme.rs:708-783).  Kept: per-pair arithmetic (f32, Hastings
// erfc A&S 7.1.26, optional cubic LJ switch), per-entry f32 sums folded
// into i64 fixed point at 2^48 with truncation, Newton's third law with the
// j-side suppressed for j-atoms that belong to the i-block, per-atom energy
// and virial split half/half.
//
// MI355X design:
//   * one workgroup of 4 waves per 64-atom i-block; lane = i-atom; the 64
//     j-atoms of an entry rotate through the wave with DPP wave_rol:1
//     (a VALU modifier, no LDS round trip) in 64 lock-step steps;
//   * i-side sums stay in registers across entries, are combined across the
//     4 waves in LDS and stored once with a plain store (no atomics, no
//     clearing pass); only the j-side uses i64 atomics;
//   * the LJ type-pair row of the i-atom is held in registers (<= 4 types)
//     and selected per pair, so the inner loop touches no memory;
//   * exclusions by key: a pair is excluded when the atoms' keys differ by
//     at most a span.  Span 0 with one key per molecule excludes every
//     intra-molecule pair of a rigid group, as the reference derives them
//     (src/forces/topology.rs:1225-1245); keys chain * stride + bead index
//     with span 2 exclude the 1-2 and 1-3 pairs of linear chains.  The
//     atom's self pair is excluded the same way;
//   * when the box is orthorhombic and wide enough, every j-atom is shifted
//     once per entry to the image nearest the i-block centre, so the inner
//     loop needs no minimum-image arithmetic (the single-periodic-copy path,
//     jit_composed.rs:1565-1600); both i and j are then held relative to
//     the block centre (shift computed in FP64, rounded once), which keeps
//     the f32 pair displacement accurate in boxes of 100+ Bohr.
#include "hd_device.h"
#include "hd_kernels.h"

#define PAIR_WAVES 4
#define PAIR_MAXT 4

struct PairParams {
  float c6[PAIR_MAXT * PAIR_MAXT];   // 4 eps sigma^6
  float c12[PAIR_MAXT * PAIR_MAXT];  // 4 eps sigma^12
  float ljcut2[PAIR_MAXT * PAIR_MAXT];
  float rsw2[PAIR_MAXT * PAIR_MAXT];
  float inv_delta[PAIR_MAXT * PAIR_MAXT];  // 1/(rc^2 - rs^2), 0 when no switch
  int ntypes;
  float kc, alpha, coul_rc2;  // coul_rc2 < 0 disables the Coulomb term
  float rmax2;
  float spc_margin;  // minimum (L/2 - half_extent) that enables the shifted path
  unsigned excl_span;  // |key_i - key_j| <= excl_span excludes the pair
  int ortho;           // the box has no tilt
  // 1: one cutoff for every interacting type pair and Coulomb, no switch;
  // 2: LJ only (no Coulomb), one cutoff for every interacting type pair, no
  // switch; 0: neither (the general loop)
  int uniform;
  float uc_rc2;  // the shared cutoff^2 of modes 1 and 2
  float cg;  // -alpha^2 log2(e): exp(-(alpha r)^2) = exp2(cg r^2)
  float pa;  // 0.3275911 alpha, the Hastings t = 1 / (1 + pa r)
};

// 32768-bit map of the i-block's keys widened by the exclusion span, used
// to skip the per-pair key test in entries with no excludable j-atom; with
// a wider span every entry runs the test
#define PAIR_KEYMAP_WORDS 1024
#define PAIR_KEYMAP_SPAN 8

// One wave's copy of the current entry's j-atoms for the regrouped loop,
// written twice over so that step t reads record lane + t with no wrap:
// positions (block-centre relative) and charges; the LJ force coefficients
// (12 c12, 6 c6) of the j-atom against each i type, so a lane reads its own
// type's row with no per-pair select; types (energy steps) and keys.
template <int NT>
struct PairJStage {
  float4 pq[128];
  float2 lj[NT > 0 ? NT : 1][128];
  unsigned type[128];
  unsigned key[128];
};

// Orders one wave's LDS writes before its lanes read each other's records.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The i-atom's row of the LJ type-pair table, held in VGPRs for the whole
// kernel.  NT = table width (1, 2 or 4).  Each value passes through an
// empty asm that pins it to a register: otherwise the compiler folds the
// per-pair select chain back into an indexed load from a private array,
// which lands in scratch memory inside the inner loop.
template <int NT>
struct LjRow {
  float c6[NT], c12[NT], f6[NT], f12[NT], cut2[NT], rsw2[NT], idl[NT];
};

// NT = 0: more than four types.  The type-pair table sits in LDS (copied
// in by every workgroup) and a lane keeps pointers to its i-type's row:
// a[t] = (12 c12, 6 c6, c12, c6), b[t] = (cut^2, r_switch^2, 1/(cut^2 -
// r_switch^2), 0) for j-type t.  One LDS read per pair replaces the
// register select chain.
#define PAIR_TAB_TYPES 32
template <>
struct LjRow<0> {
  const float4* a;
  const float4* b;
  bool valid;
};

template <int NT>
struct PairTab {
  float4 a[1], b[1];
};
template <>
struct PairTab<0> {
  float4 a[PAIR_TAB_TYPES * PAIR_TAB_TYPES], b[PAIR_TAB_TYPES * PAIR_TAB_TYPES];
};

template <int NT>
__device__ __forceinline__ float pick(const float (&v)[NT], unsigned t) {
  if (NT == 1) return v[0];
  if (NT == 2) return t ? v[1] : v[0];
  float r = v[0];
  r = (t == 1u) ? v[1] : r;
  r = (t == 2u) ? v[2] : r;
  r = (t == 3u) ? v[3] : r;
  return r;
}

// Per-pair LJ coefficients of the i-row against j-type t.
template <int NT>
__device__ __forceinline__ float lj_f12(const LjRow<NT>& R, unsigned t) {
  if constexpr (NT == 0) return R.a[t].x; else return pick<NT>(R.f12, t);
}
template <int NT>
__device__ __forceinline__ float lj_f6(const LjRow<NT>& R, unsigned t) {
  if constexpr (NT == 0) return R.a[t].y; else return pick<NT>(R.f6, t);
}
template <int NT>
__device__ __forceinline__ float lj_c12(const LjRow<NT>& R, unsigned t) {
  if constexpr (NT == 0) return R.a[t].z; else return pick<NT>(R.c12, t);
}
template <int NT>
__device__ __forceinline__ float lj_c6(const LjRow<NT>& R, unsigned t) {
  if constexpr (NT == 0) return R.a[t].w; else return pick<NT>(R.c6, t);
}
template <int NT>
__device__ __forceinline__ float lj_cut2(const LjRow<NT>& R, unsigned t) {
  if constexpr (NT == 0) return R.valid ? R.b[t].x : -1.0f; else return pick<NT>(R.cut2, t);
}
template <int NT>
__device__ __forceinline__ float lj_rsw2(const LjRow<NT>& R, unsigned t) {
  if constexpr (NT == 0) return R.b[t].y; else return pick<NT>(R.rsw2, t);
}
template <int NT>
__device__ __forceinline__ float lj_idl(const LjRow<NT>& R, unsigned t) {
  if constexpr (NT == 0) return R.b[t].z; else return pick<NT>(R.idl, t);
}

__device__ __forceinline__ void pin(float& x) { __asm__("" : "+v"(x)); }

// Copy the type-pair table (2 float4 per pair, nt <= PAIR_TAB_TYPES) into
// LDS; the caller synchronises before the first read.
__device__ __forceinline__ void load_tab(PairTab<0>& T, const float4* __restrict__ lj_tab,
                                         int nt) {
  for (unsigned k = threadIdx.x; k < (unsigned)(nt * nt); k += blockDim.x) {
    T.a[k] = lj_tab[2 * k];
    T.b[k] = lj_tab[2 * k + 1];
  }
}

template <int NT>
__device__ __forceinline__ LjRow<NT> load_row(const PairParams& P, unsigned itype, bool ivalid,
                                              const PairTab<NT>& T) {
  LjRow<NT> r;
  if constexpr (NT == 0) {
    r.a = &T.a[itype * (unsigned)P.ntypes];
    r.b = &T.b[itype * (unsigned)P.ntypes];
    r.valid = ivalid;
    return r;
  } else {
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int k = (t < P.ntypes) ? (int)itype * P.ntypes + t : 0;
    r.c6[t] = P.c6[k];
    r.c12[t] = P.c12[k];
    r.f6[t] = 6.0f * P.c6[k];
    r.f12[t] = 12.0f * P.c12[k];
    r.cut2[t] = ivalid ? P.ljcut2[k] : -1.0f;
    r.rsw2[t] = P.rsw2[k];
    r.idl[t] = P.inv_delta[k];
    pin(r.c6[t]); pin(r.c12[t]); pin(r.f6[t]); pin(r.f12[t]);
    pin(r.cut2[t]); pin(r.rsw2[t]); pin(r.idl[t]);
  }
  return r;
  }
}

// The j-atom's offset from the block centre in the image nearest it, as
// hd_rel_to_center computes it (FP64 difference less a whole number of box
// lengths, rounded once), with the image count taken from an f32 product
// with the reciprocal box length instead of an FP64 division.  On the
// shifted path every staged j-atom lies within the search radius plus the
// block half-extent of the centre, at least the skin away from half a box
// length, so the f32 count is the FP64 one.
__device__ __forceinline__ void pair_rel_to_center(float x, float y, float z, float4 cb,
                                                   const HdLattice& L, float ilx, float ily,
                                                   float ilz, float& rx, float& ry, float& rz) {
  double nx = (double)rintf((x - cb.x) * ilx);
  double ny = (double)rintf((y - cb.y) * ily);
  double nz = (double)rintf((z - cb.z) * ilz);
  rx = (float)(((double)x - (double)cb.x) - (double)L.lx * nx);
  ry = (float)(((double)y - (double)cb.y) - (double)L.ly * ny);
  rz = (float)(((double)z - (double)cb.z) - (double)L.lz * nz);
}

// One axis of a minimum-image pair displacement in an orthorhombic box.
// Within the box the f32 difference of the two coordinates is exact or
// rounds at the pair's own scale; across a face it would round at the
// box's scale (an error of up to half an ulp of L, ~1e-6 of a contact
// distance, which the r^-13 of the LJ repulsion multiplies), so such pairs
// take the difference in FP64 and round once.
__device__ __forceinline__ float pair_axis_disp(float a, float b, float len, float ilen) {
  float d = a - b;
  float k = rintf(d * ilen);
  if (k != 0.f) d = (float)(((double)a - (double)b) - (double)len * (double)k);
  return d;
}

// The same for a tilted cell: the image triple from the f32 fractional
// rounding of hd_min_image_rcp, and, when it is not (0, 0, 0), the
// difference and the lattice shift formed in FP64 and rounded once.
__device__ __forceinline__ void pair_tri_disp(float pix, float piy, float piz, float pjx,
                                              float pjy, float pjz, const HdLattice& L,
                                              float ilx, float ily, float ilz, float& dx,
                                              float& dy, float& dz) {
  dx = pix - pjx;
  dy = piy - pjy;
  dz = piz - pjz;
  float sc = dz * ilz;
  float sb = (dy - sc * L.yz) * ily;
  float sa = (dx - sb * L.xy - sc * L.xz) * ilx;
  float ka = floorf(sa + 0.5f), kb = floorf(sb + 0.5f), kc = floorf(sc + 0.5f);
  if (ka == 0.f && kb == 0.f && kc == 0.f) return;
  double a = ka, b = kb, c = kc;
  dx = (float)(((double)pix - (double)pjx) - (a * L.lx + b * L.xy + c * L.xz));
  dy = (float)(((double)piy - (double)pjy) - (b * L.ly + c * L.yz));
  dz = (float)(((double)piz - (double)pjz) - c * L.lz);
}

// Far-away coordinates for padding lanes on the shifted path: any pair
// with one of them fails both cutoffs, and |d|^2 stays finite (< 1e37).
#define PAIR_FAR_I 1.0e18f
#define PAIR_FAR_J -1.0e18f

// The 64 steps of one entry on the shifted path when one cutoff serves LJ
// and Coulomb and there is no switch.  Lane l meets j-atom (l + t) mod 64 at
// step t, read from the wave's LDS stage (jb, jt, jk point at record l); the
// j-side sums still rotate with DPP, so they meet the same j-atoms and are
// back in the loading lane after 64 steps.  Same terms as the general loop,
// regrouped for fewer instructions per pair: exp(-(alpha r)^2) as
// exp2(cg r^2); the Ewald force factor erfc(alpha r)/r^3 +
// 2 alpha/sqrt(pi) exp(-(alpha r)^2)/r^2 as gauss * (poly(t) t / r +
// 2 alpha/sqrt(pi)) / r^2; the LJ and Coulomb factors summed before the
// shared 1/r^2 and a single cutoff select.  The force arithmetic is the same
// for every EV, so energy steps move atoms exactly as force-only steps do.
// XK = false drops the key test (no j-atom of the entry can be excluded
// against the block).
template <int EV, int NT, bool XK, bool COUL>
__device__ __forceinline__ void uc_steps(const LjRow<NT>& R, float pix, float piy, float piz,
                                         unsigned ikey, unsigned span, float kqi, float rc2,
                                         float cg, float pa, float k2, const float4* jb,
                                         const float2* jl, const unsigned* jt,
                                         const unsigned* jk, float& efx,
                                         float& efy, float& efz, float& jfx, float& jfy,
                                         float& jfz, float& ee, float& je) {
#pragma unroll 8
  for (int t = 0; t < 64; ++t) {
    float4 pq = jb[t];
    float2 ljc;
    if constexpr (NT == 0) {
      float4 q = R.a[jt[t]];
      ljc = make_float2(q.x, q.y);
    } else {
      ljc = (NT > 1) ? jl[t] : make_float2(R.f12[0], R.f6[0]);
    }
    float qj = pq.w;
    float dx = pix - pq.x, dy = piy - pq.y, dz = piz - pq.z;
    float r2 = dx * dx + dy * dy + dz * dz;
    // LJ only: 1/r^2 from one reciprocal (half the rounding of squaring a
    // reciprocal square root, and one instruction fewer)
    float inv_r = COUL ? __builtin_amdgcn_rsqf(r2) : 0.f;
    float inv_r2 = COUL ? inv_r * inv_r : __builtin_amdgcn_rcpf(r2);
    float ir6 = inv_r2 * inv_r2 * inv_r2;
    float lj = ir6 * (ljc.x * ir6 - ljc.y);
    float qq = 0.f, gauss = 0.f, poly = 0.f, factor;
    if (COUL) {
      float r = r2 * inv_r;
      gauss = __builtin_amdgcn_exp2f(cg * r2);
      float tt = __builtin_amdgcn_rcpf(fmaf(pa, r, 1.0f));
      poly =
          (0.254829592f +
           (-0.284496736f + (1.421413741f + (-1.453152027f + 1.061405429f * tt) * tt) * tt) * tt) *
          tt;
      float ci = gauss * fmaf(poly, inv_r, k2);
      qq = kqi * qj;
      factor = inv_r2 * fmaf(qq, ci, lj);
    } else {
      factor = inv_r2 * lj;
    }
    bool on = r2 <= rc2;
    if (XK) on = on & ((jk[t] - ikey + span) > 2u * span);
    factor = on ? factor : 0.f;
    float nfactor = -factor;
    efx = fmaf(factor, dx, efx); efy = fmaf(factor, dy, efy); efz = fmaf(factor, dz, efz);
    jfx = fmaf(nfactor, dx, jfx); jfy = fmaf(nfactor, dy, jfy); jfz = fmaf(nfactor, dz, jfz);
    if (EV == 1) {
      // LJ energy (lj.rs:237-247) and qq erfc(alpha r) / r (spme.rs:744-770)
      unsigned jtype = (NT != 1) ? jt[t] : 0u;
      float elj = ir6 * (lj_c12(R, jtype) * ir6 - lj_c6(R, jtype));
      float energy = on ? (COUL ? fmaf(qq * gauss, poly * inv_r, elj) : elj) : 0.f;
      ee = fmaf(0.5f, energy, ee);
      je = fmaf(0.5f, energy, je);
    }
    jfx = hd_rot1(jfx); jfy = hd_rot1(jfy); jfz = hd_rot1(jfz);
    if (EV == 1) je = hd_rot1(je);
  }
}

// One wave's share of the entries of block b.  Lane = i-atom; the 64
// j-atoms of an entry rotate through the wave with DPP wave_rol:1 and the
// j-side partial sums travel with them, arriving back in the loading lane
// after 64 steps.  SPC: coordinates are relative to the block centre (no
// per-pair minimum image); otherwise general-cell minimum image per pair.
template <int EV, int NT, bool SW, bool SPC, bool UC>
__device__ __forceinline__ void pair_entries(
    const PairParams& P, const HdLattice& L, unsigned b, unsigned w, unsigned lane, float pix,
    float piy, float piz, float qi, unsigned ikey, const LjRow<NT>& R, float4 cb,
    const float4* __restrict__ tile_posq, const unsigned* __restrict__ tile_type,
    const unsigned* __restrict__ tile_key, const unsigned* __restrict__ ents, unsigned nent,
    unsigned n, unsigned long long* __restrict__ fpj, unsigned long long* __restrict__ fpj_ew,
    long long& ai_fx, long long& ai_fy, long long& ai_fz, long long& ai_e, long long& ai_w,
    const unsigned* keymap, PairJStage<NT>* stage, unsigned itype) {
  const float two_a_sqrtpi = 2.0f * P.alpha * 0.5641895835477563f;
  const float alpha = P.alpha, coul_rc2 = P.coul_rc2;
  const float kqi = P.kc * qi;
  const float ilx = 1.0f / L.lx, ily = 1.0f / L.ly, ilz = 1.0f / L.lz;
  for (unsigned e = w; e < nent; e += PAIR_WAVES) {
    unsigned jslot = ents[(size_t)e * 64 + lane];
    bool jvalid = jslot < n;
    unsigned jsafe = jvalid ? jslot : 0u;
    float4 pj = tile_posq[jsafe];
    unsigned jtype = (NT != 1) ? tile_type[jsafe] : 0u;
    unsigned jkey = tile_key[jsafe];
    bool jin = jvalid && (jslot >> 6) == b;
    float pjx = pj.x, pjy = pj.y, pjz = pj.z, qj = jvalid ? pj.w : 0.f;
    if (SPC) {
      pair_rel_to_center(pj.x, pj.y, pj.z, cb, L, ilx, ily, ilz, pjx, pjy, pjz);
      if (!jvalid) pjx = pjy = pjz = PAIR_FAR_J;
    } else {
      // validity travels in the key: padding lanes can never pass the
      // key test (i padding carries HD_SENTINEL - 1, j padding HD_SENTINEL)
      jkey = jvalid ? jkey : HD_SENTINEL;
    }
    float jfx = 0.f, jfy = 0.f, jfz = 0.f, je = 0.f, jw = 0.f;
    float efx = 0.f, efy = 0.f, efz = 0.f, ee = 0.f, ew = 0.f;
    if constexpr (UC) {
      // a j-atom can be excluded only when its key is in the block's map
      unsigned word = keymap[(jkey >> 5) & (PAIR_KEYMAP_WORDS - 1u)];
      bool hit = jvalid && ((word >> (jkey & 31u)) & 1u);
      PairJStage<NT>& S = stage[w];
      float4 rec = make_float4(pjx, pjy, pjz, qj);
      S.pq[lane] = rec; S.pq[lane + 64] = rec;
      if constexpr (NT > 1) {
        // the same products load_row forms for an i-atom of type ti
#pragma unroll
        for (int ti = 0; ti < NT; ++ti) {
          float f12[NT], f6[NT];
#pragma unroll
          for (int u = 0; u < NT; ++u) {
            int k = (ti < P.ntypes && u < P.ntypes) ? ti * P.ntypes + u : 0;
            f12[u] = 12.0f * P.c12[k];
            f6[u] = 6.0f * P.c6[k];
          }
          float2 c = make_float2(pick<NT>(f12, jtype), pick<NT>(f6, jtype));
          S.lj[ti][lane] = c; S.lj[ti][lane + 64] = c;
        }
      }
      if (EV == 1 || NT == 0) { S.type[lane] = jtype; S.type[lane + 64] = jtype; }
      S.key[lane] = jkey; S.key[lane + 64] = jkey;
      wave_lds_sync();
      const float2* jl = &S.lj[(NT > 1) ? itype : 0u][lane];
      bool xk = P.excl_span > PAIR_KEYMAP_SPAN || __any(hit);
      const unsigned* jt = &S.type[lane];
      const unsigned* jk = &S.key[lane];
      if (P.uniform == 2) {
        if (xk)
          uc_steps<EV, NT, true, false>(R, pix, piy, piz, ikey, P.excl_span, kqi, P.uc_rc2, P.cg,
                                        P.pa, two_a_sqrtpi, &S.pq[lane], jl, jt, jk, efx, efy,
                                        efz, jfx, jfy, jfz, ee, je);
        else
          uc_steps<EV, NT, false, false>(R, pix, piy, piz, ikey, P.excl_span, kqi, P.uc_rc2,
                                         P.cg, P.pa, two_a_sqrtpi, &S.pq[lane], jl, jt, jk, efx,
                                         efy, efz, jfx, jfy, jfz, ee, je);
      } else {
        if (xk)
          uc_steps<EV, NT, true, true>(R, pix, piy, piz, ikey, P.excl_span, kqi, P.uc_rc2, P.cg,
                                       P.pa, two_a_sqrtpi, &S.pq[lane], jl, jt, jk, efx, efy, efz,
                                       jfx, jfy, jfz, ee, je);
        else
          uc_steps<EV, NT, false, true>(R, pix, piy, piz, ikey, P.excl_span, kqi, P.uc_rc2, P.cg,
                                        P.pa, two_a_sqrtpi, &S.pq[lane], jl, jt, jk, efx, efy,
                                        efz, jfx, jfy, jfz, ee, je);
      }
      // every lane's reads of this entry precede the next entry's writes
      wave_lds_sync();
    } else {
#pragma unroll 4
    for (int t = 0; t < 64; ++t) {
      float dx = pix - pjx, dy = piy - pjy, dz = piz - pjz;
      if (!SPC) {
        if (P.ortho) {
          dx = pair_axis_disp(pix, pjx, L.lx, ilx);
          dy = pair_axis_disp(piy, pjy, L.ly, ily);
          dz = pair_axis_disp(piz, pjz, L.lz, ilz);
        } else {
          pair_tri_disp(pix, piy, piz, pjx, pjy, pjz, L, ilx, ily, ilz, dx, dy, dz);
        }
      }
      float r2 = dx * dx + dy * dy + dz * dz;
      float inv_r = __builtin_amdgcn_rsqf(r2);
      float inv_r2 = inv_r * inv_r;
      // LJ (lj.rs:237-247, 250-287) from the i-row, column = j type
      float ir6 = inv_r2 * inv_r2 * inv_r2;
      float flj = inv_r2 * ir6 * (lj_f12(R, jtype) * ir6 - lj_f6(R, jtype));
      float elj = 0.f;
      if (EV == 1 || SW) elj = ir6 * (lj_c12(R, jtype) * ir6 - lj_c6(R, jtype));
      if (SW) {
        float rs2 = lj_rsw2(R, jtype), idl = lj_idl(R, jtype);
        float tau = fmaxf((r2 - rs2) * idl, 0.0f);
        float omt = 1.0f - tau;
        float s = omt * omt * (1.0f + 2.0f * tau);
        float chain = 12.0f * tau * omt * idl;
        flj = s * flj + chain * elj;
        elj = s * elj;
      }
      bool lj_on = r2 <= lj_cut2(R, jtype);
      // real-space Coulomb, f32 Hastings erfc (spme.rs:744-770)
      float r = r2 * inv_r;
      float qq = kqi * qj;
      float ar = alpha * r;
      float gauss = __expf(-(ar * ar));
      float tt = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ar);
      float erfc_ar =
          (0.254829592f +
           (-0.284496736f + (1.421413741f + (-1.453152027f + 1.061405429f * tt) * tt) * tt) * tt) *
          tt * gauss;
      float ec = qq * erfc_ar * inv_r;
      float fc = qq * inv_r2 * (erfc_ar * inv_r + two_a_sqrtpi * gauss);
      bool c_on = r2 <= coul_rc2;
      // excluded when |key_j - key_i| <= span (span 0: equal molecule keys)
      bool ok = (jkey - ikey + P.excl_span) > 2u * P.excl_span;
      if (!SPC) ok = ok & (max(jkey, ikey) < HD_SENTINEL - 1u);
      float factor = (ok & lj_on) ? flj : 0.f;
      factor += (ok & c_on) ? fc : 0.f;
      // i and j sides as fused multiply-adds of +-factor * d
      float nfactor = -factor;
      efx = fmaf(factor, dx, efx); efy = fmaf(factor, dy, efy); efz = fmaf(factor, dz, efz);
      jfx = fmaf(nfactor, dx, jfx); jfy = fmaf(nfactor, dy, jfy); jfz = fmaf(nfactor, dz, jfz);
      if (EV == 1) {
        float energy = (ok & lj_on) ? elj : 0.f;
        energy += (ok & c_on) ? ec : 0.f;
        ee = fmaf(0.5f, energy, ee);
        je = fmaf(0.5f, energy, je);
      }
      if (EV && !SPC) {
        float fr2 = factor * r2;
        ew = fmaf(0.5f, fr2, ew);
        jw = fmaf(0.5f, fr2, jw);
      }
      pjx = hd_rot1(pjx); pjy = hd_rot1(pjy); pjz = hd_rot1(pjz); qj = hd_rot1(qj);
      jkey = hd_rot1_u(jkey);
      if (NT != 1) jtype = hd_rot1_u(jtype);
      jfx = hd_rot1(jfx); jfy = hd_rot1(jfy); jfz = hd_rot1(jfz);
      if (EV == 1) je = hd_rot1(je);
      if (EV && !SPC) jw = hd_rot1(jw);
    }
    }
    if (EV && SPC) {
      // the entry's pair virial sum_ij F_ij . (r_i - r_j) split by atom: both
      // positions are relative to the block centre (j in the image nearest
      // it), so the i side is r_i . F_i(entry) and the j side r_j . F_j(entry)
      // (j partials are back in their loading lanes, and so are the j
      // positions); pairs inside the block are evaluated from both sides and
      // their j sides dropped, which again leaves F_ij . r_ij once
      ew = pix * efx + piy * efy + piz * efz;
      jw = pjx * jfx + pjy * jfy + pjz * jfz;
    }
    ai_fx += hd_to_fixed48(efx);
    ai_fy += hd_to_fixed48(efy);
    ai_fz += hd_to_fixed48(efz);
    if (EV) {
      ai_e += hd_to_fixed48(ee);
      ai_w += hd_to_fixed48(ew);
    }
    // Newton's third law: the j-side sums of atoms that belong to block b
    // are dropped (their pairs were evaluated from both sides as i-pairs)
    if (jvalid && !jin) {
      unsigned long long* d = fpj + (size_t)jslot * 4;
      long long vx = hd_to_fixed48(jfx), vy = hd_to_fixed48(jfy), vz = hd_to_fixed48(jfz);
      if (vx) atomicAdd(d + 0, (unsigned long long)vx);
      if (vy) atomicAdd(d + 1, (unsigned long long)vy);
      if (vz) atomicAdd(d + 2, (unsigned long long)vz);
      if (EV) {
        long long ve = hd_to_fixed48(je), vw = hd_to_fixed48(jw);
        if (EV == 1 && ve) atomicAdd(fpj_ew + (size_t)jslot * 2, (unsigned long long)ve);
        if (vw) atomicAdd(fpj_ew + (size_t)jslot * 2 + 1, (unsigned long long)vw);
      }
    }
  }
}

template <int EV, int NT, bool SW>
__global__ void __launch_bounds__(256) k_pair(
    PairParams P, const float* __restrict__ lattice, const float4* __restrict__ tile_posq,
    const unsigned* __restrict__ tile_type, const unsigned* __restrict__ tile_key,
    const float4* __restrict__ center, const float4* __restrict__ half,
    const unsigned* __restrict__ entry_atoms, const unsigned* __restrict__ entry_count,
    const unsigned* __restrict__ block_order, unsigned cap, unsigned n,
    long long* __restrict__ fpi, long long* __restrict__ fpi_ew,
    unsigned long long* __restrict__ fpj, unsigned long long* __restrict__ fpj_ew,
    const float4* __restrict__ lj_tab) {
  __shared__ long long red[PAIR_WAVES][5][64];
  __shared__ unsigned keymap[PAIR_KEYMAP_WORDS];
  __shared__ PairJStage<NT> stage[PAIR_WAVES];
  __shared__ PairTab<NT> tab;
  if constexpr (NT == 0) {
    load_tab(tab, lj_tab, P.ntypes);
    __syncthreads();
  }
  unsigned b = block_order[blockIdx.x];  // heaviest blocks first
  unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  HdLattice L = hd_load_lattice(lattice);
  unsigned islot = b * 64 + lane;
  bool ivalid = islot < n;
  float4 pi = tile_posq[ivalid ? islot : 0u];
  unsigned itype = ivalid ? tile_type[islot] : 0u;
  unsigned ikey = ivalid ? tile_key[islot] : (HD_SENTINEL - 1u);
  float4 cb = center[b], hb = half[b];
  bool spc = (L.xy == 0.f && L.xz == 0.f && L.yz == 0.f) &&
             (0.5f * L.lx - hb.x >= P.spc_margin) && (0.5f * L.ly - hb.y >= P.spc_margin) &&
             (0.5f * L.lz - hb.z >= P.spc_margin);
  float pix = pi.x, piy = pi.y, piz = pi.z, qi = ivalid ? pi.w : 0.f;
  if (!ivalid) { pix = cb.x; piy = cb.y; piz = cb.z; }
  if (spc) {
    hd_rel_to_center(pix, piy, piz, cb.x, cb.y, cb.z, L, pix, piy, piz);
    if (!ivalid) pix = piy = piz = PAIR_FAR_I;
  }
  LjRow<NT> R = load_row<NT>(P, itype, ivalid, tab);
  unsigned nent = entry_count[b];
  const unsigned* ents = entry_atoms + (size_t)b * cap * 64;
  long long ax = 0, ay = 0, az = 0, ae = 0, aw = 0;
  // the regrouped loop: one cutoff, no switch
  constexpr bool UCOK = !SW;
  bool uc = UCOK && spc && P.uniform;  // uniform over the workgroup
  if (uc) {
    if constexpr (UCOK) {
      for (unsigned k = threadIdx.x; k < PAIR_KEYMAP_WORDS; k += 64 * PAIR_WAVES) keymap[k] = 0u;
      __syncthreads();
      if (w == 0 && ivalid && P.excl_span <= PAIR_KEYMAP_SPAN)
        for (unsigned d = 0; d <= 2u * P.excl_span; ++d) {
          unsigned key = ikey - P.excl_span + d;
          atomicOr(&keymap[(key >> 5) & (PAIR_KEYMAP_WORDS - 1u)], 1u << (key & 31u));
        }
      __syncthreads();
      pair_entries<EV, NT, SW, true, true>(P, L, b, w, lane, pix, piy, piz, qi, ikey, R, cb,
                                           tile_posq, tile_type, tile_key, ents, nent, n, fpj,
                                           fpj_ew, ax, ay, az, ae, aw, keymap, stage, itype);
    }
  } else if (spc)
    pair_entries<EV, NT, SW, true, false>(P, L, b, w, lane, pix, piy, piz, qi, ikey, R, cb,
                                          tile_posq, tile_type, tile_key, ents, nent, n, fpj,
                                          fpj_ew, ax, ay, az, ae, aw, keymap, stage, itype);
  else
    pair_entries<EV, NT, SW, false, false>(P, L, b, w, lane, pix, piy, piz, qi, ikey, R, cb,
                                           tile_posq, tile_type, tile_key, ents, nent, n, fpj,
                                           fpj_ew, ax, ay, az, ae, aw, keymap, stage, itype);
  red[w][0][lane] = ax;
  red[w][1][lane] = ay;
  red[w][2][lane] = az;
  if (EV) {
    red[w][3][lane] = ae;
    red[w][4][lane] = aw;
  }
  __syncthreads();
  if (w == 0) {
    long long s[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < PAIR_WAVES; ++k) {
      s[0] += red[k][0][lane];
      s[1] += red[k][1][lane];
      s[2] += red[k][2][lane];
      if (EV) {
        s[3] += red[k][3][lane];
        s[4] += red[k][4][lane];
      }
    }
    reinterpret_cast<longlong2*>(fpi)[2 * islot] = make_longlong2(s[0], s[1]);
    reinterpret_cast<longlong2*>(fpi)[2 * islot + 1] = make_longlong2(s[2], 0);
    if (EV) reinterpret_cast<longlong2*>(fpi_ew)[islot] = make_longlong2(s[3], s[4]);
  }
}

// Atom-list mode: one workgroup of 4 waves per 64-atom i-block, lane =
// i-atom; wave w walks rows w, w + 4, ... of the block's list region (row
// k = the k-th neighbour of every lane: coalesced index loads, gathered j
// positions).  Each wave issues the loads of ALIST_U rows before it
// evaluates them, so ALIST_U gathers per lane are in flight at once.
// Every pair is evaluated from both sides, so there is no j-side partial
// and no atomic: the i-side f32 sums are folded into i64 fixed point at
// 2^48 every 64 neighbours (the reference's per-entry fold), combined over
// the 4 waves in LDS (integer sums, so the order does not matter) and
// stored once.  Energy and virial are halved per side as in the packed
// kernel.  Exclusions were applied when the lists were built.  The
// displacement takes a per-pair minimum image (ORTHO: one rint per axis).
#define ALIST_U 4

template <int EV, int NT, bool SW, bool ORTHO>
__device__ __forceinline__ void alist_pair(const PairParams& P, const HdLattice& L, float ilx,
                                           float ily, float ilz, const LjRow<NT>& R, float4 pi,
                                           float kqi, float4 pj, unsigned jtype, float& efx,
                                           float& efy, float& efz, float& ee, float& ew) {
  const float two_a_sqrtpi = 2.0f * P.alpha * 0.5641895835477563f;
  float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
  if (ORTHO) {
    dx = pair_axis_disp(pi.x, pj.x, L.lx, ilx);
    dy = pair_axis_disp(pi.y, pj.y, L.ly, ily);
    dz = pair_axis_disp(pi.z, pj.z, L.lz, ilz);
  } else {
    pair_tri_disp(pi.x, pi.y, pi.z, pj.x, pj.y, pj.z, L, ilx, ily, ilz, dx, dy, dz);
  }
  float r2 = dx * dx + dy * dy + dz * dz;
  float inv_r = __builtin_amdgcn_rsqf(r2);
  float inv_r2 = inv_r * inv_r;
  float ir6 = inv_r2 * inv_r2 * inv_r2;
  float flj = inv_r2 * ir6 * (lj_f12(R, jtype) * ir6 - lj_f6(R, jtype));
  float elj = 0.f;
  if (EV == 1 || SW) elj = ir6 * (lj_c12(R, jtype) * ir6 - lj_c6(R, jtype));
  if (SW) {
    float rs2 = lj_rsw2(R, jtype), idl = lj_idl(R, jtype);
    float tau = fmaxf((r2 - rs2) * idl, 0.0f);
    float omt = 1.0f - tau;
    float sw = omt * omt * (1.0f + 2.0f * tau);
    float chain = 12.0f * tau * omt * idl;
    flj = sw * flj + chain * elj;
    elj = sw * elj;
  }
  bool lj_on = r2 <= lj_cut2(R, jtype);
  float r = r2 * inv_r;
  float qq = kqi * pj.w;
  float ar = P.alpha * r;
  float gauss = __expf(-(ar * ar));
  float tt = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ar);
  float erfc_ar =
      (0.254829592f +
       (-0.284496736f + (1.421413741f + (-1.453152027f + 1.061405429f * tt) * tt) * tt) * tt) *
      tt * gauss;
  float ec = qq * erfc_ar * inv_r;
  float fc = qq * inv_r2 * (erfc_ar * inv_r + two_a_sqrtpi * gauss);
  bool c_on = r2 <= P.coul_rc2;
  float factor = lj_on ? flj : 0.f;
  factor += c_on ? fc : 0.f;
  efx += factor * dx;
  efy += factor * dy;
  efz += factor * dz;
  if (EV == 1) {
    float energy = lj_on ? elj : 0.f;
    energy += c_on ? ec : 0.f;
    ee += 0.5f * energy;
  }
  if (EV) ew += 0.5f * factor * r2;
}

template <int EV, int NT, bool SW, bool ORTHO>
__global__ void __launch_bounds__(256) k_pair_alist(
    PairParams P, const float* __restrict__ lattice, const float4* __restrict__ tile_posq,
    const unsigned* __restrict__ tile_type, const unsigned* __restrict__ rows_in,
    const unsigned* __restrict__ row_count, unsigned cap, unsigned n, unsigned nblocks,
    long long* __restrict__ fpi, long long* __restrict__ fpi_ew,
    const float4* __restrict__ lj_tab) {
  __shared__ long long red[PAIR_WAVES][5][64];
  __shared__ PairTab<NT> tab;
  if constexpr (NT == 0) {
    load_tab(tab, lj_tab, P.ntypes);
    __syncthreads();
  }
  unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned b = blockIdx.x;
  HdLattice L = hd_load_lattice(lattice);
  const float ilx = 1.0f / L.lx, ily = 1.0f / L.ly, ilz = 1.0f / L.lz;
  unsigned islot = b * 64 + lane;
  bool ivalid = islot < n;
  float4 pi = tile_posq[ivalid ? islot : 0u];
  unsigned itype = ivalid ? tile_type[islot] : 0u;
  LjRow<NT> R = load_row<NT>(P, itype, ivalid, tab);
  const float kqi = P.kc * (ivalid ? pi.w : 0.f);
  const unsigned rows = row_count[b];
  const unsigned* list = rows_in + (size_t)b * cap * 64 + lane;
  long long ax = 0, ay = 0, az = 0, ae = 0, aw = 0;
  float efx = 0.f, efy = 0.f, efz = 0.f, ee = 0.f, ew = 0.f;
  unsigned since_fold = 0;
  for (unsigned k = w; k < rows; k += PAIR_WAVES * ALIST_U) {
    unsigned js[ALIST_U];
#pragma unroll
    for (int u = 0; u < ALIST_U; ++u) {
      unsigned kk = k + u * PAIR_WAVES;
      js[u] = kk < rows ? list[(size_t)kk * 64] : HD_SENTINEL;
    }
    float4 pj[ALIST_U];
    unsigned tj[ALIST_U];
#pragma unroll
    for (int u = 0; u < ALIST_U; ++u) {
      unsigned jsafe = js[u] < n ? js[u] : 0u;
      pj[u] = tile_posq[jsafe];
      tj[u] = (NT != 1) ? tile_type[jsafe] : 0u;
    }
#pragma unroll
    for (int u = 0; u < ALIST_U; ++u)
      if (js[u] < n)
        alist_pair<EV, NT, SW, ORTHO>(P, L, ilx, ily, ilz, R, pi, kqi, pj[u], tj[u], efx, efy,
                                      efz, ee, ew);
    since_fold += ALIST_U;
    if (since_fold >= 64) {
      ax += hd_to_fixed48(efx);
      ay += hd_to_fixed48(efy);
      az += hd_to_fixed48(efz);
      if (EV) {
        ae += hd_to_fixed48(ee);
        aw += hd_to_fixed48(ew);
      }
      efx = efy = efz = ee = ew = 0.f;
      since_fold = 0;
    }
  }
  ax += hd_to_fixed48(efx);
  ay += hd_to_fixed48(efy);
  az += hd_to_fixed48(efz);
  if (EV) {
    ae += hd_to_fixed48(ee);
    aw += hd_to_fixed48(ew);
  }
  red[w][0][lane] = ax;
  red[w][1][lane] = ay;
  red[w][2][lane] = az;
  if (EV) {
    red[w][3][lane] = ae;
    red[w][4][lane] = aw;
  }
  __syncthreads();
  if (w == 0 && ivalid) {
    long long sm[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < PAIR_WAVES; ++q) {
      sm[0] += red[q][0][lane];
      sm[1] += red[q][1][lane];
      sm[2] += red[q][2][lane];
      if (EV) {
        sm[3] += red[q][3][lane];
        sm[4] += red[q][4][lane];
      }
    }
    reinterpret_cast<longlong2*>(fpi)[2 * islot] = make_longlong2(sm[0], sm[1]);
    reinterpret_cast<longlong2*>(fpi)[2 * islot + 1] = make_longlong2(sm[2], 0);
    if (EV) reinterpret_cast<longlong2*>(fpi_ew)[islot] = make_longlong2(sm[3], sm[4]);
  }
}

// Harmonic angles (src/forces/angle.rs:258-313 functor, composed kernel
// src/forces/jit_composed.rs:2300-2345): one thread per angle i-j-k with
// minimum-image bond vectors, theta = atan2(|r_ij||r_kj| sin, dot),
// F_i, F_k from g = -k (theta - theta_0) / sin theta, F_j = -(F_i + F_k),
// energy and virial split in thirds.  The three forces are folded into the
// j-side fixed-point accumulators at 2^48 (integer sums, so the total is
// independent of the order in which angles land).
__global__ void k_angles(const float4* __restrict__ posq, const float* __restrict__ lattice,
                         const unsigned* __restrict__ ijk, const double* __restrict__ kth,
                         const double* __restrict__ th0, const unsigned* __restrict__ slot_of,
                         unsigned long long* __restrict__ fpj,
                         unsigned long long* __restrict__ fpj_ew, unsigned n_angles, int ev) {
  unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_angles) return;
  HdLattice L = hd_load_lattice(lattice);
  unsigned ai = ijk[3 * m], aj = ijk[3 * m + 1], ak = ijk[3 * m + 2];
  float4 pi = posq[ai], pj = posq[aj], pk = posq[ak];
  // FP64 throughout: the force is proportional to theta - theta0, a small
  // difference of two angles near 2 rad that f32 would resolve only to
  // ~1e-7 rad, i.e. to a relative 1e-4 of the force for a 1e-3 rad strain
  double ui[3] = {(double)pi.x - (double)pj.x, (double)pi.y - (double)pj.y,
                  (double)pi.z - (double)pj.z};
  double uk[3] = {(double)pk.x - (double)pj.x, (double)pk.y - (double)pj.y,
                  (double)pk.z - (double)pj.z};
  hd_min_image_f64(ui[0], ui[1], ui[2], L);
  hd_min_image_f64(uk[0], uk[1], uk[2], L);
  double li2 = ui[0] * ui[0] + ui[1] * ui[1] + ui[2] * ui[2];
  double lk2 = uk[0] * uk[0] + uk[1] * uk[1] + uk[2] * uk[2];
  double gi[3] = {0.0, 0.0, 0.0}, gk[3] = {0.0, 0.0, 0.0}, u = 0.0, w = 0.0;
  if (li2 > 0.0 && lk2 > 0.0) {
    double li = sqrt(li2), lk = sqrt(lk2);
    double cr[3] = {ui[1] * uk[2] - ui[2] * uk[1], ui[2] * uk[0] - ui[0] * uk[2],
                    ui[0] * uk[1] - ui[1] * uk[0]};
    double sn_l = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);  // |ui||uk| sin
    double dt_l = ui[0] * uk[0] + ui[1] * uk[1] + ui[2] * uk[2];        // |ui||uk| cos
    if (sn_l > 1e-12 * li * lk) {
      double theta = atan2(sn_l, dt_l);
      double dth = theta - th0[m];
      double kf = kth[m];
      // dtheta/dx_i = (cos ui/|ui|^2 - uk/(|ui||uk|)) / sin, and for k
      // alike; F = -k dth dtheta/dx
      double cs = dt_l / (li * lk), sn = sn_l / (li * lk);
      double g = -kf * dth / sn, inv_ik = 1.0 / (li * lk);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        gi[d] = g * (cs * ui[d] / li2 - uk[d] * inv_ik);
        gk[d] = g * (cs * uk[d] / lk2 - ui[d] * inv_ik);
      }
      u = 0.5 * kf * dth * dth;
      w = ui[0] * gi[0] + ui[1] * gi[1] + ui[2] * gi[2] + uk[0] * gk[0] + uk[1] * gk[1] +
          uk[2] * gk[2];
    }
  }
  float fix = (float)gi[0], fiy = (float)gi[1], fiz = (float)gi[2];
  float fkx = (float)gk[0], fky = (float)gk[1], fkz = (float)gk[2];
  float fjx = (float)-(gi[0] + gk[0]), fjy = (float)-(gi[1] + gk[1]),
        fjz = (float)-(gi[2] + gk[2]);
  unsigned si = slot_of[ai], sj = slot_of[aj], sk = slot_of[ak];
  const float f3[3][3] = {{fix, fiy, fiz}, {fjx, fjy, fjz}, {fkx, fky, fkz}};
  const unsigned sl[3] = {si, sj, sk};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      long long v = hd_to_fixed48(f3[a][d]);
      if (v) atomicAdd(fpj + (size_t)sl[a] * 4 + d, (unsigned long long)v);
    }
  if (ev) {
    long long e3 = hd_to_fixed48((float)(u / 3.0)), w3 = hd_to_fixed48((float)(w / 3.0));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (e3) atomicAdd(fpj_ew + (size_t)sl[a] * 2, (unsigned long long)e3);
      if (w3) atomicAdd(fpj_ew + (size_t)sl[a] * 2 + 1, (unsigned long long)w3);
    }
  }
}

void hd_launch_angles(hipStream_t st, const float4* posq, const float* lattice,
                      const unsigned* ijk, const double* kth, const double* th0,
                      const unsigned* slot_of, long long* fpj, long long* fpj_ew,
                      unsigned n_angles, bool ev) {
  if (n_angles == 0) return;
  hipLaunchKernelGGL(k_angles, dim3((n_angles + 255) / 256), dim3(256), 0, st, posq, lattice,
                     ijk, kth, th0, slot_of, (unsigned long long*)fpj,
                     (unsigned long long*)fpj_ew, n_angles, ev ? 1 : 0);
}

// Per-atom scatter of a bonded term's force, energy share and virial share
// into the j-side fixed-point accumulators (slot order).
__device__ __forceinline__ void bonded_scatter(unsigned long long* __restrict__ fpj,
                                               unsigned long long* __restrict__ fpj_ew,
                                               unsigned slot, double fx, double fy, double fz,
                                               long long e_share, long long w_share, int ev) {
  const float f3[3] = {(float)fx, (float)fy, (float)fz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    long long v = hd_to_fixed48(f3[d]);
    if (v) atomicAdd(fpj + (size_t)slot * 4 + d, (unsigned long long)v);
  }
  if (ev) {
    if (e_share) atomicAdd(fpj_ew + (size_t)slot * 2, (unsigned long long)e_share);
    if (w_share) atomicAdd(fpj_ew + (size_t)slot * 2 + 1, (unsigned long long)w_share);
  }
}

__device__ __forceinline__ void disp_f64(const float4& a, const float4& b, const HdLattice& L,
                                         double d[3]) {
  d[0] = (double)a.x - (double)b.x;
  d[1] = (double)a.y - (double)b.y;
  d[2] = (double)a.z - (double)b.z;
  hd_min_image_f64(d[0], d[1], d[2], L);
}

// Two-body bonded terms in FP64 (src/forces/harmonic_bond.rs:288-330,
// src/forces/morse.rs:290-320): U = k/2 (r - r0)^2, or
// U = De (1 - exp(-a (r - re)))^2; F_i = -dU/dr (r_i - r_j)/r, F_j = -F_i;
// W = F_i . (r_i - r_j); energy and virial shared half and half.
__global__ void k_bonds(const float4* __restrict__ posq, const float* __restrict__ lattice,
                        const unsigned* __restrict__ ij, const int* __restrict__ kind,
                        const double* __restrict__ prm, const unsigned* __restrict__ slot_of,
                        unsigned long long* __restrict__ fpj,
                        unsigned long long* __restrict__ fpj_ew, unsigned n_bonds, int ev) {
  unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_bonds) return;
  HdLattice L = hd_load_lattice(lattice);
  unsigned ai = ij[2 * m], aj = ij[2 * m + 1];
  double d[3];
  disp_f64(posq[ai], posq[aj], L, d);
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double r = sqrt(r2);
  double dudr = 0.0, u = 0.0;
  const double* p = prm + 3 * (size_t)m;
  if (r > 1e-7) {
    if (kind[m] == 0) {
      double dr = r - p[1];
      dudr = p[0] * dr;
      u = 0.5 * p[0] * dr * dr;
    } else {
      double e = exp(-p[1] * (r - p[2]));
      dudr = 2.0 * p[0] * p[1] * (1.0 - e) * e;
      u = p[0] * (1.0 - e) * (1.0 - e);
    }
  }
  double g = r > 1e-7 ? -dudr / r : 0.0;
  long long e2 = 0, w2 = 0;
  if (ev) {
    e2 = hd_to_fixed48((float)(0.5 * u));
    w2 = hd_to_fixed48((float)(0.5 * g * r2));
  }
  bonded_scatter(fpj, fpj_ew, slot_of[ai], g * d[0], g * d[1], g * d[2], e2, w2, ev);
  bonded_scatter(fpj, fpj_ew, slot_of[aj], -g * d[0], -g * d[1], -g * d[2], e2, w2, ev);
}

void hd_launch_bonds(hipStream_t st, const float4* posq, const float* lattice, const unsigned* ij,
                     const int* kind, const double* prm, const unsigned* slot_of, long long* fpj,
                     long long* fpj_ew, unsigned n_bonds, bool ev) {
  if (n_bonds == 0) return;
  hipLaunchKernelGGL(k_bonds, dim3((n_bonds + 255) / 256), dim3(256), 0, st, posq, lattice, ij,
                     kind, prm, slot_of, (unsigned long long*)fpj, (unsigned long long*)fpj_ew,
                     n_bonds, ev ? 1 : 0);
}

// Periodic dihedral U = k (1 + cos(n phi - phi0)) (src/forces/dihedral.rs:23,
// 360-460) in FP64.  With F = r_i - r_j, G = r_j - r_k, H = r_l - r_k,
// A = F x G and B = H x G: cos phi = A.B / (|A||B|) and, since B.G = 0,
// sin phi = (B x A).G / (|A||B||G|) = -|G| F.B / (|A||B|), the same angle
// and sign as the reference's atan2(|b2| b1.n, m.n); the gradients are
// those of Blondel and Karplus (J. Comput. Chem. 17, 1132 (1996)):
//   dphi/dr_i = -|G|/|A|^2 A,  dphi/dr_l = |G|/|B|^2 B,
//   dphi/dr_j = |G|/|A|^2 A + (F.G)/(|A|^2 |G|) A - (H.G)/(|B|^2 |G|) B,
//   dphi/dr_k = -|G|/|B|^2 B - (F.G)/(|A|^2 |G|) A + (H.G)/(|B|^2 |G|) B.
// W = sum_a (r_a - r_j) . F_a; energy and virial shared in quarters.
__global__ void k_dihedrals(const float4* __restrict__ posq, const float* __restrict__ lattice,
                            const unsigned* __restrict__ ijkl, const double* __restrict__ prm,
                            const unsigned* __restrict__ slot_of,
                            unsigned long long* __restrict__ fpj,
                            unsigned long long* __restrict__ fpj_ew, unsigned n_dih, int ev) {
  unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_dih) return;
  HdLattice L = hd_load_lattice(lattice);
  unsigned a0 = ijkl[4 * m], a1 = ijkl[4 * m + 1], a2 = ijkl[4 * m + 2], a3 = ijkl[4 * m + 3];
  float4 p0 = posq[a0], p1 = posq[a1], p2 = posq[a2], p3 = posq[a3];
  double F[3], G[3], H[3];
  disp_f64(p0, p1, L, F);
  disp_f64(p1, p2, L, G);
  disp_f64(p3, p2, L, H);
  double A[3] = {F[1] * G[2] - F[2] * G[1], F[2] * G[0] - F[0] * G[2], F[0] * G[1] - F[1] * G[0]};
  double B[3] = {H[1] * G[2] - H[2] * G[1], H[2] * G[0] - H[0] * G[2], H[0] * G[1] - H[1] * G[0]};
  double a2n = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  double b2n = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  double g2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
  double fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0}, fk[3] = {0, 0, 0}, fl[3] = {0, 0, 0};
  double u = 0.0, w = 0.0;
  if (a2n >= 1e-14 && b2n >= 1e-14 && g2 >= 1e-14) {
    double gl = sqrt(g2);
    double fb = F[0] * B[0] + F[1] * B[1] + F[2] * B[2];
    double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
    double phi = atan2(-gl * fb, ab);
    const double* p = prm + 3 * (size_t)m;
    double k = p[0], nm = p[1], phi0 = p[2];
    double delta = nm * phi - phi0;
    double sd, cd;
    sincos(delta, &sd, &cd);
    u = k * (1.0 + cd);
    double mdu = k * nm * sd;  // -dU/dphi
    double fg = F[0] * G[0] + F[1] * G[1] + F[2] * G[2];
    double hg = H[0] * G[0] + H[1] * G[1] + H[2] * G[2];
    double ca = gl / a2n, cb = gl / b2n;
    double sa = fg / (a2n * gl), sb = hg / (b2n * gl);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double gi = -ca * A[d], gl_ = cb * B[d];
      double gj = ca * A[d] + sa * A[d] - sb * B[d];
      double gk = -cb * B[d] - sa * A[d] + sb * B[d];
      fi[d] = mdu * gi;
      fj[d] = mdu * gj;
      fk[d] = mdu * gk;
      fl[d] = mdu * gl_;
    }
    // positions relative to r_j: r_i - r_j = F, r_k - r_j = -G, r_l - r_j = H - G
    for (int d = 0; d < 3; ++d) w += F[d] * fi[d] - G[d] * fk[d] + (H[d] - G[d]) * fl[d];
  }
  long long e4 = 0, w4 = 0;
  if (ev) {
    e4 = hd_to_fixed48((float)(0.25 * u));
    w4 = hd_to_fixed48((float)(0.25 * w));
  }
  bonded_scatter(fpj, fpj_ew, slot_of[a0], fi[0], fi[1], fi[2], e4, w4, ev);
  bonded_scatter(fpj, fpj_ew, slot_of[a1], fj[0], fj[1], fj[2], e4, w4, ev);
  bonded_scatter(fpj, fpj_ew, slot_of[a2], fk[0], fk[1], fk[2], e4, w4, ev);
  bonded_scatter(fpj, fpj_ew, slot_of[a3], fl[0], fl[1], fl[2], e4, w4, ev);
}

void hd_launch_dihedrals(hipStream_t st, const float4* posq, const float* lattice,
                         const unsigned* ijkl, const double* prm, const unsigned* slot_of,
                         long long* fpj, long long* fpj_ew, unsigned n_dih, bool ev) {
  if (n_dih == 0) return;
  hipLaunchKernelGGL(k_dihedrals, dim3((n_dih + 255) / 256), dim3(256), 0, st, posq, lattice,
                     ijkl, prm, slot_of, (unsigned long long*)fpj, (unsigned long long*)fpj_ew,
                     n_dih, ev ? 1 : 0);
}

// Scaled exclusions (src/forces/jit_composed.rs:1445-1458 with the LJ and
// SPME-real fragments' own scale tables, lj.rs:295-321, spme.rs:723-762):
// the pair kernel applies weight 1 to every pair within the cutoff whose
// keys differ by more than the span and 0 to the rest; this pass adds
// (scale - that weight) times each term, evaluated in FP64 with the same
// Hastings erfc and cubic LJ switch, so the pair ends at its scale.
__global__ void k_excl_scaled(const float4* __restrict__ posq, const float* __restrict__ lattice,
                              const HdExclPair* __restrict__ pairs, unsigned n_pairs, double kc,
                              double alpha, double coul_rc2, const unsigned* __restrict__ slot_of,
                              unsigned long long* __restrict__ fpj,
                              unsigned long long* __restrict__ fpj_ew, int ev) {
  unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_pairs) return;
  HdLattice L = hd_load_lattice(lattice);
  HdExclPair P = pairs[m];
  float4 pi = posq[P.i], pj = posq[P.j];
  double d[3];
  disp_f64(pi, pj, L, d);
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double factor = 0.0, energy = 0.0;
  if (r2 > 0.0) {
    double inv_r2 = 1.0 / r2, inv_r = sqrt(inv_r2), r = r2 * inv_r;
    if (P.w_lj != 0.0 && r2 <= P.cut2 && P.eps != 0.0) {
      double sr2 = P.sigma * P.sigma * inv_r2;
      double sr6 = sr2 * sr2 * sr2;
      double f = 24.0 * P.eps * inv_r2 * (2.0 * sr6 * sr6 - sr6);
      double e = 4.0 * P.eps * (sr6 * sr6 - sr6);
      double delta = P.cut2 - P.rsw2;
      if (delta > 0.0 && r2 > P.rsw2) {
        double tau = (r2 - P.rsw2) / delta, omt = 1.0 - tau;
        double s = omt * omt * (1.0 + 2.0 * tau);
        f = s * f + 12.0 * tau * omt / delta * e;
        e = s * e;
      }
      factor += P.w_lj * f;
      energy += P.w_lj * e;
    }
    if (P.w_coul != 0.0 && coul_rc2 > 0.0 && r2 <= coul_rc2) {
      double qq = kc * (double)pi.w * (double)pj.w;
      double ar = alpha * r;
      double gauss = exp(-ar * ar);
      double t = 1.0 / (1.0 + 0.3275911 * ar);
      double ec = (0.254829592 +
                   (-0.284496736 + (1.421413741 + (-1.453152027 + 1.061405429 * t) * t) * t) * t) *
                  t * gauss;
      factor += P.w_coul * qq * inv_r2 * (ec * inv_r + 2.0 * alpha * 0.5641895835477563 * gauss);
      energy += P.w_coul * qq * ec * inv_r;
    }
  }
  long long e2 = 0, w2 = 0;
  if (ev) {
    e2 = hd_to_fixed48((float)(0.5 * energy));
    w2 = hd_to_fixed48((float)(0.5 * factor * r2));
  }
  bonded_scatter(fpj, fpj_ew, slot_of[P.i], factor * d[0], factor * d[1], factor * d[2], e2, w2,
                 ev);
  bonded_scatter(fpj, fpj_ew, slot_of[P.j], -factor * d[0], -factor * d[1], -factor * d[2], e2,
                 w2, ev);
}

void hd_launch_excl_scaled(hipStream_t st, const float4* posq, const float* lattice,
                           const HdExclPair* pairs, unsigned n_pairs, double kc, double alpha,
                           double coul_rc2, const unsigned* slot_of, long long* fpj,
                           long long* fpj_ew, bool ev) {
  if (n_pairs == 0) return;
  hipLaunchKernelGGL(k_excl_scaled, dim3((n_pairs + 255) / 256), dim3(256), 0, st, posq, lattice,
                     pairs, n_pairs, kc, alpha, coul_rc2, slot_of, (unsigned long long*)fpj,
                     (unsigned long long*)fpj_ew, ev ? 1 : 0);
}

// Diagnostic: number of distinct pairs of the current list with
// r^2 <= r_max^2 that are neither excluded nor padding (the "useful" pairs
// the pair kernel evaluates).  One wave per i-block, plain loops.
__global__ void __launch_bounds__(64) k_count_pairs(
    float rmax2, unsigned excl_span, const float* __restrict__ lattice, const float4* __restrict__ tile_posq,
    const unsigned* __restrict__ tile_key, const unsigned* __restrict__ entry_atoms,
    const unsigned* __restrict__ entry_count, unsigned cap, unsigned n,
    unsigned long long* __restrict__ out) {
  unsigned b = blockIdx.x, lane = threadIdx.x;
  unsigned islot = b * 64 + lane;
  if (islot >= n) return;
  HdLattice L = hd_load_lattice(lattice);
  float4 pi = tile_posq[islot];
  unsigned ikey = tile_key[islot];
  unsigned nent = entry_count[b];
  const unsigned* ents = entry_atoms + (size_t)b * cap * 64;
  unsigned long long cnt = 0;
  for (unsigned e = 0; e < nent; ++e)
    for (unsigned t = 0; t < 64; ++t) {
      unsigned js = ents[(size_t)e * 64 + t];
      if (js >= n) continue;
      if ((js >> 6) == b && js <= islot) continue;
      if (tile_key[js] - ikey + excl_span <= 2u * excl_span) continue;
      float4 pj = tile_posq[js];
      float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      hd_min_image(dx, dy, dz, L);
      if (dx * dx + dy * dy + dz * dz <= rmax2) ++cnt;
    }
  atomicAdd(out, cnt);
}

// Atom-list mode: each pair is in both atoms' lists; count it from the
// lower slot.
__global__ void __launch_bounds__(64) k_count_pairs_alist(
    float rmax2, const float* __restrict__ lattice, const float4* __restrict__ tile_posq,
    const unsigned* __restrict__ rows_in, const unsigned* __restrict__ row_count, unsigned cap,
    unsigned n, unsigned long long* __restrict__ out) {
  unsigned b = blockIdx.x, lane = threadIdx.x;
  unsigned islot = b * 64 + lane;
  if (islot >= n) return;
  HdLattice L = hd_load_lattice(lattice);
  float4 pi = tile_posq[islot];
  const unsigned* list = rows_in + (size_t)b * cap * 64 + lane;
  unsigned long long cnt = 0;
  for (unsigned k = 0; k < row_count[b]; ++k) {
    unsigned js = list[(size_t)k * 64];
    if (js >= n || js <= islot) continue;
    float4 pj = tile_posq[js];
    float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    hd_min_image(dx, dy, dz, L);
    if (dx * dx + dy * dy + dz * dz <= rmax2) ++cnt;
  }
  atomicAdd(out, cnt);
}

void hd_launch_count_pairs(hipStream_t st, const HdPairArgs& a, unsigned long long* out) {
  unsigned nblocks = a.npad / 64;
  if (a.atom_lists)
    hipLaunchKernelGGL(k_count_pairs_alist, dim3(nblocks), dim3(64), 0, st,
                       (float)(a.rmax * a.rmax), a.lattice, a.tile_posq, a.entry_atoms,
                       a.entry_count, a.cap, a.n, out);
  else
    hipLaunchKernelGGL(k_count_pairs, dim3(nblocks), dim3(64), 0, st, (float)(a.rmax * a.rmax),
                       a.excl_span, a.lattice, a.tile_posq, a.tile_key, a.entry_atoms,
                       a.entry_count, a.cap, a.n, out);
}

template <int EV, int NT, bool SW>
static void launch_pair_t(hipStream_t st, const HdPairArgs& a, const PairParams& P) {
  unsigned nblocks = a.npad / 64;
  if (a.atom_lists) {
    dim3 grid(nblocks);
    if (P.ortho)
      hipLaunchKernelGGL((k_pair_alist<EV, NT, SW, true>), grid, dim3(64 * PAIR_WAVES), 0, st, P,
                         a.lattice, a.tile_posq, a.tile_type, a.entry_atoms, a.entry_count,
                         a.cap, a.n, nblocks, a.fpi, a.fpi_ew, a.lj_tab);
    else
      hipLaunchKernelGGL((k_pair_alist<EV, NT, SW, false>), grid, dim3(64 * PAIR_WAVES), 0, st,
                         P, a.lattice, a.tile_posq, a.tile_type, a.entry_atoms, a.entry_count,
                         a.cap, a.n, nblocks, a.fpi, a.fpi_ew, a.lj_tab);
    return;
  }
  hipLaunchKernelGGL((k_pair<EV, NT, SW>), dim3(nblocks), dim3(64 * PAIR_WAVES), 0, st, P,
                     a.lattice, a.tile_posq, a.tile_type, a.tile_key, a.center, a.half,
                     a.entry_atoms, a.entry_count, a.block_order, a.cap, a.n, a.fpi, a.fpi_ew,
                     (unsigned long long*)a.fpj, (unsigned long long*)a.fpj_ew, a.lj_tab);
}

template <int EV, int NT>
static void launch_pair_nt(hipStream_t st, const HdPairArgs& a, const PairParams& P, bool sw) {
  if (sw) launch_pair_t<EV, NT, true>(st, a, P);
  else launch_pair_t<EV, NT, false>(st, a, P);
}

template <int EV>
static void launch_pair_ev(hipStream_t st, const HdPairArgs& a, const PairParams& P, bool sw) {
  if (P.ntypes <= 1) launch_pair_nt<EV, 1>(st, a, P, sw);
  else if (P.ntypes == 2) launch_pair_nt<EV, 2>(st, a, P, sw);
  else if (P.ntypes <= PAIR_MAXT) launch_pair_nt<EV, 4>(st, a, P, sw);
  else launch_pair_nt<EV, 0>(st, a, P, sw);
}

// f32 coefficients of one type pair, as every pair kernel form uses them
struct PairCoef {
  float c6, c12, cut2, rsw2, idl;
  bool sw;  // a switch region on an interacting pair
};
static PairCoef pair_coef(double s, double e, double rc, double rs) {
  PairCoef q;
  double s6 = s * s * s * s * s * s;
  q.c6 = (float)(4.0 * e * s6);
  q.c12 = (float)(4.0 * e * s6 * s6);
  q.cut2 = (e != 0.0) ? (float)(rc * rc) : -1.0f;
  q.rsw2 = (float)(rs * rs);
  double d = rc * rc - rs * rs;
  q.idl = d > 0.0 ? (float)(1.0 / d) : 0.0f;
  q.sw = d > 0.0 && e != 0.0;
  return q;
}

int hd_pair_table_max_types() { return PAIR_TAB_TYPES; }

void hd_pair_lj_table(int nt, const double* sigma, const double* eps, const double* cut,
                      const double* rsw, float* out) {
  for (int k = 0; k < nt * nt; ++k) {
    PairCoef q = pair_coef(sigma[k], eps[k], cut[k], rsw[k]);
    float* o = out + 8 * (size_t)k;
    o[0] = 12.0f * q.c12; o[1] = 6.0f * q.c6; o[2] = q.c12; o[3] = q.c6;
    o[4] = q.cut2; o[5] = q.rsw2; o[6] = q.idl; o[7] = 0.0f;
  }
}

void hd_launch_pair(hipStream_t st, const HdPairArgs& a, int ev) {
  PairParams P;
  memset(&P, 0, sizeof(P));
  P.ntypes = a.ntypes;
  bool any_switch = false;
  const int nt2 = a.ntypes * a.ntypes;
  PairCoef q[PAIR_TAB_TYPES * PAIR_TAB_TYPES];
  for (int k = 0; k < nt2; ++k) {
    q[k] = pair_coef(a.sigma[k], a.eps[k], a.cut[k], a.rsw[k]);
    if (q[k].sw) any_switch = true;
  }
  // up to four types the rows travel in the kernel arguments; beyond that
  // the kernels read a.lj_tab (hd_pair_lj_table)
  if (a.ntypes <= PAIR_MAXT)
    for (int k = 0; k < nt2; ++k) {
      P.c6[k] = q[k].c6;
      P.c12[k] = q[k].c12;
      P.ljcut2[k] = q[k].cut2;
      P.rsw2[k] = q[k].rsw2;
      P.inv_delta[k] = q[k].idl;
    }
  P.kc = (float)a.kc;
  P.alpha = (float)a.alpha;
  P.coul_rc2 = a.coul_rc > 0.0 ? (float)(a.coul_rc * a.coul_rc) : -1.0f;
  P.rmax2 = (float)(a.rmax * a.rmax);
  P.spc_margin = (float)(a.rmax + 0.5 * a.r_skin);
  P.excl_span = a.excl_span;
  P.ortho = a.ortho;
  bool uniform = P.coul_rc2 > 0.f && !any_switch;
  for (int k = 0; k < nt2; ++k)
    if (!(q[k].c6 == 0.f && q[k].c12 == 0.f) && q[k].cut2 != P.coul_rc2) uniform = false;
  // LJ only: the interacting type pairs share one cutoff
  float ljc = -1.f;
  bool lj_uniform = !(P.coul_rc2 > 0.f) && !any_switch;
  for (int k = 0; k < nt2; ++k) {
    if (q[k].c6 == 0.f && q[k].c12 == 0.f) continue;
    if (ljc < 0.f) ljc = q[k].cut2;
    else if (q[k].cut2 != ljc) lj_uniform = false;
  }
  if (!(ljc > 0.f)) lj_uniform = false;
  P.uniform = uniform ? 1 : (lj_uniform ? 2 : 0);
  P.uc_rc2 = uniform ? P.coul_rc2 : ljc;
  P.cg = (float)(-a.alpha * a.alpha * 1.4426950408889634);
  P.pa = (float)(0.3275911 * a.alpha);
  if (ev == 1) launch_pair_ev<1>(st, a, P, any_switch);
  else if (ev == 2) launch_pair_ev<2>(st, a, P, any_switch);
  else launch_pair_ev<0>(st, a, P, any_switch);
}
