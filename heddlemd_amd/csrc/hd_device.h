// Device-side helpers shared by every HeddleMD HIP kernel (gfx950, wave64).
//
// Conventions (all kernels):
//   * Hartree atomic units; k_B = 1, k_C = 1.
//   * Lattice [lx, ly, lz, xy, xz, yz] (lower-triangular box), kept on the
//     device as float[6] so that a barostat can rescale it without a host
//     round trip.
//   * posq float4 (x, y, z, q) per atom; velocities/forces are SoA float.
//   * "slot" = position of an atom in the spatially sorted order built by
//     the neighbour list (sorted_id[slot] = atom id).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD_WAVE 64
#define HD_SENTINEL 0xFFFFFFFFu

struct HdLattice {
  float lx, ly, lz, xy, xz, yz;
};

// Workgroup index remap for the 8 XCDs of MI355X, which receive
// workgroups in turn (b to XCD b % 8): the result gives XCD x the
// contiguous range [x * per, (x + 1) * per) with per = nb / 8, and leaves
// the nb % 8 tail in place.  A bijection on [0, nb).
__device__ __forceinline__ unsigned hd_xcd_block(unsigned b, unsigned nb) {
  unsigned per = nb >> 3;
  return b < (per << 3) ? (b & 7u) * per + (b >> 3) : b;
}

__device__ __forceinline__ HdLattice hd_load_lattice(const float* L) {
  HdLattice l;
  l.lx = L[0]; l.ly = L[1]; l.lz = L[2];
  l.xy = L[3]; l.xz = L[4]; l.yz = L[5];
  return l;
}

// Cartesian -> fractional by back substitution (same operation order as
// the reference, kernels/pbc.cuh:18-27).
__device__ __forceinline__ void hd_cart_to_frac(float x, float y, float z, const HdLattice& L,
                                                float& sa, float& sb, float& sc) {
  sc = z / L.lz;
  sb = (y - sc * L.yz) / L.ly;
  sa = (x - sb * L.xy - sc * L.xz) / L.lx;
}

// In-place minimum image (kernels/pbc.cuh:30-43).
__device__ __forceinline__ void hd_min_image(float& dx, float& dy, float& dz, const HdLattice& L) {
  float sa, sb, sc;
  hd_cart_to_frac(dx, dy, dz, L, sa, sb, sc);
  float ka = floorf(sa + 0.5f), kb = floorf(sb + 0.5f), kc = floorf(sc + 0.5f);
  dx -= ka * L.lx + kb * L.xy + kc * L.xz;
  dy -= kb * L.ly + kc * L.yz;
  dz -= kc * L.lz;
}

// The same minimum image in FP64 (the lattice's f32 values widened): exact
// differences of f32 coordinates, for the constraint solvers.
__device__ __forceinline__ void hd_min_image_f64(double& dx, double& dy, double& dz,
                                                 const HdLattice& L) {
  double lx = L.lx, ly = L.ly, lz = L.lz, xy = L.xy, xz = L.xz, yz = L.yz;
  double sc = dz / lz;
  double sb = (dy - sc * yz) / ly;
  double sa = (dx - sb * xy - sc * xz) / lx;
  double ka = floor(sa + 0.5), kb = floor(sb + 0.5), kc = floor(sc + 0.5);
  dx -= ka * lx + kb * xy + kc * xz;
  dy -= kb * ly + kc * yz;
  dz -= kc * lz;
}

// Minimum image with the reciprocal box lengths precomputed (inv = 1/lx,
// 1/ly, 1/lz).  It differs from hd_min_image only by the rounding of the
// fractional coordinate, i.e. only for separations within an ulp of half a
// box length, which lie far beyond any cutoff (setup requires r_search <
// L/2).  Used by the pair kernel's general-cell path.
__device__ __forceinline__ void hd_min_image_rcp(float& dx, float& dy, float& dz,
                                                 const HdLattice& L, float ilx, float ily,
                                                 float ilz) {
  float sc = dz * ilz;
  float sb = (dy - sc * L.yz) * ily;
  float sa = (dx - sb * L.xy - sc * L.xz) * ilx;
  float ka = floorf(sa + 0.5f), kb = floorf(sb + 0.5f), kc = floorf(sc + 0.5f);
  dx -= ka * L.lx + kb * L.xy + kc * L.xz;
  dy -= kb * L.ly + kc * L.yz;
  dz -= kc * L.lz;
}

// In-place wrap into the primary cell returning the image triple
// (kernels/pbc.cuh:74-92).
__device__ __forceinline__ void hd_wrap_with_image(float& x, float& y, float& z, int& ka_o,
                                                   int& kb_o, int& kc_o, const HdLattice& L) {
  float sa, sb, sc;
  hd_cart_to_frac(x, y, z, L, sa, sb, sc);
  float ka = floorf(sa + 0.5f), kb = floorf(sb + 0.5f), kc = floorf(sc + 0.5f);
  x -= ka * L.lx + kb * L.xy + kc * L.xz;
  y -= kb * L.ly + kc * L.yz;
  z -= kc * L.lz;
  ka_o = (int)ka;
  kb_o = (int)kb;
  kc_o = (int)kc;
}

// Shift (x,y,z) to the periodic image nearest the centre (cx,cy,cz)
// (kernels/pbc.cuh:54-72).
__device__ __forceinline__ void hd_wrap_against_center(float& x, float& y, float& z, float cx,
                                                       float cy, float cz, const HdLattice& L) {
  float dx = x - cx, dy = y - cy, dz = z - cz;
  float sa, sb, sc;
  hd_cart_to_frac(dx, dy, dz, L, sa, sb, sc);
  float ka = floorf(sa + 0.5f), kb = floorf(sb + 0.5f), kc = floorf(sc + 0.5f);
  x -= ka * L.lx + kb * L.xy + kc * L.xz;
  y -= kb * L.ly + kc * L.yz;
  z -= kc * L.lz;
}

// Position relative to a centre, shifted to the nearest periodic image,
// evaluated in FP64 and rounded once: the result is small (|d| <~ r_cut +
// block extent), so it keeps ~1e-7 relative precision where the f32
// difference of two wrapped absolute coordinates would carry the rounding
// error of the large absolute value.  Orthorhombic cells only.
__device__ __forceinline__ void hd_rel_to_center(float x, float y, float z, float cx, float cy,
                                                 float cz, const HdLattice& L, float& rx,
                                                 float& ry, float& rz) {
  double dx = (double)x - (double)cx, dy = (double)y - (double)cy, dz = (double)z - (double)cz;
  double lx = L.lx, ly = L.ly, lz = L.lz;
  dx -= lx * floor(dx / lx + 0.5);
  dy -= ly * floor(dy / ly + 0.5);
  dz -= lz * floor(dz / lz + 0.5);
  rx = (float)dx;
  ry = (float)dy;
  rz = (float)dz;
}

// The same with the reciprocal box lengths (FP64) precomputed: a multiply
// instead of a correctly rounded FP64 division per axis.  The nearest image
// can differ from hd_rel_to_center only for a separation within an ulp of
// half a box length, far beyond any search radius.
__device__ __forceinline__ void hd_rel_to_center_rcp(float x, float y, float z, float cx,
                                                     float cy, float cz, const HdLattice& L,
                                                     double ilx, double ily, double ilz,
                                                     float& rx, float& ry, float& rz) {
  double dx = (double)x - (double)cx, dy = (double)y - (double)cy, dz = (double)z - (double)cz;
  dx -= (double)L.lx * floor(dx * ilx + 0.5);
  dy -= (double)L.ly * floor(dy * ily + 0.5);
  dz -= (double)L.lz * floor(dz * ilz + 0.5);
  rx = (float)dx;
  ry = (float)dy;
  rz = (float)dz;
}

// SPME spline origin of one atom: wrap into the primary cell, fractional
// coordinate + 1/2, u = s * n, primary bin floor(u) (mod n) and offset
// t = u - floor(u) (the geometry of kernels/spme_spread_gather.cuh:90-126).
// Evaluated in FP64 and rounded once, so the offset t keeps ~1e-7
// precision on 128^3 grids (f32 would lose log2(n) bits of it).
__device__ __forceinline__ void hd_spline_origin(float x, float y, float z, const HdLattice& L,
                                                 unsigned na, unsigned nb, unsigned nc,
                                                 float& ta, float& tb, float& tc, int& ga,
                                                 int& gb, int& gc) {
  double sc = (double)z / (double)L.lz;
  double sb = ((double)y - sc * (double)L.yz) / (double)L.ly;
  double sa = ((double)x - sb * (double)L.xy - sc * (double)L.xz) / (double)L.lx;
  sa -= floor(sa + 0.5);
  sb -= floor(sb + 0.5);
  sc -= floor(sc + 0.5);
  double ua = (sa + 0.5) * (double)na, ub = (sb + 0.5) * (double)nb, uc = (sc + 0.5) * (double)nc;
  double fa = floor(ua), fb = floor(ub), fc = floor(uc);
  ta = (float)(ua - fa);
  tb = (float)(ub - fb);
  tc = (float)(uc - fc);
  int la = (int)fa, lb = (int)fb, lc = (int)fc;
  ga = ((la % (int)na) + (int)na) % (int)na;
  gb = ((lb % (int)nb) + (int)nb) % (int)nb;
  gc = ((lc % (int)nc) + (int)nc) % (int)nc;
}

// Per-slot SPME spline record of a gathered position: (t_a, t_b, t_c, q)
// and the packed origin cell g_a | g_b << 10 | g_c << 20, or HD_SPME_NO_CELL
// for an uncharged atom (nothing written to spl then).
#define HD_SPME_NO_CELL 0xFFFFFFFFu
__device__ __forceinline__ void hd_spline_record(float4 pq, const HdLattice& L, unsigned na,
                                                 unsigned nb, unsigned nc, float4* spl,
                                                 unsigned* sg, unsigned s) {
  if (pq.w == 0.f) {
    sg[s] = HD_SPME_NO_CELL;
    return;
  }
  float ta, tb, tc;
  int ga, gb, gc;
  hd_spline_origin(pq.x, pq.y, pq.z, L, na, nb, nc, ta, tb, tc, ga, gb, gc);
  spl[s] = make_float4(ta, tb, tc, pq.w);
  sg[s] = (unsigned)ga | ((unsigned)gb << 10) | ((unsigned)gc << 20);
}

// Wave64 rotation through DPP: every lane receives the value held by lane
// (lane + 1) mod 64 ("wave_rol:1"; a full-rate VALU move, no LDS traffic).
__device__ __forceinline__ int hd_rot1_i(int v) {
  // every lane is written (no row or bank masked, no lane out of range), so
  // the "old" operand is never observed and bound_ctrl is irrelevant
  return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, true);
}
__device__ __forceinline__ float hd_rot1(float v) {
  return __int_as_float(hd_rot1_i(__float_as_int(v)));
}
__device__ __forceinline__ unsigned hd_rot1_u(unsigned v) {
  return (unsigned)hd_rot1_i((int)v);
}

// f32 -> i64 fixed point at scale 2^48, truncating toward zero; the scale
// is applied as two exact power-of-two products so no intermediate
// rounding occurs (src/forces/jit_composed.rs:1462-1475).
__device__ __forceinline__ long long hd_to_fixed48(float f) {
  float s = f * 16777216.0f;
  s *= 16777216.0f;
  return (long long)s;
}
// i64 -> f32 (round to nearest) then * 2^-48 (kernels/neighbor.cu:744-758).
__device__ __forceinline__ float hd_from_fixed48(long long v) {
  const float inv = (1.0f / 16777216.0f) / 16777216.0f;
  return (float)v * inv;
}

__device__ __forceinline__ unsigned long long hd_lane_mask_lt() {
  unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}
