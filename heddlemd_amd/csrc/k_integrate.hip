// Integration, constraints, reductions and the force finalize on gfx950.
//
// Arithmetic in this file is kept operation-for-operation identical to the
// reference's single-precision kernels (no FMA contraction), so the f32
// velocity-Verlet update is bit-exact against a scalar restatement:
//   * velocity Verlet kick/drift, f32 and compensated (f32, f64)
//     (kernels/integrate.cu:24-263);
//   * SETTLE positions / velocities (kernels/settle.cu:145-497);
//   * Nose-Hoover chain (src/integrator/nose_hoover_chain.rs:109-279), here
//     advanced on the device in FP64 so the step needs no host round trip;
//   * the post-force kick fragment and thermostat factor
//     (src/integrator/velocity_verlet.rs:75-124,
//      src/integrator/nose_hoover_chain.rs:282-300).
// Reductions (kinetic energy, potential energy, virial) are FP64 with a
// fixed chunking and a fixed pairwise tree, so they are bit-identical run
// to run (the reference's f32 trees: kernels/nose_hoover.cu:22-105,
// kernels/barostat.cu:11-78).
// (before the includes, so the shared helpers compile without contraction
// in this translation unit too)
#pragma clang fp contract(off)

#include "hd_device.h"
#include "hd_kernels.h"

template <bool LOSSLESS, bool SNAP>
__global__ void k_kick_drift(HdIntegrateArgs a) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  HdLattice L = hd_load_lattice(a.lattice);
  float f = a.factor ? a.factor[0] : 1.0f;
  float m = a.mass[i];
  float ax = a.fx[i] / m, ay = a.fy[i] / m, az = a.fz[i] / m;
  float half_dt = a.dt * 0.5f;
  float4 pq = a.posq[i];
  if (SNAP) a.snap[i] = pq;
  float vx = a.vx[i] * f, vy = a.vy[i] * f, vz = a.vz[i] * f;
  int nx = a.ix[i], ny = a.iy[i], nz = a.iz[i];
  float px, py, pz;
  if (LOSSLESS) {
    double ex = (double)vx + a.vxlo[i] + (double)(ax * half_dt);
    double ey = (double)vy + a.vylo[i] + (double)(ay * half_dt);
    double ez = (double)vz + a.vzlo[i] + (double)(az * half_dt);
    vx = (float)ex; vy = (float)ey; vz = (float)ez;
    double lx = ex - (double)vx, ly = ey - (double)vy, lz = ez - (double)vz;
    a.vxlo[i] = lx; a.vylo[i] = ly; a.vzlo[i] = lz;
    double dx = ((double)vx + lx) * (double)a.dt;
    double dy = ((double)vy + ly) * (double)a.dt;
    double dz = ((double)vz + lz) * (double)a.dt;
    double qx = (double)pq.x + a.xlo[i] + dx;
    double qy = (double)pq.y + a.ylo[i] + dy;
    double qz = (double)pq.z + a.zlo[i] + dz;
    px = (float)qx; py = (float)qy; pz = (float)qz;
    a.xlo[i] = qx - (double)px; a.ylo[i] = qy - (double)py; a.zlo[i] = qz - (double)pz;
  } else {
    vx = vx + ax * half_dt;
    vy = vy + ay * half_dt;
    vz = vz + az * half_dt;
    px = pq.x + vx * a.dt;
    py = pq.y + vy * a.dt;
    pz = pq.z + vz * a.dt;
  }
  int ka, kb, kc;
  hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
  a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
  a.posq[i] = make_float4(px, py, pz, pq.w);
  a.ix[i] = nx + ka; a.iy[i] = ny + kb; a.iz[i] = nz + kc;
}

void hd_launch_kick_drift(hipStream_t st, const HdIntegrateArgs& a) {
  if (a.n == 0) return;
  dim3 g((a.n + 255) / 256), b(256);
  bool ll = a.xlo != nullptr, sn = a.snap != nullptr;
  if (ll && sn) hipLaunchKernelGGL((k_kick_drift<true, true>), g, b, 0, st, a);
  else if (ll) hipLaunchKernelGGL((k_kick_drift<true, false>), g, b, 0, st, a);
  else if (sn) hipLaunchKernelGGL((k_kick_drift<false, true>), g, b, 0, st, a);
  else hipLaunchKernelGGL((k_kick_drift<false, false>), g, b, 0, st, a);
}

// ---------------------------------------------------------------------
// SETTLE
// ---------------------------------------------------------------------
__device__ __forceinline__ float settle_csqrt(float x) { return sqrtf(x > 0.f ? x : 0.f); }

// SETTLE of one water on positions: p the unconstrained positions, q the
// snapshot before the drift; the velocities are read from memory, or with
// VREG taken from vr (the values the drift has just computed).  The new
// positions, velocities and constraint virials are stored, and also
// returned in po, vo, wo; without VSTORE the velocities and virials are
// only returned (a velocity stage that follows stores them)
template <bool VREG, bool VSTORE>
__device__ __forceinline__ void settle_group(const HdSettleArgs& a, const HdLattice& L,
                                             const unsigned (&idx)[3], const float4 (&p)[3],
                                             const float4 (&q)[3], const float (&vr)[3][3],
                                             float4 (&po)[3], float (&vo)[3][3],
                                             float (&wo)[3]) {
  float cur[3][3] = {{p[0].x, p[0].y, p[0].z}, {p[1].x, p[1].y, p[1].z}, {p[2].x, p[2].y, p[2].z}};
  float snp[3][3] = {{q[0].x, q[0].y, q[0].z}, {q[1].x, q[1].y, q[1].z}, {q[2].x, q[2].y, q[2].z}};
  float xb0 = snp[1][0] - snp[0][0], yb0 = snp[1][1] - snp[0][1], zb0 = snp[1][2] - snp[0][2];
  float xc0 = snp[2][0] - snp[0][0], yc0 = snp[2][1] - snp[0][1], zc0 = snp[2][2] - snp[0][2];
  hd_min_image(xb0, yb0, zb0, L);
  hd_min_image(xc0, yc0, zc0, L);
  float xp[3][3];
  for (int k = 0; k < 3; ++k) {
    float dx = cur[k][0] - snp[k][0], dy = cur[k][1] - snp[k][1], dz = cur[k][2] - snp[k][2];
    hd_min_image(dx, dy, dz, L);
    xp[k][0] = dx; xp[k][1] = dy; xp[k][2] = dz;
  }
  float m0 = a.m_o, m1 = a.m_h, m2 = a.m_h;
  float itm = 1.0f / (m0 + m1 + m2);
  float ra = a.ra, rb = a.rb, rc = a.rc;
  float d2 = 2.0f * rc, d2sq = d2 * d2;
  float xcom = (xp[0][0] * m0 + (xb0 + xp[1][0]) * m1 + (xc0 + xp[2][0]) * m2) * itm;
  float ycom = (xp[0][1] * m0 + (yb0 + xp[1][1]) * m1 + (yc0 + xp[2][1]) * m2) * itm;
  float zcom = (xp[0][2] * m0 + (zb0 + xp[1][2]) * m1 + (zc0 + xp[2][2]) * m2) * itm;
  float xa1 = xp[0][0] - xcom, ya1 = xp[0][1] - ycom, za1 = xp[0][2] - zcom;
  float xb1 = xb0 + xp[1][0] - xcom, yb1 = yb0 + xp[1][1] - ycom, zb1 = zb0 + xp[1][2] - zcom;
  float xc1 = xc0 + xp[2][0] - xcom, yc1 = yc0 + xp[2][1] - ycom, zc1 = zc0 + xp[2][2] - zcom;
  float ez_x = yb0 * zc0 - zb0 * yc0, ez_y = zb0 * xc0 - xb0 * zc0, ez_z = xb0 * yc0 - yb0 * xc0;
  float ex_x = ya1 * ez_z - za1 * ez_y, ex_y = za1 * ez_x - xa1 * ez_z, ex_z = xa1 * ez_y - ya1 * ez_x;
  float ey_x = ez_y * ex_z - ez_z * ex_y, ey_y = ez_z * ex_x - ez_x * ex_z, ey_z = ez_x * ex_y - ez_y * ex_x;
  float exl = settle_csqrt(ex_x * ex_x + ex_y * ex_y + ex_z * ex_z);
  float eyl = settle_csqrt(ey_x * ey_x + ey_y * ey_y + ey_z * ey_z);
  float ezl = settle_csqrt(ez_x * ez_x + ez_y * ez_y + ez_z * ez_z);
  float iex = exl > 0.f ? 1.0f / exl : 0.f, iey = eyl > 0.f ? 1.0f / eyl : 0.f;
  float iez = ezl > 0.f ? 1.0f / ezl : 0.f;
  float r11 = ex_x * iex, r21 = ex_y * iex, r31 = ex_z * iex;
  float r12 = ey_x * iey, r22 = ey_y * iey, r32 = ey_z * iey;
  float r13 = ez_x * iez, r23 = ez_y * iez, r33 = ez_z * iez;
  float b0x = r11 * xb0 + r21 * yb0 + r31 * zb0, b0y = r12 * xb0 + r22 * yb0 + r32 * zb0;
  float c0x = r11 * xc0 + r21 * yc0 + r31 * zc0, c0y = r12 * xc0 + r22 * yc0 + r32 * zc0;
  float a1z = r13 * xa1 + r23 * ya1 + r33 * za1;
  float b1x = r11 * xb1 + r21 * yb1 + r31 * zb1, b1y = r12 * xb1 + r22 * yb1 + r32 * zb1;
  float b1z = r13 * xb1 + r23 * yb1 + r33 * zb1;
  float c1x = r11 * xc1 + r21 * yc1 + r31 * zc1, c1y = r12 * xc1 + r22 * yc1 + r32 * zc1;
  float c1z = r13 * xc1 + r23 * yc1 + r33 * zc1;
  float sinphi = a1z / ra;
  float cosphi = settle_csqrt(1.0f - sinphi * sinphi);
  float sinpsi = (b1z - c1z) / (2.0f * rc * cosphi);
  float cospsi = settle_csqrt(1.0f - sinpsi * sinpsi);
  float a2y = ra * cosphi;
  float b2x = -rc * cospsi;
  float b2y = -rb * cosphi - rc * sinpsi * sinphi;
  float c2y = -rb * cosphi + rc * sinpsi * sinphi;
  float b2xsq = b2x * b2x;
  float hhsq = 4.0f * b2xsq + (b2y - c2y) * (b2y - c2y) + (b1z - c1z) * (b1z - c1z);
  float delta_x = 2.0f * b2x + settle_csqrt(4.0f * b2xsq - hhsq + d2sq);
  b2x -= delta_x * 0.5f;
  float al = b2x * (b0x - c0x) + b0y * b2y + c0y * c2y;
  float be = b2x * (c0y - b0y) + b0x * b2y + c0x * c2y;
  float ga = b0x * b1y - b1x * b0y + c0x * c1y - c1x * c0y;
  float ab2 = al * al + be * be;
  float iab2 = ab2 > 0.f ? 1.0f / ab2 : 0.f;
  float sinth = (al * ga - be * settle_csqrt(ab2 - ga * ga)) * iab2;
  float costh = settle_csqrt(1.0f - sinth * sinth);
  float a3x = -a2y * sinth, a3y = a2y * costh, a3z = a1z;
  float b3x = b2x * costh - b2y * sinth, b3y = b2x * sinth + b2y * costh, b3z = b1z;
  float c3x = -b2x * costh - c2y * sinth, c3y = -b2x * sinth + c2y * costh, c3z = c1z;
  float rcom[3][3] = {
      {r11 * a3x + r12 * a3y + r13 * a3z, r21 * a3x + r22 * a3y + r23 * a3z, r31 * a3x + r32 * a3y + r33 * a3z},
      {r11 * b3x + r12 * b3y + r13 * b3z, r21 * b3x + r22 * b3y + r23 * b3z, r31 * b3x + r32 * b3y + r33 * b3z},
      {r11 * c3x + r12 * c3y + r13 * c3z, r21 * c3x + r22 * c3y + r23 * c3z, r31 * c3x + r32 * c3y + r33 * c3z}};
  float com[3] = {xcom, ycom, zcom};
  float unc[3][3] = {{xp[0][0], xp[0][1], xp[0][2]},
                     {xb0 + xp[1][0], yb0 + xp[1][1], zb0 + xp[1][2]},
                     {xc0 + xp[2][0], yc0 + xp[2][1], zc0 + xp[2][2]}};
  float mat[3] = {m0, m1, m2};
  float idt = a.dt != 0.f ? 1.0f / a.dt : 0.f;
  float idt2 = a.dt != 0.f ? 1.0f / (a.dt * a.dt) : 0.f;
  for (int k = 0; k < 3; ++k) {
    float c0 = (com[0] + rcom[k][0]) - unc[k][0];
    float c1 = (com[1] + rcom[k][1]) - unc[k][1];
    float c2 = (com[2] + rcom[k][2]) - unc[k][2];
    unsigned id = idx[k];
    float vx = VREG ? vr[k][0] : a.vx[id];
    float vy = VREG ? vr[k][1] : a.vy[id];
    float vz = VREG ? vr[k][2] : a.vz[id];
    vo[k][0] = vx + c0 * idt;
    vo[k][1] = vy + c1 * idt;
    vo[k][2] = vz + c2 * idt;
    po[k] = make_float4(cur[k][0] + c0, cur[k][1] + c1, cur[k][2] + c2, p[k].w);
    wo[k] = mat[k] * idt2 * (c0 * rcom[k][0] + c1 * rcom[k][1] + c2 * rcom[k][2]);
    a.posq[id] = po[k];
    if (VSTORE) {
      a.vx[id] = vo[k][0];
      a.vy[id] = vo[k][1];
      a.vz[id] = vo[k][2];
      a.w_cons[id] = wo[k];
    }
  }
}

__global__ void k_settle_positions(HdSettleArgs a) {
  unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_groups) return;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned idx[3] = {a.groups[3 * g], a.groups[3 * g + 1], a.groups[3 * g + 2]};
  float4 p[3] = {a.posq[idx[0]], a.posq[idx[1]], a.posq[idx[2]]};
  float4 q[3] = {a.snap[idx[0]], a.snap[idx[1]], a.snap[idx[2]]};
  float vr[3][3] = {}, vo[3][3], wo[3];
  float4 po[3];
  settle_group<false, true>(a, L, idx, p, q, vr, po, vo, wo);
}

template <bool ACC, bool REG>
__device__ __forceinline__ void settle_vel_group(const HdSettleArgs& a, const HdLattice& L,
                                                 const unsigned (&ids)[3],
                                                 const float4 (&pin)[3],
                                                 const float (&vin)[3][3], const float (&win)[3]);

// VV rescale + kick + drift + wrap fused with SETTLE on positions and then
// on velocities (with the constraint virial), for a system whose atoms all
// belong to SETTLE groups (no SHAKE, f32 state): one thread per water runs
// k_kick_drift<false, true>'s arithmetic on its three atoms, then
// k_settle_positions' and k_settle_velocities<true>'s on the values held
// in registers, so the intermediate positions, velocities, virials and the
// snapshot skip the round trips through memory.  The same operations in
// the same order as the three kernels: bit-identical results.
__global__ void k_kick_drift_settle(HdIntegrateArgs ia, HdSettleArgs a) {
  unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_groups) return;
  HdLattice L = hd_load_lattice(a.lattice);
  float f = ia.factor ? ia.factor[0] : 1.0f;
  float half_dt = ia.dt * 0.5f;
  unsigned idx[3] = {a.groups[3 * g], a.groups[3 * g + 1], a.groups[3 * g + 2]};
  float4 p[3], q[3];
  float vr[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsigned i = idx[k];
    float m = ia.mass[i];
    float ax = ia.fx[i] / m, ay = ia.fy[i] / m, az = ia.fz[i] / m;
    float4 pq = ia.posq[i];
    q[k] = pq;
    float vx = ia.vx[i] * f, vy = ia.vy[i] * f, vz = ia.vz[i] * f;
    int nx = ia.ix[i], ny = ia.iy[i], nz = ia.iz[i];
    vx = vx + ax * half_dt;
    vy = vy + ay * half_dt;
    vz = vz + az * half_dt;
    float px = pq.x + vx * ia.dt;
    float py = pq.y + vy * ia.dt;
    float pz = pq.z + vz * ia.dt;
    int ka, kb, kc;
    hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
    ia.ix[i] = nx + ka; ia.iy[i] = ny + kb; ia.iz[i] = nz + kc;
    p[k] = make_float4(px, py, pz, pq.w);
    vr[k][0] = vx; vr[k][1] = vy; vr[k][2] = vz;
  }
  float4 po[3];
  float vo[3][3], wo[3];
  settle_group<true, false>(a, L, idx, p, q, vr, po, vo, wo);
  settle_vel_group<true, true>(a, L, idx, po, vo, wo);
}

void hd_launch_kick_drift_settle(hipStream_t st, const HdIntegrateArgs& ia,
                                 const HdSettleArgs& a) {
  if (!a.n_groups) return;
  hipLaunchKernelGGL(k_kick_drift_settle, dim3((a.n_groups + 127) / 128), dim3(128), 0, st, ia,
                     a);
}

// SETTLE of one water on velocities; with REG the positions, velocities
// and constraint virials come from pin, vin, win (the values the position
// stage has just computed) instead of memory
template <bool ACC, bool REG>
__device__ __forceinline__ void settle_vel_group(const HdSettleArgs& a, const HdLattice& L,
                                                 const unsigned (&ids)[3],
                                                 const float4 (&pin)[3],
                                                 const float (&vin)[3][3], const float (&win)[3]) {
  unsigned i0 = ids[0], i1 = ids[1], i2 = ids[2];
  float4 p0 = REG ? pin[0] : a.posq[i0], p1 = REG ? pin[1] : a.posq[i1];
  float4 p2 = REG ? pin[2] : a.posq[i2];
  float px[3] = {p0.x, p1.x, p2.x}, py[3] = {p0.y, p1.y, p2.y}, pz[3] = {p0.z, p1.z, p2.z};
  for (int k = 1; k < 3; ++k) {
    float dx = px[k] - px[0], dy = py[k] - py[0], dz = pz[k] - pz[0];
    hd_min_image(dx, dy, dz, L);
    px[k] = px[0] + dx; py[k] = py[0] + dy; pz[k] = pz[0] + dz;
  }
  float m_o = a.m_o, m_h = a.m_h;
  float io = m_o > 0.f ? 1.0f / m_o : 0.f, ih = m_h > 0.f ? 1.0f / m_h : 0.f;
  float vx[3], vy[3], vz[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    vx[k] = REG ? vin[k][0] : a.vx[ids[k]];
    vy[k] = REG ? vin[k][1] : a.vy[ids[k]];
    vz[k] = REG ? vin[k][2] : a.vz[ids[k]];
  }
  float r1x = px[0] - px[1], r1y = py[0] - py[1], r1z = pz[0] - pz[1];
  float r2x = px[0] - px[2], r2y = py[0] - py[2], r2z = pz[0] - pz[2];
  float r3x = px[1] - px[2], r3y = py[1] - py[2], r3z = pz[1] - pz[2];
  float d11 = r1x * r1x + r1y * r1y + r1z * r1z;
  float d22 = r2x * r2x + r2y * r2y + r2z * r2z;
  float d33 = r3x * r3x + r3y * r3y + r3z * r3z;
  float d12 = r1x * r2x + r1y * r2y + r1z * r2z;
  float d13 = r1x * r3x + r1y * r3y + r1z * r3z;
  float d23 = r2x * r3x + r2y * r3y + r2z * r3z;
  float b1 = (vx[0] - vx[1]) * r1x + (vy[0] - vy[1]) * r1y + (vz[0] - vz[1]) * r1z;
  float b2 = (vx[0] - vx[2]) * r2x + (vy[0] - vy[2]) * r2y + (vz[0] - vz[2]) * r2z;
  float b3 = (vx[1] - vx[2]) * r3x + (vy[1] - vy[2]) * r3y + (vz[1] - vz[2]) * r3z;
  float ioh = io + ih;
  float m[3][3] = {{d11 * ioh, d12 * io, -d13 * ih},
                   {d12 * io, d22 * ioh, d23 * ih},
                   {-d13 * ih, d23 * ih, 2.0f * d33 * ih}};
  float rhs[3] = {-b1, -b2, -b3};
  float det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
              m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
              m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  float gg[3] = {0.f, 0.f, 0.f};
  if (det != 0.f) {
    float idet = 1.0f / det;
    for (int c = 0; c < 3; ++c) {
      float t[3][3];
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) t[r][q] = (q == c) ? rhs[r] : m[r][q];
      float dc = t[0][0] * (t[1][1] * t[2][2] - t[1][2] * t[2][1]) -
                 t[0][1] * (t[1][0] * t[2][2] - t[1][2] * t[2][0]) +
                 t[0][2] * (t[1][0] * t[2][1] - t[1][1] * t[2][0]);
      gg[c] = dc * idet;
    }
  }
  float g1 = gg[0], g2 = gg[1], g3 = gg[2];
  float dv[3][3] = {
      {(g1 * r1x + g2 * r2x) * io, (g1 * r1y + g2 * r2y) * io, (g1 * r1z + g2 * r2z) * io},
      {(-g1 * r1x + g3 * r3x) * ih, (-g1 * r1y + g3 * r3y) * ih, (-g1 * r1z + g3 * r3z) * ih},
      {(-g2 * r2x - g3 * r3x) * ih, (-g2 * r2y - g3 * r3y) * ih, (-g2 * r2z - g3 * r3z) * ih}};
  for (int k = 0; k < 3; ++k) {
    a.vx[ids[k]] = vx[k] + dv[k][0];
    a.vy[ids[k]] = vy[k] + dv[k][1];
    a.vz[ids[k]] = vz[k] + dv[k][2];
  }
  if (ACC && a.dt > 0.f) {
    float idt = 1.0f / a.dt;
    float tot = m_o + 2.0f * m_h;
    float iM = tot > 0.f ? 1.0f / tot : 0.f;
    float cx = (m_o * px[0] + m_h * px[1] + m_h * px[2]) * iM;
    float cy = (m_o * py[0] + m_h * py[1] + m_h * py[2]) * iM;
    float cz = (m_o * pz[0] + m_h * pz[1] + m_h * pz[2]) * iM;
    float ma[3] = {m_o, m_h, m_h};
    for (int k = 0; k < 3; ++k) {
      float sc = ma[k] * idt;
      float w0 = REG ? win[k] : a.w_cons[ids[k]];
      a.w_cons[ids[k]] = w0 + sc * (dv[k][0] * (px[k] - cx) + dv[k][1] * (py[k] - cy) +
                                    dv[k][2] * (pz[k] - cz));
    }
  } else if (REG) {
    // the position stage left these to this stage
    for (int k = 0; k < 3; ++k) a.w_cons[ids[k]] = win[k];
  }
}

template <bool ACC>
__global__ void k_settle_velocities(HdSettleArgs a) {
  unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_groups) return;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned ids[3] = {a.groups[3 * g], a.groups[3 * g + 1], a.groups[3 * g + 2]};
  float4 pin[3] = {};
  float vin[3][3] = {}, win[3] = {};
  settle_vel_group<ACC, false>(a, L, ids, pin, vin, win);
}

void hd_launch_settle_positions(hipStream_t st, const HdSettleArgs& a) {
  if (!a.n_groups) return;
  hipLaunchKernelGGL(k_settle_positions, dim3((a.n_groups + 127) / 128), dim3(128), 0, st, a);
}

void hd_launch_settle_velocities(hipStream_t st, const HdSettleArgs& a, bool acc) {
  if (!a.n_groups) return;
  if (acc)
    hipLaunchKernelGGL(k_settle_velocities<true>, dim3((a.n_groups + 127) / 128), dim3(128), 0,
                       st, a);
  else
    hipLaunchKernelGGL(k_settle_velocities<false>, dim3((a.n_groups + 127) / 128), dim3(128), 0,
                       st, a);
}

// ---------------------------------------------------------------------
// Deterministic FP64 reductions
// ---------------------------------------------------------------------
__device__ __forceinline__ double block_tree_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (unsigned s = 1; s < blockDim.x; s <<= 1) {
    if ((threadIdx.x % (2 * s)) == 0) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// Block b owns the contiguous chunk [b*chunk, (b+1)*chunk).
__global__ void __launch_bounds__(256) k_ke_partials(unsigned n, const float* __restrict__ vx,
                                                     const float* __restrict__ vy,
                                                     const float* __restrict__ vz,
                                                     const float* __restrict__ mass,
                                                     double* __restrict__ partials,
                                                     HdDispArgs d) {
  __shared__ double sh[256];
  unsigned chunk = (n + gridDim.x - 1) / gridDim.x;
  unsigned b0 = blockIdx.x * chunk, b1 = min(b0 + chunk, n);
  double s = 0.0;
  bool over = false;
  HdLattice L;
  float dl = 0.f;
  if (d.status) {
    L = hd_load_lattice(d.lattice);
#pragma unroll
    for (int k = 0; k < 6; ++k) dl = fmaxf(dl, fabsf(d.lattice[k] - d.ref_lattice[k]));
  }
  for (unsigned i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    double x = vx[i], y = vy[i], z = vz[i];
    s += 0.5 * (double)mass[i] * (x * x + y * y + z * z);
    if (d.status) {
      // the displacement test of k_disp_check
      float4 p = d.posq[i], r = d.ref[i];
      float dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
      hd_min_image(dx, dy, dz, L);
      over |= sqrtf(dx * dx + dy * dy + dz * dz) + 1.7320508f * dl > d.half_skin;
    }
  }
  if (d.status && __any(over) && (threadIdx.x & 63) == 0) atomicOr(d.status, NB_STATUS_REBUILD);
  double t = block_tree_sum(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// One Nose-Hoover chain sub-step (nose_hoover_chain.rs:109-180), FP64.
// The forward sweep's scale for link j reads pxi[j + 1] before that link is
// updated again, i.e. the value the backward sweep left: its exp is the
// backward sweep's, reused (chain length <= 16).  M > 0 fixes the chain
// length at compile time, so the unrolled sweeps keep the chain in
// registers; M == 0 takes the runtime m.
template <int M>
__device__ __forceinline__ double nhc_sub_step(int m_rt, double* xi, double* pxi,
                                               const double* q, double dt, double k,
                                               double g_dof, double kt) {
  const int m = M > 0 ? M : m_rt;
  constexpr int UNROLL = M > 0 ? 16 : 1;  // full unroll of the fixed lengths only
  double sc[M > 0 ? M : 16];
#pragma unroll UNROLL
  for (int j = m - 1; j >= 0; --j) {
    double s = (j == m - 1) ? 1.0 : exp(-dt / 8.0 * pxi[j + 1] / q[j + 1]);
    sc[j] = s;
    pxi[j] *= s;
    double gj = (j == 0) ? k - g_dof * kt : pxi[j - 1] * pxi[j - 1] / q[j - 1] - kt;
    pxi[j] += dt / 4.0 * gj;
    pxi[j] *= s;
  }
  double factor = exp(-dt / 2.0 * pxi[0] / q[0]);
  k *= factor * factor;
#pragma unroll UNROLL
  for (int j = 0; j < m; ++j) xi[j] += dt / 2.0 * pxi[j] / q[j];
#pragma unroll UNROLL
  for (int j = 0; j < m; ++j) {
    double s = sc[j];
    pxi[j] *= s;
    double gj = (j == 0) ? k - g_dof * kt : pxi[j - 1] * pxi[j - 1] / q[j - 1] - kt;
    pxi[j] += dt / 4.0 * gj;
    pxi[j] *= s;
  }
  return factor;
}

// One chain half step: n_yoshida x n_resp sub-steps; returns the
// cumulative velocity factor
template <int M>
__device__ __forceinline__ double nhc_half_step(const HdNhcArgs& nhc, float dt, double k,
                                                double* xi, double* pxi, const double* q) {
  double cum = 1.0;
  for (int y = 0; y < nhc.n_yoshida; ++y)
    for (int r = 0; r < nhc.n_resp; ++r) {
      double delta = nhc.w[y] * (double)dt / (2.0 * nhc.n_resp);
      double f = nhc_sub_step<M>(nhc.chain, xi, pxi, q, delta, 2.0 * k, nhc.g_dof, nhc.kt);
      cum *= f;
      k *= f * f;
    }
  return cum;
}

// The same with the chain state (xi, pxi, q) copied into registers for the
// half step and xi, pxi stored back: the sweeps' loads and stores otherwise
// go to memory one after another through the state buffer
template <int M>
__device__ double nhc_half_step_local(const HdNhcArgs& nhc, float dt, double k) {
  double xi[M], pxi[M], q[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    xi[j] = nhc.state[j];
    pxi[j] = nhc.state[M + j];
    q[j] = nhc.state[2 * M + j];
  }
  double cum = nhc_half_step<M>(nhc, dt, k, xi, pxi, q);
#pragma unroll
  for (int j = 0; j < M; ++j) {
    nhc.state[j] = xi[j];
    nhc.state[M + j] = pxi[j];
  }
  return cum;
}

__global__ void __launch_bounds__(256) k_ke_final(const double* __restrict__ partials,
                                                  unsigned np, double* __restrict__ ke,
                                                  HdNhcArgs nhc, int use_nhc, float dt,
                                                  float* __restrict__ factor) {
  __shared__ double sh[256];
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < np; i += blockDim.x) s += partials[i];
  double k = block_tree_sum(s, sh);
  if (threadIdx.x != 0) return;
  ke[0] = k;
  if (!use_nhc) return;
  double cum;
  switch (nhc.chain) {
    case 1: cum = nhc_half_step_local<1>(nhc, dt, k); break;
    case 2: cum = nhc_half_step_local<2>(nhc, dt, k); break;
    case 3: cum = nhc_half_step_local<3>(nhc, dt, k); break;
    case 4: cum = nhc_half_step_local<4>(nhc, dt, k); break;
    case 5: cum = nhc_half_step_local<5>(nhc, dt, k); break;
    default: {
      int m = nhc.chain;
      cum = nhc_half_step<0>(nhc, dt, k, nhc.state, nhc.state + m, nhc.state + 2 * m);
    }
  }
  factor[0] = (float)cum;
}

void hd_launch_ke(hipStream_t st, unsigned n, const float* vx, const float* vy, const float* vz,
                  const float* mass, double* partials, double* ke, const HdNhcArgs* nhc,
                  float dt, float* factor, const HdDispArgs* disp) {
  HdDispArgs d;
  memset(&d, 0, sizeof(d));
  if (disp) d = *disp;
  hipLaunchKernelGGL(k_ke_partials, dim3(HD_REDUCE_BLOCKS), dim3(256), 0, st, n, vx, vy, vz,
                     mass, partials, d);
  HdNhcArgs na;
  memset(&na, 0, sizeof(na));
  if (nhc) na = *nhc;
  hipLaunchKernelGGL(k_ke_final, dim3(1), dim3(256), 0, st, partials, (unsigned)HD_REDUCE_BLOCKS,
                     ke, na, nhc ? 1 : 0, dt, factor);
}


__global__ void __launch_bounds__(256) k_sum2_final(const double* __restrict__ partials,
                                                    unsigned np, double* __restrict__ out) {
  __shared__ double sh[256];
  double sa = 0.0, sb = 0.0;
  for (unsigned i = threadIdx.x; i < np; i += blockDim.x) {
    sa += partials[2 * i];
    sb += partials[2 * i + 1];
  }
  double ta = block_tree_sum(sa, sh);
  __syncthreads();
  double tb = block_tree_sum(sb, sh);
  if (threadIdx.x == 0) {
    out[0] = ta;
    out[1] = tb;
  }
}

void hd_launch_sum2_final(hipStream_t st, const double* partials, unsigned np, double* out) {
  hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(256), 0, st, partials, np, out);
}


// Time reversal: v -> -v including the compensated residuals (lossless
// mode), so a following run retraces the trajectory
// (rqm/pipeline-reversibility.md, tests/integration_lossless.rs:302-395).
__global__ void k_negate_velocities(float* vx, float* vy, float* vz, double* vxlo, double* vylo,
                                    double* vzlo, unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = -vx[i];
  vy[i] = -vy[i];
  vz[i] = -vz[i];
  if (vxlo) {
    vxlo[i] = -vxlo[i];
    vylo[i] = -vylo[i];
    vzlo[i] = -vzlo[i];
  }
}
void hd_launch_negate_velocities(hipStream_t st, float* vx, float* vy, float* vz, double* vxlo,
                                 double* vylo, double* vzlo, unsigned n) {
  if (!n) return;
  hipLaunchKernelGGL(k_negate_velocities, dim3((n + 255) / 256), dim3(256), 0, st, vx, vy, vz,
                     vxlo, vylo, vzlo, n);
}

__global__ void k_fill_f32(float* p, float v, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}
void hd_launch_fill_f32(hipStream_t st, float* p, float v, size_t count) {
  if (!count) return;
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, v,
                     count);
}

// ---------------------------------------------------------------------
// Force finalize: pair fixed point + SPME gather (+ constraint virial),
// scattered back to atom order, optionally fused with the post-force kick.
// ---------------------------------------------------------------------
#define GATHER_P 4
__device__ __forceinline__ void bspline_wd(float x, float& w, float& dw) {
  // M_{p-1}(x) and M_{p-1}(x-1) by Cox-de Boor, then M_p (spme_spread_gather.cuh:56-86)
  float v[GATHER_P + 1];
#pragma unroll
  for (int i = 0; i < GATHER_P + 1; ++i) {
    float xi = x - (float)i;
    v[i] = (xi >= 0.f && xi < 1.f) ? 1.f : 0.f;
  }
#pragma unroll
  for (int order = 2; order < GATHER_P; ++order) {
    float inv = 1.0f / (float)(order - 1);
#pragma unroll
    for (int i = 0; i < GATHER_P - order + 1; ++i) {
      float xi = x - (float)i;
      v[i] = xi * inv * v[i] + ((float)order - xi) * inv * v[i + 1];
    }
  }
  dw = v[0] - v[1];
  float inv = 1.0f / (float)(GATHER_P - 1);
  w = x * inv * v[0] + ((float)GATHER_P - x) * inv * v[1];
}

template <bool EV, bool GATHER, bool KICK, bool LOSSLESS>
__global__ void __launch_bounds__(256) k_finalize(HdFinalizeArgs a) {
  // With the SPME gather four lanes share a slot: lane q of the quad reads
  // the a-plane da = q of the 4x4x4 stencil (16 grid values), so four
  // times as many reads are in flight; the quad's partial sums are combined
  // and its first lane finishes the slot.
  //
  // With a.partials set (energy or virial launches), the workgroup also sums
  // its slots' final per-atom energy and virial in FP64, in slot order, and
  // writes them to partials[2 * wg + {0, 1}]; every thread reaches that
  // reduction.
  //
  // Workgroups are handed to the 8 XCDs in turn (workgroup b to XCD b % 8);
  // wg gives each XCD one contiguous range of slot blocks (atoms close in
  // space), so the grid values a range gathers stay in that XCD's L2.
  unsigned wg = hd_xcd_block(blockIdx.x, gridDim.x);
  unsigned t = wg * blockDim.x + threadIdx.x;
  unsigned s = GATHER ? (t >> 2) : t;
  unsigned q = GATHER ? (t & 3u) : 0u;
  bool live = s < a.n;  // whole quads
  unsigned gcell = (GATHER && live) ? a.sg[s] : 0xFFFFFFFFu;
  float phi = 0.f, gra = 0.f, grb = 0.f, grc = 0.f, qi = 0.f;
  if (GATHER && gcell != 0xFFFFFFFFu) {
    // spline record of this slot (k_spline_setup): (t_a, t_b, t_c, q) and
    // the packed origin cell; uncharged atoms have none and no SPME terms
    float4 rec = a.spl[s];
    qi = rec.w;
    int ga0 = (int)(gcell & 1023u), gb0 = (int)((gcell >> 10) & 1023u), gc0 = (int)(gcell >> 20);
    float wa, dwa, wb[GATHER_P], dwb[GATHER_P], wc[GATHER_P], dwc[GATHER_P];
    int gib[GATHER_P], gic[GATHER_P];
    bspline_wd((float)q + rec.x, wa, dwa);
#pragma unroll
    for (int d = 0; d < GATHER_P; ++d) {
      bspline_wd((float)d + rec.y, wb[d], dwb[d]);
      bspline_wd((float)d + rec.z, wc[d], dwc[d]);
      // periodic wrap; the grid sides are powers of two
      gib[d] = (gb0 - d) & ((int)a.nb - 1);
      gic[d] = (gc0 - d) & ((int)a.nc - 1);
    }
    int gia = (ga0 - (int)q) & ((int)a.na - 1);
#pragma unroll
    for (int db = 0; db < GATHER_P; ++db) {
      const float* row = a.vgrid + ((size_t)gia * a.nb + gib[db]) * a.nc;
#pragma unroll
      for (int dc = 0; dc < GATHER_P; ++dc) {
        float v = row[gic[dc]];
        phi += v * wa * wb[db] * wc[dc];
        gra += v * dwa * wb[db] * wc[dc];
        grb += v * wa * dwb[db] * wc[dc];
        grc += v * wa * wb[db] * dwc[dc];
      }
    }
  }
  if (GATHER) {
    // planes 0 + 1 and 2 + 3, then the two halves
    phi += __shfl_xor(phi, 1); gra += __shfl_xor(gra, 1);
    grb += __shfl_xor(grb, 1); grc += __shfl_xor(grc, 1);
    phi += __shfl_xor(phi, 2); gra += __shfl_xor(gra, 2);
    grb += __shfl_xor(grb, 2); grc += __shfl_xor(grc, 2);
  }
  float pe = 0.f, vir = 0.f;
  if (live && q == 0) {
  unsigned id = a.sorted_id[s];
  const longlong2* fi = reinterpret_cast<const longlong2*>(a.fpi) + 2 * s;
  const longlong2* fj = reinterpret_cast<const longlong2*>(a.fpj) + 2 * s;
  longlong2 i01 = fi[0], i2x = fi[1], j01 = fj[0], j2x = fj[1];
  float fx = hd_from_fixed48(i01.x + j01.x);
  float fy = hd_from_fixed48(i01.y + j01.y);
  float fz = hd_from_fixed48(i2x.x + j2x.x);
  if (EV) {
    longlong2 ie = reinterpret_cast<const longlong2*>(a.fpi_ew)[s];
    longlong2 je = reinterpret_cast<const longlong2*>(a.fpj_ew)[s];
    pe = hd_from_fixed48(ie.x + je.x);
    vir = hd_from_fixed48(ie.y + je.y);
  }
  if (GATHER && gcell != 0xFFFFFFFFu) {
    HdLattice L = hd_load_lattice(a.lattice);
    float ilx = 1.0f / L.lx, ily = 1.0f / L.ly, ilz = 1.0f / L.lz;
    float bax = ilx, bay = -L.xy * ilx * ily, baz = (L.xy * L.yz - L.xz * L.ly) * ilx * ily * ilz;
    float bby = ily, bbz = -L.yz * ily * ilz, bcz = ilz;
    float g_a = (float)a.na * gra, g_b = (float)a.nb * grb, g_c = (float)a.nc * grc;
    float rfx = -qi * (g_a * bax);
    float rfy = -qi * (g_a * bay + g_b * bby);
    float rfz = -qi * (g_a * baz + g_b * bbz + g_c * bcz);
    fx = fx + rfx;
    fy = fy + rfy;
    fz = fz + rfz;
    if (EV) {
      float self = (float)(a.kc * a.alpha * 0.5641895835477563) * qi * qi;
      pe = pe + (0.5f * qi * phi - self);
    }
  }
  // every atom, charged or not, carries W_recip / N (spme.rs:815-823)
  if (GATHER && EV) vir = vir + a.w_share[0];
  a.fx[id] = fx;
  a.fy[id] = fy;
  a.fz[id] = fz;
  if (EV) {
    if (a.w_cons) vir = vir + a.w_cons[id];
    a.pe[id] = pe;
    a.vir[id] = vir;
  }
  if (KICK) {
    float m = a.mass[id];
    float ax = fx / m, ay = fy / m, az = fz / m;
    float half_dt = a.dt * 0.5f;
    float f = a.factor ? a.factor[0] : 1.0f;
    if (LOSSLESS) {
      double ex = (double)a.vx[id] + a.vxlo[id] + (double)(ax * half_dt);
      double ey = (double)a.vy[id] + a.vylo[id] + (double)(ay * half_dt);
      double ez = (double)a.vz[id] + a.vzlo[id] + (double)(az * half_dt);
      float nx = (float)ex, ny = (float)ey, nz = (float)ez;
      a.vxlo[id] = ex - (double)nx;
      a.vylo[id] = ey - (double)ny;
      a.vzlo[id] = ez - (double)nz;
      a.vx[id] = nx * f;
      a.vy[id] = ny * f;
      a.vz[id] = nz * f;
    } else {
      a.vx[id] = (a.vx[id] + ax * half_dt) * f;
      a.vy[id] = (a.vy[id] + ay * half_dt) * f;
      a.vz[id] = (a.vz[id] + az * half_dt) * f;
    }
  }
  }
  if (EV && a.partials) {
    __shared__ double sh[256];
    double sp = block_tree_sum((double)pe, sh);
    __syncthreads();
    double sv = block_tree_sum((double)vir, sh);
    if (threadIdx.x == 0) {
      a.partials[2 * wg] = sp;
      a.partials[2 * wg + 1] = sv;
    }
  }
}

unsigned hd_finalize_blocks(unsigned n, bool gather) {
  // four lanes per slot with the SPME gather
  return ((gather ? 4u : 1u) * n + 255) / 256;
}

void hd_launch_finalize(hipStream_t st, const HdFinalizeArgs& a, bool ev) {
  if (!a.n) return;
  bool ga = a.vgrid != nullptr, k = a.kick, ll = a.vxlo != nullptr;
  dim3 g(hd_finalize_blocks(a.n, ga)), b(256);
#define HD_FIN(E, G, K, LL) \
  if (ev == E && ga == G && k == K && ll == LL) { \
    hipLaunchKernelGGL((k_finalize<E, G, K, LL>), g, b, 0, st, a); \
    return; \
  }
  HD_FIN(false, false, false, false) HD_FIN(false, false, true, false)
  HD_FIN(false, true, false, false) HD_FIN(false, true, true, false)
  HD_FIN(true, false, false, false) HD_FIN(true, false, true, false)
  HD_FIN(true, true, false, false) HD_FIN(true, true, true, false)
  HD_FIN(false, false, true, true) HD_FIN(false, true, true, true)
  HD_FIN(true, false, true, true) HD_FIN(true, true, true, true)
  HD_FIN(false, false, false, true) HD_FIN(false, true, false, true)
  HD_FIN(true, false, false, true) HD_FIN(true, true, false, true)
#undef HD_FIN
}

// ---------------------------------------------------------------------
// Philox-4x32-10 (Salmon et al., SC11; kernels/philox.cuh:20-70) and the
// C-rescale barostat (kernels/barostat.cu:144-232).
// ---------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(unsigned k0, unsigned k1, unsigned c0, unsigned c1,
                                              unsigned c2, unsigned c3, unsigned out[4]) {
  for (int r = 0; r < 10; ++r) {
    unsigned hi0 = __umulhi(c0, 0xD2511F53u), lo0 = c0 * 0xD2511F53u;
    unsigned hi2 = __umulhi(c2, 0xCD9E8D57u), lo2 = c2 * 0xCD9E8D57u;
    unsigned n0 = hi2 ^ c1 ^ k0, n1 = lo2, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ double philox_gaussian_f64(unsigned k0, unsigned k1, unsigned c0,
                                                      unsigned c1, unsigned c2, unsigned c3) {
  unsigned o[4];
  philox4x32_10(k0, k1, c0, c1, c2, c3, o);
  const double sc = 1.0 / 4294967296.0;
  double u1 = ((double)o[0] + 0.5) * sc, u2 = ((double)o[1] + 0.5) * sc;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ---------------------------------------------------------------------
// SHAKE / RATTLE for general constraint groups (kernels/shake.cu:96-490,
// src/integrator/shake.rs): one thread per group, Gauss-Seidel sweeps in
// topology order with the reference's f32 operation order; the group's
// atoms and constraints are staged in LDS at a fixed per-thread stride,
// which lifts the reference's register-bound caps (8 atoms, 12
// constraints) to HD_SHAKE_MAX_ATOMS / HD_SHAKE_MAX_CONS so that whole
// polymer chains can be one group.
// ---------------------------------------------------------------------
unsigned hd_shake_lds_per_thread(unsigned max_atoms, unsigned max_cons) {
  return 4u * (11u * max_atoms + 6u * max_cons);
}

__device__ __forceinline__ void shake_min_image_to(float ax, float ay, float az, float& bx,
                                                   float& by, float& bz, const HdLattice& L) {
  float dx = bx - ax, dy = by - ay, dz = bz - az;
  hd_min_image(dx, dy, dz, L);
  bx = ax + dx; by = ay + dy; bz = az + dz;
}

__device__ __forceinline__ void shake_com(const float* x, const float* y, const float* z,
                                          const float* m, unsigned n, float& cx, float& cy,
                                          float& cz) {
  float total = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
  for (unsigned a = 0; a < n; ++a) {
    sx += m[a] * x[a]; sy += m[a] * y[a]; sz += m[a] * z[a];
    total += m[a];
  }
  float inv = total > 0.f ? 1.0f / total : 0.f;
  cx = sx * inv; cy = sy * inv; cz = sz * inv;
}

__device__ __forceinline__ void shake_stats(unsigned* st, unsigned it, int max_iter) {
  atomicAdd(&st[0], it);
  atomicMax(&st[1], it);
  if ((int)it >= max_iter) atomicAdd(&st[2], 1u);
}

// Groups within the reference's 8-atom cap: one thread per group, the
// reference's sequential sweep.
__global__ void k_shake_positions(HdShakeArgs a) {
  extern __shared__ float sh[];
  unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_short) return;
  unsigned g = a.slist[t];
  const unsigned MA = a.max_atoms, MC = a.max_cons;
  float* base = sh + (size_t)threadIdx.x * (11u * MA + 6u * MC);
  float *x0 = base, *y0 = x0 + MA, *z0 = y0 + MA;
  float *xu = z0 + MA, *yu = xu + MA, *zu = yu + MA;
  float *xc = zu + MA, *yc = xc + MA, *zc = yc + MA;
  float *m = zc + MA, *im = m + MA;
  float *gx = im + MA, *gy = gx + MC, *gz = gy + MC, *ip = gz + MC;
  unsigned* cij = (unsigned*)(ip + MC);
  float* lam = ip + 2u * MC;  // accumulated multiplier per constraint
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
  unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
  for (unsigned k = 0; k < na; ++k) {
    unsigned i = a.gatoms[a0 + k];
    float4 p = a.posq[i], q = a.snap[i];
    xu[k] = p.x; yu[k] = p.y; zu[k] = p.z;
    x0[k] = q.x; y0[k] = q.y; z0[k] = q.z;
    float mk = a.mass[i];
    m[k] = mk;
    im[k] = mk > 0.f ? 1.0f / mk : 0.f;
  }
  // one image for the whole group: relative to atom 0 (shake.cu:213-219)
  for (unsigned k = 1; k < na; ++k) {
    shake_min_image_to(x0[0], y0[0], z0[0], x0[k], y0[k], z0[k], L);
    shake_min_image_to(xu[0], yu[0], zu[0], xu[k], yu[k], zu[k], L);
  }
  for (unsigned k = 0; k < na; ++k) { xc[k] = xu[k]; yc[k] = yu[k]; zc[k] = zu[k]; }
  for (unsigned k = 0; k < nc; ++k) {
    unsigned pk = a.cij[c0 + k];
    unsigned li = pk & 0xFFu, lj = (pk >> 8) & 0xFFu;
    cij[k] = pk;
    gx[k] = x0[li] - x0[lj];
    gy[k] = y0[li] - y0[lj];
    gz[k] = z0[li] - z0[lj];
    ip[k] = im[li] + im[lj];
    lam[k] = 0.f;
  }
  if (a.lam0) {
    // warm start: move along every constraint gradient by a predicted
    // multiplier, in topology order, before the Gauss-Seidel sweeps
    for (unsigned k = 0; k < nc; ++k) {
      unsigned pk = cij[k];
      unsigned li = pk & 0xFFu, lj = (pk >> 8) & 0xFFu;
      float lambda = ip[k] > 0.f ? a.lam0[c0 + k] / ip[k] : 0.f;
      float ki = lambda * im[li], kj = lambda * im[lj];
      xc[li] -= ki * gx[k]; yc[li] -= ki * gy[k]; zc[li] -= ki * gz[k];
      xc[lj] += kj * gx[k]; yc[lj] += kj * gy[k]; zc[lj] += kj * gz[k];
      lam[k] = lambda;
    }
  }
  int it = 0;
  for (; it < a.max_iter; ++it) {
    bool converged = true;
    for (unsigned k = 0; k < nc; ++k) {
      unsigned pk = cij[k];
      unsigned li = pk & 0xFFu, lj = (pk >> 8) & 0xFFu;
      float dx = xc[li] - xc[lj], dy = yc[li] - yc[lj], dz = zc[li] - zc[lj];
      float dist2 = dx * dx + dy * dy + dz * dz;
      float sigma = dist2 - a.cr2[c0 + k];
      if (fabsf(sigma) > a.tol) {
        converged = false;
        float ddot = dx * gx[k] + dy * gy[k] + dz * gz[k];
        float denom = 2.0f * ddot * ip[k];
        if (denom == 0.f) continue;
        float lambda = sigma / denom;
        float ki = lambda * im[li], kj = lambda * im[lj];
        xc[li] -= ki * gx[k]; yc[li] -= ki * gy[k]; zc[li] -= ki * gz[k];
        xc[lj] += kj * gx[k]; yc[lj] += kj * gy[k]; zc[lj] += kj * gz[k];
        lam[k] += lambda;
      }
    }
    if (converged) break;
  }
  if (a.stats) shake_stats(a.stats, (unsigned)it, a.max_iter);
  if (a.lam_out)
    for (unsigned k = 0; k < nc; ++k) a.lam_out[c0 + k] = lam[k] * ip[k];
  float cxg, cyg, czg;
  shake_com(xc, yc, zc, m, na, cxg, cyg, czg);
  float inv_dt = a.dt != 0.f ? 1.0f / a.dt : 0.f;
  float inv_dt2 = a.dt != 0.f ? 1.0f / (a.dt * a.dt) : 0.f;
  for (unsigned k = 0; k < na; ++k) {
    unsigned i = a.gatoms[a0 + k];
    float dxg = xc[k] - xu[k], dyg = yc[k] - yu[k], dzg = zc[k] - zu[k];
    a.vx[i] += dxg * inv_dt; a.vy[i] += dyg * inv_dt; a.vz[i] += dzg * inv_dt;
    float4 p = a.posq[i];
    if (k == 0) { p.x = xc[0]; p.y = yc[0]; p.z = zc[0]; }
    else { p.x += dxg; p.y += dyg; p.z += dzg; }
    a.posq[i] = p;
    float rx = xc[k] - cxg, ry = yc[k] - cyg, rz = zc[k] - czg;
    float scale = m[k] * inv_dt2;
    a.w_cons[i] = scale * (dxg * rx + dyg * ry + dzg * rz);
  }
}

template <bool ACC>
__global__ void k_rattle_velocities(HdShakeArgs a) {
  extern __shared__ float sh[];
  unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_short) return;
  unsigned g = a.slist[t];
  const unsigned MA = a.max_atoms, MC = a.max_cons;
  float* base = sh + (size_t)threadIdx.x * (11u * MA + 6u * MC);
  float *px = base, *py = px + MA, *pz = py + MA;
  float *v0 = pz + MA, *v1 = v0 + MA, *v2 = v1 + MA;
  float *dv0 = v2 + MA, *dv1 = dv0 + MA, *dv2 = dv1 + MA;
  float *m = dv2 + MA, *im = m + MA;
  float *dx = im + MA, *dy = dx + MC, *dz = dy + MC, *d2 = dz + MC;
  float* ipc = d2 + MC;  // 1/m_i + 1/m_j per constraint (5th constraint slot)
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
  unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
  for (unsigned k = 0; k < na; ++k) {
    unsigned i = a.gatoms[a0 + k];
    float4 p = a.posq[i];
    px[k] = p.x; py[k] = p.y; pz[k] = p.z;
    v0[k] = a.vx[i]; v1[k] = a.vy[i]; v2[k] = a.vz[i];
    dv0[k] = dv1[k] = dv2[k] = 0.f;
    float mk = a.mass[i];
    m[k] = mk;
    im[k] = mk > 0.f ? 1.0f / mk : 0.f;
  }
  for (unsigned k = 1; k < na; ++k) shake_min_image_to(px[0], py[0], pz[0], px[k], py[k], pz[k], L);
  for (unsigned k = 0; k < nc; ++k) {
    unsigned pk = a.cij[c0 + k];
    unsigned li = pk & 0xFFu, lj = (pk >> 8) & 0xFFu;
    dx[k] = px[li] - px[lj];
    dy[k] = py[li] - py[lj];
    dz[k] = pz[li] - pz[lj];
    d2[k] = dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k];
    ipc[k] = im[li] + im[lj];
  }
  int it = 0;
  for (; it < a.max_iter; ++it) {
    bool converged = true;
    for (unsigned k = 0; k < nc; ++k) {
      unsigned pk = a.cij[c0 + k];
      unsigned li = pk & 0xFFu, lj = (pk >> 8) & 0xFFu;
      float vxr = v0[li] - v0[lj], vyr = v1[li] - v1[lj], vzr = v2[li] - v2[lj];
      float vrel = vxr * dx[k] + vyr * dy[k] + vzr * dz[k];
      if (fabsf(vrel) > a.rattle_tol) {
        converged = false;
        float denom = d2[k] * ipc[k];
        if (denom == 0.f) continue;
        float mu = vrel / denom;
        float ki = mu * im[li], kj = mu * im[lj];
        v0[li] -= ki * dx[k]; v1[li] -= ki * dy[k]; v2[li] -= ki * dz[k];
        v0[lj] += kj * dx[k]; v1[lj] += kj * dy[k]; v2[lj] += kj * dz[k];
        dv0[li] -= ki * dx[k]; dv1[li] -= ki * dy[k]; dv2[li] -= ki * dz[k];
        dv0[lj] += kj * dx[k]; dv1[lj] += kj * dy[k]; dv2[lj] += kj * dz[k];
      }
    }
    if (converged) break;
  }
  if (a.stats) shake_stats(a.stats + 3, (unsigned)it, a.max_iter);
  if (ACC && a.dt > 0.f) {
    float inv_dt = 1.0f / a.dt;
    float cx, cy, cz;
    shake_com(px, py, pz, m, na, cx, cy, cz);
    for (unsigned k = 0; k < na; ++k) {
      float rx = px[k] - cx, ry = py[k] - cy, rz = pz[k] - cz;
      float scale = m[k] * inv_dt;
      unsigned i = a.gatoms[a0 + k];
      a.w_cons[i] += scale * (dv0[k] * rx + dv1[k] * ry + dv2[k] * rz);
    }
  }
  for (unsigned k = 0; k < na; ++k) {
    unsigned i = a.gatoms[a0 + k];
    a.vx[i] = v0[k]; a.vy[i] = v1[k]; a.vz[i] = v2[k];
  }
}

// Groups beyond the reference's 8-atom cap (chains): one wave per group.
// Lane k owns atom k and constraint k (a group holds at most 64 of each).
//   * Frame: lane k takes the bond vector d_k = minimg(x_k - x_(k-1)), an
//     inclusive wave scan gives S_k, and the group is solved in
//     coordinates S_k - S_mid relative to the middle atom, so a chain longer
//     than half the box does not fold back and |r|^2 - d^2 resolves to the
//     f32 spacing of the group's extent rather than of its place in the box.
//   * Sweeps: the host colours the constraints so that constraints of one
//     colour share no atom (a chain alternates two colours); a sweep visits
//     the colours in order and the lanes of one colour update their
//     disjoint atoms at once, which is a sequential Gauss-Seidel sweep in
//     (colour, index) order.
//   * Every atom moves by its delta; centre-of-mass sums are xor butterflies
//     read from lane 0.
// The CPU oracle restates the scan and butterfly orders.

// inclusive scan over the wave: lane k gets v_0 + ... + v_k, grouped by
// the offsets 1, 2, 4, ..., 32
__device__ __forceinline__ float chain_scan(float v, unsigned lane) {
#pragma unroll
  for (unsigned off = 1; off < 64; off <<= 1) {
    float u = __shfl_up(v, off);
    if (lane >= off) v = u + v;
  }
  return v;
}

// sum over the wave as lane 0 of an xor butterfly, broadcast
__device__ __forceinline__ float chain_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
  return __shfl(v, 0);
}

// the group frame of the wrapped coordinates (x, y, z) of lane k's atom
__device__ __forceinline__ void chain_frame(float& x, float& y, float& z, unsigned lane,
                                            unsigned na, const HdLattice& L) {
  float px = __shfl_up(x, 1u), py = __shfl_up(y, 1u), pz = __shfl_up(z, 1u);
  float dx = 0.f, dy = 0.f, dz = 0.f;
  if (lane >= 1 && lane < na) {
    dx = x - px; dy = y - py; dz = z - pz;
    hd_min_image(dx, dy, dz, L);
  }
  dx = chain_scan(dx, lane);
  dy = chain_scan(dy, lane);
  dz = chain_scan(dz, lane);
  unsigned mid = na / 2;
  x = dx - __shfl(dx, (int)mid);
  y = dy - __shfl(dy, (int)mid);
  z = dz - __shfl(dz, (int)mid);
}

__device__ __forceinline__ void chain_com(float x, float y, float z, float m, float& cx,
                                          float& cy, float& cz) {
  float total = chain_sum(m);
  float inv = total > 0.f ? 1.0f / total : 0.f;
  cx = chain_sum(m * x) * inv;
  cy = chain_sum(m * y) * inv;
  cz = chain_sum(m * z) * inv;
}

__global__ __launch_bounds__(64) void k_shake_positions_long(HdShakeArgs a) {
  __shared__ float x0[64], y0[64], z0[64], xc[64], yc[64], zc[64], ims[64];
  const unsigned lane = threadIdx.x;
  const unsigned g = a.llist[blockIdx.x];
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
  unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
  unsigned ia = 0;
  float ux = 0.f, uy = 0.f, uz = 0.f, sx = 0.f, sy = 0.f, sz = 0.f, mk = 0.f, imk = 0.f;
  if (lane < na) {
    ia = a.gatoms[a0 + lane];
    float4 p = a.posq[ia], q = a.snap[ia];
    ux = p.x; uy = p.y; uz = p.z;
    sx = q.x; sy = q.y; sz = q.z;
    mk = a.mass[ia];
    imk = mk > 0.f ? 1.0f / mk : 0.f;
  }
  chain_frame(sx, sy, sz, lane, na, L);
  chain_frame(ux, uy, uz, lane, na, L);
  x0[lane] = sx; y0[lane] = sy; z0[lane] = sz;
  xc[lane] = ux; yc[lane] = uy; zc[lane] = uz;
  ims[lane] = imk;
  __syncthreads();
  unsigned li = 0, lj = 0, col = 0xFFFFFFFFu;
  float gx = 0.f, gy = 0.f, gz = 0.f, ip = 0.f, r2 = 0.f, lam = 0.f;
  if (lane < nc) {
    unsigned pk = a.cij[c0 + lane];
    li = pk & 0xFFu;
    lj = (pk >> 8) & 0xFFu;
    col = a.color[c0 + lane];
    gx = x0[li] - x0[lj];
    gy = y0[li] - y0[lj];
    gz = z0[li] - z0[lj];
    ip = ims[li] + ims[lj];
    r2 = a.cr2[c0 + lane];
  }
  const float imi = ims[li], imj = ims[lj];
  if (a.lam0) {
    for (unsigned q = 0; q < a.n_colors; ++q) {
      if (col == q) {
        float lambda = ip > 0.f ? a.lam0[c0 + lane] / ip : 0.f;
        float ki = lambda * imi, kj = lambda * imj;
        xc[li] -= ki * gx; yc[li] -= ki * gy; zc[li] -= ki * gz;
        xc[lj] += kj * gx; yc[lj] += kj * gy; zc[lj] += kj * gz;
        lam = lambda;
      }
      __syncthreads();
    }
  }
  int it = 0;
  for (; it < a.max_iter; ++it) {
    int bad = 0;
    for (unsigned q = 0; q < a.n_colors; ++q) {
      if (col == q) {
        float dx = xc[li] - xc[lj], dy = yc[li] - yc[lj], dz = zc[li] - zc[lj];
        float dist2 = dx * dx + dy * dy + dz * dz;
        float sigma = dist2 - r2;
        if (fabsf(sigma) > a.tol) {
          bad = 1;
          float ddot = dx * gx + dy * gy + dz * gz;
          float denom = 2.0f * ddot * ip;
          if (denom != 0.f) {
            float lambda = sigma / denom;
            float ki = lambda * imi, kj = lambda * imj;
            xc[li] -= ki * gx; yc[li] -= ki * gy; zc[li] -= ki * gz;
            xc[lj] += kj * gx; yc[lj] += kj * gy; zc[lj] += kj * gz;
            lam += lambda;
          }
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(bad)) break;
  }
  if (lane == 0 && a.stats) shake_stats(a.stats, (unsigned)it, a.max_iter);
  if (a.lam_out && lane < nc) a.lam_out[c0 + lane] = lam * ip;
  const float cxl = xc[lane], cyl = yc[lane], czl = zc[lane];
  float cmx, cmy, cmz;
  chain_com(cxl, cyl, czl, mk, cmx, cmy, cmz);
  if (lane < na) {
    float inv_dt = a.dt != 0.f ? 1.0f / a.dt : 0.f;
    float inv_dt2 = a.dt != 0.f ? 1.0f / (a.dt * a.dt) : 0.f;
    float dxg = cxl - ux, dyg = cyl - uy, dzg = czl - uz;
    a.vx[ia] += dxg * inv_dt; a.vy[ia] += dyg * inv_dt; a.vz[ia] += dzg * inv_dt;
    float4 p = a.posq[ia];
    p.x += dxg; p.y += dyg; p.z += dzg;
    a.posq[ia] = p;
    float rx = cxl - cmx, ry = cyl - cmy, rz = czl - cmz;
    float scale = mk * inv_dt2;
    a.w_cons[ia] = scale * (dxg * rx + dyg * ry + dzg * rz);
  }
}

template <bool ACC>
__global__ __launch_bounds__(64) void k_rattle_velocities_long(HdShakeArgs a) {
  __shared__ float px[64], py[64], pz[64], v0[64], v1[64], v2[64];
  __shared__ float dv0[64], dv1[64], dv2[64], ims[64];
  const unsigned lane = threadIdx.x;
  const unsigned g = a.llist[blockIdx.x];
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
  unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
  unsigned ia = 0;
  float x = 0.f, y = 0.f, z = 0.f, mk = 0.f, imk = 0.f;
  v0[lane] = v1[lane] = v2[lane] = 0.f;
  if (lane < na) {
    ia = a.gatoms[a0 + lane];
    float4 p = a.posq[ia];
    x = p.x; y = p.y; z = p.z;
    v0[lane] = a.vx[ia]; v1[lane] = a.vy[ia]; v2[lane] = a.vz[ia];
    mk = a.mass[ia];
    imk = mk > 0.f ? 1.0f / mk : 0.f;
  }
  chain_frame(x, y, z, lane, na, L);
  px[lane] = x; py[lane] = y; pz[lane] = z;
  dv0[lane] = dv1[lane] = dv2[lane] = 0.f;
  ims[lane] = imk;
  __syncthreads();
  unsigned li = 0, lj = 0, col = 0xFFFFFFFFu;
  float dx = 0.f, dy = 0.f, dz = 0.f, d2 = 0.f, ipc = 0.f;
  if (lane < nc) {
    unsigned pk = a.cij[c0 + lane];
    li = pk & 0xFFu;
    lj = (pk >> 8) & 0xFFu;
    col = a.color[c0 + lane];
    dx = px[li] - px[lj];
    dy = py[li] - py[lj];
    dz = pz[li] - pz[lj];
    d2 = dx * dx + dy * dy + dz * dz;
    ipc = ims[li] + ims[lj];
  }
  const float imi = ims[li], imj = ims[lj];
  int it = 0;
  for (; it < a.max_iter; ++it) {
    int bad = 0;
    for (unsigned q = 0; q < a.n_colors; ++q) {
      if (col == q) {
        float vxr = v0[li] - v0[lj], vyr = v1[li] - v1[lj], vzr = v2[li] - v2[lj];
        float vrel = vxr * dx + vyr * dy + vzr * dz;
        if (fabsf(vrel) > a.rattle_tol) {
          bad = 1;
          float denom = d2 * ipc;
          if (denom != 0.f) {
            float mu = vrel / denom;
            float ki = mu * imi, kj = mu * imj;
            v0[li] -= ki * dx; v1[li] -= ki * dy; v2[li] -= ki * dz;
            v0[lj] += kj * dx; v1[lj] += kj * dy; v2[lj] += kj * dz;
            dv0[li] -= ki * dx; dv1[li] -= ki * dy; dv2[li] -= ki * dz;
            dv0[lj] += kj * dx; dv1[lj] += kj * dy; dv2[lj] += kj * dz;
          }
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(bad)) break;
  }
  if (lane == 0 && a.stats) shake_stats(a.stats + 3, (unsigned)it, a.max_iter);
  if (ACC && a.dt > 0.f) {
    float cx, cy, cz;
    chain_com(x, y, z, mk, cx, cy, cz);
    if (lane < na) {
      float rx = x - cx, ry = y - cy, rz = z - cz;
      float scale = mk * (1.0f / a.dt);
      a.w_cons[ia] += scale * (dv0[lane] * rx + dv1[lane] * ry + dv2[lane] * rz);
    }
  }
  if (lane < na) {
    a.vx[ia] = v0[lane]; a.vy[ia] = v1[lane]; a.vz[ia] = v2[lane];
  }
}

// Inputs of the SHAKE multiplier predictor, one thread per constraint,
// all dimensionless.  For constraint k between atoms i, j with snapshot
// bond g and unconstrained bond r (both minimum-imaged):
//   nu1_k    = (|r|^2 - d^2) / (2 g.r), the first Gauss-Seidel multiplier;
//   kappa_kl = c_kl (r_k.g_l) / (ip_l g_k.r_k), the coupling of the
//              linearised system nu_k + sum_l kappa_kl nu_l = nu1_k, where
//              c_kl = +-1/m of the atom that k and l share (0 if none) and
//              ip = 1/m_i + 1/m_j;
//   rho_k    = g.r / d^2 - 1;
//   nuprev_k = the converged nu of the constraint in the previous solve
//              (a.lam_out before this step's SHAKE overwrites it).
// Row k: [nu1_k, nu1_k-1, nu1_k+1, nuprev_k, nu1_k-2 + nu1_k+2,
// kappa_k,k-1, kappa_k,k+1, rho_k], neighbours in the group's constraint
// order, zero past the group's ends.
struct ShakeBond {
  float gx, gy, gz, rx, ry, rz, imi, imj, ip, gr, nu1;
  unsigned li, lj;
};

__device__ __forceinline__ ShakeBond shake_bond(const HdShakeArgs& a, const HdLattice& L,
                                                unsigned a0, unsigned c) {
  ShakeBond b;
  unsigned pk = a.cij[c];
  b.li = pk & 0xFFu;
  b.lj = (pk >> 8) & 0xFFu;
  unsigned i = a.gatoms[a0 + b.li], j = a.gatoms[a0 + b.lj];
  float4 si = a.snap[i], sj = a.snap[j], pi = a.posq[i], pj = a.posq[j];
  b.gx = si.x - sj.x; b.gy = si.y - sj.y; b.gz = si.z - sj.z;
  hd_min_image(b.gx, b.gy, b.gz, L);
  b.rx = pi.x - pj.x; b.ry = pi.y - pj.y; b.rz = pi.z - pj.z;
  hd_min_image(b.rx, b.ry, b.rz, L);
  float mi = a.mass[i], mj = a.mass[j];
  b.imi = mi > 0.f ? 1.0f / mi : 0.f;
  b.imj = mj > 0.f ? 1.0f / mj : 0.f;
  b.ip = b.imi + b.imj;
  b.gr = b.gx * b.rx + b.gy * b.ry + b.gz * b.rz;
  float sigma = (b.rx * b.rx + b.ry * b.ry + b.rz * b.rz) - a.cr2[c];
  b.nu1 = b.gr != 0.f ? sigma / (2.0f * b.gr) : 0.f;
  return b;
}

__device__ __forceinline__ float shake_coupling(const ShakeBond& k, const ShakeBond& l) {
  float c = 0.f;
  if (k.li == l.li) c += l.imi;
  if (k.li == l.lj) c -= l.imj;
  if (k.lj == l.li) c -= l.imi;
  if (k.lj == l.lj) c += l.imj;
  float den = l.ip * k.gr;
  if (c == 0.f || den == 0.f) return 0.f;
  float num = c * (k.rx * l.gx + k.ry * l.gy + k.rz * l.gz);
  return num / den;
}

__global__ void k_shake_features(HdShakeArgs a) {
  unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cons) return;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned g = a.cgrp[c];
  unsigned a0 = a.goff[g], cb = a.coff[g], ce = a.coff[g + 1];
  ShakeBond k = shake_bond(a, L, a0, c);
  float f[HD_SHAKE_NFEAT];
  f[0] = k.nu1;
  f[1] = f[2] = f[3] = f[4] = f[5] = f[6] = 0.f;
  if (c >= cb + 1) {
    ShakeBond l = shake_bond(a, L, a0, c - 1);
    f[1] = l.nu1;
    f[5] = shake_coupling(k, l);
  }
  if (c + 1 < ce) {
    ShakeBond l = shake_bond(a, L, a0, c + 1);
    f[2] = l.nu1;
    f[6] = shake_coupling(k, l);
  }
  f[3] = a.lam_out ? a.lam_out[c] : 0.f;
  float f2l = c >= cb + 2 ? shake_bond(a, L, a0, c - 2).nu1 : 0.f;
  float f2r = c + 2 < ce ? shake_bond(a, L, a0, c + 2).nu1 : 0.f;
  f[4] = f2l + f2r;
  f[7] = k.gr / a.cr2[c] - 1.0f;
  float4* out = reinterpret_cast<float4*>(a.feat + (size_t)c * HD_SHAKE_NFEAT);
  out[0] = make_float4(f[0], f[1], f[2], f[3]);
  out[1] = make_float4(f[4], f[5], f[6], f[7]);
}

void hd_launch_shake_features(hipStream_t st, const HdShakeArgs& a) {
  if (a.n_cons == 0) return;
  hipLaunchKernelGGL(k_shake_features, dim3((a.n_cons + 255) / 256), dim3(256), 0, st, a);
}

void hd_launch_shake_positions(hipStream_t st, const HdShakeArgs& a) {
  if (a.n_short) {
    unsigned T = a.threads;
    size_t lds = (size_t)T * hd_shake_lds_per_thread(a.max_atoms, a.max_cons);
    hipLaunchKernelGGL(k_shake_positions, dim3((a.n_short + T - 1) / T), dim3(T), lds, st, a);
  }
  if (a.n_long)
    hipLaunchKernelGGL(k_shake_positions_long, dim3(a.n_long), dim3(64), 0, st, a);
}

void hd_launch_rattle_velocities(hipStream_t st, const HdShakeArgs& a, bool acc) {
  if (a.n_short) {
    unsigned T = a.threads;
    size_t lds = (size_t)T * hd_shake_lds_per_thread(a.max_atoms, a.max_cons);
    dim3 grid((a.n_short + T - 1) / T);
    if (acc)
      hipLaunchKernelGGL(k_rattle_velocities<true>, grid, dim3(T), lds, st, a);
    else
      hipLaunchKernelGGL(k_rattle_velocities<false>, grid, dim3(T), lds, st, a);
  }
  if (a.n_long) {
    if (acc)
      hipLaunchKernelGGL(k_rattle_velocities_long<true>, dim3(a.n_long), dim3(64), 0, st, a);
    else
      hipLaunchKernelGGL(k_rattle_velocities_long<false>, dim3(a.n_long), dim3(64), 0, st, a);
  }
}

// ---------------------------------------------------------------------
// Langevin BAOAB (src/integrator/langevin_baoab.rs: B, A, O, A, F, B with
// the last B fused into the force finalize).  B + A and O + A are fused
// pairs of the reference's vv_kick / lan_drift_half / lan_ou_step kernels
// (integrate.cu, langevin.cu:13-80) with the same f32 operations in the
// same order.  The draw counter lives on the device: the B + A kernel
// advances it and the O + A kernel of the same step uses counter - 1, so a
// captured step graph replays with fresh draws.
// ---------------------------------------------------------------------
template <bool SNAP>
__global__ void k_lan_kick_drift(HdIntegrateArgs a, HdLangevinArgs l) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) l.counter[0] += 1ull;
  if (i >= a.n) return;
  HdLattice L = hd_load_lattice(a.lattice);
  float m = a.mass[i];
  float ax = a.fx[i] / m, ay = a.fy[i] / m, az = a.fz[i] / m;
  float half_dt = a.dt * 0.5f;
  float vx = a.vx[i] + ax * half_dt;
  float vy = a.vy[i] + ay * half_dt;
  float vz = a.vz[i] + az * half_dt;
  float4 pq = a.posq[i];
  if (SNAP) a.snap[i] = pq;
  float px = pq.x + vx * half_dt, py = pq.y + vy * half_dt, pz = pq.z + vz * half_dt;
  int ka, kb, kc;
  hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
  a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
  a.posq[i] = make_float4(px, py, pz, pq.w);
  a.ix[i] += ka; a.iy[i] += kb; a.iz[i] += kc;
}

template <bool SNAP>
__global__ void k_lan_ou_drift(HdIntegrateArgs a, HdLangevinArgs l) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned long long ctr = l.counter[0] - 1ull;
  unsigned clo = (unsigned)(ctr & 0xFFFFFFFFull), chi = (unsigned)(ctr >> 32);
  float m = a.mass[i];
  float sigma = sqrtf((1.0f - l.alpha * l.alpha) * l.kt / m);
  float xi0 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 0u);
  float xi1 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 1u);
  float xi2 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 2u);
  float vx = l.alpha * a.vx[i] + sigma * xi0;
  float vy = l.alpha * a.vy[i] + sigma * xi1;
  float vz = l.alpha * a.vz[i] + sigma * xi2;
  float half_dt = a.dt * 0.5f;
  float4 pq = a.posq[i];
  if (SNAP) a.snap[i] = pq;
  float px = pq.x + vx * half_dt, py = pq.y + vy * half_dt, pz = pq.z + vz * half_dt;
  int ka, kb, kc;
  hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
  a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
  a.posq[i] = make_float4(px, py, pz, pq.w);
  a.ix[i] += ka; a.iy[i] += kb; a.iz[i] += kc;
}

void hd_launch_lan_kick_drift(hipStream_t st, const HdIntegrateArgs& a, const HdLangevinArgs& l) {
  dim3 g(a.n ? (a.n + 255) / 256 : 1), b(256);
  if (a.snap) hipLaunchKernelGGL((k_lan_kick_drift<true>), g, b, 0, st, a, l);
  else hipLaunchKernelGGL((k_lan_kick_drift<false>), g, b, 0, st, a, l);
}

void hd_launch_lan_ou_drift(hipStream_t st, const HdIntegrateArgs& a, const HdLangevinArgs& l) {
  if (a.n == 0) return;
  dim3 g((a.n + 255) / 256), b(256);
  if (a.snap) hipLaunchKernelGGL((k_lan_ou_drift<true>), g, b, 0, st, a, l);
  else hipLaunchKernelGGL((k_lan_ou_drift<false>), g, b, 0, st, a, l);
}

// One thread: P = (2K + W)/(3V), mu^3 = 1 - kappa (dt/tau)(P0 - P)
// + sqrt(2 kappa kT dt/(tau V)) r, clamped, lattice *= (f32) mu; the
// draw counter advances by one.  K and W are read from FP64 scalars.
// One thread: mu from K and W, the lattice scaled, the diagnostics and
// the draw counter advanced
__device__ void crescale_mu_update(double k, double w, float* lattice, float* mu_out,
                                   double* diag, unsigned long long* counter,
                                   const HdBarostatArgs& b) {
  unsigned long long ctr = *counter;
  double lx = lattice[0], ly = lattice[1], lz = lattice[2];
  double v_pre = lx * ly * lz;
  double p = (2.0 * k + w) / (3.0 * v_pre);
  double r = philox_gaussian_f64(b.seed_lo, b.seed_hi, (unsigned)(ctr & 0xFFFFFFFFull),
                                 (unsigned)(ctr >> 32), 0u, 0u);
  double det = -b.compressibility * (b.dt / b.tau) * (b.pressure - p);
  double amp = sqrt(2.0 * b.compressibility * b.kt * b.dt / (b.tau * v_pre));
  double mu3 = 1.0 + det + amp * r;
  double mu3c = mu3 > b.mu3_min ? mu3 : b.mu3_min;
  double mu = cbrt(mu3c);
  double v_post = v_pre * mu3c;
  float muf = (float)mu;
  mu_out[0] = muf;
  for (int i = 0; i < 6; ++i) lattice[i] *= muf;
  diag[0] = p;
  diag[1] = v_post;
  diag[2] += b.pressure * (v_post - v_pre);
  *counter = ctr + 1ull;
}

// k_sum2_final's reduction, then its thread 0 runs the C-rescale update
// (mu from K = k[0] and the virial it has just reduced)
__global__ void __launch_bounds__(256) k_sum2_final_crescale(
    const double* __restrict__ partials, unsigned np, double* __restrict__ out,
    const double* __restrict__ k, float* lattice, float* mu_out, double* diag,
    unsigned long long* counter, HdBarostatArgs b) {
  __shared__ double sh[256];
  double sa = 0.0, sb = 0.0;
  for (unsigned i = threadIdx.x; i < np; i += blockDim.x) {
    sa += partials[2 * i];
    sb += partials[2 * i + 1];
  }
  double ta = block_tree_sum(sa, sh);
  __syncthreads();
  double tb = block_tree_sum(sb, sh);
  if (threadIdx.x == 0) {
    out[0] = ta;
    out[1] = tb;
    crescale_mu_update(k[0], tb, lattice, mu_out, diag, counter, b);
  }
}

__global__ void k_scale_positions(float4* __restrict__ posq, const float* __restrict__ mu,
                                  unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f = mu[0];
  float4 p = posq[i];
  p.x *= f;
  p.y *= f;
  p.z *= f;
  posq[i] = p;
}

void hd_launch_scale_positions(hipStream_t st, float4* posq, const float* mu, unsigned n) {
  if (n)
    hipLaunchKernelGGL(k_scale_positions, dim3((n + 255) / 256), dim3(256), 0, st, posq, mu, n);
}

void hd_launch_sum2_final_crescale(hipStream_t st, const double* partials, unsigned np,
                                   double* out, const double* k, float* lattice, float* mu,
                                   double* diag, unsigned long long* counter,
                                   const HdBarostatArgs& b) {
  hipLaunchKernelGGL(k_sum2_final_crescale, dim3(1), dim3(256), 0, st, partials, np, out, k,
                     lattice, mu, diag, counter, b);
}

// In-bounds check of the DPP wave_rol:1 lane layout: out[l] = value of
// lane (l+1) mod 64 after one rotation.
__global__ void k_selftest_dpp(int* out) {
  int v = (int)threadIdx.x * 7 + 3;
  out[threadIdx.x] = hd_rot1_i(v);
}
void hd_launch_selftest_dpp(hipStream_t st, int* out) {
  hipLaunchKernelGGL(k_selftest_dpp, dim3(1), dim3(64), 0, st, out);
}
