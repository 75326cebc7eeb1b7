// Integration, constraints, reductions and the force finalize on gfx950.
//
// Arithmetic in this file is kept operation-for-operation identical to the
// reference's single-precision kernels (no FMA contraction), so the f32
// velocity-Verlet update is bit-exact against a scalar restatement:
//   * velocity Verlet kick/drift, f32 and compensated (f32, f64)
//     (kernels/integrate.cu:24-263);
//   * the constraint contract of SETTLE (kernels/settle.cu:145-497):
//     rigid water from the pre-drift snapshot, then zero bond velocities,
//     with the constraint virial (solved here by FP64 Newton / direct
//     solves, see below);
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
// SETTLE: rigid three-site water (Miyamoto & Kollman, J. Comput. Chem. 13,
// 952-962 (1992)).
//
// Positions.  With u_k the drifted (unconstrained) positions and g_c the
// three bond vectors of the step's start (the snapshot), the constrained
// positions are
//     x_k = u_k + (1/m_k) sum_c e_kc lambda_c g_c,
//     |x_i(c) - x_j(c)|^2 = d_c^2  for the bonds c = OH1, OH2, H1H2,
// with e_kc = +1 (k = i(c)), -1 (k = j(c)), 0 otherwise: displacements
// along the constraint forces of the start of the step.  Miyamoto and
// Kollman give the closed-form solution of these three equations; here
// they are solved by Newton's method on (lambda_1, lambda_2, lambda_3) in
// FP64, in a frame at the snapshot oxygen, which converges quadratically
// to the same point from lambda = 0 and loses nothing to the f32 storage:
// the correction is rounded once, into the drifted f32 positions.
//
// Velocities.  The bond-relative velocities along the new bonds r_c must
// vanish: with v_k' = v_k + (1/m_k) sum_c e_kc mu_c r_c the conditions
// r_c . (v'_i(c) - v'_j(c)) = 0 are linear in mu, one 3x3 system per
// water, solved directly (adjugate, FP64).
//
// Both stages use one coupling matrix form: d(x_i(c) - x_j(c))/d(lambda_c')
// = K_cc' b_c' with K_cc' = e_i(c)c'/m_i(c) - e_j(c)c'/m_j(c).
// Constraint virial per atom: m_k dx_k . (x_k - X_com) / dt^2 for the
// position correction dx_k, plus m_k dv_k . (x_k - X_com) / dt for the
// velocity correction dv_k (the totals are frame independent: the
// constraint forces of a water sum to zero).
// ---------------------------------------------------------------------
#define HD_SETTLE_NEWTON_MAX 8
// the minimizer's projection (dt = 0) starts from moves of up to max_step,
// far from the manifold: Newton gets more iterations there
#define HD_SETTLE_PROJECT_MAX 64

// bond c joins water atoms tri_i(c) and tri_j(c): O-H1, O-H2, H1-H2
__device__ __forceinline__ int tri_i(int c) { return c == 2 ? 1 : 0; }
__device__ __forceinline__ int tri_j(int c) { return c == 0 ? 1 : 2; }
// e_kc: how atom k enters bond c
__device__ __forceinline__ double tri_sign(int k, int c) {
  return k == tri_i(c) ? 1.0 : (k == tri_j(c) ? -1.0 : 0.0);
}

// x = A^-1 b through the adjugate; a singular A leaves x = 0
__device__ __forceinline__ void solve3_adj(const double (&A)[3][3], const double (&b)[3],
                                           double (&x)[3]) {
  double cof[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int r1 = (r + 1) % 3, r2 = (r + 2) % 3, q1 = (q + 1) % 3, q2 = (q + 2) % 3;
      cof[r][q] = A[r1][q1] * A[r2][q2] - A[r1][q2] * A[r2][q1];
    }
  double det = A[0][0] * cof[0][0] + A[0][1] * cof[0][1] + A[0][2] * cof[0][2];
  double inv = det != 0.0 ? 1.0 / det : 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) x[r] = (cof[0][r] * b[0] + cof[1][r] * b[1] + cof[2][r] * b[2]) * inv;
}

// coupling coefficients K_cc' of the three bonds for atom inverse masses im
__device__ __forceinline__ void tri_coupling(const double (&im)[3], double (&K)[3][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      K[c][q] = tri_sign(tri_i(c), q) * im[tri_i(c)] - tri_sign(tri_j(c), q) * im[tri_j(c)];
}

__device__ __forceinline__ double dot3(const double (&a)[3], const double (&b)[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// minimum-image difference a - b of two f32 positions, exact in FP64
__device__ __forceinline__ void rel_f64(float4 a, float4 b, const HdLattice& L, double (&d)[3]) {
  d[0] = (double)a.x - (double)b.x;
  d[1] = (double)a.y - (double)b.y;
  d[2] = (double)a.z - (double)b.z;
  hd_min_image_f64(d[0], d[1], d[2], L);
}

struct SettleOut {
  float4 p[3];
  float v[3][3];
  float w[3];
};

// Position stage of one water: p the drifted positions, q the snapshot,
// vin the velocities after the kick (f32).  Returns the constrained
// positions, the velocities with the correction / dt, and the position
// virials.
// the geometry of group g: kind 0 from the scalar fields, others from the
// kind table
struct SettleGeom {
  double d_oh, d_hh;
  float m_o, m_h;
};
__device__ __forceinline__ SettleGeom settle_geom(const HdSettleArgs& a, unsigned g) {
  unsigned k = a.gkind ? (unsigned)a.gkind[g] : 0u;
  if (k == 0) return SettleGeom{a.d_oh, a.d_hh, a.m_o, a.m_h};
  return SettleGeom{a.kd_oh[k], a.kd_hh[k], a.km_o[k], a.km_h[k]};
}

__device__ __forceinline__ SettleOut settle_water_positions(const SettleGeom& a,
                                                            const HdLattice& L,
                                                            const float4 (&p)[3],
                                                            const float4 (&q)[3],
                                                            const float (&vin)[3][3],
                                                            float dt) {
  const double mass[3] = {(double)a.m_o, (double)a.m_h, (double)a.m_h};
  const double im[3] = {1.0 / mass[0], 1.0 / mass[1], 1.0 / mass[2]};
  const double target[3] = {a.d_oh * a.d_oh, a.d_oh * a.d_oh, a.d_hh * a.d_hh};
  // snapshot frame at the old oxygen: s_k; drifted positions u_k = s_k + (p_k - q_k)
  double s[3][3] = {{0.0, 0.0, 0.0}}, u[3][3];
  rel_f64(q[1], q[0], L, s[1]);
  rel_f64(q[2], q[0], L, s[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double step[3];
    rel_f64(p[k], q[k], L, step);
#pragma unroll
    for (int d = 0; d < 3; ++d) u[k][d] = s[k][d] + step[d];
  }
  double bond0[3][3];  // g_c
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) bond0[c][d] = s[tri_i(c)][d] - s[tri_j(c)][d];
  double K[3][3];
  tri_coupling(im, K);
  double lam[3] = {0.0, 0.0, 0.0};
  double x[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) x[k][d] = u[k][d];
  const int n_newton = dt != 0.f ? HD_SETTLE_NEWTON_MAX : HD_SETTLE_PROJECT_MAX;
  for (int it = 0; it < n_newton; ++it) {
    double bond[3][3], res[3], worst = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int d = 0; d < 3; ++d) bond[c][d] = x[tri_i(c)][d] - x[tri_j(c)][d];
      res[c] = dot3(bond[c], bond[c]) - target[c];
      worst = fmax(worst, fabs(res[c]) / target[c]);
    }
    if (worst <= 1e-14) break;
    // Jacobian of the residuals in lambda: 2 K_cc' (r_c . g_c')
    double J[3][3], rhs[3], dl[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int q2 = 0; q2 < 3; ++q2) J[c][q2] = 2.0 * K[c][q2] * dot3(bond[c], bond0[q2]);
      rhs[c] = -res[c];
    }
    solve3_adj(J, rhs, dl);
#pragma unroll
    for (int c = 0; c < 3; ++c) lam[c] += dl[c];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double sh = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) sh += tri_sign(k, c) * lam[c] * bond0[c][d];
        x[k][d] = u[k][d] + im[k] * sh;
      }
  }
  double msum = mass[0] + mass[1] + mass[2], com[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) com[d] = (mass[0] * x[0][d] + mass[1] * x[1][d] + mass[2] * x[2][d]) / msum;
  double idt = dt != 0.f ? 1.0 / (double)dt : 0.0;
  SettleOut o;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double shift[3] = {x[k][0] - u[k][0], x[k][1] - u[k][1], x[k][2] - u[k][2]};
    o.p[k] = make_float4((float)((double)p[k].x + shift[0]), (float)((double)p[k].y + shift[1]),
                         (float)((double)p[k].z + shift[2]), p[k].w);
#pragma unroll
    for (int d = 0; d < 3; ++d) o.v[k][d] = (float)((double)vin[k][d] + shift[d] * idt);
    double arm[3] = {x[k][0] - com[0], x[k][1] - com[1], x[k][2] - com[2]};
    o.w[k] = (float)(mass[k] * idt * idt * dot3(shift, arm));
  }
  return o;
}

// Velocity stage of one water at its final positions p (f32 as stored):
// the velocities v are projected; with ACC (dt > 0) the velocity virial is
// added to w.
template <bool ACC>
__device__ __forceinline__ void settle_water_velocities(const SettleGeom& a, const HdLattice& L,
                                                        const float4 (&p)[3], float (&v)[3][3],
                                                        float (&w)[3], float dt) {
  const double mass[3] = {(double)a.m_o, (double)a.m_h, (double)a.m_h};
  const double im[3] = {1.0 / mass[0], 1.0 / mass[1], 1.0 / mass[2]};
  double y[3][3] = {{0.0, 0.0, 0.0}};
  rel_f64(p[1], p[0], L, y[1]);
  rel_f64(p[2], p[0], L, y[2]);
  double bond[3][3], rel[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double dv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      bond[c][d] = y[tri_i(c)][d] - y[tri_j(c)][d];
      dv[d] = (double)v[tri_i(c)][d] - (double)v[tri_j(c)][d];
    }
    rel[c] = -dot3(bond[c], dv);
  }
  double K[3][3], A[3][3], mu[3];
  tri_coupling(im, K);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q2 = 0; q2 < 3; ++q2) A[c][q2] = K[c][q2] * dot3(bond[c], bond[q2]);
  solve3_adj(A, rel, mu);
  double msum = mass[0] + mass[1] + mass[2], com[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) com[d] = (mass[0] * y[0][d] + mass[1] * y[1][d] + mass[2] * y[2][d]) / msum;
  double idt = (ACC && dt > 0.f) ? 1.0 / (double)dt : 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double kick[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double sh = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) sh += tri_sign(k, c) * mu[c] * bond[c][d];
      kick[d] = im[k] * sh;
      v[k][d] = (float)((double)v[k][d] + kick[d]);
    }
    if (ACC && dt > 0.f) {
      double arm[3] = {y[k][0] - com[0], y[k][1] - com[1], y[k][2] - com[2]};
      w[k] = (float)((double)w[k] + mass[k] * idt * dot3(kick, arm));
    }
  }
}

__device__ __forceinline__ void settle_load_ids(const HdSettleArgs& a, unsigned g,
                                                unsigned (&ids)[3]) {
  ids[0] = a.groups[3 * g];
  ids[1] = a.groups[3 * g + 1];
  ids[2] = a.groups[3 * g + 2];
}

__global__ void __launch_bounds__(128) k_settle_positions(HdSettleArgs a) {
  unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_groups) return;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned ids[3];
  settle_load_ids(a, g, ids);
  float4 p[3], q[3];
  float v[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p[k] = a.posq[ids[k]];
    q[k] = a.snap[ids[k]];
    v[k][0] = a.vx[ids[k]]; v[k][1] = a.vy[ids[k]]; v[k][2] = a.vz[ids[k]];
  }
  SettleOut o = settle_water_positions(settle_geom(a, g), L, p, q, v, a.dt);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a.posq[ids[k]] = o.p[k];
    a.vx[ids[k]] = o.v[k][0]; a.vy[ids[k]] = o.v[k][1]; a.vz[ids[k]] = o.v[k][2];
    a.w_cons[ids[k]] = o.w[k];
  }
}

template <bool ACC>
__global__ void __launch_bounds__(128) k_settle_velocities(HdSettleArgs a) {
  unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_groups) return;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned ids[3];
  settle_load_ids(a, g, ids);
  float4 p[3];
  float v[3][3], w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p[k] = a.posq[ids[k]];
    v[k][0] = a.vx[ids[k]]; v[k][1] = a.vy[ids[k]]; v[k][2] = a.vz[ids[k]];
    w[k] = ACC ? a.w_cons[ids[k]] : 0.f;
  }
  settle_water_velocities<ACC>(settle_geom(a, g), L, p, v, w, a.dt);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a.vx[ids[k]] = v[k][0]; a.vy[ids[k]] = v[k][1]; a.vz[ids[k]] = v[k][2];
    if (ACC) a.w_cons[ids[k]] = w[k];
  }
}

// VV rescale + kick + drift + wrap fused with SETTLE on positions and then
// on velocities (with the constraint virial), for a system whose atoms all
// belong to SETTLE groups (no SHAKE, f32 state): one thread per water runs
// k_kick_drift<false, true>'s arithmetic on its three atoms, then the two
// SETTLE stages on the f32 values held in registers, the same values the
// unfused kernels pass through memory, so the results are bit-identical
// and the intermediate state and the snapshot skip the round trips.
__global__ void __launch_bounds__(128) k_kick_drift_settle(HdIntegrateArgs ia, HdSettleArgs a) {
  unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_groups) return;
  HdLattice L = hd_load_lattice(a.lattice);
  float f = ia.factor ? ia.factor[0] : 1.0f;
  float half_dt = ia.dt * 0.5f;
  unsigned ids[3];
  settle_load_ids(a, g, ids);
  float4 p[3], q[3];
  float v[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsigned i = ids[k];
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
    v[k][0] = vx; v[k][1] = vy; v[k][2] = vz;
  }
  SettleGeom G = settle_geom(a, g);
  SettleOut o = settle_water_positions(G, L, p, q, v, a.dt);
  settle_water_velocities<true>(G, L, o.p, o.v, o.w, a.dt);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsigned i = ids[k];
    a.posq[i] = o.p[k];
    a.vx[i] = o.v[k][0]; a.vy[i] = o.v[k][1]; a.vz[i] = o.v[k][2];
    a.w_cons[i] = o.w[k];
  }
}

void hd_launch_kick_drift_settle(hipStream_t st, const HdIntegrateArgs& ia,
                                 const HdSettleArgs& a) {
  if (!a.n_groups) return;
  hipLaunchKernelGGL(k_kick_drift_settle, dim3((a.n_groups + 127) / 128), dim3(128), 0, st, ia,
                     a);
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

// MTK-NPT cell-coupled updates, isotropic (kernels/mtk.cu:1-70,
// src/integrator/mtk_npt.rs:392-420).  The host forms the scalars in FP64
// and rounds them once:
//   kick:  v <- e v + phi (F / m),  e = exp(-alpha dt/2),
//          phi = (dt/2) sinh(alpha dt/4)/(alpha dt/4) exp(-alpha dt/4),
//          the exact solution of dv/dt = F/m - alpha v over dt/2;
//   drift: x <- e x + phi v,  e = exp(beta dt),
//          phi = dt sinh(beta dt/2)/(beta dt/2) exp(beta dt/2),
//          the exact solution of dx/dt = v + beta x over dt, after which
//          the box is scaled by e.  Positions are then wrapped into the
//          scaled cell with their image counts (scaling keeps fractional
//          coordinates, so only atoms that crossed a face move).
__global__ void k_mtk_kick(float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
                           const float* __restrict__ fx, const float* __restrict__ fy,
                           const float* __restrict__ fz, const float* __restrict__ mass, float e,
                           float phi, unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float inv_m = 1.0f / mass[i];
  vx[i] = e * vx[i] + phi * (fx[i] * inv_m);
  vy[i] = e * vy[i] + phi * (fy[i] * inv_m);
  vz[i] = e * vz[i] + phi * (fz[i] * inv_m);
}

__global__ void k_mtk_drift(float4* __restrict__ posq, int* __restrict__ ix, int* __restrict__ iy,
                            int* __restrict__ iz, const float* __restrict__ vx,
                            const float* __restrict__ vy, const float* __restrict__ vz,
                            const float* __restrict__ lattice, float e, float phi, unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  HdLattice L = hd_load_lattice(lattice);
  float4 pq = posq[i];
  float px = e * pq.x + phi * vx[i];
  float py = e * pq.y + phi * vy[i];
  float pz = e * pq.z + phi * vz[i];
  int ka, kb, kc;
  hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
  posq[i] = make_float4(px, py, pz, pq.w);
  ix[i] += ka;
  iy[i] += kb;
  iz[i] += kc;
}

__global__ void k_scale_velocities(float* __restrict__ vx, float* __restrict__ vy,
                                   float* __restrict__ vz, float f, unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] *= f;
  vy[i] *= f;
  vz[i] *= f;
}

void hd_launch_mtk_kick(hipStream_t st, float* vx, float* vy, float* vz, const float* fx,
                        const float* fy, const float* fz, const float* mass, float e, float phi,
                        unsigned n) {
  if (!n) return;
  hipLaunchKernelGGL(k_mtk_kick, dim3((n + 255) / 256), dim3(256), 0, st, vx, vy, vz, fx, fy, fz,
                     mass, e, phi, n);
}

void hd_launch_mtk_drift(hipStream_t st, float4* posq, int* ix, int* iy, int* iz, const float* vx,
                         const float* vy, const float* vz, const float* lattice, float e,
                         float phi, unsigned n) {
  if (!n) return;
  hipLaunchKernelGGL(k_mtk_drift, dim3((n + 255) / 256), dim3(256), 0, st, posq, ix, iy, iz, vx,
                     vy, vz, lattice, e, phi, n);
}

void hd_launch_scale_velocities(hipStream_t st, float* vx, float* vy, float* vz, float f,
                                unsigned n) {
  if (!n) return;
  hipLaunchKernelGGL(k_scale_velocities, dim3((n + 255) / 256), dim3(256), 0, st, vx, vy, vz, f,
                     n);
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
// Steepest descent (src/minimizer/steepest_descent.rs, kernels/minimize.cu):
// the trial move x += F * (step / F_max) keeps the pre-move positions for a
// restore; with a constraint frame it also writes the moved positions as
// the frame the position-only projection takes its directions from.
__global__ void k_sd_move(float4* __restrict__ posq, const float* __restrict__ fx,
                          const float* __restrict__ fy, const float* __restrict__ fz,
                          float4* __restrict__ save, float4* __restrict__ frame, float scale,
                          unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = posq[i];
  save[i] = p;
  p.x = p.x + fx[i] * scale;
  p.y = p.y + fy[i] * scale;
  p.z = p.z + fz[i] * scale;
  posq[i] = p;
  if (frame) frame[i] = p;
}

__global__ void k_sd_restore(float4* __restrict__ posq, const float4* __restrict__ save,
                             unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) posq[i] = save[i];
}

// max_i |F_i|^2 in FP64 by one workgroup (a max is exact in any order)
__global__ void __launch_bounds__(1024) k_fmax2(const float* __restrict__ fx,
                                                const float* __restrict__ fy,
                                                const float* __restrict__ fz, unsigned n,
                                                double* __restrict__ out) {
  __shared__ double sh[1024];
  double m = 0.0;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    double x = fx[i], y = fy[i], z = fz[i];
    m = fmax(m, x * x + y * y + z * z);
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (unsigned s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// Monte-Carlo barostat trial move (src/integrator/mc_barostat.rs:218-330,
// kernels/mc_barostat.cu): every molecule is shifted by (scale - 1) times
// its centre of mass (taken relative to its first atom through minimum
// images, in FP64), so molecules keep their shapes while their centres
// scale with the box.  mol_atoms == null: every atom is its own molecule.
__global__ void k_mc_scale_com(float4* __restrict__ posq, const unsigned* __restrict__ mol_off,
                               const unsigned* __restrict__ mol_atoms,
                               const float* __restrict__ mass, const float* __restrict__ lattice,
                               float scale, unsigned n_mol) {
  unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_mol) return;
  unsigned lo = mol_atoms ? mol_off[m] : m, hi = mol_atoms ? mol_off[m + 1] : m + 1;
  unsigned a0 = mol_atoms ? mol_atoms[lo] : m;
  float4 ref = posq[a0];
  HdLattice L = hd_load_lattice(lattice);
  double cx = 0.0, cy = 0.0, cz = 0.0, mt = 0.0;
  for (unsigned k = lo; k < hi; ++k) {
    unsigned a = mol_atoms ? mol_atoms[k] : k;
    float4 p = posq[a];
    double dx = (double)p.x - (double)ref.x, dy = (double)p.y - (double)ref.y,
           dz = (double)p.z - (double)ref.z;
    hd_min_image_f64(dx, dy, dz, L);
    double w = mass[a];
    cx += w * dx;
    cy += w * dy;
    cz += w * dz;
    mt += w;
  }
  double f = (double)scale - 1.0;
  float sx = (float)(f * ((double)ref.x + cx / mt));
  float sy = (float)(f * ((double)ref.y + cy / mt));
  float sz = (float)(f * ((double)ref.z + cz / mt));
  for (unsigned k = lo; k < hi; ++k) {
    unsigned a = mol_atoms ? mol_atoms[k] : k;
    float4 p = posq[a];
    p.x += sx;
    p.y += sy;
    p.z += sz;
    posq[a] = p;
  }
}

void hd_launch_mc_scale_com(hipStream_t st, float4* posq, const unsigned* mol_off,
                            const unsigned* mol_atoms, const float* mass, const float* lattice,
                            float scale, unsigned n_mol) {
  if (n_mol == 0) return;
  hipLaunchKernelGGL(k_mc_scale_com, dim3((n_mol + 255) / 256), dim3(256), 0, st, posq, mol_off,
                     mol_atoms, mass, lattice, scale, n_mol);
}

void hd_launch_sd_move(hipStream_t st, float4* posq, const float* fx, const float* fy,
                       const float* fz, float4* save, float4* frame, float scale, unsigned n) {
  hipLaunchKernelGGL(k_sd_move, dim3((n + 255) / 256), dim3(256), 0, st, posq, fx, fy, fz, save,
                     frame, scale, n);
}
void hd_launch_sd_restore(hipStream_t st, float4* posq, const float4* save, unsigned n) {
  hipLaunchKernelGGL(k_sd_restore, dim3((n + 255) / 256), dim3(256), 0, st, posq, save, n);
}
void hd_launch_fmax2(hipStream_t st, const float* fx, const float* fy, const float* fz,
                     unsigned n, double* out) {
  hipLaunchKernelGGL(k_fmax2, dim3(1), dim3(1024), 0, st, fx, fy, fz, n, out);
}

// ---------------------------------------------------------------------
// Force finalize: pair fixed point + SPME gather (+ constraint virial),
// scattered back to atom order, optionally fused with the post-force kick.
// ---------------------------------------------------------------------
template <int GATHER_P>
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

template <bool EV, bool GATHER, bool KICK, bool LOSSLESS, int GATHER_P = 4>
__global__ void __launch_bounds__(256) k_finalize(HdFinalizeArgs a) {
  // With the SPME gather four lanes share a slot: lane q of the quad reads
  // the a-planes da = q and q + 4 (< GATHER_P) of the P x P x P stencil
  // (P^2 grid values each), so four times as many reads are in flight; the
  // quad's partial sums are combined and its first lane finishes the slot.
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
    float wb[GATHER_P], dwb[GATHER_P], wc[GATHER_P], dwc[GATHER_P];
    int gib[GATHER_P], gic[GATHER_P];
#pragma unroll
    for (int d = 0; d < GATHER_P; ++d) {
      bspline_wd<GATHER_P>((float)d + rec.y, wb[d], dwb[d]);
      bspline_wd<GATHER_P>((float)d + rec.z, wc[d], dwc[d]);
      // periodic wrap (the origin cell lies in [0, n))
      gib[d] = gb0 - d;
      gib[d] += gib[d] < 0 ? (int)a.nb : 0;
      gic[d] = gc0 - d;
      gic[d] += gic[d] < 0 ? (int)a.nc : 0;
    }
#pragma unroll
    for (int h = 0; h < (GATHER_P + 3) / 4; ++h) {
      int da = (int)q + 4 * h;
      if (da >= GATHER_P) break;
      float wa, dwa;
      bspline_wd<GATHER_P>((float)da + rec.x, wa, dwa);
      int gia = ga0 - da;
      gia += gia < 0 ? (int)a.na : 0;
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
  if (ga && a.order == 5) {
#define HD_FIN5(E, K, LL) \
  if (ev == E && k == K && ll == LL) { \
    hipLaunchKernelGGL((k_finalize<E, true, K, LL, 5>), g, b, 0, st, a); \
    return; \
  }
    HD_FIN5(false, false, false) HD_FIN5(false, true, false) HD_FIN5(true, false, false)
    HD_FIN5(true, true, false) HD_FIN5(false, false, true) HD_FIN5(false, true, true)
    HD_FIN5(true, false, true) HD_FIN5(true, true, true)
#undef HD_FIN5
  }
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
// CSVR: K_new = c K + (K_target/g)(1 - c)(sum_{i=1}^{g-1} xi_i^2 + r^2)
//   + 2 r sqrt(c (1 - c) K K_target / g)           (Bussi et al. 2007, A7)
// with xi_i the Philox normal (seed, counter, i, 0) and r that of index 0
// (kernels/nose_hoover.cu:145-255): the g - 1 squares are summed in
// HD_CSVR_BLOCKS contiguous chunks (one block each, fixed tree), the chunks
// by the final kernel, so the sum is the same on every run.
__global__ void __launch_bounds__(256) k_csvr_partials(HdRescaleArgs t) {
  __shared__ double sh[256];
  unsigned long long ctr = *t.counter;
  unsigned clo = (unsigned)(ctr & 0xFFFFFFFFull), chi = (unsigned)(ctr >> 32);
  unsigned m = t.g_dof > 0u ? t.g_dof - 1u : 0u;
  unsigned chunk = (m + gridDim.x - 1u) / gridDim.x;
  unsigned start = blockIdx.x * chunk, end = min(start + chunk, m);
  double s = 0.0;
  for (unsigned j = start + threadIdx.x; j < end; j += blockDim.x) {
    double xi = philox_gaussian_f64(t.seed_lo, t.seed_hi, clo, chi, j + 1u, 0u);
    s += xi * xi;
  }
  double tot = block_tree_sum(s, sh);
  if (threadIdx.x == 0) t.partials[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_ke_final_rescale(const double* __restrict__ partials,
                                                          unsigned np, double* __restrict__ ke,
                                                          HdRescaleArgs t,
                                                          float* __restrict__ factor) {
  __shared__ double sh[256];
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < np; i += blockDim.x) s += partials[i];
  double k = block_tree_sum(s, sh);
  double sx = 0.0;
  if (t.kind == 2) {
    __syncthreads();
    double a = 0.0;
    for (unsigned i = threadIdx.x; i < HD_CSVR_BLOCKS; i += blockDim.x) a += t.partials[i];
    sx = block_tree_sum(a, sh);
  }
  if (threadIdx.x != 0) return;
  ke[0] = k;
  double f = 1.0;
  if (t.kind == 2) {
    unsigned long long ctr = *t.counter;
    double r = philox_gaussian_f64(t.seed_lo, t.seed_hi, (unsigned)(ctr & 0xFFFFFFFFull),
                                   (unsigned)(ctr >> 32), 0u, 0u);
    double cross = k > 0.0 ? 2.0 * r * sqrt(t.c * t.one_minus_c * k * t.kt_per_dof) : 0.0;
    double kn = t.c * k + t.kt_per_dof * t.one_minus_c * (sx + r * r) + cross;
    if (!isfinite(kn) || kn <= 0.0) kn = k;
    if (k > 0.0 && kn != k) f = sqrt(kn / k);
    t.injection[0] += kn - k;
    *t.counter = ctr + 1ull;
  } else if (k > 0.0) {
    double l2 = 1.0 + t.dt_over_tau * (t.k_target / k - 1.0);
    if (!(l2 > 0.0)) l2 = 0.0;
    f = sqrt(l2);
    t.injection[0] += k * (l2 - 1.0);
  }
  factor[0] = (float)f;
}

void hd_launch_ke_rescale(hipStream_t st, unsigned n, const float* vx, const float* vy,
                          const float* vz, const float* mass, double* partials, double* ke,
                          const HdRescaleArgs& t, float* factor, const HdDispArgs* disp) {
  HdDispArgs d;
  memset(&d, 0, sizeof(d));
  if (disp) d = *disp;
  hipLaunchKernelGGL(k_ke_partials, dim3(HD_REDUCE_BLOCKS), dim3(256), 0, st, n, vx, vy, vz,
                     mass, partials, d);
  if (t.kind == 2)
    hipLaunchKernelGGL(k_csvr_partials, dim3(HD_CSVR_BLOCKS), dim3(256), 0, st, t);
  hipLaunchKernelGGL(k_ke_final_rescale, dim3(1), dim3(256), 0, st, partials,
                     (unsigned)HD_REDUCE_BLOCKS, ke, t, factor);
}

// ---------------------------------------------------------------------
// Andersen thermostat (src/integrator/andersen.rs, kernels/andersen.cu):
// after the post-force kick, atom i collides with probability p: u from
// Philox (seed, counter, i, 3) lane 0 mapped to (0, 1); on a collision its
// velocity is redrawn as sqrt(kT/m) (f32) times the Philox normals of
// (seed, counter, i, 0..2).  The counter advances once per step, after.
__global__ void k_andersen(float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
                           const float* __restrict__ mass, const unsigned long long* counter,
                           unsigned seed_lo, unsigned seed_hi, float p, float kt, unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned long long ctr = *counter;
  unsigned clo = (unsigned)(ctr & 0xFFFFFFFFull), chi = (unsigned)(ctr >> 32);
  unsigned o[4];
  philox4x32_10(seed_lo, seed_hi, clo, chi, i, 3u, o);
  double u = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
  if (u >= (double)p) return;
  float sigma = sqrtf(kt / mass[i]);
  vx[i] = sigma * (float)philox_gaussian_f64(seed_lo, seed_hi, clo, chi, i, 0u);
  vy[i] = sigma * (float)philox_gaussian_f64(seed_lo, seed_hi, clo, chi, i, 1u);
  vz[i] = sigma * (float)philox_gaussian_f64(seed_lo, seed_hi, clo, chi, i, 2u);
}

__global__ void k_counter_advance(unsigned long long* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter += 1ull;
}

void hd_launch_andersen(hipStream_t st, float* vx, float* vy, float* vz, const float* mass,
                        unsigned long long* counter, unsigned long long seed, float p, float kt,
                        unsigned n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_andersen, dim3((n + 255) / 256), dim3(256), 0, st, vx, vy, vz, mass,
                     counter, (unsigned)(seed & 0xFFFFFFFFull), (unsigned)(seed >> 32), p, kt, n);
  hipLaunchKernelGGL(k_counter_advance, dim3(1), dim3(64), 0, st, counter);
}

// ---------------------------------------------------------------------
// SHAKE / RATTLE for general constraint groups (the constraint contract of
// kernels/shake.cu:96-490 and src/integrator/shake.rs: Gauss-Seidel
// sweeps, gradients from the pre-drift snapshot, the same tolerances and
// sweep cap, the position- and velocity-level constraint virial).
//
// One wave per group; lane k owns atom k and constraint k (a group holds at
// most 64 of each, well beyond the reference's 8 atoms / 12 constraints,
// so whole polymer chains are one group).
//   * Frame: lane k takes the bond vector b_k = minimg(x_k - x_(k-1)), an
//     inclusive wave scan gives P_k = b_1 + ... + b_k, and the group is
//     solved in coordinates P_k - P_mid relative to its middle atom.  A
//     chain longer than half the box does not fold back onto one atom's
//     image, and |r|^2 - d^2 resolves to the f32 spacing of the group's
//     extent rather than of its place in the box.
//   * Sweeps: the host colours the constraints so that constraints of one
//     colour share no atom; a sweep visits the colours in order and the
//     lanes of one colour update their disjoint atoms at once, which is a
//     sequential Gauss-Seidel sweep in (colour, index) order.
//   * Every atom moves by its correction; centre-of-mass sums are xor
//     butterflies read from lane 0.
// The CPU oracle restates the scan, sweep and butterfly orders.
// ---------------------------------------------------------------------

__device__ __forceinline__ void shake_stats(unsigned* st, unsigned it, int max_iter) {
  atomicAdd(&st[0], it);
  atomicMax(&st[1], it);
  if ((int)it >= max_iter) atomicAdd(&st[2], 1u);
}

// inclusive scan over the wave: lane k gets v_0 + ... + v_k, grouped by
// the offsets 1, 2, 4, ..., 32
__device__ __forceinline__ float wave_scan(float v, unsigned lane) {
#pragma unroll
  for (unsigned off = 1; off < 64; off <<= 1) {
    float u = __shfl_up(v, off);
    if (lane >= off) v = u + v;
  }
  return v;
}

// sum over the wave as lane 0 of an xor butterfly, broadcast
__device__ __forceinline__ float wave_total(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
  return __shfl(v, 0);
}

// the group frame of the wrapped coordinates (x, y, z) of lane k's atom
__device__ __forceinline__ void group_frame(float& x, float& y, float& z, unsigned lane,
                                            unsigned na, const HdLattice& L) {
  float xl = __shfl_up(x, 1u), yl = __shfl_up(y, 1u), zl = __shfl_up(z, 1u);
  float bx = 0.f, by = 0.f, bz = 0.f;
  if (lane >= 1 && lane < na) {
    bx = x - xl; by = y - yl; bz = z - zl;
    hd_min_image(bx, by, bz, L);
  }
  bx = wave_scan(bx, lane);
  by = wave_scan(by, lane);
  bz = wave_scan(bz, lane);
  unsigned mid = na / 2;
  x = bx - __shfl(bx, (int)mid);
  y = by - __shfl(by, (int)mid);
  z = bz - __shfl(bz, (int)mid);
}

__device__ __forceinline__ void wave_com(float x, float y, float z, float m, float& cx,
                                         float& cy, float& cz) {
  float mtot = wave_total(m);
  float inv = mtot > 0.f ? 1.0f / mtot : 0.f;
  cx = wave_total(m * x) * inv;
  cy = wave_total(m * y) * inv;
  cz = wave_total(m * z) * inv;
}

// One constraint's end atoms and colour, from the packed (i | j << 8) form
struct GroupBond {
  unsigned ai, aj, colour;
};

__device__ __forceinline__ GroupBond group_bond(const HdShakeArgs& a, unsigned c0, unsigned lane,
                                                unsigned nc) {
  GroupBond b{0u, 0u, 0xFFFFFFFFu};
  if (lane < nc) {
    unsigned packed = a.cij[c0 + lane];
    b.ai = packed & 0xFFu;
    b.aj = (packed >> 8) & 0xFFu;
    b.colour = a.color[c0 + lane];
  }
  return b;
}

// Positions: the drifted atoms (pos*) are moved along the snapshot bonds
// (g*) until every |r|^2 - d^2 is within the tolerance.
__global__ __launch_bounds__(64) void k_shake_group(HdShakeArgs a) {
  __shared__ float sx[64], sy[64], sz[64], posx[64], posy[64], posz[64], winv[64];
  const unsigned lane = threadIdx.x;
  const unsigned g = a.big ? a.big[blockIdx.x] : blockIdx.x;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
  unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
  unsigned atom = 0;
  float ux = 0.f, uy = 0.f, uz = 0.f, ox = 0.f, oy = 0.f, oz = 0.f, mk = 0.f, imk = 0.f;
  if (lane < na) {
    atom = a.gatoms[a0 + lane];
    float4 p = a.posq[atom], q = a.snap[atom];
    ux = p.x; uy = p.y; uz = p.z;
    ox = q.x; oy = q.y; oz = q.z;
    mk = a.mass[atom];
    imk = mk > 0.f ? 1.0f / mk : 0.f;
  }
  group_frame(ox, oy, oz, lane, na, L);
  group_frame(ux, uy, uz, lane, na, L);
  sx[lane] = ox; sy[lane] = oy; sz[lane] = oz;
  posx[lane] = ux; posy[lane] = uy; posz[lane] = uz;
  winv[lane] = imk;
  __syncthreads();
  const GroupBond b = group_bond(a, c0, lane, nc);
  float gx = 0.f, gy = 0.f, gz = 0.f, wsum = 0.f, dd = 0.f, acc = 0.f;
  if (lane < nc) {
    gx = sx[b.ai] - sx[b.aj];
    gy = sy[b.ai] - sy[b.aj];
    gz = sz[b.ai] - sz[b.aj];
    wsum = winv[b.ai] + winv[b.aj];
    dd = a.cr2[c0 + lane];
  }
  const float wi = winv[b.ai], wj = winv[b.aj];
  if (a.lam0) {
    // warm start: one move by the predicted multiplier, in sweep order
    for (unsigned col = 0; col < a.n_colors; ++col) {
      if (b.colour == col) {
        float lam = wsum > 0.f ? a.lam0[c0 + lane] / wsum : 0.f;
        float si = lam * wi, sj = lam * wj;
        posx[b.ai] -= si * gx; posy[b.ai] -= si * gy; posz[b.ai] -= si * gz;
        posx[b.aj] += sj * gx; posy[b.aj] += sj * gy; posz[b.aj] += sj * gz;
        acc = lam;
      }
      __syncthreads();
    }
  }
  int sweeps = 0;
  for (; sweeps < a.max_iter; ++sweeps) {
    int open = 0;
    for (unsigned col = 0; col < a.n_colors; ++col) {
      if (b.colour == col) {
        float rx = posx[b.ai] - posx[b.aj], ry = posy[b.ai] - posy[b.aj],
              rz = posz[b.ai] - posz[b.aj];
        float rr = rx * rx + ry * ry + rz * rz;
        float resid = rr - dd;
        if (fabsf(resid) > a.tol) {
          open = 1;
          float proj = rx * gx + ry * gy + rz * gz;
          float slope = 2.0f * proj * wsum;
          if (slope != 0.f) {
            float lam = resid / slope;
            float si = lam * wi, sj = lam * wj;
            posx[b.ai] -= si * gx; posy[b.ai] -= si * gy; posz[b.ai] -= si * gz;
            posx[b.aj] += sj * gx; posy[b.aj] += sj * gy; posz[b.aj] += sj * gz;
            acc += lam;
          }
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(open)) break;
  }
  if (lane == 0 && a.stats) shake_stats(a.stats, (unsigned)sweeps, a.max_iter);
  if (a.lam_out && lane < nc) a.lam_out[c0 + lane] = acc * wsum;
  const float fx = posx[lane], fy = posy[lane], fz = posz[lane];
  float cmx, cmy, cmz;
  wave_com(fx, fy, fz, mk, cmx, cmy, cmz);
  if (lane < na) {
    float rdt = a.dt != 0.f ? 1.0f / a.dt : 0.f;
    float rdt2 = a.dt != 0.f ? 1.0f / (a.dt * a.dt) : 0.f;
    float mx = fx - ux, my = fy - uy, mz = fz - uz;
    a.vx[atom] += mx * rdt; a.vy[atom] += my * rdt; a.vz[atom] += mz * rdt;
    float4 p = a.posq[atom];
    p.x += mx; p.y += my; p.z += mz;
    a.posq[atom] = p;
    float armx = fx - cmx, army = fy - cmy, armz = fz - cmz;
    a.w_cons[atom] = mk * rdt2 * (mx * armx + my * army + mz * armz);
  }
}

// Velocities: relative velocities along the final bonds are removed until
// every |v_rel . r| is within the tolerance; with ACC the velocity-level
// virial is added to the position-level one.
template <bool ACC>
__global__ __launch_bounds__(64) void k_rattle_group(HdShakeArgs a) {
  __shared__ float px[64], py[64], pz[64], vxs[64], vys[64], vzs[64];
  __shared__ float kx[64], ky[64], kz[64], winv[64];
  const unsigned lane = threadIdx.x;
  const unsigned g = a.big ? a.big[blockIdx.x] : blockIdx.x;
  HdLattice L = hd_load_lattice(a.lattice);
  unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
  unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
  unsigned atom = 0;
  float x = 0.f, y = 0.f, z = 0.f, mk = 0.f, imk = 0.f;
  vxs[lane] = vys[lane] = vzs[lane] = 0.f;
  if (lane < na) {
    atom = a.gatoms[a0 + lane];
    float4 p = a.posq[atom];
    x = p.x; y = p.y; z = p.z;
    vxs[lane] = a.vx[atom]; vys[lane] = a.vy[atom]; vzs[lane] = a.vz[atom];
    mk = a.mass[atom];
    imk = mk > 0.f ? 1.0f / mk : 0.f;
  }
  group_frame(x, y, z, lane, na, L);
  px[lane] = x; py[lane] = y; pz[lane] = z;
  kx[lane] = ky[lane] = kz[lane] = 0.f;
  winv[lane] = imk;
  __syncthreads();
  const GroupBond b = group_bond(a, c0, lane, nc);
  float rx = 0.f, ry = 0.f, rz = 0.f, rr = 0.f, wsum = 0.f;
  if (lane < nc) {
    rx = px[b.ai] - px[b.aj];
    ry = py[b.ai] - py[b.aj];
    rz = pz[b.ai] - pz[b.aj];
    rr = rx * rx + ry * ry + rz * rz;
    wsum = winv[b.ai] + winv[b.aj];
  }
  const float wi = winv[b.ai], wj = winv[b.aj];
  int sweeps = 0;
  for (; sweeps < a.max_iter; ++sweeps) {
    int open = 0;
    for (unsigned col = 0; col < a.n_colors; ++col) {
      if (b.colour == col) {
        float ux = vxs[b.ai] - vxs[b.aj], uy = vys[b.ai] - vys[b.aj], uz = vzs[b.ai] - vzs[b.aj];
        float along = ux * rx + uy * ry + uz * rz;
        if (fabsf(along) > a.rattle_tol) {
          open = 1;
          float slope = rr * wsum;
          if (slope != 0.f) {
            float mu = along / slope;
            float si = mu * wi, sj = mu * wj;
            vxs[b.ai] -= si * rx; vys[b.ai] -= si * ry; vzs[b.ai] -= si * rz;
            vxs[b.aj] += sj * rx; vys[b.aj] += sj * ry; vzs[b.aj] += sj * rz;
            kx[b.ai] -= si * rx; ky[b.ai] -= si * ry; kz[b.ai] -= si * rz;
            kx[b.aj] += sj * rx; ky[b.aj] += sj * ry; kz[b.aj] += sj * rz;
          }
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(open)) break;
  }
  if (lane == 0 && a.stats) shake_stats(a.stats + 3, (unsigned)sweeps, a.max_iter);
  if (ACC && a.dt > 0.f) {
    float cmx, cmy, cmz;
    wave_com(x, y, z, mk, cmx, cmy, cmz);
    if (lane < na) {
      float armx = x - cmx, army = y - cmy, armz = z - cmz;
      a.w_cons[atom] += mk * (1.0f / a.dt) * (kx[lane] * armx + ky[lane] * army + kz[lane] * armz);
    }
  }
  if (lane < na) {
    a.vx[atom] = vxs[lane]; a.vy[atom] = vys[lane]; a.vz[atom] = vzs[lane];
  }
}

// Small groups (at most SHAKE_SMALL_ATOMS atoms and SHAKE_SMALL_CONS
// constraints: waters as SHAKE groups, short fragments) run one group per
// lane instead of one per wave.  Each lane performs the wave kernels'
// f32 operations in the same order: the frame scan is the wave scan
// restricted to its first four lanes (offsets 1 and 2; the lanes beyond
// the group hold zero bonds), a colour's constraints share no atom so they
// are applied one after another in index order, and the centre-of-mass
// sums are the xor butterfly's (v0 + v2) + (v1 + v3) (the other lanes add
// exact zeros).  The sweep statistics are reduced over the wave before
// the atomics.
#define SHAKE_SMALL_ATOMS 4
#define SHAKE_SMALL_CONS 6

__device__ __forceinline__ void small_frame(float (&x)[SHAKE_SMALL_ATOMS],
                                            float (&y)[SHAKE_SMALL_ATOMS],
                                            float (&z)[SHAKE_SMALL_ATOMS], unsigned na,
                                            const HdLattice& L) {
  float b[3][SHAKE_SMALL_ATOMS];
#pragma unroll
  for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
    float bx = 0.f, by = 0.f, bz = 0.f;
    if (k >= 1 && (unsigned)k < na) {
      bx = x[k] - x[k - 1]; by = y[k] - y[k - 1]; bz = z[k] - z[k - 1];
      hd_min_image(bx, by, bz, L);
    }
    b[0][k] = bx; b[1][k] = by; b[2][k] = bz;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float s1[SHAKE_SMALL_ATOMS], s2[SHAKE_SMALL_ATOMS];
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) s1[k] = k >= 1 ? b[d][k - 1] + b[d][k] : b[d][k];
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) s2[k] = k >= 2 ? s1[k - 2] + s1[k] : s1[k];
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) b[d][k] = s2[k];
  }
  unsigned mid = na / 2;
  float px = b[0][0], py = b[1][0], pz = b[2][0];
#pragma unroll
  for (int k = 1; k < SHAKE_SMALL_ATOMS; ++k)
    if ((unsigned)k == mid) { px = b[0][k]; py = b[1][k]; pz = b[2][k]; }
#pragma unroll
  for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
    x[k] = b[0][k] - px; y[k] = b[1][k] - py; z[k] = b[2][k] - pz;
  }
}

// the wave butterfly's total of four lane values with zeros elsewhere
__device__ __forceinline__ float small_total(const float (&v)[SHAKE_SMALL_ATOMS]) {
  return (v[0] + v[2]) + (v[1] + v[3]);
}

__device__ __forceinline__ void small_stats(unsigned* st, bool live, unsigned it, int max_iter) {
  unsigned sum = live ? it : 0u, mx = sum, capped = (live && (int)it >= max_iter) ? 1u : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sum += (unsigned)__shfl_xor((int)sum, off);
    mx = max(mx, (unsigned)__shfl_xor((int)mx, off));
    capped += (unsigned)__shfl_xor((int)capped, off);
  }
  if ((threadIdx.x & 63u) == 0) {
    if (sum) atomicAdd(&st[0], sum);
    atomicMax(&st[1], mx);
    if (capped) atomicAdd(&st[2], capped);
  }
}

__global__ __launch_bounds__(256) void k_shake_small(HdShakeArgs a) {
  unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = t < a.n_small;
  unsigned g = live ? a.small[t] : 0u;
  HdLattice L = hd_load_lattice(a.lattice);
  int sweeps = 0;
  if (live) {
    unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
    unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
    unsigned atom[SHAKE_SMALL_ATOMS];
    float ux[SHAKE_SMALL_ATOMS], uy[SHAKE_SMALL_ATOMS], uz[SHAKE_SMALL_ATOMS];
    float sx[SHAKE_SMALL_ATOMS], sy[SHAKE_SMALL_ATOMS], sz[SHAKE_SMALL_ATOMS];
    float mk[SHAKE_SMALL_ATOMS], wv[SHAKE_SMALL_ATOMS];
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
      atom[k] = 0u; ux[k] = uy[k] = uz[k] = sx[k] = sy[k] = sz[k] = mk[k] = wv[k] = 0.f;
      if ((unsigned)k < na) {
        atom[k] = a.gatoms[a0 + k];
        float4 p = a.posq[atom[k]], q = a.snap[atom[k]];
        ux[k] = p.x; uy[k] = p.y; uz[k] = p.z;
        sx[k] = q.x; sy[k] = q.y; sz[k] = q.z;
        mk[k] = a.mass[atom[k]];
        wv[k] = mk[k] > 0.f ? 1.0f / mk[k] : 0.f;
      }
    }
    small_frame(sx, sy, sz, na, L);
    small_frame(ux, uy, uz, na, L);
    float px[SHAKE_SMALL_ATOMS], py[SHAKE_SMALL_ATOMS], pz[SHAKE_SMALL_ATOMS];
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) { px[k] = ux[k]; py[k] = uy[k]; pz[k] = uz[k]; }
    unsigned ci[SHAKE_SMALL_CONS], cj[SHAKE_SMALL_CONS], col[SHAKE_SMALL_CONS];
    float gx[SHAKE_SMALL_CONS], gy[SHAKE_SMALL_CONS], gz[SHAKE_SMALL_CONS];
    float ws[SHAKE_SMALL_CONS], dd[SHAKE_SMALL_CONS], acc[SHAKE_SMALL_CONS];
#pragma unroll
    for (int c = 0; c < SHAKE_SMALL_CONS; ++c) {
      ci[c] = cj[c] = 0u; col[c] = 0xFFFFFFFFu;
      gx[c] = gy[c] = gz[c] = ws[c] = dd[c] = acc[c] = 0.f;
      if ((unsigned)c < nc) {
        unsigned pk = a.cij[c0 + c];
        ci[c] = pk & 0xFFu; cj[c] = (pk >> 8) & 0xFFu;
        col[c] = a.color[c0 + c];
        float sxi = 0.f, syi = 0.f, szi = 0.f, sxj = 0.f, syj = 0.f, szj = 0.f, wi = 0.f, wj = 0.f;
#pragma unroll
        for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
          if ((unsigned)k == ci[c]) { sxi = sx[k]; syi = sy[k]; szi = sz[k]; wi = wv[k]; }
          if ((unsigned)k == cj[c]) { sxj = sx[k]; syj = sy[k]; szj = sz[k]; wj = wv[k]; }
        }
        gx[c] = sxi - sxj; gy[c] = syi - syj; gz[c] = szi - szj;
        ws[c] = wi + wj;
        dd[c] = a.cr2[c0 + c];
      }
    }
    // one constraint's move of its two atoms
    auto move = [&](int c, float lam) {
      float wi = 0.f, wj = 0.f;
#pragma unroll
      for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
        if ((unsigned)k == ci[c]) wi = wv[k];
        if ((unsigned)k == cj[c]) wj = wv[k];
      }
      float si = lam * wi, sj = lam * wj;
#pragma unroll
      for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
        if ((unsigned)k == ci[c]) { px[k] -= si * gx[c]; py[k] -= si * gy[c]; pz[k] -= si * gz[c]; }
        if ((unsigned)k == cj[c]) { px[k] += sj * gx[c]; py[k] += sj * gy[c]; pz[k] += sj * gz[c]; }
      }
    };
    if (a.lam0) {
      for (unsigned q = 0; q < a.n_colors; ++q)
#pragma unroll
        for (int c = 0; c < SHAKE_SMALL_CONS; ++c)
          if (col[c] == q) {
            float lam = ws[c] > 0.f ? a.lam0[c0 + c] / ws[c] : 0.f;
            move(c, lam);
            acc[c] = lam;
          }
    }
    for (; sweeps < a.max_iter; ++sweeps) {
      int open = 0;
      for (unsigned q = 0; q < a.n_colors; ++q)
#pragma unroll
        for (int c = 0; c < SHAKE_SMALL_CONS; ++c)
          if (col[c] == q) {
            float xi = 0.f, yi = 0.f, zi = 0.f, xj = 0.f, yj = 0.f, zj = 0.f;
#pragma unroll
            for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
              if ((unsigned)k == ci[c]) { xi = px[k]; yi = py[k]; zi = pz[k]; }
              if ((unsigned)k == cj[c]) { xj = px[k]; yj = py[k]; zj = pz[k]; }
            }
            float rx = xi - xj, ry = yi - yj, rz = zi - zj;
            float rr = rx * rx + ry * ry + rz * rz;
            float resid = rr - dd[c];
            if (fabsf(resid) > a.tol) {
              open = 1;
              float proj = rx * gx[c] + ry * gy[c] + rz * gz[c];
              float slope = 2.0f * proj * ws[c];
              if (slope != 0.f) {
                float lam = resid / slope;
                move(c, lam);
                acc[c] += lam;
              }
            }
          }
      if (!open) break;
    }
    if (a.lam_out)
#pragma unroll
      for (int c = 0; c < SHAKE_SMALL_CONS; ++c)
        if ((unsigned)c < nc) a.lam_out[c0 + c] = acc[c] * ws[c];
    float mx[SHAKE_SMALL_ATOMS], my[SHAKE_SMALL_ATOMS], mz[SHAKE_SMALL_ATOMS];
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
      mx[k] = mk[k] * px[k]; my[k] = mk[k] * py[k]; mz[k] = mk[k] * pz[k];
    }
    float mtot = small_total(mk);
    float inv = mtot > 0.f ? 1.0f / mtot : 0.f;
    float cmx = small_total(mx) * inv, cmy = small_total(my) * inv, cmz = small_total(mz) * inv;
    float rdt = a.dt != 0.f ? 1.0f / a.dt : 0.f;
    float rdt2 = a.dt != 0.f ? 1.0f / (a.dt * a.dt) : 0.f;
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
      if ((unsigned)k >= na) continue;
      float dx = px[k] - ux[k], dy = py[k] - uy[k], dz = pz[k] - uz[k];
      a.vx[atom[k]] += dx * rdt; a.vy[atom[k]] += dy * rdt; a.vz[atom[k]] += dz * rdt;
      float4 p = a.posq[atom[k]];
      p.x += dx; p.y += dy; p.z += dz;
      a.posq[atom[k]] = p;
      float armx = px[k] - cmx, army = py[k] - cmy, armz = pz[k] - cmz;
      a.w_cons[atom[k]] = mk[k] * rdt2 * (dx * armx + dy * army + dz * armz);
    }
  }
  if (a.stats) small_stats(a.stats, live, (unsigned)sweeps, a.max_iter);
}

template <bool ACC>
__global__ __launch_bounds__(256) void k_rattle_small(HdShakeArgs a) {
  unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = t < a.n_small;
  unsigned g = live ? a.small[t] : 0u;
  HdLattice L = hd_load_lattice(a.lattice);
  int sweeps = 0;
  if (live) {
    unsigned a0 = a.goff[g], na = a.goff[g + 1] - a0;
    unsigned c0 = a.coff[g], nc = a.coff[g + 1] - c0;
    unsigned atom[SHAKE_SMALL_ATOMS];
    float x[SHAKE_SMALL_ATOMS], y[SHAKE_SMALL_ATOMS], z[SHAKE_SMALL_ATOMS];
    float vx[SHAKE_SMALL_ATOMS], vy[SHAKE_SMALL_ATOMS], vz[SHAKE_SMALL_ATOMS];
    float kx[SHAKE_SMALL_ATOMS], ky[SHAKE_SMALL_ATOMS], kz[SHAKE_SMALL_ATOMS];
    float mk[SHAKE_SMALL_ATOMS], wv[SHAKE_SMALL_ATOMS];
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
      atom[k] = 0u;
      x[k] = y[k] = z[k] = vx[k] = vy[k] = vz[k] = kx[k] = ky[k] = kz[k] = mk[k] = wv[k] = 0.f;
      if ((unsigned)k < na) {
        atom[k] = a.gatoms[a0 + k];
        float4 p = a.posq[atom[k]];
        x[k] = p.x; y[k] = p.y; z[k] = p.z;
        vx[k] = a.vx[atom[k]]; vy[k] = a.vy[atom[k]]; vz[k] = a.vz[atom[k]];
        mk[k] = a.mass[atom[k]];
        wv[k] = mk[k] > 0.f ? 1.0f / mk[k] : 0.f;
      }
    }
    small_frame(x, y, z, na, L);
    unsigned ci[SHAKE_SMALL_CONS], cj[SHAKE_SMALL_CONS], col[SHAKE_SMALL_CONS];
    float rx[SHAKE_SMALL_CONS], ry[SHAKE_SMALL_CONS], rz[SHAKE_SMALL_CONS];
    float rr[SHAKE_SMALL_CONS], ws[SHAKE_SMALL_CONS];
#pragma unroll
    for (int c = 0; c < SHAKE_SMALL_CONS; ++c) {
      ci[c] = cj[c] = 0u; col[c] = 0xFFFFFFFFu;
      rx[c] = ry[c] = rz[c] = rr[c] = ws[c] = 0.f;
      if ((unsigned)c < nc) {
        unsigned pk = a.cij[c0 + c];
        ci[c] = pk & 0xFFu; cj[c] = (pk >> 8) & 0xFFu;
        col[c] = a.color[c0 + c];
        float xi = 0.f, yi = 0.f, zi = 0.f, xj = 0.f, yj = 0.f, zj = 0.f, wi = 0.f, wj = 0.f;
#pragma unroll
        for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
          if ((unsigned)k == ci[c]) { xi = x[k]; yi = y[k]; zi = z[k]; wi = wv[k]; }
          if ((unsigned)k == cj[c]) { xj = x[k]; yj = y[k]; zj = z[k]; wj = wv[k]; }
        }
        rx[c] = xi - xj; ry[c] = yi - yj; rz[c] = zi - zj;
        rr[c] = rx[c] * rx[c] + ry[c] * ry[c] + rz[c] * rz[c];
        ws[c] = wi + wj;
      }
    }
    for (; sweeps < a.max_iter; ++sweeps) {
      int open = 0;
      for (unsigned q = 0; q < a.n_colors; ++q)
#pragma unroll
        for (int c = 0; c < SHAKE_SMALL_CONS; ++c)
          if (col[c] == q) {
            float vxi = 0.f, vyi = 0.f, vzi = 0.f, vxj = 0.f, vyj = 0.f, vzj = 0.f, wi = 0.f,
                  wj = 0.f;
#pragma unroll
            for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
              if ((unsigned)k == ci[c]) { vxi = vx[k]; vyi = vy[k]; vzi = vz[k]; wi = wv[k]; }
              if ((unsigned)k == cj[c]) { vxj = vx[k]; vyj = vy[k]; vzj = vz[k]; wj = wv[k]; }
            }
            float ux = vxi - vxj, uy = vyi - vyj, uz = vzi - vzj;
            float along = ux * rx[c] + uy * ry[c] + uz * rz[c];
            if (fabsf(along) > a.rattle_tol) {
              open = 1;
              float slope = rr[c] * ws[c];
              if (slope != 0.f) {
                float mu = along / slope;
                float si = mu * wi, sj = mu * wj;
#pragma unroll
                for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
                  if ((unsigned)k == ci[c]) {
                    vx[k] -= si * rx[c]; vy[k] -= si * ry[c]; vz[k] -= si * rz[c];
                    kx[k] -= si * rx[c]; ky[k] -= si * ry[c]; kz[k] -= si * rz[c];
                  }
                  if ((unsigned)k == cj[c]) {
                    vx[k] += sj * rx[c]; vy[k] += sj * ry[c]; vz[k] += sj * rz[c];
                    kx[k] += sj * rx[c]; ky[k] += sj * ry[c]; kz[k] += sj * rz[c];
                  }
                }
              }
            }
          }
      if (!open) break;
    }
    if (ACC && a.dt > 0.f) {
      float mx[SHAKE_SMALL_ATOMS], my[SHAKE_SMALL_ATOMS], mz[SHAKE_SMALL_ATOMS];
#pragma unroll
      for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
        mx[k] = mk[k] * x[k]; my[k] = mk[k] * y[k]; mz[k] = mk[k] * z[k];
      }
      float mtot = small_total(mk);
      float inv = mtot > 0.f ? 1.0f / mtot : 0.f;
      float cmx = small_total(mx) * inv, cmy = small_total(my) * inv, cmz = small_total(mz) * inv;
#pragma unroll
      for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
        if ((unsigned)k >= na) continue;
        float armx = x[k] - cmx, army = y[k] - cmy, armz = z[k] - cmz;
        a.w_cons[atom[k]] += mk[k] * (1.0f / a.dt) * (kx[k] * armx + ky[k] * army + kz[k] * armz);
      }
    }
#pragma unroll
    for (int k = 0; k < SHAKE_SMALL_ATOMS; ++k) {
      if ((unsigned)k >= na) continue;
      a.vx[atom[k]] = vx[k]; a.vy[atom[k]] = vy[k]; a.vz[atom[k]] = vz[k];
    }
  }
  if (a.stats) small_stats(a.stats + 3, live, (unsigned)sweeps, a.max_iter);
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
  unsigned nb = a.big ? a.n_big : a.n_groups;
  if (nb) hipLaunchKernelGGL(k_shake_group, dim3(nb), dim3(64), 0, st, a);
  if (a.big && a.n_small)
    hipLaunchKernelGGL(k_shake_small, dim3((a.n_small + 255) / 256), dim3(256), 0, st, a);
}

void hd_launch_rattle_velocities(hipStream_t st, const HdShakeArgs& a, bool acc) {
  unsigned nb = a.big ? a.n_big : a.n_groups;
  if (nb) {
    if (acc)
      hipLaunchKernelGGL(k_rattle_group<true>, dim3(nb), dim3(64), 0, st, a);
    else
      hipLaunchKernelGGL(k_rattle_group<false>, dim3(nb), dim3(64), 0, st, a);
  }
  if (a.big && a.n_small) {
    dim3 g((a.n_small + 255) / 256);
    if (acc) hipLaunchKernelGGL(k_rattle_small<true>, g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_rattle_small<false>, g, dim3(256), 0, st, a);
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

// B + A, SETTLE on positions, O + A, SETTLE on positions and SETTLE on
// velocities (with the constraint virial) in one kernel, for a system
// whose atoms all belong to SETTLE groups: one thread per water runs
// k_lan_kick_drift's and k_lan_ou_drift's arithmetic on its three atoms
// and the SETTLE stages on the f32 values held in registers, the values
// the five separate kernels pass through memory, so the results are
// bit-identical.  Every block reads the draw counter before it takes a
// ticket (counter[1]); the block that takes the last ticket advances the
// counter and returns the ticket to 0, so all blocks draw with the value
// the O + A kernel would see and a captured step replays unchanged.
__global__ void __launch_bounds__(128) k_lan_settle(HdIntegrateArgs ia, HdSettleArgs a,
                                                    HdLangevinArgs l) {
  __shared__ unsigned long long s_ctr;
  if (threadIdx.x == 0) s_ctr = l.counter[0];
  __syncthreads();
  unsigned long long ctr = s_ctr;
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned long long t = atomicAdd(l.counter + 1, 1ull);
    if (t == (unsigned long long)gridDim.x - 1ull) {
      atomicExch(l.counter, ctr + 1ull);
      atomicExch(l.counter + 1, 0ull);
    }
  }
  unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_groups) return;
  HdLattice L = hd_load_lattice(a.lattice);
  float half_dt = ia.dt * 0.5f;
  unsigned ids[3];
  settle_load_ids(a, g, ids);
  int img[3][3];
  float4 p[3], q[3];
  float v[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsigned i = ids[k];
    float m = ia.mass[i];
    float ax = ia.fx[i] / m, ay = ia.fy[i] / m, az = ia.fz[i] / m;
    float vx = ia.vx[i] + ax * half_dt;
    float vy = ia.vy[i] + ay * half_dt;
    float vz = ia.vz[i] + az * half_dt;
    float4 pq = ia.posq[i];
    q[k] = pq;
    float px = pq.x + vx * half_dt, py = pq.y + vy * half_dt, pz = pq.z + vz * half_dt;
    int ka, kb, kc;
    hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
    img[k][0] = ia.ix[i] + ka; img[k][1] = ia.iy[i] + kb; img[k][2] = ia.iz[i] + kc;
    p[k] = make_float4(px, py, pz, pq.w);
    v[k][0] = vx; v[k][1] = vy; v[k][2] = vz;
  }
  SettleGeom G = settle_geom(a, g);
  SettleOut o = settle_water_positions(G, L, p, q, v, a.dt);
  unsigned clo = (unsigned)(ctr & 0xFFFFFFFFull), chi = (unsigned)(ctr >> 32);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsigned i = ids[k];
    float m = ia.mass[i];
    float sigma = sqrtf((1.0f - l.alpha * l.alpha) * l.kt / m);
    float xi0 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 0u);
    float xi1 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 1u);
    float xi2 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 2u);
    float vx = l.alpha * o.v[k][0] + sigma * xi0;
    float vy = l.alpha * o.v[k][1] + sigma * xi1;
    float vz = l.alpha * o.v[k][2] + sigma * xi2;
    float4 pq = o.p[k];
    q[k] = pq;
    float px = pq.x + vx * half_dt, py = pq.y + vy * half_dt, pz = pq.z + vz * half_dt;
    int ka, kb, kc;
    hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
    img[k][0] += ka; img[k][1] += kb; img[k][2] += kc;
    p[k] = make_float4(px, py, pz, pq.w);
    v[k][0] = vx; v[k][1] = vy; v[k][2] = vz;
  }
  o = settle_water_positions(G, L, p, q, v, a.dt);
  settle_water_velocities<true>(G, L, o.p, o.v, o.w, a.dt);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsigned i = ids[k];
    a.posq[i] = o.p[k];
    a.vx[i] = o.v[k][0]; a.vy[i] = o.v[k][1]; a.vz[i] = o.v[k][2];
    a.w_cons[i] = o.w[k];
    ia.ix[i] = img[k][0]; ia.iy[i] = img[k][1]; ia.iz[i] = img[k][2];
  }
}

// B + A and O + A in one kernel for a system without constraints: one
// thread per atom runs k_lan_kick_drift's and k_lan_ou_drift's arithmetic
// in registers (bit-identical to the two kernels); the draw counter is
// advanced by the block that takes the last ticket, as in k_lan_settle.
__global__ void __launch_bounds__(256) k_lan_free(HdIntegrateArgs a, HdLangevinArgs l) {
  __shared__ unsigned long long s_ctr;
  if (threadIdx.x == 0) s_ctr = l.counter[0];
  __syncthreads();
  unsigned long long ctr = s_ctr;
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned long long t = atomicAdd(l.counter + 1, 1ull);
    if (t == (unsigned long long)gridDim.x - 1ull) {
      atomicExch(l.counter, ctr + 1ull);
      atomicExch(l.counter + 1, 0ull);
    }
  }
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  HdLattice L = hd_load_lattice(a.lattice);
  float m = a.mass[i];
  float ax = a.fx[i] / m, ay = a.fy[i] / m, az = a.fz[i] / m;
  float half_dt = a.dt * 0.5f;
  float vx = a.vx[i] + ax * half_dt;
  float vy = a.vy[i] + ay * half_dt;
  float vz = a.vz[i] + az * half_dt;
  float4 pq = a.posq[i];
  float px = pq.x + vx * half_dt, py = pq.y + vy * half_dt, pz = pq.z + vz * half_dt;
  int ka, kb, kc;
  hd_wrap_with_image(px, py, pz, ka, kb, kc, L);
  int nx = a.ix[i] + ka, ny = a.iy[i] + kb, nz = a.iz[i] + kc;
  unsigned clo = (unsigned)(ctr & 0xFFFFFFFFull), chi = (unsigned)(ctr >> 32);
  float sigma = sqrtf((1.0f - l.alpha * l.alpha) * l.kt / m);
  float xi0 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 0u);
  float xi1 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 1u);
  float xi2 = (float)philox_gaussian_f64(l.seed_lo, l.seed_hi, clo, chi, i, 2u);
  vx = l.alpha * vx + sigma * xi0;
  vy = l.alpha * vy + sigma * xi1;
  vz = l.alpha * vz + sigma * xi2;
  float qx = px + vx * half_dt, qy = py + vy * half_dt, qz = pz + vz * half_dt;
  hd_wrap_with_image(qx, qy, qz, ka, kb, kc, L);
  a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
  a.posq[i] = make_float4(qx, qy, qz, pq.w);
  a.ix[i] = nx + ka; a.iy[i] = ny + kb; a.iz[i] = nz + kc;
}

void hd_launch_lan_free(hipStream_t st, const HdIntegrateArgs& a, const HdLangevinArgs& l) {
  // one block at least: the counter advances as with the separate kernels
  hipLaunchKernelGGL(k_lan_free, dim3(a.n ? (a.n + 255) / 256 : 1), dim3(256), 0, st, a, l);
}

void hd_launch_lan_settle(hipStream_t st, const HdIntegrateArgs& ia, const HdSettleArgs& a,
                          const HdLangevinArgs& l) {
  if (!a.n_groups) return;
  hipLaunchKernelGGL(k_lan_settle, dim3((a.n_groups + 127) / 128), dim3(128), 0, st, ia, a, l);
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
  // Berendsen (barostat.cu:228-275): the same step without the noise term,
  // the draw or the P0 dV account
  double r = b.stochastic ? philox_gaussian_f64(b.seed_lo, b.seed_hi,
                                                (unsigned)(ctr & 0xFFFFFFFFull),
                                                (unsigned)(ctr >> 32), 0u, 0u)
                          : 0.0;
  double det = -b.compressibility * (b.dt / b.tau) * (b.pressure - p);
  double amp = b.stochastic ? sqrt(2.0 * b.compressibility * b.kt * b.dt / (b.tau * v_pre)) : 0.0;
  double mu3 = 1.0 + det + amp * r;
  double mu3c = mu3 > b.mu3_min ? mu3 : b.mu3_min;
  double mu = cbrt(mu3c);
  double v_post = v_pre * mu3c;
  float muf = (float)mu;
  mu_out[0] = muf;
  for (int i = 0; i < 6; ++i) lattice[i] *= muf;
  diag[0] = p;
  diag[1] = v_post;
  if (b.stochastic) {
    diag[2] += b.pressure * (v_post - v_pre);
    *counter = ctr + 1ull;
  }
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
