// Host-side launch wrappers for the HeddleMD HIP kernels.  Every wrapper
// enqueues on the caller's stream and never synchronises.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

struct HdNeighborArgs {
  unsigned n, npad, nvox, cap;
  unsigned nv[3];
  float rs2, half_skin;
  const float4* posq;
  const float* lattice;
  float* ref_lattice;
  float4* ref_pos;
  const unsigned* voxel_rank;
  unsigned *atom_key, *voxel_count, *voxel_off, *voxel_cursor;
  unsigned* voxel_part;  // hd_voxel_scan_parts(nvox) + 1 chunk sums
  unsigned *sorted_id, *slot_of;
  const unsigned *type, *key;
  unsigned *tile_type, *tile_key;
  float4* tile_posq;
  float4 *center, *half;
  unsigned *entry_atoms, *entry_count, *max_count;
  unsigned *status, *rebuild_count;
  // atom-list mode: entry_atoms holds per-block rows of per-atom neighbour
  // lists (row k = k-th neighbour of each of the block's 64 atoms) and
  // entry_count the block's row count; exclusions are applied at build
  int atom_lists;
  unsigned excl_span;
  unsigned long long* block_mask;  // nblocks x hd_block_pair_chunks(nblocks) candidate masks
  unsigned* block_order;           // blocks by entry count, heaviest first (rebuild steps)
  unsigned* build_ticket;          // zeroed word for the folded rebuild-clear
};
unsigned hd_block_pair_chunks(unsigned nblocks);

// check = false when the displacement check already ran this step (KE pass)
void hd_launch_neighbor_update(hipStream_t st, const HdNeighborArgs& a, bool check = true);
// the displacement check alone (sets NB_STATUS_REBUILD)
void hd_launch_disp_check(hipStream_t st, const HdNeighborArgs& a);
// number of chunk sums of the voxel-count scan (the buffer needs one more)
unsigned hd_voxel_scan_parts(unsigned nvox);
// block_pairs = false when hd_launch_neighbor_small already ran the
// block-pair box tests of this step; fold_clear: the packed-list build's
// last workgroup orders the blocks and clears the rebuild request (no
// separate launch; needs build_ticket)
void hd_launch_neighbor_build(hipStream_t st, const HdNeighborArgs& a, bool block_pairs = true,
                              bool fold_clear = false);
void hd_launch_tile_gather(hipStream_t st, const float4* posq, const unsigned* sorted_id,
                           const float* lattice, float4* tile_posq, long long* fpj,
                           long long* fpj_ew, unsigned n, unsigned npad);
// the same, and on rebuild steps the slot types and keys and the blocks'
// bounding boxes (which the neighbour build then uses)
void hd_launch_tile_gather_rebuild(hipStream_t st, const HdNeighborArgs& a, long long* fpj,
                                   long long* fpj_ew);
// Systems of at most 8192 atoms and 8192 voxels: the displacement check
// (when check), the gated voxel binning, the tile gather with its
// rebuild-step extras and the block-pair box tests in one one-workgroup
// launch, with the results of hd_launch_neighbor_update +
// hd_launch_tile_gather_rebuild + the box tests of hd_launch_neighbor_build
bool hd_neighbor_small(const HdNeighborArgs& a);
// spl != nullptr: the gather also writes the SPME spline records of a grid
// of na x nb x nc cells (k_spline_setup's output)
void hd_launch_neighbor_small(hipStream_t st, const HdNeighborArgs& a, bool check, long long* fpj,
                              long long* fpj_ew, float4* spl = nullptr, unsigned* sg = nullptr,
                              unsigned na = 0, unsigned nb = 0, unsigned nc = 0);

struct HdPairArgs {
  unsigned n, npad, cap;
  int ntypes;
  const double *sigma, *eps, *cut, *rsw;  // host arrays, ntypes^2
  double kc, alpha, coul_rc, rmax, r_skin;
  const float* lattice;
  const float4* tile_posq;
  const unsigned *tile_type, *tile_key;
  unsigned excl_span;  // pairs with |key_i - key_j| <= excl_span are excluded
  const float4 *center, *half;
  const unsigned *entry_atoms, *entry_count;
  long long *fpi, *fpi_ew, *fpj, *fpj_ew;
  int atom_lists;  // per-atom full lists (no j side) instead of the packed cluster list
  int ortho;       // no box tilt
  const unsigned* block_order;  // workgroup -> block for the packed-list kernel
  const float4* lj_tab;  // ntypes > 4: the type-pair table (hd_pair_lj_table), device
};
// Largest type count of the LDS table form of the pair kernels.
int hd_pair_table_max_types();
// The type-pair table read by the pair kernels beyond four types: 8 floats
// per pair k = i nt + j: 12 c12, 6 c6, c12, c6, cut^2 (-1 when eps = 0),
// r_switch^2, 1/(cut^2 - r_switch^2) (0 without a switch), 0.
void hd_pair_lj_table(int nt, const double* sigma, const double* eps, const double* cut,
                      const double* rsw, float* out);
// ev: 0 forces only, 1 + per-atom energy and virial, 2 + per-atom virial
// only (no pair energy is accumulated)
void hd_launch_pair(hipStream_t st, const HdPairArgs& a, int ev);
void hd_launch_angles(hipStream_t st, const float4* posq, const float* lattice,
                      const unsigned* ijk, const double* kth, const double* th0,
                      const unsigned* slot_of, long long* fpj, long long* fpj_ew,
                      unsigned n_angles, bool ev);

// Two-body bonded terms: kind 0 harmonic (k, r0), 1 Morse (De, a, re);
// three parameters per bond.
void hd_launch_bonds(hipStream_t st, const float4* posq, const float* lattice,
                     const unsigned* ij, const int* kind, const double* prm,
                     const unsigned* slot_of, long long* fpj, long long* fpj_ew,
                     unsigned n_bonds, bool ev);
// Periodic dihedrals: k, n, phi0 per dihedral.
void hd_launch_dihedrals(hipStream_t st, const float4* posq, const float* lattice,
                         const unsigned* ijkl, const double* prm, const unsigned* slot_of,
                         long long* fpj, long long* fpj_ew, unsigned n_dihedrals, bool ev);

// Scaled exclusions: for each listed pair the pair kernel's contribution
// (weight 1, or 0 when the keys exclude the pair) is moved to the pair's
// scale by adding w times the LJ and real-space Coulomb terms.
struct HdExclPair {
  unsigned i, j;
  double sigma, eps, cut2, rsw2;  // LJ of the pair's types
  double w_lj, w_coul;            // scale minus the pair kernel's weight
};
void hd_launch_excl_scaled(hipStream_t st, const float4* posq, const float* lattice,
                           const HdExclPair* pairs, unsigned n_pairs, double kc, double alpha,
                           double coul_rc2, const unsigned* slot_of, long long* fpj,
                           long long* fpj_ew, bool ev);

struct HdSpmeArgs {
  unsigned n, na, nb, nc;
  double kc, alpha;
  const float* lattice;
  const float4* tile_posq;
  long long* rho_fixed;
  float2* cgrid;
  float* vgrid;
  const float2 *tw_a, *tw_b, *tw_c;
  const float *bf_a, *bf_b, *bf_c;
  double2* partials;
  double* recip;   // [E_recip, W_recip]
  float* w_share;  // W_recip / N
  float4* spl;   // per slot (t_a, t_b, t_c, q) of the spline origin
  unsigned* sg;  // per slot packed origin cell, 0xFFFFFFFF when q == 0
  // voxel binning of the last neighbour rebuild (slot ranges per voxel)
  const unsigned *voxel_rank, *voxel_off;
  unsigned nv[3];
  int mv[3];  // voxel margin per axis: >= half the skin in voxel units, plus one
  int order;  // B-spline order, 4 or 5
  int splines_done;  // spl / sg already hold this evaluation's records
  int no_plane;      // FFT: separate c and b passes even where a plane fits in LDS
};
void hd_launch_spme(hipStream_t st, const HdSpmeArgs& a);
// the spread half of hd_launch_spme (spline records + charge grid) alone
void hd_launch_spme_spread(hipStream_t st, const HdSpmeArgs& a);
void hd_launch_recip_reduce(hipStream_t st, const HdSpmeArgs& a, unsigned n_partials);

// Split-precision FP16 matrix-core DFT of the reciprocal grid (k_dft.hip)
struct HdDftArgs {
  unsigned na, nb, nc, nch;
  const long long* rho;
  float2* cgrid;
  float* vgrid;
  const float *bfa, *bfb, *bfc;
  const float* lattice;
  float kc, alpha;
  double2* partials;
};
// DFT matrices, FP16 high and 2^12-scaled low parts, [M][K] row-major:
// c-axis forward (M = 2 nch rounded up to 32, K = nc), c-axis inverse
// (M = nc, K = 2 nch rounded up to 8), b and a axes forward / inverse
// (M = K = 2n, the [[Cr, -Ci], [Ci, Cr]] block form)
struct HdDftMats {
  const _Float16 *cf_hi, *cf_lo, *ci_hi, *ci_lo;
  unsigned cf_m, cf_k, ci_m, ci_k;
  const _Float16 *bf_hi, *bf_lo, *bi_hi, *bi_lo, *af_hi, *af_lo, *ai_hi, *ai_lo;
};
void hd_launch_dft(hipStream_t st, const HdDftArgs& a, const HdDftMats& m);
unsigned hd_dft_partials_count(unsigned nb, unsigned nc);
unsigned hd_spme_partials_count(unsigned nb, unsigned nc);
// brick shape of the owner-computes spread; returns the number of bricks
unsigned hd_spread_bricks(unsigned na, unsigned nb, unsigned nc, unsigned brick[3]);

struct HdFinalizeArgs {
  unsigned n, npad;
  const unsigned* sorted_id;
  const float4* tile_posq;
  const long long *fpi, *fpi_ew, *fpj, *fpj_ew;
  // SPME gather (vgrid == nullptr disables), from the spline records
  const float* vgrid;
  const float4* spl;
  const unsigned* sg;
  unsigned na, nb, nc;
  int order;  // B-spline order of the gather, 4 or 5
  double kc, alpha;
  const float* w_share;
  const float* lattice;
  const float* w_cons;  // constraint virial per atom (nullptr = none)
  float *fx, *fy, *fz, *pe, *vir;
  // fused post-force kick (kick == false disables)
  bool kick;
  float dt;
  float *vx, *vy, *vz;
  double *vxlo, *vylo, *vzlo;  // lossless residuals (nullptr = f32 mode)
  const float* mass;
  const float* factor;  // thermostat factor applied after the kick
  // per-workgroup FP64 sums of the per-atom energy and virial (nullptr =
  // none), 2 * hd_finalize_blocks() doubles
  double* partials;
};
unsigned hd_finalize_blocks(unsigned n, bool gather);
void hd_launch_finalize(hipStream_t st, const HdFinalizeArgs& a, bool ev);
// sum of np (a, b) partial pairs into out[0], out[1], in fixed order
void hd_launch_sum2_final(hipStream_t st, const double* partials, unsigned np, double* out);

struct HdIntegrateArgs {
  unsigned n;
  float dt;
  float4* posq;
  int *ix, *iy, *iz;
  float *vx, *vy, *vz;
  double *xlo, *ylo, *zlo, *vxlo, *vylo, *vzlo;
  const float *fx, *fy, *fz, *mass;
  const float* lattice;
  const float* factor;  // thermostat factor applied before the kick
  float4* snap;         // pre-drift positions (nullptr = no snapshot)
};
void hd_launch_kick_drift(hipStream_t st, const HdIntegrateArgs& a);

// Langevin BAOAB half-steps (kernels/langevin.cu:13-80).  Uses the fields
// n, dt, posq, ix..iz, vx..vz, fx..fz, mass, lattice, snap of `a`.
struct HdLangevinArgs {
  float alpha, kt;                    // exp(-gamma dt) and k_B T, f32
  unsigned seed_lo, seed_hi;          // Philox key
  unsigned long long* counter;        // device draw counter; counter[1] is
                                      // hd_launch_lan_settle's block ticket,
                                      // 0 between launches
};
// B + A: kick with the current forces, then x += v dt/2 (advances the
// draw counter by one for the O step that follows)
void hd_launch_lan_kick_drift(hipStream_t st, const HdIntegrateArgs& a, const HdLangevinArgs& l);
// O + A: Ornstein-Uhlenbeck velocity update, then x += v dt/2
void hd_launch_lan_ou_drift(hipStream_t st, const HdIntegrateArgs& a, const HdLangevinArgs& l);
// B + A and O + A in one kernel for a system without constraints,
// bit-identical to the two separate kernels
void hd_launch_lan_free(hipStream_t st, const HdIntegrateArgs& a, const HdLangevinArgs& l);

#define HD_SETTLE_KINDS 8
struct HdSettleArgs {
  unsigned n_groups;
  const unsigned* groups;  // 3 per group: O, H1, H2
  double d_oh, d_hh;  // constrained O-H and H-H distances (kind 0)
  float m_o, m_h, dt;  // masses of kind 0
  // further geometries: kind k >= 1 of group g = gkind[g] (nullptr: every
  // group is kind 0)
  const unsigned char* gkind;
  double kd_oh[HD_SETTLE_KINDS], kd_hh[HD_SETTLE_KINDS];
  float km_o[HD_SETTLE_KINDS], km_h[HD_SETTLE_KINDS];
  float4* posq;
  const float4* snap;
  float *vx, *vy, *vz;
  float* w_cons;  // per-atom constraint virial (written)
  const float* lattice;
};
void hd_launch_settle_positions(hipStream_t st, const HdSettleArgs& a);
// hd_launch_kick_drift (f32 state), hd_launch_settle_positions and
// hd_launch_settle_velocities (with the virial) in one kernel; valid when
// the SETTLE groups hold every atom
void hd_launch_kick_drift_settle(hipStream_t st, const HdIntegrateArgs& ia,
                                 const HdSettleArgs& a);
void hd_launch_settle_velocities(hipStream_t st, const HdSettleArgs& a, bool accumulate_virial);
// B + A, SETTLE (positions), O + A, SETTLE (positions, then velocities with
// the virial) in one kernel, bit-identical to the five separate kernels;
// valid when the SETTLE groups hold every atom and no SHAKE group exists
void hd_launch_lan_settle(hipStream_t st, const HdIntegrateArgs& ia, const HdSettleArgs& a,
                          const HdLangevinArgs& l);

// General SHAKE / RATTLE groups (kernels/shake.cu): atoms goff[g] ..
// goff[g+1] of gatoms, constraints coff[g] .. coff[g+1] with group-local
// indices packed (i | j << 8) and r2 = (f32 d)^2.  Groups within the
// reference's 8-atom cap: one thread per group, its atoms and constraints
// in LDS at a fixed per-thread stride.  Longer groups: one wave per group,
// constraints swept colour by colour.
#define HD_SHAKE_MAX_ATOMS 64
#define HD_SHAKE_MAX_CONS 64
#define HD_SHAKE_REF_ATOMS 8   // MAX_GROUP_ATOMS (src/integrator/shake.rs:21)
#define HD_SHAKE_MAX_COLORS 16
struct HdShakeArgs {
  // one wave per group (at most 64 atoms and 64 constraints each)
  unsigned n_groups;
  const unsigned char* color;  // constraint colour (sweep order)
  unsigned n_colors;           // colours used by any group
  const unsigned *goff, *gatoms, *coff, *cij;
  const float* cr2;
  float4* posq;
  const float4* snap;
  float *vx, *vy, *vz;
  const float* mass;
  float* w_cons;
  const float* lattice;
  float dt, tol, rattle_tol;
  int max_iter;
  unsigned* stats;  // [sum sweeps, max sweeps, unconverged] for SHAKE then RATTLE
  // multipliers in the dimensionless form nu = lambda (1/m_i + 1/m_j): the
  // fraction of the snapshot bond vector that the constraint moves
  const float* lam0;  // per-constraint nu applied before the sweeps (or null)
  float* lam_out;     // per-constraint total nu of this call (or null)
  // with `big` set, the wave kernels take the groups big[0 .. n_big) and
  // the lane-per-group kernels the groups small[0 .. n_small) (at most 4
  // atoms and 6 constraints); without it every group runs on a wave
  const unsigned *big, *small;
  unsigned n_big, n_small;
  const unsigned* cgrp;  // constraint -> group (feature kernel)
  float* feat;           // n_cons x HD_SHAKE_NFEAT predictor inputs (feature kernel)
  unsigned n_cons;
};
// LDS bytes per thread of a group with the given maxima
void hd_launch_shake_positions(hipStream_t st, const HdShakeArgs& a);
void hd_launch_rattle_velocities(hipStream_t st, const HdShakeArgs& a, bool accumulate_virial);
// Predictor inputs of every constraint (HD_SHAKE_NFEAT floats each) into
// a.feat, from posq (after the drift) and snap.
#define HD_SHAKE_NFEAT 8
void hd_launch_shake_features(hipStream_t st, const HdShakeArgs& a);

// Multiplier predictor: an 8 -> 64 -> 64 -> 1 tanh network on the scaled
// features, predicting the correction to the first-order multiplier:
//   u = x / s (columns 0-4 and rho; the couplings 5, 6 unscaled),
//   h1 = tanh(W1 u + b1), h2 = tanh(W2 h1 + b2), o = w3.h2 + b3,
//   nu = x0 + s o.
// Parameter pack (floats): W1[64][8], b1[64], W2[64][64], b2[64], w3[64],
// b3, s.  Training minimises the mean of (o - (t - x0) / s)^2.
#define HD_MLP_IN 8
#define HD_MLP_H 64
#define HD_MLP_W1 0
#define HD_MLP_B1 (HD_MLP_W1 + HD_MLP_H * HD_MLP_IN)
#define HD_MLP_W2 (HD_MLP_B1 + HD_MLP_H)
#define HD_MLP_B2 (HD_MLP_W2 + HD_MLP_H * HD_MLP_H)
#define HD_MLP_W3 (HD_MLP_B2 + HD_MLP_H)
#define HD_MLP_B3 (HD_MLP_W3 + HD_MLP_H)
#define HD_MLP_NTRAIN (HD_MLP_B3 + 1)
#define HD_MLP_SCALE HD_MLP_NTRAIN
#define HD_MLP_NPARAM (HD_MLP_NTRAIN + 1)
#define HD_MLP_PSTRIDE 4864  // floats per block partial: gradient, then the loss sum
#define HD_MLP_SAMPLES 64    // samples per block of the gradient kernel
#define HD_MLP_NOALLOC_ROWS 16384  // rows per hd_mlp_gradient_noalloc call
// nu[c] = prediction for feature row c
void hd_launch_mlp_predict(hipStream_t st, const float* params, const float* feat, unsigned n,
                           float* nu);
// Gradient partials of one mini-batch: samples order(first + q), q < count,
// with order(m) = (mul * m + add) mod n; block b writes the gradient sum of
// its 64 samples and their loss sum to partials[b * HD_MLP_PSTRIDE ...].
// inv_count scales the gradient (1 / samples in the mini-batch).
// Returns the number of blocks.
unsigned hd_launch_mlp_grad(hipStream_t st, const float* params, const float* x,
                            const float* t, uint64_t n, uint64_t first, unsigned count,
                            uint64_t mul, uint64_t add, float inv_count, float* partials);
// Sums the partials of nb blocks in block order; grad_out (or null)
// receives the gradient; with adam non-null, one Adam step:
//   m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, p -= lr_t m / (sqrt(v) + eps)
// (lr_t carries the bias correction); loss_acc[0] += the loss sum.
struct HdAdam {
  float b1, b2, lr_t, eps;
  float *m, *v;
};
void hd_launch_mlp_reduce(hipStream_t st, const float* partials, unsigned nb, float* params,
                          float* grad_out, const HdAdam* adam, double* loss_acc);
// params[HD_MLP_SCALE] = rms of column 0 of x (n rows), 1 if that is 0.
void hd_launch_mlp_scale(hipStream_t st, const float* x, uint64_t n, float* params);
// Appends this step's rows (feat, nu) at the slot it claims by advancing
// counter[0]; slots at or beyond cap (n rows each) are not written.
void hd_launch_shake_log(hipStream_t st, const float* feat, const float* nu, unsigned n,
                         float* log_x, float* log_t, unsigned long long* counter,
                         unsigned long long cap);
// Device view of a context's SHAKE log (synchronises the context stream):
// rows = logged steps x constraints (host runtime, used by the trainer).
struct hd_ctx;
extern "C" int hd_ctx_shake_log_view(hd_ctx* c, int* device, const float** x,
                                     const float** t, uint64_t* rows);

// Stochastic velocity rescaling (kind 2: CSVR, Bussi, Donadio and
// Parrinello, J. Chem. Phys. 126, 014101 (2007); src/integrator/csvr.rs,
// kernels/nose_hoover.cu:145-255) and the weak-coupling rescale (kind 3:
// Berendsen, src/integrator/berendsen.rs, nose_hoover.cu:390-417), applied
// as the post-force velocity factor.
#define HD_CSVR_BLOCKS 64
// Steepest descent: save posq, posq += F * scale (and copy the moved
// positions to `frame` when non-null); restore; max |F|^2 in FP64.
void hd_launch_sd_move(hipStream_t st, float4* posq, const float* fx, const float* fy,
                       const float* fz, float4* save, float4* frame, float scale, unsigned n);
void hd_launch_sd_restore(hipStream_t st, float4* posq, const float4* save, unsigned n);
// MC barostat: molecule centres scaled by `scale` (rigid shift per molecule)
void hd_launch_mc_scale_com(hipStream_t st, float4* posq, const unsigned* mol_off,
                            const unsigned* mol_atoms, const float* mass, const float* lattice,
                            float scale, unsigned n_mol);
void hd_launch_fmax2(hipStream_t st, const float* fx, const float* fy, const float* fz,
                     unsigned n, double* out);
// Andersen collisions after the post-force kick, then the draw counter + 1
void hd_launch_andersen(hipStream_t st, float* vx, float* vy, float* vz, const float* mass,
                        unsigned long long* counter, unsigned long long seed, float p, float kt,
                        unsigned n);
struct HdDispArgs;
struct HdRescaleArgs {
  int kind;                        // 2 CSVR, 3 Berendsen
  unsigned seed_lo, seed_hi, g_dof;
  double c, one_minus_c, kt_per_dof;  // CSVR: exp(-dt/tau), 1 - c, K_target / g
  double k_target, dt_over_tau;       // Berendsen
  unsigned long long* counter;  // CSVR Philox draw counter (one per step)
  double* injection;            // cumulative K_new - K_old
  double* partials;             // CSVR: HD_CSVR_BLOCKS sums of xi^2
};
// The kinetic energy (and the displacement check) as hd_launch_ke, then the
// rescale factor of `t` into factor[0].
void hd_launch_ke_rescale(hipStream_t st, unsigned n, const float* vx, const float* vy,
                          const float* vz, const float* mass, double* partials, double* ke,
                          const HdRescaleArgs& t, float* factor, const HdDispArgs* disp);

struct HdNhcArgs {
  int chain, n_yoshida, n_resp;
  double w[7];
  double kt, g_dof;
  double* state;  // xi[chain], pxi[chain], q[chain]
};
// Neighbour status word bits (device-side rebuild request and list state).
#define NB_STATUS_REBUILD 1u
#define NB_STATUS_HIGHWATER 2u
#define NB_STATUS_OVERFLOW 8u
#define NB_STATUS_FORCE 0x100u

// Optional neighbour displacement check folded into the KE pass (the
// positions final for the step): sets NB_STATUS_REBUILD in *status when
// an atom has moved more than half_skin (plus the lattice change) since the
// last rebuild, as k_disp_check does.
struct HdDispArgs {
  const float4* posq;
  const float4* ref;
  const float* lattice;
  const float* ref_lattice;
  float half_skin;
  unsigned* status;
};
// KE reduction in FP64 (fixed chunking and tree) into ke[0]; when nhc is
// non-null also advances the chain by one half step and writes the
// cumulative velocity factor (f32) to factor[0].
void hd_launch_ke(hipStream_t st, unsigned n, const float* vx, const float* vy, const float* vz,
                  const float* mass, double* partials, double* ke, const HdNhcArgs* nhc,
                  float dt, float* factor, const HdDispArgs* disp = nullptr);
struct HdBarostatArgs {
  unsigned seed_lo, seed_hi;
  double pressure, tau, compressibility, kt, dt, mu3_min;
  int stochastic;  // 1 C-rescale (noise, draw counter, P0 dV), 0 Berendsen
};
// C-rescale: the scalar reduction of hd_launch_sum2_final whose last
// thread also computes mu from K = k[0], W = the reduced virial (out[1])
// and a Philox normal, and scales the device lattice by mu;
// hd_launch_scale_positions then scales every position by mu
void hd_launch_sum2_final_crescale(hipStream_t st, const double* partials, unsigned np,
                                   double* out, const double* k, float* lattice, float* mu,
                                   double* diag, unsigned long long* counter,
                                   const HdBarostatArgs& b);
void hd_launch_scale_positions(hipStream_t st, float4* posq, const float* mu, unsigned n);
// Pairs within r_c (both atoms valid, not excluded) in the current list.
void hd_launch_count_pairs(hipStream_t st, const HdPairArgs& a, unsigned long long* out);
void hd_launch_fill_f32(hipStream_t st, float* p, float v, size_t count);
// MTK-NPT cell-coupled half kick and drift (the drift wraps into the
// lattice it is given, the already scaled one), and a plain velocity scale
void hd_launch_mtk_kick(hipStream_t st, float* vx, float* vy, float* vz, const float* fx,
                        const float* fy, const float* fz, const float* mass, float e, float phi,
                        unsigned n);
void hd_launch_mtk_drift(hipStream_t st, float4* posq, int* ix, int* iy, int* iz, const float* vx,
                         const float* vy, const float* vz, const float* lattice, float e,
                         float phi, unsigned n);
void hd_launch_scale_velocities(hipStream_t st, float* vx, float* vy, float* vz, float f,
                                unsigned n);
void hd_launch_negate_velocities(hipStream_t st, float* vx, float* vy, float* vz, double* vxlo,
                                 double* vylo, double* vzlo, unsigned n);
void hd_launch_selftest_mfma(hipStream_t st, float* out);
void hd_launch_selftest_mfma_k16(hipStream_t st, float* out, bool bf16);
void hd_launch_selftest_dpp(hipStream_t st, int* out);

#define HD_REDUCE_BLOCKS 256
