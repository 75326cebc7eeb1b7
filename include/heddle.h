/*
 * heddle.h -- C ABI of the MI355X-native HeddleMD MD step (libheddle.so).
 *
 * The reference binds its GPU work from Rust through cudarc on one CUDA
 * stream (src/gpu/device.rs:34-73, src/gpu/kernels.rs) behind the
 * Potential / Integrator / Thermostat / Constraint slot traits
 * (src/forces/mod.rs:79-292, src/integrator/mod.rs:233-967,
 * src/integrator/constraint.rs:43).  This header is the boundary that
 * replaces that binding: plain pointers and sizes, no C++ or torch types,
 * host arrays owned by the caller, device memory owned by the context, all
 * calls ordered on one HIP stream per context (inside a force evaluation
 * the SPME pipeline forks onto a second stream of the context and joins
 * back before the forces are finalized, so callers see one ordered
 * stream).  A context is not thread safe; separate contexts are
 * independent.
 *
 * Units: Hartree atomic units (Bohr, electron mass, e, Hartree, atomic
 * time; k_B = k_C = 1) as the reference uses internally
 * (src/units/atomic_constants.rs).  Lattice = [lx, ly, lz, xy, xz, yz].
 *
 * Every int-returning call returns HD_OK (0) or a negative HD_ERR_* code;
 * hd_last_error() then holds a message.
 */
#ifndef HEDDLE_H
#define HEDDLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HD_OK 0
#define HD_ERR_ARG -1
#define HD_ERR_HIP -2
#define HD_ERR_STATE -3
#define HD_ERR_UNSUPPORTED -4
#define HD_ERR_OVERFLOW -5

typedef struct hd_ctx hd_ctx;

/* Library version string. */
const char *hd_version(void);

/* Create a context on HIP device `device` (the reference hard-wires device 0
 * and one stream: src/gpu/device.rs:66-73). */
int hd_ctx_create(int device, hd_ctx **out);
void hd_ctx_destroy(hd_ctx *ctx);
/* Message of the last failed call on this context ("" if none). */
const char *hd_last_error(const hd_ctx *ctx);

/* Particle state upload (src/gpu/buffers.rs:11-44 ParticleBuffers::new).
 * posq: 4n floats (x, y, z, q); mass, type: n each; excl_key: n molecule
 * keys (pairs with equal keys are excluded from the pair terms, the
 * constraint-group exclusions of src/forces/topology.rs:1225-1245), or
 * NULL for "no exclusions". */
int hd_set_system(hd_ctx *ctx, uint64_t n, const float *posq, const float *mass,
                  const uint32_t *type, const uint32_t *excl_key);
/* Harmonic angles (src/forces/angle.rs): atoms ijk[3a..3a+2] (j the apex),
 * force constant k_theta[a] (energy / rad^2) and rest angle theta0[a]
 * (rad).  Before hd_setup. */
int hd_set_angles(hd_ctx *ctx, uint64_t n_angles, const uint32_t *ijk, const double *k_theta,
                  const double *theta0);
/* Exclusion rule: a pair is excluded from the LJ and real-space terms when
 * |excl_key_i - excl_key_j| <= span (default 0: equal keys, i.e. whole
 * molecules).  Keys chain * stride + bead with span 2 give the 1-2 / 1-3
 * exclusions of linear chains (the CSR lists of src/forces/topology.rs
 * for these topologies).  Before hd_setup. */
int hd_set_exclusion_span(hd_ctx *ctx, uint32_t span);
/* Two-body bonded terms (src/forces/harmonic_bond.rs, src/forces/morse.rs;
 * the [bonds] section of src/forces/topology.rs): atoms ij[2b], ij[2b+1];
 * kind[b] 0 harmonic U = k/2 (r - r0)^2 with prm[3b..] = (k, r0, unused),
 * 1 Morse U = De (1 - exp(-a (r - re)))^2 with prm[3b..] = (De, a, re).
 * FP64 arithmetic.  Before hd_setup. */
int hd_set_bonds(hd_ctx *ctx, uint64_t n_bonds, const uint32_t *ij, const int *kind,
                 const double *prm);
/* Periodic dihedrals U = k (1 + cos(n phi - phi0)) (src/forces/dihedral.rs:23,
 * 360-460): atoms ijkl[4d..4d+3], IUPAC angle (trans = pi).  FP64.  Before
 * hd_setup. */
int hd_set_dihedrals(hd_ctx *ctx, uint64_t n_dihedrals, const uint32_t *ijkl,
                     const double *k_phi, const double *n_mult, const double *phi0);
/* The effective exclusion list of src/forces/topology.rs:1184-1290 (explicit
 * rows, then bond, angle, constraint-group and scaled 1-4 pairs): pair p is
 * (ij[2p], ij[2p+1]) with LJ scale s_lj[p] and real-space Coulomb scale
 * s_coul[p] in [0, 1] (lj.rs:295-321, spme.rs:723-762; unlisted pairs have
 * scale 1).  Replaces the keys of hd_set_system and the span.  The pair
 * kernel leaves out pairs by key: when the fully excluded pairs form whole
 * molecules (cliques) or runs of consecutive atoms excluded up to a fixed
 * index distance, exactly those; otherwise every pair inside each connected
 * molecule, and the molecule's pairs that are not fully excluded are added
 * back in FP64 by a correction pass (at most 8 Mi such pairs, else
 * HD_ERR_UNSUPPORTED).  Scaled pairs are any.  After hd_set_system, before
 * hd_setup. */
int hd_set_exclusions(hd_ctx *ctx, uint64_t n_pairs, const uint32_t *ij, const double *s_lj,
                      const double *s_coul);
/* Box (src/pbc.rs SimulationBox lattice). */
int hd_set_lattice(hd_ctx *ctx, const double lattice[6]);
int hd_set_velocities(hd_ctx *ctx, const float *vx, const float *vy, const float *vz);
/* Image counters (io/init_state.rs InitImages); NULL resets to zero. */
int hd_set_images(hd_ctx *ctx, const int32_t *ix, const int32_t *iy, const int32_t *iz);

/* Lennard-Jones type-pair tables, ntypes <= 32: up to four types the pair
 * kernels hold the i-row in registers, beyond that they read the table from
 * LDS (src/forces/lj.rs:22-335). */
int hd_set_lj(hd_ctx *ctx, int ntypes, const double *sigma, const double *epsilon,
              const double *cutoff, const double *r_switch);
/* SPME: real-space erfc pair term + reciprocal grid
 * (src/forces/spme.rs:36-52 SpmeParameters).  Grid sides of 2^a 3^b 5^c
 * points (the small primes the reference's cuFFT plans take), at least 8,
 * at most 256 along a and b and at most 512 (even) along c; B-spline order
 * 4 or 5. */
int hd_set_spme(hd_ctx *ctx, double alpha, double r_cut, int na, int nb, int nc, int order,
                double k_c);
/* Reciprocal-space transform: mode 0 = FP32 radix FFT in LDS (default;
 * where an a-plane fits in LDS the c and b passes run per plane in one
 * kernel each way), mode 1 = split-precision FP16 matrix-core DFT (each 1D
 * stage a GEMM with high/low FP16 operands accumulated in FP32;
 * power-of-two axes of 32 to 128 points), mode 2 = the FFT of mode 0 with
 * every pass a launch of its own (the same transform, equal to f32
 * rounding; for comparison).  The
 * reference uses cuFFT in FP32 (src/gpu/cufft.rs); mode 1 has no
 * counterpart there.  May be switched between phases. */
int hd_set_spme_dft(hd_ctx *ctx, int mode);
/* Cell-list neighbour skin (src/io/config.rs:1232, default 0.3 * r_cut). */
int hd_set_neighbor(hd_ctx *ctx, double r_skin);
/* Rigid three-site water via SETTLE (src/integrator/settle.rs):
 * groups = 3 * n_groups atom ids (O, H1, H2). */
int hd_set_settle(hd_ctx *ctx, uint64_t n_groups, const uint32_t *groups, double d_oh,
                  double d_hh, double m_o, double m_h);
/* A further set of SETTLE groups with their own geometry (another water
 * model, or another constraint type in the reference's topology: settle.rs
 * keeps d_OH, d_HH and the O/H masses per group, settle.rs:196-240); groups
 * with an already known geometry join its kind.  At most 8 geometries per
 * context; the first call behaves as hd_set_settle. */
int hd_add_settle(hd_ctx *ctx, uint64_t n_groups, const uint32_t *groups, double d_oh,
                  double d_hh, double m_o, double m_h);
/* General SHAKE / RATTLE constraint groups (src/integrator/shake.rs,
 * kernels/shake.cu): group g holds atoms group_atoms[group_offsets[g] ..
 * group_offsets[g+1]) and constraints con_offsets[g] .. con_offsets[g+1]
 * between group-local atoms con_i/con_j at distance con_d.  Up to 64 atoms
 * and 64 constraints per group (the reference caps groups at 8 atoms and
 * 12 constraints, shake.rs:21-22); an atom belongs to at most one
 * constraint group (SETTLE or SHAKE). */
int hd_set_shake(hd_ctx *ctx, uint64_t n_groups, const uint32_t *group_offsets,
                 const uint32_t *group_atoms, const uint32_t *con_offsets, const uint8_t *con_i,
                 const uint8_t *con_j, const double *con_d);
/* SHAKE tolerance on |r^2 - d^2| (Bohr^2), RATTLE tolerance on |v_rel . r|
 * and the sweep cap; the defaults are the reference's 3.57e-6, 8.63e-17
 * and 32 (shake.cu:251-256, 446). */
int hd_set_shake_tolerance(hd_ctx *ctx, double tol, double rattle_tol, int max_iter);
/* Sweep counts of the last hd_step call (or of hd_constrain_velocities):
 * out[0..2] SHAKE total sweeps over all groups and calls, the largest
 * count, groups that hit the cap; out[3..5] the same for RATTLE. */
int hd_get_constraint_stats(hd_ctx *ctx, uint64_t out[6]);
/* Pair list form.  0: the packed 64-atom cluster list of the reference's
 * neighbour pipeline (kernels/neighbor.cu:416-631), half list with j-side
 * fixed-point atomics.  1: per-atom full lists, each pair listed from both
 * atoms and evaluated from both sides with no atomics (no counterpart in
 * the reference).  Both give the same forces to f32 rounding.  May be
 * changed after hd_setup (the list is rebuilt). */
int hd_set_pair_list(hd_ctx *ctx, int mode);
/* Velocity Verlet (src/integrator/velocity_verlet.rs); lossless = the
 * compensated (f32, f64) mode. */
int hd_set_integrator(hd_ctx *ctx, double dt, int lossless);
/* Nose-Hoover chain thermostat (src/integrator/nose_hoover_chain.rs:17-31);
 * kt <= 0 removes it.  n_constraints feeds the thermal DOF 3N - n_c - 3. */
int hd_set_nhc(hd_ctx *ctx, double kt, double tau, int chain, int yoshida, int n_resp,
               uint64_t n_constraints);
/* CSVR stochastic velocity rescaling (src/integrator/csvr.rs,
 * kernels/nose_hoover.cu:145-255): after the drift, K_new from K, the
 * relaxation c = exp(-dt/tau) and Philox normals (key seed, counter = the
 * device draw counter, one per step) sums the g - 1 squares of draw
 * indices 1..g-1 and the normal of index 0, g = 3N - n_constraints - 3;
 * the post-force half kick is scaled by sqrt(K_new / K).  One thermostat
 * per context; after hd_setup only re-parameterisable. */
int hd_set_csvr(hd_ctx *ctx, double kt, double tau, uint64_t seed, uint64_t n_constraints);
/* Berendsen weak-coupling thermostat (src/integrator/berendsen.rs,
 * nose_hoover.cu:390-417): factor sqrt(max(0, 1 + dt/tau (K_target/K - 1)))
 * on the post-force half kick, K_target = g kT / 2. */
int hd_set_berendsen(hd_ctx *ctx, double kt, double tau, uint64_t n_constraints);
/* Andersen thermostat (src/integrator/andersen.rs, kernels/andersen.cu):
 * after the post-force kick every atom collides with probability
 * p = clamp(rate dt, 0, 1) (f32) and then takes a fresh Maxwell-Boltzmann
 * velocity; Philox key seed, counter = the device draw counter (one per
 * step, read by hd_get_thermostat out[1]).  The conserved-quantity column
 * is K + U.  One thermostat per context. */
int hd_set_andersen(hd_ctx *ctx, double kt, double collision_rate, uint64_t seed);
/* CSVR / Berendsen state: out[0] cumulative energy the rescales injected
 * (the conserved quantity is K + U - out[0]), out[1] CSVR draw counter. */
int hd_get_thermostat(hd_ctx *ctx, double out[2]);

/* MTK-NPT integrator (src/integrator/mtk_npt.rs:88-470, kernels/mtk.cu):
 * isotropic Martyna-Tobias-Klein NPT with a Nose-Hoover chain on the
 * particles and one on the cell, in the reference's symmetric Trotter
 * splitting; kt = k_B T, pressure the target, tau_t / tau_p the chain and
 * cell periods, chain_length >= 1, yoshida_order 1, 3, 5 or 7, n_resp >= 1,
 * n_constraints for the degrees of freedom.  It owns the thermostat and the
 * barostat (no other one may be set) and, like the reference's, takes no
 * constraints and no lossless mode; steps run eagerly (the chains and the
 * cell momentum live on the host).  Set before hd_setup. */
int hd_set_mtk(hd_ctx *ctx, double kt, double pressure, double tau_t, double tau_p,
               int chain_length, int yoshida_order, int n_resp, uint64_t n_constraints);
/* MTK-NPT state of the last step: out[0] pressure (2K + W)/(3V), out[1]
 * volume, out[2] the conserved quantity less K + U (P V, the cell and chain
 * energies; mtk_npt.rs:246-268), out[3] p_eps, out[4] eps = ln(V/V0)/3. */
int hd_get_mtk(hd_ctx *ctx, double out[5]);
/* Langevin BAOAB integrator (src/integrator/langevin_baoab.rs,
 * kernels/langevin.cu): friction gamma (1/time), k_B*T and the Philox seed
 * of the Ornstein-Uhlenbeck draws.  Replaces velocity Verlet for the whole
 * context and owns the thermostat (no Nose-Hoover chain alongside); the
 * draw counter (one per step) lives on the device.  After hd_setup it can
 * only be re-parameterised. */
int hd_set_langevin(hd_ctx *ctx, double friction, double kt, uint64_t seed);
/* Number of Ornstein-Uhlenbeck draws made so far (steps taken). */
int hd_get_langevin_counter(hd_ctx *ctx, uint64_t *counter);

/* C-rescale barostat (src/integrator/c_rescale_barostat.rs,
 * kernels/barostat.cu:144-232): target pressure, tau, isothermal
 * compressibility (atomic units), k_B*T of the noise term and the Philox
 * seed.  tau <= 0 removes it.  Forces scalars (virial) on every step. */
int hd_set_crescale(hd_ctx *ctx, double pressure, double tau, double compressibility, double kt,
                    uint64_t seed);
/* Molecule partition for the Monte-Carlo barostat (MoleculeList,
 * src/forces/topology.rs:1466-1560: connected components of bonds and
 * constraint groups): molecule m holds atoms[offsets[m] .. offsets[m+1]).
 * Without it every atom is its own molecule.  Before hd_setup. */
int hd_set_molecules(hd_ctx *ctx, uint64_t n_mol, const uint32_t *offsets,
                     const uint32_t *atoms);
/* One Monte-Carlo volume move (src/integrator/mc_barostat.rs:218-330,
 * kernels/mc_barostat.cu): U_old from a force evaluation at the current
 * state; u1, u2 from Philox (key seed, counter (lo, hi, 0, 0)), the caller
 * advancing the counter by one per move as the reference does; dV =
 * max_dv (2 u1 - 1); rejected at once when V + dV <= 0 or the scaled box
 * would be narrower than twice the search radius; otherwise the box and
 * every molecule's centre of mass scale by ((V + dV)/V)^(1/3), U_new is
 * evaluated and the move is kept when w = dU + P dV - N_mol kT ln(V'/V) <= 0
 * or u2 < exp(-w/kT); a rejected move restores positions, forces, scalars
 * and box.  out[0] 1 accepted / 0 not, out[1] V before, out[2] V after,
 * out[3] U_new - U_old (0 when rejected early).  The adaptive step size,
 * the move cadence and the acceptance statistics stay with the caller. */
int hd_mc_barostat_move(hd_ctx *ctx, double pressure, double kt, double max_dv, uint64_t seed,
                        uint64_t counter, double out[4]);
/* Berendsen barostat (src/integrator/berendsen_barostat.rs,
 * kernels/barostat.cu:228-275): the C-rescale step without the noise term
 * (mu^3 = 1 - kappa dt/tau (P0 - P)), no draw and no P0 dV account.
 * Replaces a C-rescale barostat; tau <= 0 removes it. */
int hd_set_berendsen_barostat(hd_ctx *ctx, double pressure, double tau, double compressibility);

/* Allocate device state and build the first neighbour list; must follow
 * the hd_set_* calls and precede any evaluation. */
int hd_setup(hd_ctx *ctx);
/* One force evaluation at the current positions (ForceField::step,
 * src/forces/mod.rs:662-757); want_scalars also fills per-atom energy and
 * virial and the totals. */
int hd_force_eval(hd_ctx *ctx, int want_scalars);
/* Steepest-descent minimizer (src/minimizer/steepest_descent.rs,
 * kernels/minimize.cu).  hd_minimize_begin evaluates the forces and the
 * potential energy (SteepestDescentMinimizer::initial_state) and sets the
 * step length; each hd_minimize_step (::step) moves x += F step / F_max,
 * projects SETTLE / SHAKE groups back onto their constraints (positions
 * only, directions at the moved positions), evaluates the trial energy and
 * keeps the move if it is lower (step *= step_increase, at most max_step)
 * or restores the positions (step *= step_decrease).  out[0] energy after
 * the call, out[1] max |F_i| of the current forces, out[2] step length
 * used, out[3] 1 accepted / 0 rejected.  Velocities are not touched;
 * convergence tests are the caller's (check_convergence). */
int hd_minimize_begin(hd_ctx *ctx, double initial_step, double out[4]);
int hd_minimize_step(hd_ctx *ctx, double max_step, double step_increase, double step_decrease,
                     double out[4]);
/* Replace the positions (and charges) after hd_setup: 4n floats x, y, z, q
 * (the position upload of src/gpu/buffers.rs ParticleBuffers).  The next
 * evaluation recomputes the forces; the neighbour list is rebuilt when the
 * displacement check asks for it. */
int hd_set_positions(hd_ctx *ctx, const float *posq);
/* Neighbour pre-step (NeighborListState::pre_step,
 * src/forces/neighbor_list.rs:1265-1364, and neighbor_displacement_check_flag,
 * kernels/neighbor.cu:50-75): runs the displacement check of the current
 * positions against those of the last rebuild and returns the status word:
 * bit 0 rebuild due (some |x - x_ref| plus sqrt(3) max|dL| exceeds
 * r_skin / 2), bit 1 high water, bit 3 overflow.  The rebuild itself runs
 * inside the next force evaluation, gated on bit 0 on the device. */
int hd_neighbor_pre_step(hd_ctx *ctx, uint32_t *status_out);
/* Project velocities onto the constraint manifold
 * (Constraint::apply_initial_velocity_projection, settle.rs:406-430). */
int hd_constrain_velocities(hd_ctx *ctx);
/* Advance n_steps MD steps (the reference runner's per-step sequence,
 * src/runner.rs:2585-2687); the last step evaluates scalars when
 * scalars_last != 0.  Uses a captured hipGraph per step when enabled. */
int hd_step(hd_ctx *ctx, int n_steps, int scalars_last);
/* Time reversal v -> -v, including the FP64 residuals of the lossless mode
 * (the velocity flip of tests/integration_lossless.rs:302-395). */
int hd_negate_velocities(hd_ctx *ctx);
int hd_use_graph(hd_ctx *ctx, int enable);
/* Fused integrator kernels (default on): when every atom belongs to a
 * SETTLE water, the drift and both SETTLE stages run as one kernel (under
 * Langevin BAOAB: both half steps and the three SETTLE stages, five
 * kernels); under Langevin BAOAB without constraints the two half-step
 * kernels run as one.  0 runs the separate kernels, with bit-identical
 * results.  No reference counterpart. */
int hd_set_fused_settle(hd_ctx *ctx, int enable);
/* Systems of at most 8192 atoms run the neighbour displacement check, the
 * rebuild's voxel binning and the slot-order gather in one one-workgroup
 * kernel (default on); 0 takes the general multi-kernel path, with
 * identical results.  No reference counterpart. */
int hd_set_small_neighbor(hd_ctx *ctx, int enable);
int hd_sync(hd_ctx *ctx);

/* Batched replicas.  A batch advances many independent, set-up contexts
 * (same device) with one hipGraph launch per step: the graph forks into
 * every context's stream, runs that context's own step there and joins,
 * so each replica's trajectory is byte-identical to its solo hd_step run.
 * hd_batch_step synchronises at its end and checks every replica's
 * neighbour status (errors name the replica).  The contexts stay owned by
 * the caller and must outlive the batch.  No counterpart in the reference,
 * which runs one system per process (src/runner.rs). */
typedef struct hd_batch hd_batch;
int hd_batch_create(hd_ctx *const *ctxs, int n, hd_batch **out);
void hd_batch_destroy(hd_batch *batch);
int hd_batch_step(hd_batch *batch, int n_steps, int scalars_last);
/* Shape of the step graph: `lanes` parallel chains of replicas (replica i
 * runs after replica i - lanes); 0 = every replica its own branch.  Does
 * not change any replica's arithmetic. */
int hd_batch_set_lanes(hd_batch *batch, int lanes);
const char *hd_batch_last_error(const hd_batch *batch);

/* Learned initial multipliers for SHAKE.  No counterpart in the reference
 * (whose SHAKE starts every solve from zero, shake.cu:251-330).
 *
 * Per constraint the device forms 8 dimensionless features: the first
 * Gauss-Seidel multiplier nu1 = (|r|^2 - d^2) / (2 g.r) of the constraint
 * and of its neighbours k-1 and k+1 in the group's constraint order, the
 * constraint's converged multiplier of the previous solve, the sum of nu1
 * over the second neighbours k-2 and k+2, the linearised couplings to k-1
 * and k+1, and g.r / d^2 - 1 (g: snapshot bond, r: unconstrained bond).
 * Multipliers are exchanged in
 * the form nu = lambda (1/m_i + 1/m_j).  A predictor (hd_mlp) maps the
 * features to nu; SHAKE then starts from the predicted nu and runs its
 * usual sweeps to the same tolerance.
 *
 * hd_set_shake_predictor: params = HD_MLP_NPARAM floats (hd_mlp_get_params)
 * or NULL to switch prediction off.  Needs SHAKE groups and hd_setup. */
int hd_set_shake_predictor(hd_ctx *ctx, const float *params);
/* Predicted (nu0, zero without a predictor) and converged (nu) multipliers
 * of the last SHAKE solve, one per constraint in hd_set_shake order; either
 * pointer may be NULL. */
int hd_get_shake_multipliers(hd_ctx *ctx, float *nu0, float *nu);
/* Step log of SHAKE solves, device-resident: every step appends its
 * feature rows and converged nu until max_steps steps are held (later steps
 * are not logged).  Start resets the log. */
int hd_shake_log_start(hd_ctx *ctx, uint64_t max_steps);
int hd_shake_log_stop(hd_ctx *ctx);
int hd_shake_log_size(hd_ctx *ctx, uint64_t *steps, uint64_t *rows_per_step);
/* Copies logged steps first_step .. first_step + n_steps - 1: x gets
 * 8 floats per row, t one. */
int hd_shake_log_read(hd_ctx *ctx, uint64_t first_step, uint64_t n_steps, float *x, float *t);

/* The predictor: an 8 -> 64 -> 64 -> 1 tanh network (4,801 trained
 * parameters plus the feature scale s) on one device.  Scaled inputs
 * u = x / s (features 0-4 and 7) and x (5, 6); prediction
 * nu = x0 + s (w3 . tanh(W2 tanh(W1 u + b1) + b2) + b3).  Parameter pack:
 * W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3, s.  Training
 * minimises the mean of ((nu - t) / s)^2 with Adam (0.9, 0.999, 1e-8);
 * s <= 0 is replaced by the rms of column 0 of the training rows.
 * Training is deterministic: fixed-order sums, seeded row order. */
typedef struct hd_mlp hd_mlp;
int hd_mlp_param_count(void);
int hd_mlp_create(int device, uint64_t seed, hd_mlp **out);
void hd_mlp_destroy(hd_mlp *mlp);
const char *hd_mlp_last_error(const hd_mlp *mlp);
/* Setting parameters also resets the Adam state. */
int hd_mlp_set_params(hd_mlp *mlp, const float *params);
int hd_mlp_get_params(hd_mlp *mlp, float *params);
/* nu[i] for the feature rows x[8 i .. 8 i + 7]. */
int hd_mlp_predict(hd_mlp *mlp, const float *x, uint64_t n, float *nu);
/* Gradient of the loss over all n rows (HD_MLP_NTRAIN = param_count - 1
 * floats, the scale excluded) and the loss itself; no update. */
int hd_mlp_gradient(hd_mlp *mlp, const float *x, const float *t, uint64_t n, float *grad,
                    double *loss);
/* hd_mlp_gradient on device memory, allocating nothing.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  x (8 n
 * floats), t (n), grad (param_count - 1 floats) and loss (one double) are
 * device pointers on the predictor's device; loss receives the loss SUM
 * over the rows (hd_mlp_gradient reports the mean: divide by n); the
 * gradient is the same as hd_mlp_gradient's, bit for bit.  n <= 16384
 * (HD_ERR_ARG beyond).  The per-workgroup partials live in one static
 * device buffer per GPU (about 5 MB, held until the process ends): calls
 * that overlap on one GPU race on it.  The call synchronises the
 * predictor's stream before it returns. */
int hd_mlp_gradient_noalloc(hd_mlp *mlp, const float *x, const float *t, uint64_t n,
                            float *grad, double *loss);
/* Adam over `epochs` passes of mini-batches of `batch` rows (a multiple of
 * 64); loss[e] = mean loss during epoch e (may be NULL). */
int hd_mlp_train(hd_mlp *mlp, const float *x, const float *t, uint64_t n, int epochs, int batch,
                 float lr, uint64_t seed, double *loss);
/* The same on a context's SHAKE log, read in place on the device. */
int hd_mlp_train_log(hd_mlp *mlp, hd_ctx *ctx, int epochs, int batch, float lr, uint64_t seed,
                     double *loss);

/* Downloads (caller-owned host arrays of n entries; posq has 4n). */
int hd_get_positions(hd_ctx *ctx, float *posq, int32_t *ix, int32_t *iy, int32_t *iz);
int hd_get_velocities(hd_ctx *ctx, float *vx, float *vy, float *vz);
int hd_get_forces(hd_ctx *ctx, float *fx, float *fy, float *fz);
/* Per-atom energy and virial of the last scalar evaluation.  Pair energies
 * are split half and half between the two atoms (the reference's
 * convention); the pair virial is split as r_i . F_i over pairs measured
 * from the i-block's centre when the box has no tilt (the same total, a
 * different split), half and half otherwise. */
int hd_get_per_atom(hd_ctx *ctx, float *pe, float *virial);
int hd_get_lattice(hd_ctx *ctx, double lattice[6]);
/* Scalars of the last scalar evaluation, FP64:
 * out[0] kinetic energy, out[1] potential energy, out[2] virial,
 * out[3] E_recip (without self term), out[4] W_recip,
 * out[5] thermostat term of the conserved quantity (Nose-Hoover chain
 * energy, or minus the energy CSVR / Berendsen rescales injected),
 * out[6] neighbour rebuild count, out[7] neighbour status word. */
int hd_get_scalars(hd_ctx *ctx, double out[8]);
/* Canonical pair set of the current packed list: unique (min id, max id)
 * pairs with min < max, excluding sentinels (neighbor_list.rs:1366-1506).
 * Returns the count in *n_pairs; writes at most cap pairs (2*cap u32). */
int hd_get_pairs(hd_ctx *ctx, uint32_t *out, uint64_t cap, uint64_t *n_pairs);
/* Number of pairs with r <= r_cut (max cutoff) in the current list that
 * are neither excluded nor padding: the pair kernel's useful work. */
int hd_count_pairs(hd_ctx *ctx, uint64_t *count);
/* Barostat state: out[0] last mu, out[1] pressure, out[2] volume after the
 * move, out[3] cumulative P0*dV, out[4] Philox draw counter. */
int hd_get_barostat(hd_ctx *ctx, double out[5]);
/* out[0] blocks, out[1] entry capacity per block, out[2] max entries in a
 * block, out[3] total entries, out[4] rebuilds, out[5] status. */
int hd_get_neighbor_stats(hd_ctx *ctx, uint64_t out[6]);
/* SPME grids of the last evaluation: rho (charge grid, na*nb*nc, from the
 * fixed-point spread) and V (potential grid after the convolution). */
int hd_get_spme_grids(hd_ctx *ctx, float *rho, float *v);
/* Device self test of a matrix-core lane layout (v_mfma_f32_32x32x8_f16):
 * out[l*16 + v] = element v of lane l of D = A B for the integer test
 * matrices documented in k_dft.hip.  No counterpart in the reference. */
int hd_selftest_mfma(hd_ctx *ctx, float out[1024]);
/* The same for the gfx950 doubled-K forms the split-precision DFT runs on,
 * v_mfma_f32_32x32x16_f16 (bf16 = 0) or _bf16 (bf16 = 1), with the test
 * matrices A[i][k] = (16 i + k) mod 61 - 30, B[k][j] = (32 k + j) mod 23 - 11.
 * No counterpart in the reference. */
int hd_selftest_mfma_k16(hd_ctx *ctx, int bf16, float out[1024]);
/* Device self test of the wave64 rotation: out[l] = 7*((l+1)%64)+3. */
int hd_selftest_dpp(hd_ctx *ctx, int32_t out[64]);
/* Mean device time (ms) per occurrence of each stage over the steps of the
 * last timed hd_step call (hipEvents on the context stream; stage names
 * follow the reference's .timings table, src/timings.rs); returns the
 * number of stages written (<= cap). */
int hd_stage_times(hd_ctx *ctx, char *names, int name_len, double *ms, int cap);
/* Enable per-stage event timing for non-graph steps. */
int hd_set_timing(hd_ctx *ctx, int enable);

#ifdef __cplusplus
}
#endif
#endif
