// Host runtime of libheddle: context, device buffers, step orchestration
// and hipGraph capture behind the C ABI in include/heddle.h.
//
// Step sequence (one MD step, velocity Verlet + SETTLE + Nose-Hoover chain;
// the reference runner's order, src/runner.rs:2585-2687 and
// src/integrator/mod.rs:344-429):
//   1. thermostat pre half step: KE -> chain -> factor      (apply_pre)
//   2. rescale + kick + drift + wrap (+ SETTLE snapshot)    (vv_kick_drift)
//   3. SETTLE positions                                     (apply_after_drift)
//   4. SETTLE velocities + constraint virial                (apply_after_kick)
//   5. thermostat post half step: KE -> chain -> factor     (apply_post)
//   6. force evaluation: displacement check, gated rebuild, tile gather,
//      pair kernel, SPME, finalize                          (ForceField::step)
//   7. post-force kick * thermostat factor (fused into the finalize)
// Steps 4 and 5 read only velocities and constrained positions, which the
// force evaluation does not change, so running them before it (the
// reference runs them after) gives the same numbers and lets the post-force
// kick fuse into the force finalize.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/heddle.h"
#include "hd_kernels.h"

#define HD_VERSION "heddlemd-amd 0.1.0 (gfx950)"
#define HD_SENTINEL_KEY_LIMIT 0xFFFFFFFEu

namespace {

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
};

struct Stage {
  std::string name;
  hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace

struct hd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream for the SPME pipeline, forked from and joined back into
  // `stream` inside each force evaluation (default priority)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // status-word polling inside hd_step: two pinned 4-word slots, each
  // filled by an async copy after a graph launch and marked by its event
  unsigned* h_stat = nullptr;
  hipEvent_t ev_stat[2] = {nullptr, nullptr};
  // off while a batch captures this context's step: the batch graph already
  // runs the replicas side by side, and per-replica forks slowed it
  bool spme_fork = true;
  std::string err;

  // configuration
  uint64_t n = 0;
  std::vector<float> h_posq, h_mass, h_vx, h_vy, h_vz;
  std::vector<uint32_t> h_type, h_key;
  std::vector<int32_t> h_ix, h_iy, h_iz;
  double lattice[6] = {0, 0, 0, 0, 0, 0};
  int ntypes = 0;
  std::vector<double> lj_sigma, lj_eps, lj_cut, lj_rsw;
  bool have_spme = false;
  double alpha = 0, coul_rc = 0, kc = 1.0;
  int grid[3] = {0, 0, 0};
  int spline_order = 4;
  double r_skin = -1.0;
  unsigned excl_span = 0;
  int atom_lists = 0;  // 1: per-atom full neighbour lists instead of the packed cluster list
  std::vector<uint32_t> h_angles;  // i, j, k per angle
  std::vector<double> h_ang_k, h_ang_t0;
  unsigned* d_angles = nullptr;
  double *d_ang_k = nullptr, *d_ang_t0 = nullptr;
  std::vector<uint32_t> h_bonds;  // i, j per bond
  std::vector<int> h_bond_kind;
  std::vector<double> h_bond_prm;  // 3 per bond
  unsigned* d_bonds = nullptr;
  int* d_bond_kind = nullptr;
  double* d_bond_prm = nullptr;
  std::vector<uint32_t> h_dih;  // i, j, k, l per dihedral
  std::vector<double> h_dih_prm;  // k, n, phi0 per dihedral
  unsigned* d_dih = nullptr;
  double* d_dih_prm = nullptr;
  // scaled exclusions: (i, j, w_lj, w_coul) with w = scale - the pair
  // kernel's weight; the LJ columns are filled at hd_setup
  std::vector<HdExclPair> h_excl;
  HdExclPair* d_excl = nullptr;
  std::vector<uint32_t> h_groups;
  double d_oh = 0, d_hh = 0, m_o = 0, m_h = 0;
  // SETTLE geometries beyond the first (hd_add_settle): kind per group and
  // the kinds' (d_oh, d_hh, m_o, m_h); kind 0 is the scalar fields above
  std::vector<uint8_t> h_gkind;
  std::vector<std::array<double, 4>> settle_kinds;
  unsigned char* d_gkind = nullptr;
  double dt = 0;
  bool lossless = false;
  bool have_nhc = false;
  double nhc_kt = 0, nhc_tau = 0;
  int nhc_chain = 3, nhc_yoshida = 3, nhc_resp = 1;
  uint64_t nhc_nc = 0;
  bool have_baro = false;
  double baro_p = 0, baro_tau = 0, baro_kappa = 0, baro_kt = 0;
  uint64_t baro_seed = 0;
  bool have_lan = false;  // Langevin BAOAB instead of velocity Verlet
  // velocity-rescale thermostats: 0 none, 2 CSVR, 3 Berendsen
  int rescale = 0;
  // Andersen collisions (a thermostat of its own kind)
  bool have_andersen = false;
  double and_kt = 0, and_rate = 0;
  uint64_t and_seed = 0;
  unsigned long long* and_counter = nullptr;
  double rs_kt = 0, rs_tau = 0;
  uint64_t rs_seed = 0, rs_nc = 0;
  unsigned long long* rs_counter = nullptr;  // CSVR draw counter
  double* rs_injection = nullptr;            // cumulative K_new - K_old
  double* rs_partials = nullptr;             // CSVR partial sums
  bool baro_stochastic = true;  // C-rescale; false: Berendsen barostat
  // Monte-Carlo barostat: molecule partition and the move's snapshots
  std::vector<uint32_t> h_mol_off, h_mol_atoms;
  unsigned *d_mol_off = nullptr, *d_mol_atoms = nullptr;
  float4* mc_pos = nullptr;
  float* mc_f = nullptr;     // 3n saved forces
  double* mc_scal = nullptr;  // saved scalars
  // steepest descent
  float4* sd_save = nullptr;
  double* sd_fmax2 = nullptr;
  double sd_step = 0, sd_energy = 0;
  bool sd_begun = false;
  double lan_friction = 0, lan_kt = 0;
  uint64_t lan_seed = 0;
  bool use_graph = true;
  bool fused_settle = true;  // all-water systems: drift and SETTLE in one kernel
  bool small_neighbor = true;  // small systems: check, binning and gather in one workgroup
  bool timing = false;
  bool ready = false;
  bool forces_valid = false;

  // derived
  unsigned npad = 0, nblocks = 0, cap = 0, nvox = 0;
  unsigned nv[3] = {0, 0, 0};
  double rmax = 0;

  // device: atom order
  float4* posq = nullptr;
  float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  float *fx = nullptr, *fy = nullptr, *fz = nullptr;
  float *mass = nullptr, *pe = nullptr, *vir = nullptr, *w_cons = nullptr;
  int *ix = nullptr, *iy = nullptr, *iz = nullptr;
  unsigned *type = nullptr, *key = nullptr;
  double *xlo = nullptr, *ylo = nullptr, *zlo = nullptr, *vxlo = nullptr, *vylo = nullptr,
         *vzlo = nullptr;
  float4 *snap = nullptr, *ref_pos = nullptr;
  float* d_lattice = nullptr;
  float4* d_lj_tab = nullptr;  // more than four LJ types: the pair kernels' table
  // MTK-NPT: fused thermostat + isotropic barostat, driven from the host
  // (the reference's mtk_npt.rs is not graph-compatible either)
  bool have_mtk = false;
  double mtk_kt = 0, mtk_p = 0, mtk_tau_t = 0, mtk_tau_p = 0, mtk_g = 1, mtk_w = 0;
  double mtk_peps = 0, mtk_eps = 0, mtk_last_p = 0, mtk_last_v = 0;
  int mtk_chain = 3, mtk_yoshida = 3, mtk_nresp = 1;
  uint64_t mtk_nc = 0;
  std::vector<double> mtk_qp, mtk_xp, mtk_pp, mtk_qc, mtk_xc, mtk_pc;
  float* ref_lattice = nullptr;
  float* baro_mu = nullptr;
  double* baro_diag = nullptr;
  unsigned long long* baro_counter = nullptr;
  unsigned long long* lan_counter = nullptr;
  // device: neighbour list
  unsigned *voxel_rank = nullptr, *atom_key = nullptr, *voxel_count = nullptr,
           *voxel_off = nullptr, *voxel_cursor = nullptr, *voxel_part = nullptr;
  unsigned *sorted_id = nullptr, *slot_of = nullptr, *tile_type = nullptr, *tile_key = nullptr;
  float4 *tile_posq = nullptr, *center = nullptr, *half = nullptr;
  unsigned *entry_atoms = nullptr, *entry_count = nullptr;
  unsigned long long* block_mask = nullptr;  // block-pair candidate masks
  unsigned* block_order = nullptr;           // pair workgroup -> block
  unsigned* ctrl = nullptr;  // [0] status, [1] rebuild count, [2] max entries, [3] build ticket
  long long *fpi = nullptr, *fpi_ew = nullptr, *fpj = nullptr, *fpj_ew = nullptr;
  // device: SPME
  long long* rho_fixed = nullptr;
  float2* cgrid = nullptr;
  float* vgrid = nullptr;
  float4* spl = nullptr;  // SPME spline records per slot
  int dft_mode = 0;        // 0 FP32 radix FFT, 1 split-precision FP16 MFMA DFT, 2 FFT, no plane fusion
  _Float16* dft_mats = nullptr;
  HdDftMats dft;
  unsigned* sg = nullptr;
  float2 *tw_a = nullptr, *tw_b = nullptr, *tw_c = nullptr;
  float *bf_a = nullptr, *bf_b = nullptr, *bf_c = nullptr;
  double2* recip_partials = nullptr;
  double* recip = nullptr;
  float* w_share = nullptr;
  // device: scalars / thermostat
  double* red_partials = nullptr;
  double* fin_partials = nullptr;  // finalize's per-workgroup energy/virial sums
  double* scal = nullptr;  // [0] ke, [1] pe, [2] vir, [3] ke scratch
  float* factors = nullptr;  // [0] pre, [1] post
  double* nhc_state = nullptr;
  unsigned* d_groups = nullptr;
  // general SHAKE / RATTLE groups
  std::vector<uint32_t> h_sh_goff, h_sh_atoms, h_sh_coff, h_sh_cij;
  std::vector<float> h_sh_r2;
  std::vector<uint8_t> h_sh_color;              // constraint colours (long groups)
  unsigned sh_n_colors = 0;
  uint8_t* d_sh_color = nullptr;
  double sh_tol = 3.57e-6, sh_rattle_tol = 8.63e-17;
  int sh_max_iter = 32;
  unsigned *d_sh_goff = nullptr, *d_sh_atoms = nullptr, *d_sh_coff = nullptr,
           *d_sh_cij = nullptr, *d_sh_stats = nullptr;
  float* d_sh_r2 = nullptr;
  // groups of at most 4 atoms and 6 constraints run one per lane
  // (d_sh_small), the others one per wave (d_sh_big); both null when no
  // group is small
  unsigned *d_sh_small = nullptr, *d_sh_big = nullptr, sh_n_small = 0, sh_n_big = 0;
  // SHAKE multiplier predictor and the step log of plain-SHAKE solutions
  unsigned* d_sh_cgrp = nullptr;  // constraint -> group
  float *d_sh_feat = nullptr, *d_sh_nu0 = nullptr, *d_sh_nu = nullptr;
  float* d_mlp = nullptr;  // predictor parameters (HD_MLP_NPARAM floats)
  bool sh_predict = false, sh_logging = false;
  float *d_log_x = nullptr, *d_log_t = nullptr;
  unsigned long long* d_log_ctr = nullptr;
  unsigned long long log_cap = 0;  // steps the log holds

  // graphs
  hipGraphExec_t graph_f = nullptr, graph_fev = nullptr;
  // HD_GRAPH_STEPS consecutive steps without scalars in one graph: one
  // launch, and no gap between graphs, per that many steps
  hipGraphExec_t graph_fk = nullptr;
  unsigned graph_gen = 0;  // bumped whenever captured work goes stale

  std::vector<Stage> stages;  // event pool; the first n_stage_used are live
  size_t n_stage_used = 0;
  std::vector<std::string> stage_names;  // aggregated by name
  std::vector<double> stage_ms;          // mean ms per occurrence
  std::vector<long> stage_count;

  std::vector<void*> allocs;
};

namespace {

int fail(hd_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, expr)                                                                   \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return fail((c), HD_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                    \
  } while (0)

template <typename T>
int dalloc(hd_ctx* c, T** p, size_t count) {
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess)
    return fail(c, HD_ERR_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  // zeroed on the context stream and waited for: a null-stream memset is not
  // ordered with the context's non-blocking stream and could land after
  // kernels that already use the buffer
  e = hipMemsetAsync(*p, 0, bytes, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, HD_ERR_HIP, "zeroing a buffer failed: %s", hipGetErrorString(e));
  c->allocs.push_back((void*)*p);
  return HD_OK;
}

#define DALLOC(c, p, count)                  \
  do {                                       \
    int _r = dalloc((c), &(p), (count));     \
    if (_r != HD_OK) return _r;              \
  } while (0)

void dfree(hd_ctx* c, void* p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

bool is_pow2(int v) { return v >= 2 && (v & (v - 1)) == 0; }
// n = 2^a 3^b 5^c: the FFT's radices
bool is_smooth235(int v) {
  if (v < 1) return false;
  for (int p : {2, 3, 5})
    while (v % p == 0) v /= p;
  return v == 1;
}

// 3D Hilbert index of cell (a, b, c) on a 2^bits cube, by Skilling's
// transpose algorithm ("Programming the Hilbert curve", AIP Conf. Proc. 707,
// 2004).  Consecutive indices are face-adjacent cells, so a run of atoms
// in this order is spatially compact (a Morton order jumps across the box
// at octant boundaries, which gives some blocks box-sized extents).
uint64_t hilbert3(uint32_t a, uint32_t b, uint32_t c, int bits) {
  uint32_t X[3] = {a, b, c};
  uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t code = 0;
  for (int bit = bits - 1; bit >= 0; --bit)
    for (int i = 0; i < 3; ++i) code = (code << 1) | ((X[i] >> bit) & 1u);
  return code;
}

std::vector<float2> twiddles(int n) {
  std::vector<float2> t(std::max(n / 2, 1));
  for (int k = 0; k < n / 2; ++k) {
    double ang = -2.0 * M_PI * k / n;
    t[k] = make_float2((float)cos(ang), (float)sin(ang));
  }
  return t;
}

double bspline_host(int p, double x) {
  double v[16];
  for (int i = 0; i < p; ++i) {
    double xi = x - i;
    v[i] = (xi >= 0.0 && xi < 1.0) ? 1.0 : 0.0;
  }
  for (int order = 2; order <= p; ++order) {
    double inv = 1.0 / (order - 1.0);
    for (int i = 0; i < p - order + 1; ++i) {
      double xi = x - i;
      v[i] = xi * inv * v[i] + (order - xi) * inv * v[i + 1];
    }
  }
  return v[0];
}

// |b(k)|^2 corrections, src/forces/spme.rs:524-585.
std::vector<float> b_factors(int n, int p) {
  std::vector<double> ms(p - 1), mod(n);
  for (int j = 1; j < p; ++j) ms[j - 1] = bspline_host(p, (double)j);
  for (int k = 0; k < n; ++k) {
    double th = 2.0 * M_PI * k / n, re = 0, im = 0;
    for (int j = 0; j < p - 1; ++j) {
      re += ms[j] * cos(th * j);
      im += ms[j] * sin(th * j);
    }
    mod[k] = re * re + im * im;
  }
  for (int k = 0; k < n; ++k)
    if (mod[k] < 1e-7) mod[k] = 0.5 * (mod[(k + n - 1) % n] + mod[(k + 1) % n]);
  std::vector<float> out(n);
  for (int k = 0; k < n; ++k) out[k] = mod[k] > 0 ? (float)(1.0 / mod[k]) : 0.f;
  return out;
}

template <typename T>
int upload(hd_ctx* c, T* dst, const T* src, size_t count) {
  if (!count) return HD_OK;
  HIPCHK(c, hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return HD_OK;
}

HdNeighborArgs nb_args(hd_ctx* c) {
  HdNeighborArgs a;
  memset(&a, 0, sizeof(a));
  a.n = (unsigned)c->n;
  a.npad = c->npad;
  a.nvox = c->nvox;
  a.cap = c->cap;
  for (int d = 0; d < 3; ++d) a.nv[d] = c->nv[d];
  double rs = c->rmax + c->r_skin;
  a.rs2 = (float)(rs * rs);
  a.half_skin = (float)(0.5 * c->r_skin);
  a.ref_lattice = c->ref_lattice;
  a.posq = c->posq;
  a.lattice = c->d_lattice;
  a.ref_pos = c->ref_pos;
  a.voxel_rank = c->voxel_rank;
  a.atom_key = c->atom_key;
  a.voxel_count = c->voxel_count;
  a.voxel_off = c->voxel_off;
  a.voxel_cursor = c->voxel_cursor;
  a.voxel_part = c->voxel_part;
  a.sorted_id = c->sorted_id;
  a.slot_of = c->slot_of;
  a.type = c->type;
  a.key = c->key;
  a.tile_type = c->tile_type;
  a.tile_key = c->tile_key;
  a.tile_posq = c->tile_posq;
  a.center = c->center;
  a.half = c->half;
  a.entry_atoms = c->entry_atoms;
  a.entry_count = c->entry_count;
  a.max_count = c->ctrl + 2;
  a.status = c->ctrl;
  a.rebuild_count = c->ctrl + 1;
  a.build_ticket = c->ctrl + 3;
  a.atom_lists = c->atom_lists;
  a.excl_span = c->excl_span;
  a.block_mask = c->block_mask;
  a.block_order = c->block_order;
  return a;
}

// Norm of row d of the inverse box matrix (a = (lx,0,0), b = (xy,ly,0),
// c = (xz,yz,lz)): a Cartesian move of length r changes fractional
// coordinate d by at most r times it, and its inverse is the box's width
// perpendicular to the other two box vectors.
double inv_row_norm(const double* L, int d) {
  double ilx = 1.0 / L[0], ily = 1.0 / L[1], ilz = 1.0 / L[2];
  if (d == 0) {
    double a1 = -L[3] * ilx * ily, a2 = (L[3] * L[5] - L[4] * L[1]) * ilx * ily * ilz;
    return sqrt(ilx * ilx + a1 * a1 + a2 * a2);
  }
  if (d == 1) {
    double b2 = -L[5] * ily * ilz;
    return sqrt(ily * ily + b2 * b2);
  }
  return ilz;
}

bool box_ortho(const hd_ctx* c) {
  return c->lattice[3] == 0.0 && c->lattice[4] == 0.0 && c->lattice[5] == 0.0;
}

HdNhcArgs nhc_args(hd_ctx* c) {
  HdNhcArgs a;
  memset(&a, 0, sizeof(a));
  static const double y1[1] = {1.0};
  static const double y3[3] = {1.3512071919596577, -1.7024143839193155, 1.3512071919596577};
  static const double y5[5] = {0.41449077179437574, 0.41449077179437574, -0.6579630871775030,
                               0.41449077179437574, 0.41449077179437574};
  static const double y7[7] = {0.7845136104775573, 0.2355732133593582, -1.1776799841788710,
                               1.3151863206839023, -1.1776799841788710, 0.2355732133593582,
                               0.7845136104775573};
  const double* w = c->nhc_yoshida == 1 ? y1 : c->nhc_yoshida == 3 ? y3 : c->nhc_yoshida == 5 ? y5 : y7;
  a.chain = c->nhc_chain;
  a.n_yoshida = c->nhc_yoshida;
  a.n_resp = c->nhc_resp;
  for (int i = 0; i < c->nhc_yoshida; ++i) a.w[i] = w[i];
  a.kt = c->nhc_kt;
  long long dof = 3LL * (long long)c->n - (long long)c->nhc_nc - 3LL;
  a.g_dof = (double)std::max(dof, 0LL);
  a.state = c->nhc_state;
  return a;
}

void stage_begin(hd_ctx* c, const char* name, bool on) {
  if (!on) return;
  if (c->n_stage_used == c->stages.size()) {
    Stage s;
    (void)hipEventCreate(&s.a);
    (void)hipEventCreate(&s.b);
    c->stages.push_back(s);
  }
  Stage& s = c->stages[c->n_stage_used++];
  s.name = name;
  (void)hipEventRecord(s.a, c->stream);
}
void stage_end(hd_ctx* c, bool on) {
  if (!on) return;
  (void)hipEventRecord(c->stages[c->n_stage_used - 1].b, c->stream);
}

// The SPME reciprocal pipeline of one force evaluation, on stream sst.
static void enqueue_spme(hd_ctx* c, hipStream_t sst, bool timed, bool splines_done) {
  HdSpmeArgs sa;
  memset(&sa, 0, sizeof(sa));
  sa.splines_done = splines_done ? 1 : 0;
  sa.no_plane = c->dft_mode == 2 ? 1 : 0;
  sa.n = (unsigned)c->n;
  sa.order = c->spline_order;
  sa.na = c->grid[0];
  sa.nb = c->grid[1];
  sa.nc = c->grid[2];
  sa.kc = c->kc;
  sa.alpha = c->alpha;
  sa.lattice = c->d_lattice;
  sa.tile_posq = c->tile_posq;
  sa.rho_fixed = c->rho_fixed;
  sa.cgrid = c->cgrid;
  sa.vgrid = c->vgrid;
  sa.spl = c->spl;
  sa.sg = c->sg;
  sa.tw_a = c->tw_a;
  sa.tw_b = c->tw_b;
  sa.tw_c = c->tw_c;
  sa.bf_a = c->bf_a;
  sa.bf_b = c->bf_b;
  sa.bf_c = c->bf_c;
  sa.partials = c->recip_partials;
  sa.recip = c->recip;
  sa.w_share = c->w_share;
  sa.voxel_rank = c->voxel_rank;
  sa.voxel_off = c->voxel_off;
  for (int d = 0; d < 3; ++d) {
    sa.nv[d] = c->nv[d];
    // atoms move at most r_skin/2 between rebuilds; 20 % slack for box
    // changes since set-up, plus one voxel for the binning's rounding
    sa.mv[d] = (int)floor(0.6 * c->r_skin * (double)c->nv[d] * inv_row_norm(c->lattice, d)) + 1;
  }
  stage_begin(c, "spme_recip", timed);
  if (c->dft_mode == 1) {
    hd_launch_spme_spread(sst, sa);
    HdDftArgs da;
    memset(&da, 0, sizeof(da));
    da.na = c->grid[0];
    da.nb = c->grid[1];
    da.nc = c->grid[2];
    da.nch = c->grid[2] / 2 + 1;
    da.rho = c->rho_fixed;
    da.cgrid = c->cgrid;
    da.vgrid = c->vgrid;
    da.bfa = c->bf_a;
    da.bfb = c->bf_b;
    da.bfc = c->bf_c;
    da.lattice = c->d_lattice;
    da.kc = (float)c->kc;
    da.alpha = (float)c->alpha;
    da.partials = c->recip_partials;
    hd_launch_dft(sst, da, c->dft);
    hd_launch_recip_reduce(sst, sa, hd_dft_partials_count(c->grid[1], c->grid[2]));
  } else {
    hd_launch_spme(sst, sa);
  }
  stage_end(c, timed);
}

// Force evaluation; kick = fuse the post-force half kick with factor
// factors[1].  Enqueue only.
bool has_shake(const hd_ctx* c) { return c->h_sh_goff.size() > 1; }
bool has_constraints(const hd_ctx* c) { return !c->h_groups.empty() || has_shake(c); }

// ev: 0 forces only; 1 energy and virial into scal[1], scal[2]; 2 virial
// only (what a barostat consumes on steps that log nothing), reduced into
// the scratch pair scal[6], scal[7] so the last logged energy stays intact
// disp_done: the displacement check already ran on this step's final
// positions (folded into the step's last KE pass)
HdBarostatArgs baro_args(hd_ctx* c);

// baro_mu: a C-rescale step follows (ev != 0): the scalar reduction also
// computes its mu, and enqueue_crescale then scales the positions
void enqueue_forces(hd_ctx* c, int ev, bool kick, bool timed, bool disp_done = false,
                    bool baro_mu = false) {
  hipStream_t st = c->stream;
  HdNeighborArgs na = nb_args(c);
  const bool small_nb = c->small_neighbor && hd_neighbor_small(na);
  if (small_nb) {
    stage_begin(c, "neighbor_update", timed);
    if (c->have_spme)
      hd_launch_neighbor_small(st, na, !disp_done, c->fpj, ev ? c->fpj_ew : nullptr, c->spl,
                               c->sg, c->grid[0], c->grid[1], c->grid[2]);
    else
      hd_launch_neighbor_small(st, na, !disp_done, c->fpj, ev ? c->fpj_ew : nullptr);
    stage_end(c, timed);
  } else {
    stage_begin(c, "neighbor_update", timed);
    hd_launch_neighbor_update(st, na, !disp_done);
    stage_end(c, timed);
    stage_begin(c, "tile_gather", timed);
    hd_launch_tile_gather_rebuild(st, na, c->fpj, ev ? c->fpj_ew : nullptr);
    stage_end(c, timed);
  }
  // The SPME pipeline reads the tile positions, the lattice and the voxel
  // tables, all final here, and nothing on the build and pair path writes
  // what it reads: it runs on the second stream, concurrent with them, and
  // joins before the finalize.  The stage-timing pass keeps one stream so
  // that every stage is timed alone.
  bool fork = c->have_spme && !timed && c->spme_fork;
  if (fork) {
    (void)hipEventRecord(c->ev_fork, st);
    (void)hipStreamWaitEvent(c->side, c->ev_fork, 0);
    enqueue_spme(c, c->side, false, small_nb);
    (void)hipEventRecord(c->ev_join, c->side);
  }
  stage_begin(c, "neighbor_build", timed);
  hd_launch_neighbor_build(st, na, !small_nb, small_nb);
  stage_end(c, timed);
  HdPairArgs pa;
  memset(&pa, 0, sizeof(pa));
  pa.n = (unsigned)c->n;
  pa.npad = c->npad;
  pa.cap = c->cap;
  pa.ntypes = c->ntypes;
  pa.sigma = c->lj_sigma.data();
  pa.eps = c->lj_eps.data();
  pa.cut = c->lj_cut.data();
  pa.rsw = c->lj_rsw.data();
  pa.kc = c->kc;
  pa.alpha = c->have_spme ? c->alpha : 0.0;
  pa.coul_rc = c->have_spme ? c->coul_rc : 0.0;
  pa.rmax = c->rmax;
  pa.r_skin = c->r_skin;
  pa.lattice = c->d_lattice;
  pa.tile_posq = c->tile_posq;
  pa.tile_type = c->tile_type;
  pa.tile_key = c->tile_key;
  pa.excl_span = c->excl_span;
  pa.center = c->center;
  pa.half = c->half;
  pa.entry_atoms = c->entry_atoms;
  pa.entry_count = c->entry_count;
  pa.block_order = c->block_order;
  pa.fpi = c->fpi;
  pa.fpi_ew = c->fpi_ew;
  pa.fpj = c->fpj;
  pa.fpj_ew = c->fpj_ew;
  pa.atom_lists = c->atom_lists;
  pa.ortho = box_ortho(c) ? 1 : 0;
  pa.lj_tab = c->d_lj_tab;
  stage_begin(c, "pair_force", timed);
  hd_launch_pair(st, pa, c->atom_lists && ev == 2 ? 1 : ev);
  if (!c->h_ang_k.empty())
    hd_launch_angles(st, c->posq, c->d_lattice, c->d_angles, c->d_ang_k, c->d_ang_t0, c->slot_of,
                     c->fpj, c->fpj_ew, (unsigned)c->h_ang_k.size(), ev);
  hd_launch_bonds(st, c->posq, c->d_lattice, c->d_bonds, c->d_bond_kind, c->d_bond_prm,
                  c->slot_of, c->fpj, c->fpj_ew, (unsigned)c->h_bond_kind.size(), ev);
  hd_launch_dihedrals(st, c->posq, c->d_lattice, c->d_dih, c->d_dih_prm, c->slot_of, c->fpj,
                      c->fpj_ew, (unsigned)(c->h_dih_prm.size() / 3), ev);
  hd_launch_excl_scaled(st, c->posq, c->d_lattice, c->d_excl, (unsigned)c->h_excl.size(), c->kc,
                        c->have_spme ? c->alpha : 0.0,
                        c->have_spme ? c->coul_rc * c->coul_rc : 0.0, c->slot_of, c->fpj,
                        c->fpj_ew, ev);
  stage_end(c, timed);
  if (fork)
    (void)hipStreamWaitEvent(st, c->ev_join, 0);
  else if (c->have_spme)
    enqueue_spme(c, st, timed, small_nb);
  HdFinalizeArgs fa;
  memset(&fa, 0, sizeof(fa));
  fa.n = (unsigned)c->n;
  fa.npad = c->npad;
  fa.sorted_id = c->sorted_id;
  fa.tile_posq = c->tile_posq;
  fa.fpi = c->fpi;
  fa.fpi_ew = c->fpi_ew;
  fa.fpj = c->fpj;
  fa.fpj_ew = c->fpj_ew;
  fa.vgrid = c->have_spme ? c->vgrid : nullptr;
  fa.spl = c->spl;
  fa.sg = c->sg;
  fa.order = c->spline_order;
  fa.na = c->grid[0];
  fa.nb = c->grid[1];
  fa.nc = c->grid[2];
  fa.kc = c->kc;
  fa.alpha = c->alpha;
  fa.w_share = c->w_share;
  fa.lattice = c->d_lattice;
  fa.w_cons = has_constraints(c) ? c->w_cons : nullptr;
  fa.fx = c->fx;
  fa.fy = c->fy;
  fa.fz = c->fz;
  fa.pe = c->pe;
  fa.vir = c->vir;
  fa.kick = kick;
  fa.dt = (float)c->dt;
  fa.vx = c->vx;
  fa.vy = c->vy;
  fa.vz = c->vz;
  fa.vxlo = c->lossless ? c->vxlo : nullptr;
  fa.vylo = c->lossless ? c->vylo : nullptr;
  fa.vzlo = c->lossless ? c->vzlo : nullptr;
  fa.mass = c->mass;
  fa.factor = (c->have_nhc || c->rescale) ? c->factors + 1 : nullptr;
  fa.partials = ev ? c->fin_partials : nullptr;
  stage_begin(c, kick ? "finalize_kick" : "finalize", timed);
  hd_launch_finalize(st, fa, ev);
  stage_end(c, timed);
  if (ev) {
    stage_begin(c, "reduce_scalars", timed);
    unsigned np = hd_finalize_blocks((unsigned)c->n, c->have_spme);
    if (baro_mu) {
      // C-rescale: K of the constrained, not yet kicked velocities (scal[3])
      HdBarostatArgs b = baro_args(c);
      hd_launch_sum2_final_crescale(st, c->fin_partials, np, c->scal + (ev == 1 ? 1 : 6),
                                    c->scal + 3, c->d_lattice, c->baro_mu, c->baro_diag,
                                    c->baro_counter, b);
    } else {
      hd_launch_sum2_final(st, c->fin_partials, np, c->scal + (ev == 1 ? 1 : 6));
    }
    stage_end(c, timed);
  }
}

HdSettleArgs settle_args(hd_ctx* c);

HdShakeArgs shake_args(hd_ctx* c, float dt) {
  HdShakeArgs a;
  memset(&a, 0, sizeof(a));
  a.n_groups = (unsigned)(c->h_sh_goff.size() - 1);
  a.goff = c->d_sh_goff;
  a.gatoms = c->d_sh_atoms;
  a.coff = c->d_sh_coff;
  a.cij = c->d_sh_cij;
  a.cr2 = c->d_sh_r2;
  a.posq = c->posq;
  a.snap = c->snap;
  a.vx = c->vx;
  a.vy = c->vy;
  a.vz = c->vz;
  a.mass = c->mass;
  a.w_cons = c->w_cons;
  a.lattice = c->d_lattice;
  a.dt = dt;
  a.tol = (float)c->sh_tol;
  a.rattle_tol = (float)c->sh_rattle_tol;
  a.max_iter = c->sh_max_iter;
  a.stats = c->d_sh_stats;
  a.color = c->d_sh_color;
  a.n_colors = c->sh_n_colors;
  a.n_cons = (unsigned)c->h_sh_cij.size();
  a.cgrp = c->d_sh_cgrp;
  a.feat = c->d_sh_feat;
  a.lam0 = c->sh_predict ? c->d_sh_nu0 : nullptr;
  a.lam_out = c->d_sh_nu;
  a.big = c->d_sh_big;
  a.small = c->d_sh_small;
  a.n_big = c->sh_n_big;
  a.n_small = c->sh_n_small;
  return a;
}

// constraint hooks of run_step (src/integrator/mod.rs:344-429) for both
// constraint kinds: positions after a drift, velocities after the kick
void enqueue_constrain_positions(hd_ctx* c, bool timed) {
  if (!c->h_groups.empty()) {
    HdSettleArgs sa = settle_args(c);
    stage_begin(c, "settle_positions", timed);
    hd_launch_settle_positions(c->stream, sa);
    stage_end(c, timed);
  }
  if (has_shake(c)) {
    HdShakeArgs sa = shake_args(c, (float)c->dt);
    if (c->sh_predict || c->sh_logging) {
      stage_begin(c, "shake_predict", timed);
      hd_launch_shake_features(c->stream, sa);
      if (c->sh_predict)
        hd_launch_mlp_predict(c->stream, c->d_mlp, c->d_sh_feat, sa.n_cons, c->d_sh_nu0);
      stage_end(c, timed);
    }
    stage_begin(c, "shake_positions", timed);
    hd_launch_shake_positions(c->stream, sa);
    stage_end(c, timed);
    if (c->sh_logging)
      hd_launch_shake_log(c->stream, c->d_sh_feat, c->d_sh_nu, sa.n_cons, c->d_log_x,
                          c->d_log_t, c->d_log_ctr, c->log_cap);
  }
}

void enqueue_constrain_velocities(hd_ctx* c, bool timed) {
  if (!c->h_groups.empty()) {
    HdSettleArgs sa = settle_args(c);
    stage_begin(c, "settle_velocities", timed);
    hd_launch_settle_velocities(c->stream, sa, true);
    stage_end(c, timed);
  }
  if (has_shake(c)) {
    stage_begin(c, "rattle_velocities", timed);
    hd_launch_rattle_velocities(c->stream, shake_args(c, (float)c->dt), true);
    stage_end(c, timed);
  }
}

HdIntegrateArgs integrate_args(hd_ctx* c) {
  HdIntegrateArgs ia;
  memset(&ia, 0, sizeof(ia));
  ia.n = (unsigned)c->n;
  ia.dt = (float)c->dt;
  ia.posq = c->posq;
  ia.ix = c->ix;
  ia.iy = c->iy;
  ia.iz = c->iz;
  ia.vx = c->vx;
  ia.vy = c->vy;
  ia.vz = c->vz;
  ia.fx = c->fx;
  ia.fy = c->fy;
  ia.fz = c->fz;
  ia.mass = c->mass;
  ia.lattice = c->d_lattice;
  ia.snap = has_constraints(c) ? c->snap : nullptr;
  return ia;
}

static void settle_kinds_args(const hd_ctx* c, HdSettleArgs& sa) {
  sa.gkind = c->d_gkind;
  for (size_t k = 0; k < c->settle_kinds.size(); ++k) {
    sa.kd_oh[k] = c->settle_kinds[k][0];
    sa.kd_hh[k] = c->settle_kinds[k][1];
    sa.km_o[k] = (float)c->settle_kinds[k][2];
    sa.km_h[k] = (float)c->settle_kinds[k][3];
  }
}

HdSettleArgs settle_args(hd_ctx* c) {
  HdSettleArgs sa;
  memset(&sa, 0, sizeof(sa));
  sa.n_groups = (unsigned)(c->h_groups.size() / 3);
  sa.groups = c->d_groups;
  sa.d_oh = c->d_oh;
  sa.d_hh = c->d_hh;
  sa.m_o = (float)c->m_o;
  sa.m_h = (float)c->m_h;
  settle_kinds_args(c, sa);
  sa.dt = (float)c->dt;
  sa.posq = c->posq;
  sa.snap = c->snap;
  sa.vx = c->vx;
  sa.vy = c->vy;
  sa.vz = c->vz;
  sa.w_cons = c->w_cons;
  sa.lattice = c->d_lattice;
  return sa;
}

HdBarostatArgs baro_args(hd_ctx* c) {
  HdBarostatArgs b;
  memset(&b, 0, sizeof(b));
  b.seed_lo = (unsigned)(c->baro_seed & 0xFFFFFFFFull);
  b.seed_hi = (unsigned)(c->baro_seed >> 32);
  b.pressure = c->baro_p;
  b.tau = c->baro_tau;
  b.compressibility = c->baro_kappa;
  b.kt = c->baro_kt;
  b.dt = c->dt;
  b.mu3_min = 1e-18;
  b.stochastic = c->baro_stochastic ? 1 : 0;
  return b;
}

// C-rescale (src/runner.rs:2664-2668, barostat.cu:144-232): mu comes from
// K of the constrained, not yet kicked velocities (scal[3]) and W of this
// evaluation, computed by the scalar reduction of enqueue_forces (baro_mu);
// positions are scaled after the kick
void enqueue_crescale(hd_ctx* c, bool timed) {
  stage_begin(c, "c_rescale", timed);
  hd_launch_scale_positions(c->stream, c->posq, c->baro_mu, (unsigned)c->n);
  stage_end(c, timed);
}

// Langevin BAOAB (src/integrator/langevin_baoab.rs plan B, A, O, A, F, B;
// the last B is the fused post-force kick).  With SETTLE the constraint
// hooks of run_step wrap each drift (snapshot before, positions after,
// with the full dt as the reference passes it) and the velocity
// constraint follows the force evaluation, before the fused kick
// (src/integrator/mod.rs:344-429); it commutes with the force
// evaluation, so it is issued before it.
void enqueue_langevin_step(hd_ctx* c, bool ev, bool timed) {
  hipStream_t st = c->stream;
  unsigned n = (unsigned)c->n;
  HdIntegrateArgs ia = integrate_args(c);
  HdLangevinArgs la;
  la.alpha = expf(-(float)c->lan_friction * (float)c->dt);
  la.kt = (float)c->lan_kt;
  la.seed_lo = (unsigned)(c->lan_seed & 0xFFFFFFFFull);
  la.seed_hi = (unsigned)(c->lan_seed >> 32);
  la.counter = c->lan_counter;
  if (c->fused_settle && !has_shake(c) && c->h_groups.size() == (size_t)c->n) {
    // every atom in a SETTLE water: both half steps and the three SETTLE
    // stages in one kernel
    stage_begin(c, "langevin_settle", timed);
    hd_launch_lan_settle(st, ia, settle_args(c), la);
    stage_end(c, timed);
  } else if (c->fused_settle && !has_constraints(c)) {
    stage_begin(c, "langevin_kick_drift_ou_drift", timed);
    hd_launch_lan_free(st, ia, la);
    stage_end(c, timed);
  } else {
    stage_begin(c, "langevin_kick_drift", timed);
    hd_launch_lan_kick_drift(st, ia, la);
    stage_end(c, timed);
    enqueue_constrain_positions(c, timed);
    stage_begin(c, "langevin_ou_drift", timed);
    hd_launch_lan_ou_drift(st, ia, la);
    stage_end(c, timed);
    enqueue_constrain_positions(c, timed);
    enqueue_constrain_velocities(c, timed);
  }
  if (c->have_baro)
    hd_launch_ke(st, n, c->vx, c->vy, c->vz, c->mass, c->red_partials, c->scal + 3, nullptr,
                 0.f, nullptr);
  // a barostat consumes the virial every step; the energies only when logged
  int mode = ev ? 1 : (c->have_baro ? 2 : 0);
  enqueue_forces(c, mode, true, timed, false, c->have_baro);
  if (c->have_baro) enqueue_crescale(c, timed);
  if (ev) hd_launch_ke(st, n, c->vx, c->vy, c->vz, c->mass, c->red_partials, c->scal + 0,
                       nullptr, 0.f, nullptr);
}

// CSVR (src/integrator/csvr.rs:188-240) and Berendsen
// (src/integrator/berendsen.rs:139-170) in thermal degrees of freedom
// g = 3N - n_constraints - 3 (at least 1)
HdRescaleArgs rescale_args(hd_ctx* c) {
  HdRescaleArgs t;
  memset(&t, 0, sizeof(t));
  t.kind = c->rescale;
  long long g = std::max(3LL * (long long)c->n - (long long)c->rs_nc - 3LL, 1LL);
  t.g_dof = (unsigned)g;
  t.seed_lo = (unsigned)(c->rs_seed & 0xFFFFFFFFull);
  t.seed_hi = (unsigned)(c->rs_seed >> 32);
  t.c = exp(-(double)(float)c->dt / c->rs_tau);
  t.one_minus_c = 1.0 - t.c;
  double k_target = 0.5 * (double)g * c->rs_kt;
  t.kt_per_dof = k_target / (double)g;
  t.k_target = k_target;
  t.dt_over_tau = (double)(float)c->dt / c->rs_tau;
  t.counter = c->rs_counter;
  t.injection = c->rs_injection;
  t.partials = c->rs_partials;
  return t;
}

void enqueue_step(hd_ctx* c, bool ev, bool timed) {
  hipStream_t st = c->stream;
  unsigned n = (unsigned)c->n;
  HdNhcArgs nhc = nhc_args(c);
  if (c->have_nhc) {
    stage_begin(c, "nhc_pre", timed);
    hd_launch_ke(st, n, c->vx, c->vy, c->vz, c->mass, c->red_partials, c->scal + 3, &nhc,
                 (float)c->dt, c->factors + 0);
    stage_end(c, timed);
  }
  if (c->have_lan) {
    enqueue_langevin_step(c, ev, timed);
    return;
  }
  HdIntegrateArgs ia;
  memset(&ia, 0, sizeof(ia));
  ia.n = n;
  ia.dt = (float)c->dt;
  ia.posq = c->posq;
  ia.ix = c->ix;
  ia.iy = c->iy;
  ia.iz = c->iz;
  ia.vx = c->vx;
  ia.vy = c->vy;
  ia.vz = c->vz;
  if (c->lossless) {
    ia.xlo = c->xlo; ia.ylo = c->ylo; ia.zlo = c->zlo;
    ia.vxlo = c->vxlo; ia.vylo = c->vylo; ia.vzlo = c->vzlo;
  }
  ia.fx = c->fx;
  ia.fy = c->fy;
  ia.fz = c->fz;
  ia.mass = c->mass;
  ia.lattice = c->d_lattice;
  ia.factor = c->have_nhc ? c->factors + 0 : nullptr;
  ia.snap = has_constraints(c) ? c->snap : nullptr;
  if (c->fused_settle && !c->lossless && !has_shake(c) && c->h_groups.size() == (size_t)c->n) {
    // every atom in a SETTLE water (3 indices per water, disjoint): the
    // drift and SETTLE on positions and velocities in one kernel
    stage_begin(c, "vv_kick_drift_settle", timed);
    hd_launch_kick_drift_settle(st, ia, settle_args(c));
    stage_end(c, timed);
  } else {
    stage_begin(c, "vv_kick_drift", timed);
    hd_launch_kick_drift(st, ia);
    stage_end(c, timed);
    enqueue_constrain_positions(c, timed);
    enqueue_constrain_velocities(c, timed);
  }
  // the positions are final for the step: the KE pass also runs the
  // neighbour displacement check
  HdDispArgs disp;
  disp.posq = c->posq;
  disp.ref = c->ref_pos;
  disp.lattice = c->d_lattice;
  disp.ref_lattice = c->ref_lattice;
  disp.half_skin = (float)(0.5 * c->r_skin);
  disp.status = c->ctrl;
  bool disp_done = c->have_nhc || c->rescale || c->have_baro;
  if (c->rescale) {
    stage_begin(c, c->rescale == 2 ? "csvr" : "berendsen", timed);
    hd_launch_ke_rescale(st, n, c->vx, c->vy, c->vz, c->mass, c->red_partials, c->scal + 3,
                         rescale_args(c), c->factors + 1, &disp);
    stage_end(c, timed);
  } else if (c->have_nhc) {
    stage_begin(c, "nhc_post", timed);
    hd_launch_ke(st, n, c->vx, c->vy, c->vz, c->mass, c->red_partials, c->scal + 3, &nhc,
                 (float)c->dt, c->factors + 1, &disp);
    stage_end(c, timed);
  } else if (c->have_baro) {
    hd_launch_ke(st, n, c->vx, c->vy, c->vz, c->mass, c->red_partials, c->scal + 3, nullptr,
                 0.f, nullptr, &disp);
  }
  // a barostat consumes the virial every step; the energies only when logged
  int mode = ev ? 1 : (c->have_baro ? 2 : 0);
  enqueue_forces(c, mode, true, timed, disp_done, c->have_baro);
  if (c->have_baro) enqueue_crescale(c, timed);
  if (c->have_andersen) {
    // the collision probability is formed in f32 (andersen.rs:160)
    float p = (float)std::min(std::max(c->and_rate * (double)(float)c->dt, 0.0), 1.0);
    stage_begin(c, "andersen", timed);
    hd_launch_andersen(st, c->vx, c->vy, c->vz, c->mass, c->and_counter, c->and_seed, p,
                       (float)c->and_kt, n);
    stage_end(c, timed);
  }
  if (ev) hd_launch_ke(st, n, c->vx, c->vy, c->vz, c->mass, c->red_partials, c->scal + 0,
                       nullptr, 0.f, nullptr);
}

}  // namespace

namespace {
template <typename T>
static int download(hd_ctx* c, T* dst, const T* src, size_t count) {
  if (!dst) return HD_OK;
  HIPCHK(c, hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  return HD_OK;
}

}  // namespace

extern "C" {

const char* hd_version(void) { return HD_VERSION; }

int hd_ctx_create(int device, hd_ctx** out) {
  if (!out) return HD_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return HD_ERR_HIP;
  if (device < 0 || device >= count) return HD_ERR_ARG;
  hd_ctx* c = new hd_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return HD_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_stat[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_stat[1], hipEventDisableTiming) != hipSuccess ||
      hipHostMalloc((void**)&c->h_stat, 8 * sizeof(unsigned), hipHostMallocDefault) != hipSuccess) {
    if (c->h_stat) (void)hipHostFree(c->h_stat);
    for (hipEvent_t e : c->ev_stat)
      if (e) (void)hipEventDestroy(e);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HD_ERR_HIP;
  }
  *out = c;
  return HD_OK;
}

void hd_ctx_destroy(hd_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->side);
  if (c->graph_f) (void)hipGraphExecDestroy(c->graph_f);
  if (c->graph_fev) (void)hipGraphExecDestroy(c->graph_fev);
  if (c->graph_fk) (void)hipGraphExecDestroy(c->graph_fk);
  for (auto& s : c->stages) {
    (void)hipEventDestroy(s.a);
    (void)hipEventDestroy(s.b);
  }
  for (void* p : c->allocs) (void)hipFree(p);
  (void)hipEventDestroy(c->ev_fork);
  (void)hipEventDestroy(c->ev_join);
  (void)hipEventDestroy(c->ev_stat[0]);
  (void)hipEventDestroy(c->ev_stat[1]);
  (void)hipHostFree(c->h_stat);
  (void)hipStreamDestroy(c->side);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* hd_last_error(const hd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int hd_set_system(hd_ctx* c, uint64_t n, const float* posq, const float* mass,
                  const uint32_t* type, const uint32_t* key) {
  if (!c || !posq || !mass || !type) return fail(c, HD_ERR_ARG, "hd_set_system: null argument");
  if (n == 0 || n >= (1ull << 31)) return fail(c, HD_ERR_ARG, "hd_set_system: bad n %llu",
                                               (unsigned long long)n);
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_system after hd_setup");
  c->n = n;
  c->h_posq.assign(posq, posq + 4 * n);
  c->h_mass.assign(mass, mass + n);
  c->h_type.assign(type, type + n);
  c->h_key.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    c->h_key[i] = key ? key[i] : (uint32_t)i;
    if (c->h_key[i] >= HD_SENTINEL_KEY_LIMIT)
      return fail(c, HD_ERR_ARG, "exclusion key %u too large", c->h_key[i]);
    if (!(mass[i] > 0.f)) return fail(c, HD_ERR_ARG, "mass[%llu] must be positive",
                                      (unsigned long long)i);
  }
  c->h_vx.assign(n, 0.f);
  c->h_vy.assign(n, 0.f);
  c->h_vz.assign(n, 0.f);
  c->h_ix.assign(n, 0);
  c->h_iy.assign(n, 0);
  c->h_iz.assign(n, 0);
  return HD_OK;
}

int hd_set_lattice(hd_ctx* c, const double* L) {
  if (!c || !L) return HD_ERR_ARG;
  if (!(L[0] > 0 && L[1] > 0 && L[2] > 0)) return fail(c, HD_ERR_ARG, "lattice lengths must be > 0");
  memcpy(c->lattice, L, sizeof(double) * 6);
  if (c->ready) {
    float lf[6];
    for (int i = 0; i < 6; ++i) lf[i] = (float)L[i];
    HIPCHK(c, hipMemcpy(c->d_lattice, lf, sizeof(lf), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->ctrl, 0, sizeof(unsigned)));
    unsigned force = 0x100u;
    HIPCHK(c, hipMemcpy(c->ctrl, &force, sizeof(unsigned), hipMemcpyHostToDevice));
  }
  return HD_OK;
}

int hd_set_velocities(hd_ctx* c, const float* vx, const float* vy, const float* vz) {
  if (!c || !vx || !vy || !vz) return HD_ERR_ARG;
  if (!c->n) return fail(c, HD_ERR_STATE, "hd_set_velocities before hd_set_system");
  if (c->ready) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int r;
    if ((r = upload(c, c->vx, vx, c->n))) return r;
    if ((r = upload(c, c->vy, vy, c->n))) return r;
    if ((r = upload(c, c->vz, vz, c->n))) return r;
    if (c->lossless) {
      HIPCHK(c, hipMemset(c->vxlo, 0, sizeof(double) * c->n));
      HIPCHK(c, hipMemset(c->vylo, 0, sizeof(double) * c->n));
      HIPCHK(c, hipMemset(c->vzlo, 0, sizeof(double) * c->n));
    }
    return HD_OK;
  }
  c->h_vx.assign(vx, vx + c->n);
  c->h_vy.assign(vy, vy + c->n);
  c->h_vz.assign(vz, vz + c->n);
  return HD_OK;
}

int hd_set_images(hd_ctx* c, const int32_t* ix, const int32_t* iy, const int32_t* iz) {
  if (!c || !c->n) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_images after hd_setup");
  for (uint64_t i = 0; i < c->n; ++i) {
    c->h_ix[i] = ix ? ix[i] : 0;
    c->h_iy[i] = iy ? iy[i] : 0;
    c->h_iz[i] = iz ? iz[i] : 0;
  }
  return HD_OK;
}

int hd_set_lj(hd_ctx* c, int ntypes, const double* sigma, const double* eps, const double* cut,
              const double* rsw) {
  if (!c || !sigma || !eps || !cut || !rsw) return HD_ERR_ARG;
  if (ntypes < 1 || ntypes > hd_pair_table_max_types())
    return fail(c, HD_ERR_UNSUPPORTED, "ntypes=%d: the pair kernels take at most %d types",
                ntypes, hd_pair_table_max_types());
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_lj after hd_setup");
  int m = ntypes * ntypes;
  for (int k = 0; k < m; ++k) {
    if (!(cut[k] > 0) || rsw[k] > cut[k] || rsw[k] < 0 || sigma[k] < 0 || eps[k] < 0)
      return fail(c, HD_ERR_ARG, "invalid LJ entry %d", k);
    int i = k / ntypes, j = k % ntypes;
    int t = j * ntypes + i;
    if (sigma[k] != sigma[t] || eps[k] != eps[t] || cut[k] != cut[t] || rsw[k] != rsw[t])
      return fail(c, HD_ERR_ARG, "LJ tables must be symmetric");
  }
  c->ntypes = ntypes;
  c->lj_sigma.assign(sigma, sigma + m);
  c->lj_eps.assign(eps, eps + m);
  c->lj_cut.assign(cut, cut + m);
  c->lj_rsw.assign(rsw, rsw + m);
  return HD_OK;
}

int hd_set_spme(hd_ctx* c, double alpha, double r_cut, int na, int nb, int nc, int order,
                double k_c) {
  if (!c) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_spme after hd_setup");
  if (order != 4 && order != 5)
    return fail(c, HD_ERR_UNSUPPORTED, "spline order %d: orders 4 and 5 are built", order);
  // the LDS FFT takes lines of 2^a 3^b 5^c points: up to 256 along a and
  // b, and an even count up to 512 along c (packed as nc/2 complex points)
  if (!is_smooth235(na) || !is_smooth235(nb) || !is_smooth235(nc) || nc % 2 || na < 8 ||
      nb < 8 || nc < 8 || na > 256 || nb > 256 || nc > 512)
    return fail(c, HD_ERR_UNSUPPORTED,
                "SPME grid %dx%dx%d: each side must be 2^a 3^b 5^c, at least 8, at most 256 "
                "(a, b) or 512 and even (c)",
                na, nb, nc);
  if (!(alpha > 0) || !(r_cut > 0)) return fail(c, HD_ERR_ARG, "alpha and r_cut must be > 0");
  c->have_spme = true;
  c->alpha = alpha;
  c->coul_rc = r_cut;
  c->grid[0] = na;
  c->grid[1] = nb;
  c->grid[2] = nc;
  c->spline_order = order;
  c->kc = k_c;
  return HD_OK;
}

int hd_set_neighbor(hd_ctx* c, double r_skin) {
  if (!c || !(r_skin >= 0)) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_neighbor after hd_setup");
  c->r_skin = r_skin;
  return HD_OK;
}

int hd_set_settle(hd_ctx* c, uint64_t ng, const uint32_t* groups, double d_oh, double d_hh,
                  double m_o, double m_h) {
  if (!c || (ng && !groups)) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_settle after hd_setup");
  if (!(d_oh > 0 && d_hh > 0 && d_hh < 2 * d_oh && m_o > 0 && m_h > 0))
    return fail(c, HD_ERR_ARG, "invalid SETTLE geometry");
  std::vector<char> seen(c->n, 0);
  for (uint64_t k = 0; k < 3 * ng; ++k) {
    if (groups[k] >= c->n) return fail(c, HD_ERR_ARG, "SETTLE atom %u out of range", groups[k]);
    if (seen[groups[k]]) return fail(c, HD_ERR_ARG, "atom %u in two SETTLE groups", groups[k]);
    seen[groups[k]] = 1;
  }
  c->h_groups.assign(groups, groups + 3 * ng);
  c->d_oh = d_oh;
  c->d_hh = d_hh;
  c->m_o = m_o;
  c->m_h = m_h;
  c->h_gkind.assign(ng, 0);
  c->settle_kinds.assign(1, {d_oh, d_hh, m_o, m_h});
  return HD_OK;
}

int hd_add_settle(hd_ctx* c, uint64_t ng, const uint32_t* groups, double d_oh, double d_hh,
                  double m_o, double m_h) {
  if (!c || (ng && !groups)) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_add_settle after hd_setup");
  if (c->h_groups.empty()) return hd_set_settle(c, ng, groups, d_oh, d_hh, m_o, m_h);
  if (!(d_oh > 0 && d_hh > 0 && d_hh < 2 * d_oh && m_o > 0 && m_h > 0))
    return fail(c, HD_ERR_ARG, "invalid SETTLE geometry");
  std::vector<char> seen(c->n, 0);
  for (uint32_t a : c->h_groups) seen[a] = 1;
  for (uint64_t k = 0; k < 3 * ng; ++k) {
    if (groups[k] >= c->n) return fail(c, HD_ERR_ARG, "SETTLE atom %u out of range", groups[k]);
    if (seen[groups[k]]) return fail(c, HD_ERR_ARG, "atom %u in two SETTLE groups", groups[k]);
    seen[groups[k]] = 1;
  }
  std::array<double, 4> g = {d_oh, d_hh, m_o, m_h};
  size_t kind = 0;
  while (kind < c->settle_kinds.size() && c->settle_kinds[kind] != g) ++kind;
  if (kind == c->settle_kinds.size()) {
    if (kind >= HD_SETTLE_KINDS)
      return fail(c, HD_ERR_UNSUPPORTED, "more than %d SETTLE geometries", HD_SETTLE_KINDS);
    c->settle_kinds.push_back(g);
  }
  c->h_groups.insert(c->h_groups.end(), groups, groups + 3 * ng);
  c->h_gkind.insert(c->h_gkind.end(), ng, (uint8_t)kind);
  return HD_OK;
}

int hd_set_angles(hd_ctx* c, uint64_t n_angles, const uint32_t* ijk, const double* k_theta,
                  const double* theta0) {
  if (!c || (n_angles && (!ijk || !k_theta || !theta0))) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_angles after hd_setup");
  for (uint64_t a = 0; a < 3 * n_angles; ++a)
    if (ijk[a] >= c->n) return fail(c, HD_ERR_ARG, "angle atom %u out of range", ijk[a]);
  c->h_angles.assign(ijk, ijk + 3 * n_angles);
  c->h_ang_k.assign(k_theta, k_theta + n_angles);
  c->h_ang_t0.assign(theta0, theta0 + n_angles);
  return HD_OK;
}

int hd_set_bonds(hd_ctx* c, uint64_t n_bonds, const uint32_t* ij, const int* kind,
                 const double* prm) {
  if (!c || (n_bonds && (!ij || !kind || !prm))) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_bonds after hd_setup");
  for (uint64_t b = 0; b < n_bonds; ++b) {
    if (ij[2 * b] >= c->n || ij[2 * b + 1] >= c->n || ij[2 * b] == ij[2 * b + 1])
      return fail(c, HD_ERR_ARG, "bond %llu: bad atoms %u %u", (unsigned long long)b, ij[2 * b],
                  ij[2 * b + 1]);
    if (kind[b] != 0 && kind[b] != 1)
      return fail(c, HD_ERR_ARG, "bond %llu: kind %d (0 harmonic, 1 Morse)",
                  (unsigned long long)b, kind[b]);
  }
  c->h_bonds.assign(ij, ij + 2 * n_bonds);
  c->h_bond_kind.assign(kind, kind + n_bonds);
  c->h_bond_prm.assign(prm, prm + 3 * n_bonds);
  return HD_OK;
}

int hd_set_dihedrals(hd_ctx* c, uint64_t n_dih, const uint32_t* ijkl, const double* k_phi,
                     const double* n_mult, const double* phi0) {
  if (!c || (n_dih && (!ijkl || !k_phi || !n_mult || !phi0))) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_dihedrals after hd_setup");
  c->h_dih.assign(ijkl, ijkl + 4 * n_dih);
  c->h_dih_prm.resize(3 * n_dih);
  for (uint64_t d = 0; d < n_dih; ++d) {
    for (int a = 0; a < 4; ++a)
      if (ijkl[4 * d + a] >= c->n)
        return fail(c, HD_ERR_ARG, "dihedral %llu: atom %u out of range", (unsigned long long)d,
                    ijkl[4 * d + a]);
    c->h_dih_prm[3 * d] = k_phi[d];
    c->h_dih_prm[3 * d + 1] = n_mult[d];
    c->h_dih_prm[3 * d + 2] = phi0[d];
  }
  return HD_OK;
}

// Full exclusions (both scales 0) become a key rule the pair kernel applies
// per pair: either every component of the full-exclusion graph is a clique
// (molecules; one key per component, span 0), or the components are runs
// of consecutive atoms in which exactly the pairs at most s apart are
// excluded (chains with 1-2 ... 1-(s+1) exclusions; key = index plus s per
// earlier component, span s).  Scaled pairs keep their listed scales
// through the correction pass.
int hd_set_exclusions(hd_ctx* c, uint64_t n_pairs, const uint32_t* ij, const double* s_lj,
                      const double* s_coul) {
  if (!c || (n_pairs && (!ij || !s_lj || !s_coul))) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_exclusions after hd_setup");
  if (!c->n) return fail(c, HD_ERR_STATE, "hd_set_exclusions before hd_set_system");
  uint64_t n = c->n;
  std::vector<std::pair<uint64_t, uint64_t>> full;
  std::set<std::pair<uint32_t, uint32_t>> seen;
  for (uint64_t p = 0; p < n_pairs; ++p) {
    uint32_t a = ij[2 * p], b = ij[2 * p + 1];
    if (a >= n || b >= n || a == b)
      return fail(c, HD_ERR_ARG, "exclusion %llu: bad atoms %u %u", (unsigned long long)p, a, b);
    if (!(s_lj[p] >= 0 && s_lj[p] <= 1 && s_coul[p] >= 0 && s_coul[p] <= 1))
      return fail(c, HD_ERR_ARG, "exclusion %llu: scales must lie in [0, 1]",
                  (unsigned long long)p);
    auto key = std::make_pair(std::min(a, b), std::max(a, b));
    if (!seen.insert(key).second)
      return fail(c, HD_ERR_ARG, "exclusion (%u, %u) listed twice", key.first, key.second);
    if (s_lj[p] == 0.0 && s_coul[p] == 0.0) full.emplace_back(key.first, key.second);
  }
  // components of the full-exclusion graph
  std::vector<uint64_t> parent(n);
  for (uint64_t i = 0; i < n; ++i) parent[i] = i;
  std::function<uint64_t(uint64_t)> root = [&](uint64_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (auto& e : full) {
    uint64_t ra = root(e.first), rb = root(e.second);
    if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
  }
  std::vector<uint64_t> size(n, 0), edges(n, 0), lo(n, UINT64_MAX), hi(n, 0);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t r = root(i);
    size[r]++;
    lo[r] = std::min(lo[r], i);
    hi[r] = std::max(hi[r], i);
  }
  uint64_t span = 0;
  for (auto& e : full) {
    edges[root(e.first)]++;
    span = std::max(span, e.second - e.first);
  }
  bool cliques = true;
  for (uint64_t i = 0; i < n; ++i)
    if (root(i) == i && edges[i] != size[i] * (size[i] - 1) / 2) cliques = false;
  // runs of consecutive atoms excluded exactly up to index distance span
  bool chains = !cliques && span <= (1u << 20) &&
                (uint64_t)n * (span + 1) < HD_SENTINEL_KEY_LIMIT;
  for (uint64_t i = 0; chains && i < n; ++i) {
    uint64_t r = root(i);
    if (hi[r] - lo[r] + 1 != size[r]) chains = false;
    if (chains && r == i) {
      uint64_t m = size[r], k = std::min(span, m - 1);
      if (edges[r] != m * k - k * (k + 1) / 2) chains = false;
    }
  }
  std::vector<uint32_t> key(n);
  unsigned new_span = 0;
  if (chains) {
    uint64_t comp = 0;
    for (uint64_t i = 0; i < n; ++i) {
      if (i > 0 && root(i) != root(i - 1)) ++comp;
      key[i] = (uint32_t)(i + comp * span);
    }
    new_span = (unsigned)span;
  } else {
    // one key per component (numbered in order of first atom): the pair
    // kernel leaves out every pair inside a component; in a component that
    // is not a clique the pairs that are not fully excluded are added back
    // below with their weights (1, or the listed scales)
    std::vector<uint32_t> id(n, 0);
    uint32_t next = 0;
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t r = root(i);
      if (r == i) id[r] = next++;
      key[i] = id[r];
    }
  }
  // scaled pairs: weight relative to what the key rule leaves
  std::vector<HdExclPair> sc;
  std::map<std::pair<uint32_t, uint32_t>, std::pair<double, double>> scaled;
  for (uint64_t p = 0; p < n_pairs; ++p) {
    if (s_lj[p] == 0.0 && s_coul[p] == 0.0) continue;
    uint32_t a = ij[2 * p], b = ij[2 * p + 1];
    scaled[std::make_pair(std::min(a, b), std::max(a, b))] = {s_lj[p], s_coul[p]};
  }
  for (auto& kv : scaled) {
    uint32_t a = kv.first.first, b = kv.first.second;
    uint32_t d = key[a] > key[b] ? key[a] - key[b] : key[b] - key[a];
    double base = d <= new_span ? 0.0 : 1.0;
    HdExclPair e;
    memset(&e, 0, sizeof(e));
    e.i = a;
    e.j = b;
    e.w_lj = kv.second.first - base;
    e.w_coul = kv.second.second - base;
    if (e.w_lj != 0.0 || e.w_coul != 0.0) sc.push_back(e);
  }
  if (!chains && !cliques) {
    // add-back pairs of the components that are not cliques
    std::vector<std::vector<uint32_t>> members(n);
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t r = root(i);
      if (edges[r] != size[r] * (size[r] - 1) / 2) members[r].push_back((uint32_t)i);
    }
    for (uint64_t r = 0; r < n; ++r) total += members[r].size() * (members[r].size() - (members[r].empty() ? 0 : 1)) / 2;
    if (total > (8u << 20))
      return fail(c, HD_ERR_UNSUPPORTED,
                  "exclusions: %llu intramolecular pairs to add back (limit 8 Mi); the molecules "
                  "are neither cliques nor chains",
                  (unsigned long long)total);
    std::set<std::pair<uint32_t, uint32_t>> full_set;
    for (auto& e : full) full_set.insert(std::make_pair((uint32_t)e.first, (uint32_t)e.second));
    for (uint64_t r = 0; r < n; ++r) {
      const auto& m = members[r];
      for (size_t x = 0; x < m.size(); ++x)
        for (size_t y = x + 1; y < m.size(); ++y) {
          auto kp = std::make_pair(m[x], m[y]);
          if (full_set.count(kp) || scaled.count(kp)) continue;
          HdExclPair e;
          memset(&e, 0, sizeof(e));
          e.i = m[x];
          e.j = m[y];
          e.w_lj = 1.0;
          e.w_coul = 1.0;
          sc.push_back(e);
        }
    }
  }
  c->h_key = key;
  c->excl_span = new_span;
  c->h_excl = sc;
  return HD_OK;
}

int hd_set_exclusion_span(hd_ctx* c, uint32_t span) {
  if (!c) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_exclusion_span after hd_setup");
  if (span > (1u << 20)) return fail(c, HD_ERR_ARG, "exclusion span %u too large", span);
  c->excl_span = span;
  return HD_OK;
}

static void drop_graphs(hd_ctx* c);

// Split one DFT matrix (M x K row-major, M a multiple of 32, K of 16) into
// FP16 high and 2^12-scaled low parts, stored in the fragment order of the
// 32x32x16 matrix-core step: element e of lane l of (row tile mt, K step
// ks) at ((mt (K/16) + ks) 64 + l) 8 + e holds W[32 mt + l % 32][16 ks +
// 8 (l / 32) + e].
static void dft_split_host(const std::vector<double>& w, unsigned M, unsigned K, _Float16* hi,
                           _Float16* lo) {
  unsigned nks = K / 16;
  for (unsigned mt = 0; mt < M / 32; ++mt)
    for (unsigned ks = 0; ks < nks; ++ks)
      for (unsigned l = 0; l < 64; ++l)
        for (unsigned e = 0; e < 8; ++e) {
          size_t dst = (((size_t)mt * nks + ks) * 64 + l) * 8 + e;
          float x = (float)w[(size_t)(32 * mt + l % 32) * K + 16 * ks + 8 * (l / 32) + e];
          _Float16 h = (_Float16)x;
          hi[dst] = h;
          lo[dst] = (_Float16)((x - (float)h) * 4096.0f);
        }
}

// DFT matrices of the split-precision path (layouts in hd_kernels.h)
static int build_dft_mats(hd_ctx* c) {
  if (c->dft_mats) return HD_OK;
  unsigned na = c->grid[0], nb = c->grid[1], nc = c->grid[2], nch = nc / 2 + 1;
  if (2 * na > 256 || 2 * nb > 256 || nc > 256 || na < 32 || nb < 32 || nc < 32 ||
      !is_pow2((int)na) || !is_pow2((int)nb) || !is_pow2((int)nc))
    return fail(c, HD_ERR_UNSUPPORTED,
                "the matrix-core DFT handles power-of-two axes of 32 to 128 points");
  unsigned cf_m = (2 * nch + 31) / 32 * 32, cf_k = nc;
  unsigned ci_m = nc, ci_k = (2 * nch + 15) / 16 * 16;
  size_t sz_cf = (size_t)cf_m * cf_k, sz_ci = (size_t)ci_m * ci_k;
  size_t sz_b = (size_t)(2 * nb) * (2 * nb), sz_a = (size_t)(2 * na) * (2 * na);
  size_t total = 2 * (sz_cf + sz_ci + 2 * sz_b + 2 * sz_a);
  std::vector<_Float16> h(total);
  const double two_pi = 2.0 * M_PI;
  std::vector<double> w;
  size_t off = 0;
  auto put = [&](const std::vector<double>& m, unsigned M, unsigned K, const _Float16** hi,
                 const _Float16** lo) {
    dft_split_host(m, M, K, h.data() + off, h.data() + off + m.size());
    *hi = (const _Float16*)(uintptr_t)off;  // offsets for now, pointers after upload
    *lo = (const _Float16*)(uintptr_t)(off + m.size());
    off += 2 * m.size();
  };
  // c-axis forward: rows k' < nch -> cos, nch <= m < 2 nch -> -sin
  w.assign(sz_cf, 0.0);
  for (unsigned m = 0; m < 2 * nch; ++m)
    for (unsigned x = 0; x < nc; ++x) {
      unsigned k = m < nch ? m : m - nch;
      double th = two_pi * (double)((k * x) % nc) / nc;
      w[(size_t)m * cf_k + x] = m < nch ? cos(th) : -sin(th);
    }
  put(w, cf_m, cf_k, &c->dft.cf_hi, &c->dft.cf_lo);
  // c-axis inverse over the Hermitian half spectrum
  w.assign(sz_ci, 0.0);
  for (unsigned x = 0; x < nc; ++x)
    for (unsigned k = 0; k < 2 * nch; ++k) {
      unsigned kk = k < nch ? k : k - nch;
      double wt = (kk == 0 || 2 * kk == nc) ? 1.0 : 2.0;
      double th = two_pi * (double)((kk * x) % nc) / nc;
      w[(size_t)x * ci_k + k] = k < nch ? wt * cos(th) : -wt * sin(th);
    }
  put(w, ci_m, ci_k, &c->dft.ci_hi, &c->dft.ci_lo);
  // complex axes, [[Cr, -Ci], [Ci, Cr]] with C = exp(sign 2 pi i m k / n)
  auto cplx = [&](unsigned n, double sign) {
    std::vector<double> m2((size_t)(2 * n) * (2 * n), 0.0);
    for (unsigned m = 0; m < n; ++m)
      for (unsigned k = 0; k < n; ++k) {
        double th = two_pi * (double)((m * k) % n) / n;
        double cr = cos(th), ci = sign * sin(th);
        m2[(size_t)m * 2 * n + k] = cr;
        m2[(size_t)m * 2 * n + n + k] = -ci;
        m2[(size_t)(n + m) * 2 * n + k] = ci;
        m2[(size_t)(n + m) * 2 * n + n + k] = cr;
      }
    return m2;
  };
  put(cplx(nb, -1.0), 2 * nb, 2 * nb, &c->dft.bf_hi, &c->dft.bf_lo);
  put(cplx(nb, +1.0), 2 * nb, 2 * nb, &c->dft.bi_hi, &c->dft.bi_lo);
  put(cplx(na, -1.0), 2 * na, 2 * na, &c->dft.af_hi, &c->dft.af_lo);
  put(cplx(na, +1.0), 2 * na, 2 * na, &c->dft.ai_hi, &c->dft.ai_lo);
  c->dft.cf_m = cf_m; c->dft.cf_k = cf_k; c->dft.ci_m = ci_m; c->dft.ci_k = ci_k;
  DALLOC(c, c->dft_mats, total);
  int r = upload(c, c->dft_mats, h.data(), total);
  if (r) return r;
  const _Float16** ptrs[] = {&c->dft.cf_hi, &c->dft.cf_lo, &c->dft.ci_hi, &c->dft.ci_lo,
                             &c->dft.bf_hi, &c->dft.bf_lo, &c->dft.bi_hi, &c->dft.bi_lo,
                             &c->dft.af_hi, &c->dft.af_lo, &c->dft.ai_hi, &c->dft.ai_lo};
  for (auto p : ptrs) *p = c->dft_mats + (size_t)(uintptr_t)(*p);
  return HD_OK;
}

int hd_set_spme_dft(hd_ctx* c, int mode) {
  if (!c || mode < 0 || mode > 2) return HD_ERR_ARG;
  if (!c->have_spme) return fail(c, HD_ERR_STATE, "no SPME configured");
  if (c->ready) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    if (mode == 1) {
      int r = build_dft_mats(c);
      if (r) return r;
    }
  }
  c->dft_mode = mode;
  return HD_OK;
}

int hd_set_shake(hd_ctx* c, uint64_t ng, const uint32_t* goff, const uint32_t* gatoms,
                 const uint32_t* coff, const uint8_t* ci, const uint8_t* cj, const double* d) {
  if (!c || (ng && (!goff || !gatoms || !coff || !ci || !cj || !d))) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_shake after hd_setup");
  if (ng && (goff[0] != 0 || coff[0] != 0))
    return fail(c, HD_ERR_ARG, "SHAKE group offsets must start at 0");
  std::vector<char> seen(c->n, 0);
  for (uint32_t a : c->h_groups) seen[a] = 1;
  std::vector<uint8_t> color(ng ? coff[ng] : 0, 0);
  unsigned ncol = 0;
  for (uint64_t g = 0; g < ng; ++g) {
    if (goff[g + 1] <= goff[g] || coff[g + 1] < coff[g])
      return fail(c, HD_ERR_ARG, "SHAKE group %llu has decreasing offsets", (unsigned long long)g);
    unsigned na = goff[g + 1] - goff[g], nc = coff[g + 1] - coff[g];
    if (na > HD_SHAKE_MAX_ATOMS || nc > HD_SHAKE_MAX_CONS)
      return fail(c, HD_ERR_UNSUPPORTED, "SHAKE group %llu: %u atoms / %u constraints exceed %d / %d",
                  (unsigned long long)g, na, nc, HD_SHAKE_MAX_ATOMS, HD_SHAKE_MAX_CONS);
    for (unsigned k = goff[g]; k < goff[g + 1]; ++k) {
      if (gatoms[k] >= c->n) return fail(c, HD_ERR_ARG, "SHAKE atom %u out of range", gatoms[k]);
      if (seen[gatoms[k]]) return fail(c, HD_ERR_ARG, "atom %u in two constraint groups", gatoms[k]);
      seen[gatoms[k]] = 1;
    }
    for (unsigned k = coff[g]; k < coff[g + 1]; ++k)
      if (ci[k] >= na || cj[k] >= na || ci[k] == cj[k] || !(d[k] > 0))
        return fail(c, HD_ERR_ARG, "SHAKE constraint %u of group %llu is invalid", k,
                    (unsigned long long)g);
    // greedy colouring in topology order: the lowest colour that no earlier
    // constraint on either atom holds
    uint32_t used[HD_SHAKE_MAX_ATOMS] = {0};
    for (unsigned k = coff[g]; k < coff[g + 1]; ++k) {
      uint32_t busy = used[ci[k]] | used[cj[k]];
      unsigned q = 0;
      while (busy & (1u << q)) ++q;
      if (q >= HD_SHAKE_MAX_COLORS)
        return fail(c, HD_ERR_UNSUPPORTED, "SHAKE group %llu needs more than %d constraint colours",
                    (unsigned long long)g, HD_SHAKE_MAX_COLORS);
      color[k] = (uint8_t)q;
      used[ci[k]] |= 1u << q;
      used[cj[k]] |= 1u << q;
      ncol = std::max(ncol, q + 1);
    }
  }
  c->h_sh_goff.assign(goff, goff + (ng ? ng + 1 : 0));
  c->h_sh_atoms.assign(gatoms, gatoms + (ng ? goff[ng] : 0));
  c->h_sh_coff.assign(coff, coff + (ng ? ng + 1 : 0));
  unsigned ncons = ng ? coff[ng] : 0;
  c->h_sh_cij.resize(ncons);
  c->h_sh_r2.resize(ncons);
  for (unsigned k = 0; k < ncons; ++k) {
    c->h_sh_cij[k] = (uint32_t)ci[k] | ((uint32_t)cj[k] << 8);
    float df = (float)d[k];  // r2 = (Real d)^2 (shake.rs:301-302)
    c->h_sh_r2[k] = df * df;
  }
  c->h_sh_color = color;
  c->sh_n_colors = ncol;
  return HD_OK;
}

int hd_set_shake_tolerance(hd_ctx* c, double tol, double rattle_tol, int max_iter) {
  if (!c || !(tol > 0) || !(rattle_tol > 0) || max_iter < 1) return HD_ERR_ARG;
  if (c->ready) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  c->sh_tol = tol;
  c->sh_rattle_tol = rattle_tol;
  c->sh_max_iter = max_iter;
  return HD_OK;
}

int hd_get_constraint_stats(hd_ctx* c, uint64_t* out) {
  if (!c || !c->ready || !out) return fail(c, HD_ERR_STATE, "not set up");
  unsigned st[6] = {0, 0, 0, 0, 0, 0};
  if (c->d_sh_stats) {
    HIPCHK(c, hipMemcpyAsync(st, c->d_sh_stats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  for (int k = 0; k < 6; ++k) out[k] = st[k];
  return HD_OK;
}

static void drop_graphs(hd_ctx* c);

// ---- SHAKE multiplier predictor and step log -------------------------

static int ensure_shake_features(hd_ctx* c) {
  size_t nc = c->h_sh_cij.size();
  if (!c->d_sh_feat || !c->d_sh_nu0) drop_graphs(c);
  if (!c->d_sh_feat) DALLOC(c, c->d_sh_feat, nc * HD_SHAKE_NFEAT);
  if (!c->d_sh_nu0) DALLOC(c, c->d_sh_nu0, nc);
  return HD_OK;
}

int hd_set_shake_predictor(hd_ctx* c, const float* params) {
  if (!c) return HD_ERR_ARG;
  if (!c->ready || !has_shake(c)) return fail(c, HD_ERR_STATE, "no SHAKE groups set up");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (params) {
    for (int k = 0; k < HD_MLP_NPARAM; ++k)
      if (!std::isfinite(params[k])) return fail(c, HD_ERR_ARG, "predictor parameter %d not finite", k);
    int r = ensure_shake_features(c);
    if (r) return r;
    if (!c->d_mlp) {
      drop_graphs(c);
      DALLOC(c, c->d_mlp, HD_MLP_NPARAM);
    }
    if ((r = upload(c, c->d_mlp, params, HD_MLP_NPARAM))) return r;
  }
  bool on = params != nullptr;
  if (on != c->sh_predict) {
    c->sh_predict = on;
    drop_graphs(c);
  }
  return HD_OK;
}

int hd_get_shake_multipliers(hd_ctx* c, float* nu0, float* nu) {
  if (!c) return HD_ERR_ARG;
  if (!c->ready || !has_shake(c)) return fail(c, HD_ERR_STATE, "no SHAKE groups set up");
  size_t nc = c->h_sh_cij.size();
  HIPCHK(c, hipSetDevice(c->device));
  if (nu0) {
    if (c->d_sh_nu0)
      HIPCHK(c, hipMemcpyAsync(nu0, c->d_sh_nu0, nc * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream));
    else
      memset(nu0, 0, nc * sizeof(float));
  }
  if (nu)
    HIPCHK(c, hipMemcpyAsync(nu, c->d_sh_nu, nc * sizeof(float), hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

int hd_shake_log_start(hd_ctx* c, uint64_t max_steps) {
  if (!c || max_steps == 0) return HD_ERR_ARG;
  if (!c->ready || !has_shake(c)) return fail(c, HD_ERR_STATE, "no SHAKE groups set up");
  size_t nc = c->h_sh_cij.size();
  if ((double)max_steps * (double)nc >= 4294967296.0)
    return fail(c, HD_ERR_ARG, "log of %llu steps x %zu constraints exceeds 2^32 rows",
                (unsigned long long)max_steps, nc);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int r = ensure_shake_features(c);
  if (r) return r;
  bool fresh = !c->sh_logging;
  if (c->log_cap != max_steps) {
    // captured step graphs hold the old log pointers: drop them with the
    // buffers
    drop_graphs(c);
    dfree(c, c->d_log_x);
    dfree(c, c->d_log_t);
    c->d_log_x = c->d_log_t = nullptr;
    c->log_cap = 0;
    DALLOC(c, c->d_log_x, (size_t)max_steps * nc * HD_SHAKE_NFEAT);
    DALLOC(c, c->d_log_t, (size_t)max_steps * nc);
    c->log_cap = max_steps;
  }
  if (!c->d_log_ctr) {
    drop_graphs(c);
    DALLOC(c, c->d_log_ctr, 1);
  }
  HIPCHK(c, hipMemsetAsync(c->d_log_ctr, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sh_logging = true;
  if (fresh) drop_graphs(c);
  return HD_OK;
}

int hd_shake_log_stop(hd_ctx* c) {
  if (!c) return HD_ERR_ARG;
  if (c->sh_logging) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->sh_logging = false;
    drop_graphs(c);
  }
  return HD_OK;
}

static int log_steps(hd_ctx* c, uint64_t* steps) {
  unsigned long long k = 0;
  if (c->d_log_ctr) {
    HIPCHK(c, hipMemcpyAsync(&k, c->d_log_ctr, sizeof(k), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  *steps = std::min<unsigned long long>(k, c->log_cap);  // the counter runs on past a full log
  return HD_OK;
}

int hd_shake_log_size(hd_ctx* c, uint64_t* steps, uint64_t* rows_per_step) {
  if (!c || !steps) return HD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  int r = log_steps(c, steps);
  if (r) return r;
  if (rows_per_step) *rows_per_step = c->h_sh_cij.size();
  return HD_OK;
}

int hd_shake_log_read(hd_ctx* c, uint64_t first_step, uint64_t n_steps, float* x, float* t) {
  if (!c || (n_steps && (!x || !t))) return HD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  uint64_t have = 0;
  int r = log_steps(c, &have);
  if (r) return r;
  if (first_step + n_steps > have)
    return fail(c, HD_ERR_ARG, "log holds %llu steps, asked for %llu..%llu",
                (unsigned long long)have, (unsigned long long)first_step,
                (unsigned long long)(first_step + n_steps));
  size_t nc = c->h_sh_cij.size();
  if (n_steps) {
    HIPCHK(c, hipMemcpyAsync(x, c->d_log_x + first_step * nc * HD_SHAKE_NFEAT,
                             n_steps * nc * HD_SHAKE_NFEAT * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(t, c->d_log_t + first_step * nc, n_steps * nc * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return HD_OK;
}

int hd_ctx_shake_log_view(hd_ctx* c, int* device, const float** x, const float** t,
                          uint64_t* rows) {
  if (!c || !c->d_log_ctr) return fail(c, HD_ERR_STATE, "no SHAKE log (hd_shake_log_start)");
  HIPCHK(c, hipSetDevice(c->device));
  uint64_t steps = 0;
  int r = log_steps(c, &steps);
  if (r) return r;
  *device = c->device;
  *x = c->d_log_x;
  *t = c->d_log_t;
  *rows = steps * c->h_sh_cij.size();
  return HD_OK;
}

int hd_set_integrator(hd_ctx* c, double dt, int lossless) {
  if (!c || !(dt > 0)) return HD_ERR_ARG;
  if (c->ready) {
    // a phase change of the time step keeps the state; the mode is fixed
    if ((lossless != 0) != c->lossless)
      return fail(c, HD_ERR_STATE, "lossless mode is fixed at hd_setup");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dt = dt;
    drop_graphs(c);
    return HD_OK;
  }
  if (lossless && c->have_lan) return fail(c, HD_ERR_UNSUPPORTED, "Langevin BAOAB has no lossless mode");
  c->dt = dt;
  c->lossless = lossless != 0;
  return HD_OK;
}

static int upload_nhc_state(hd_ctx* c) {
  int m = c->nhc_chain;
  std::vector<double> st(3 * m, 0.0);
  long long dof = std::max(3LL * (long long)c->n - (long long)c->nhc_nc - 3LL, 0LL);
  double t2 = c->nhc_tau * c->nhc_tau;
  st[2 * m] = (double)dof * c->nhc_kt * t2;
  for (int j = 1; j < m; ++j) st[2 * m + j] = c->nhc_kt * t2;
  HIPCHK(c, hipMemcpy(c->nhc_state, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice));
  return HD_OK;
}

// Yoshida-Suzuki weights of the chain sub-steps (nose_hoover_chain.rs:60-94)
static const double* yoshida_w(int order) {
  static const double w1[1] = {1.0};
  static const double w3[3] = {1.3512071919596577, -1.7024143839193155, 1.3512071919596577};
  static const double w5[5] = {0.41449077179437574, 0.41449077179437574, -0.6579630871775030,
                               0.41449077179437574, 0.41449077179437574};
  static const double w7[7] = {0.7845136104775573, 0.2355732133593582, -1.1776799841788710,
                               1.3151863206839023, -1.1776799841788710, 0.2355732133593582,
                               0.7845136104775573};
  return order == 1 ? w1 : order == 3 ? w3 : order == 5 ? w5 : w7;
}

// One Nose-Hoover chain sub-step of length dt on the host, the
// Martyna-Tuckerman-Klein chain (Mol. Phys. 87, 1117 (1996)) in the
// reference's form (nose_hoover_chain.rs:109-172): quarter-step kicks of
// the chain momenta from the top of the chain down, each scaled by the
// next link's exp(-dt/8 p/Q) before and after; the thermalised degrees of
// freedom scaled by exp(-dt/2 p_1/Q_1); the chain positions moved by
// dt/2 p/Q; the kicks again from the bottom up.  k2 = the thermalised
// kinetic energy doubled (2K for the particles, p_eps^2/W for the cell);
// returns the scale factor.
static double chain_sub_step(std::vector<double>& xi, std::vector<double>& p,
                             const std::vector<double>& q, double dt, double k2, double g,
                             double kt) {
  int m = (int)xi.size();
  if (m == 0) return 1.0;
  double k = k2;
  for (int j = m - 1; j >= 0; --j) {
    double sc = j == m - 1 ? 1.0 : exp(-dt / 8.0 * p[j + 1] / q[j + 1]);
    p[j] *= sc;
    double gj = j == 0 ? k - g * kt : p[j - 1] * p[j - 1] / q[j - 1] - kt;
    p[j] += dt / 4.0 * gj;
    p[j] *= sc;
  }
  double factor = exp(-dt / 2.0 * p[0] / q[0]);
  k *= factor * factor;
  for (int j = 0; j < m; ++j) xi[j] += dt / 2.0 * p[j] / q[j];
  for (int j = 0; j < m; ++j) {
    double sc = j == m - 1 ? 1.0 : exp(-dt / 8.0 * p[j + 1] / q[j + 1]);
    p[j] *= sc;
    double gj = j == 0 ? k - g * kt : p[j - 1] * p[j - 1] / q[j - 1] - kt;
    p[j] += dt / 4.0 * gj;
    p[j] *= sc;
  }
  return factor;
}

int hd_set_mtk(hd_ctx* c, double kt, double pressure, double tau_t, double tau_p, int chain,
               int yoshida, int n_resp, uint64_t n_constraints) {
  if (!c) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_mtk after hd_setup");
  if (!(kt > 0) || !std::isfinite(kt) || !std::isfinite(pressure) || !(tau_t > 0) ||
      !std::isfinite(tau_t) || !(tau_p > 0) || !std::isfinite(tau_p) || chain < 1 ||
      chain > 16 || n_resp < 1 || !(yoshida == 1 || yoshida == 3 || yoshida == 5 || yoshida == 7))
    return fail(c, HD_ERR_ARG, "invalid MTK-NPT parameters");
  if (c->have_nhc || c->rescale || c->have_andersen || c->have_lan || c->have_baro)
    return fail(c, HD_ERR_STATE,
                "mtk-npt owns the thermostat and the barostat: no other one beside it");
  c->have_mtk = true;
  c->mtk_kt = kt;
  c->mtk_p = pressure;
  c->mtk_tau_t = tau_t;
  c->mtk_tau_p = tau_p;
  c->mtk_chain = chain;
  c->mtk_yoshida = yoshida;
  c->mtk_nresp = n_resp;
  c->mtk_nc = n_constraints;
  return HD_OK;
}

// masses and a fresh state (mtk_npt.rs:124-188); at hd_setup, once n is known
static void mtk_init(hd_ctx* c) {
  int m = c->mtk_chain;
  long long g = (long long)(3 * c->n) - (long long)c->mtk_nc - 3;
  c->mtk_g = (double)std::max(g, 1LL);
  double kt = c->mtk_kt, t2 = c->mtk_tau_t * c->mtk_tau_t, p2 = c->mtk_tau_p * c->mtk_tau_p;
  c->mtk_qp.assign(m, kt * t2);
  c->mtk_qp[0] = c->mtk_g * kt * t2;
  c->mtk_qc.assign(m, kt * t2);
  c->mtk_w = (c->mtk_g + 3.0) * kt * p2;
  c->mtk_xp.assign(m, 0.0);
  c->mtk_pp.assign(m, 0.0);
  c->mtk_xc.assign(m, 0.0);
  c->mtk_pc.assign(m, 0.0);
  c->mtk_peps = c->mtk_eps = c->mtk_last_p = c->mtk_last_v = 0.0;
}

int hd_set_nhc(hd_ctx* c, double kt, double tau, int chain, int yoshida, int n_resp,
               uint64_t n_constraints) {
  if (!c) return HD_ERR_ARG;
  if (c->ready && !c->have_nhc)
    return fail(c, HD_ERR_STATE, "a thermostat can only be re-parameterised after hd_setup");
  if (kt <= 0) {
    if (c->ready) return fail(c, HD_ERR_STATE, "cannot remove the thermostat after hd_setup");
    c->have_nhc = false;
    return HD_OK;
  }
  if (c->have_lan || c->rescale || c->have_andersen || c->have_mtk)
    return fail(c, HD_ERR_STATE, "one thermostat per context: no Nose-Hoover chain beside it");
  if (!(tau > 0) || chain < 1 || chain > 16 || n_resp < 1 ||
      !(yoshida == 1 || yoshida == 3 || yoshida == 5 || yoshida == 7))
    return fail(c, HD_ERR_ARG, "invalid Nose-Hoover chain parameters");
  if (c->ready) {
    // new phase: fresh chain (xi = p_xi = 0) with the new masses
    if (chain != c->nhc_chain) return fail(c, HD_ERR_STATE, "chain length is fixed at hd_setup");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nhc_kt = kt;
    c->nhc_tau = tau;
    c->nhc_yoshida = yoshida;
    c->nhc_resp = n_resp;
    c->nhc_nc = n_constraints;
    drop_graphs(c);
    return upload_nhc_state(c);
  }
  c->have_nhc = true;
  c->nhc_kt = kt;
  c->nhc_tau = tau;
  c->nhc_chain = chain;
  c->nhc_yoshida = yoshida;
  c->nhc_resp = n_resp;
  c->nhc_nc = n_constraints;
  return HD_OK;
}

static int set_rescale(hd_ctx* c, int kind, double kt, double tau, uint64_t seed,
                       uint64_t n_constraints) {
  if (!c) return HD_ERR_ARG;
  if (!(kt > 0) || !(tau > 0) || !std::isfinite(kt) || !std::isfinite(tau))
    return fail(c, HD_ERR_ARG, "thermostat temperature and tau must be finite and > 0");
  if (c->have_nhc || c->have_lan || c->have_andersen || c->have_mtk)
    return fail(c, HD_ERR_STATE, "one thermostat per context: a Nose-Hoover chain, Andersen, "
                                 "Langevin BAOAB or MTK-NPT is already set");
  if (c->ready) {
    // a new phase re-parameterises the same kind; the counter and the
    // injection account carry on
    if (c->rescale != kind)
      return fail(c, HD_ERR_STATE, "the thermostat kind is fixed at hd_setup");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  c->rescale = kind;
  c->rs_kt = kt;
  c->rs_tau = tau;
  c->rs_seed = seed;
  c->rs_nc = n_constraints;
  return HD_OK;
}

int hd_set_andersen(hd_ctx* c, double kt, double collision_rate, uint64_t seed) {
  if (!c) return HD_ERR_ARG;
  if (!(kt > 0) || !(collision_rate >= 0) || !std::isfinite(kt) || !std::isfinite(collision_rate))
    return fail(c, HD_ERR_ARG, "Andersen temperature must be > 0 and the collision rate >= 0");
  if (c->have_nhc || c->have_lan || c->rescale || c->have_mtk)
    return fail(c, HD_ERR_STATE, "one thermostat per context");
  if (c->ready) {
    if (!c->have_andersen) return fail(c, HD_ERR_STATE, "the thermostat kind is fixed at hd_setup");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  c->have_andersen = true;
  c->and_kt = kt;
  c->and_rate = collision_rate;
  c->and_seed = seed;
  return HD_OK;
}

int hd_set_csvr(hd_ctx* c, double kt, double tau, uint64_t seed, uint64_t n_constraints) {
  return set_rescale(c, 2, kt, tau, seed, n_constraints);
}

int hd_set_berendsen(hd_ctx* c, double kt, double tau, uint64_t n_constraints) {
  return set_rescale(c, 3, kt, tau, 0, n_constraints);
}

int hd_get_thermostat(hd_ctx* c, double out[2]) {
  if (!c || !c->ready || !out) return fail(c, HD_ERR_STATE, "not set up");
  out[0] = out[1] = 0.0;
  if (c->have_andersen) {
    unsigned long long ctr = 0;
    HIPCHK(c, hipMemcpyAsync(&ctr, c->and_counter, sizeof(ctr), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[1] = (double)ctr;
    return HD_OK;
  }
  if (!c->rescale) return HD_OK;
  unsigned long long ctr = 0;
  HIPCHK(c, hipMemcpyAsync(out, c->rs_injection, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(&ctr, c->rs_counter, sizeof(ctr), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[1] = (double)ctr;
  return HD_OK;
}

int hd_set_langevin(hd_ctx* c, double friction, double kt, uint64_t seed) {
  if (!c) return HD_ERR_ARG;
  if (!(friction > 0) || !(kt > 0) || !std::isfinite(friction) || !std::isfinite(kt))
    return fail(c, HD_ERR_ARG, "Langevin friction and temperature must be finite and > 0");
  if (c->have_nhc || c->rescale || c->have_andersen || c->have_mtk)
    return fail(c, HD_ERR_STATE, "Langevin BAOAB owns the thermostat: remove the other thermostat");
  if (c->lossless) return fail(c, HD_ERR_UNSUPPORTED, "Langevin BAOAB has no lossless mode");
  if (c->ready) {
    if (!c->have_lan)
      return fail(c, HD_ERR_STATE, "the integrator kind is fixed at hd_setup");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  c->have_lan = true;
  c->lan_friction = friction;
  c->lan_kt = kt;
  c->lan_seed = seed;
  return HD_OK;
}

int hd_get_langevin_counter(hd_ctx* c, uint64_t* counter) {
  if (!c || !c->ready || !counter) return fail(c, HD_ERR_STATE, "not set up");
  unsigned long long v = 0;
  HIPCHK(c, hipMemcpyAsync(&v, c->lan_counter, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *counter = v;
  return HD_OK;
}

int hd_set_crescale(hd_ctx* c, double pressure, double tau, double compressibility, double kt,
                    uint64_t seed) {
  if (!c) return HD_ERR_ARG;
  if (c->lossless && tau > 0)
    return fail(c, HD_ERR_UNSUPPORTED, "lossless mode does not support a barostat");
  if (c->ready) {
    // a new phase may switch the barostat on or off or change it
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  if (tau <= 0) {
    c->have_baro = false;
    return HD_OK;
  }
  if (!(compressibility > 0) || !(kt >= 0)) return fail(c, HD_ERR_ARG, "invalid C-rescale parameters");
  if (c->have_mtk) return fail(c, HD_ERR_STATE, "mtk-npt owns the barostat");
  c->have_baro = true;
  c->baro_stochastic = true;
  c->baro_p = pressure;
  c->baro_tau = tau;
  c->baro_kappa = compressibility;
  c->baro_kt = kt;
  c->baro_seed = seed;
  return HD_OK;
}

int hd_set_berendsen_barostat(hd_ctx* c, double pressure, double tau, double compressibility) {
  if (!c) return HD_ERR_ARG;
  if (c->lossless && tau > 0)
    return fail(c, HD_ERR_UNSUPPORTED, "lossless mode does not support a barostat");
  if (c->ready) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  if (tau <= 0) {
    c->have_baro = false;
    return HD_OK;
  }
  if (!(compressibility > 0) || !std::isfinite(pressure))
    return fail(c, HD_ERR_ARG, "invalid Berendsen barostat parameters");
  if (c->have_mtk) return fail(c, HD_ERR_STATE, "mtk-npt owns the barostat");
  c->have_baro = true;
  c->baro_stochastic = false;
  c->baro_p = pressure;
  c->baro_tau = tau;
  c->baro_kappa = compressibility;
  c->baro_kt = 0.0;
  c->baro_seed = 0;
  return HD_OK;
}

static int build_list(hd_ctx* c) {
  // force a rebuild and run the neighbour pipeline once
  unsigned force = 0x100u;
  HIPCHK(c, hipMemcpyAsync(c->ctrl, &force, sizeof(unsigned), hipMemcpyHostToDevice, c->stream));
  unsigned zero = 0;
  HIPCHK(c, hipMemcpyAsync(c->ctrl + 2, &zero, sizeof(unsigned), hipMemcpyHostToDevice, c->stream));
  HdNeighborArgs na = nb_args(c);
  hd_launch_neighbor_update(c->stream, na);
  hd_launch_tile_gather_rebuild(c->stream, na, c->fpj, c->fpj_ew);
  hd_launch_neighbor_build(c->stream, na);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

static int alloc_entries(hd_ctx* c, unsigned cap) {
  if (c->entry_atoms) dfree(c, c->entry_atoms);
  c->entry_atoms = nullptr;
  c->cap = cap;
  DALLOC(c, c->entry_atoms, (size_t)c->nblocks * cap * 64);
  return HD_OK;
}

// Probe the list size, then size the per-block regions with headroom.
static int size_list(hd_ctx* c) {
  int r;
  if ((r = alloc_entries(c, 64))) return r;
  if ((r = build_list(c))) return r;
  unsigned ctrl[4];
  HIPCHK(c, hipMemcpy(ctrl, c->ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
  unsigned need = ctrl[2];
  unsigned cap = std::max(4u, (unsigned)ceil(need * 1.5) + 2);
  if (cap != c->cap) {
    if ((r = alloc_entries(c, cap))) return r;
    if ((r = build_list(c))) return r;
  }
  HIPCHK(c, hipMemcpy(ctrl, c->ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
  if (ctrl[0] & 8u) return fail(c, HD_ERR_OVERFLOW, "neighbour list overflow while sizing");
  unsigned zero = 0;
  HIPCHK(c, hipMemcpy(c->ctrl, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
  return HD_OK;
}

static void drop_graphs(hd_ctx* c);

int hd_set_pair_list(hd_ctx* c, int mode) {
  if (!c || (mode != 0 && mode != 1)) return HD_ERR_ARG;
  if (!c->ready) {
    c->atom_lists = mode;
    return HD_OK;
  }
  if (mode == c->atom_lists) return HD_OK;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->atom_lists = mode;
  // the rebuild count is kept across the switch
  unsigned ctrl[4];
  HIPCHK(c, hipMemcpy(ctrl, c->ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
  int r = size_list(c);
  if (r) return r;
  HIPCHK(c, hipMemcpy(c->ctrl + 1, &ctrl[1], sizeof(unsigned), hipMemcpyHostToDevice));
  c->forces_valid = false;
  return HD_OK;
}

static void drop_graphs(hd_ctx* c) {
  ++c->graph_gen;
  // an executable graph is destroyed only once no launch of it can still
  // be running on either stream of the context
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->side);
  if (c->graph_f) (void)hipGraphExecDestroy(c->graph_f);
  if (c->graph_fev) (void)hipGraphExecDestroy(c->graph_fev);
  if (c->graph_fk) (void)hipGraphExecDestroy(c->graph_fk);
  c->graph_f = c->graph_fev = c->graph_fk = nullptr;
}

int hd_setup(hd_ctx* c) {
  if (!c) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_setup called twice");
  if (!c->n) return fail(c, HD_ERR_STATE, "hd_setup before hd_set_system");
  if (!c->ntypes) return fail(c, HD_ERR_STATE, "hd_setup before hd_set_lj");
  if (!(c->lattice[0] > 0)) return fail(c, HD_ERR_STATE, "hd_setup before hd_set_lattice");
  for (uint64_t i = 0; i < c->n; ++i)
    if (c->h_type[i] >= (uint32_t)c->ntypes) return fail(c, HD_ERR_ARG, "type[%llu] out of range",
                                                         (unsigned long long)i);
  // tilted cells as the reference's lower-triangular lattice (pbc.cuh):
  // each tilt at most half the box length it tilts along, so fractional
  // rounding finds the nearest image of any separation below half a width
  if (fabs(c->lattice[3]) > 0.5 * c->lattice[0] || fabs(c->lattice[4]) > 0.5 * c->lattice[0] ||
      fabs(c->lattice[5]) > 0.5 * c->lattice[1])
    return fail(c, HD_ERR_UNSUPPORTED, "box tilts must be at most half the box lengths "
                                       "(|xy|, |xz| <= lx/2, |yz| <= ly/2)");
  // the compensated mode carries FP64 residuals of x and v that constraint
  // projections and barostat scaling would leave stale; the reference
  // refuses the combination (src/integrator/velocity_verlet.rs:222-229)
  if (c->lossless && has_constraints(c))
    return fail(c, HD_ERR_UNSUPPORTED, "lossless mode does not support constraints (SETTLE/SHAKE)");
  if (c->have_mtk && (has_constraints(c) || c->lossless))
    return fail(c, HD_ERR_UNSUPPORTED,
                "mtk-npt does not support constraints (SETTLE/SHAKE) or lossless mode");
  if (c->have_mtk) mtk_init(c);
  if (c->lossless && c->have_baro)
    return fail(c, HD_ERR_UNSUPPORTED, "lossless mode does not support a barostat");
  HIPCHK(c, hipSetDevice(c->device));
  c->rmax = c->have_spme ? c->coul_rc : 0.0;
  for (double v : c->lj_cut) c->rmax = std::max(c->rmax, v);
  if (c->r_skin < 0) c->r_skin = 0.3 * c->rmax;
  double rs = c->rmax + c->r_skin;
  for (int d = 0; d < 3; ++d)
    if (rs > 0.5 / inv_row_norm(c->lattice, d))
      return fail(c, HD_ERR_UNSUPPORTED,
                  "r_cut + r_skin = %g exceeds half the box width (%g): minimum image breaks", rs,
                  0.5 / inv_row_norm(c->lattice, d));
  unsigned n = (unsigned)c->n;
  c->nblocks = (n + 63) / 64;
  c->npad = c->nblocks * 64;
  // voxel grid in fractional coordinates (so the order survives box
  // rescaling): a power of two per axis, about two atoms per voxel, ranked
  // along a Hilbert curve over the enclosing 2^bits cube
  double vol = c->lattice[0] * c->lattice[1] * c->lattice[2];
  double edge = cbrt(2.0 * vol / (double)n);
  int bits = 1;
  for (int d = 0; d < 3; ++d) {
    double want = std::max(1.0, c->lattice[d] / edge);
    int k = (int)std::lround(std::log2(want));
    k = std::max(0, std::min(k, 10));
    c->nv[d] = 1u << k;
    bits = std::max(bits, std::max(k, 1));
  }
  c->nvox = c->nv[0] * c->nv[1] * c->nv[2];
  std::vector<std::pair<uint64_t, uint32_t>> codes(c->nvox);
  for (unsigned a = 0; a < c->nv[0]; ++a)
    for (unsigned b = 0; b < c->nv[1]; ++b)
      for (unsigned cc = 0; cc < c->nv[2]; ++cc) {
        unsigned lin = (a * c->nv[1] + b) * c->nv[2] + cc;
        codes[lin] = {hilbert3(a, b, cc, bits), lin};
      }
  std::sort(codes.begin(), codes.end());
  std::vector<unsigned> rank(c->nvox);
  for (unsigned r = 0; r < c->nvox; ++r) rank[codes[r].second] = r;

  DALLOC(c, c->posq, n);
  DALLOC(c, c->vx, n); DALLOC(c, c->vy, n); DALLOC(c, c->vz, n);
  DALLOC(c, c->fx, n); DALLOC(c, c->fy, n); DALLOC(c, c->fz, n);
  DALLOC(c, c->mass, n); DALLOC(c, c->pe, n); DALLOC(c, c->vir, n); DALLOC(c, c->w_cons, n);
  DALLOC(c, c->ix, n); DALLOC(c, c->iy, n); DALLOC(c, c->iz, n);
  DALLOC(c, c->type, n); DALLOC(c, c->key, n);
  DALLOC(c, c->snap, n); DALLOC(c, c->ref_pos, n);
  if (c->lossless) {
    DALLOC(c, c->xlo, n); DALLOC(c, c->ylo, n); DALLOC(c, c->zlo, n);
    DALLOC(c, c->vxlo, n); DALLOC(c, c->vylo, n); DALLOC(c, c->vzlo, n);
  }
  DALLOC(c, c->d_lattice, 6);
  if (c->ntypes > 4) {
    std::vector<float> tab(8 * (size_t)c->ntypes * c->ntypes);
    hd_pair_lj_table(c->ntypes, c->lj_sigma.data(), c->lj_eps.data(), c->lj_cut.data(),
                     c->lj_rsw.data(), tab.data());
    DALLOC(c, c->d_lj_tab, tab.size() / 4);
    int r = upload(c, c->d_lj_tab, reinterpret_cast<const float4*>(tab.data()), tab.size() / 4);
    if (r) return r;
  }
  DALLOC(c, c->ref_lattice, 6);
  DALLOC(c, c->baro_mu, 1);
  DALLOC(c, c->baro_diag, 4);
  DALLOC(c, c->baro_counter, 1);
  DALLOC(c, c->lan_counter, 2);  // draw counter, fused-kernel block ticket
  DALLOC(c, c->voxel_rank, c->nvox); DALLOC(c, c->atom_key, n);
  DALLOC(c, c->voxel_count, c->nvox); DALLOC(c, c->voxel_off, c->nvox + 1);
  DALLOC(c, c->voxel_cursor, c->nvox);
  DALLOC(c, c->voxel_part, hd_voxel_scan_parts(c->nvox) + 1);
  DALLOC(c, c->sorted_id, c->npad); DALLOC(c, c->slot_of, n);
  DALLOC(c, c->tile_type, c->npad); DALLOC(c, c->tile_key, c->npad);
  DALLOC(c, c->tile_posq, c->npad);
  DALLOC(c, c->center, c->nblocks); DALLOC(c, c->half, c->nblocks);
  DALLOC(c, c->entry_count, c->nblocks);
  DALLOC(c, c->block_mask, (size_t)c->nblocks * hd_block_pair_chunks(c->nblocks));
  DALLOC(c, c->block_order, c->nblocks);
  {
    // identity until the first rebuild orders the blocks
    std::vector<unsigned> ident(c->nblocks);
    for (unsigned b = 0; b < c->nblocks; ++b) ident[b] = b;
    if (hipMemcpy(c->block_order, ident.data(), ident.size() * sizeof(unsigned),
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(c, HD_ERR_HIP, "block order upload failed");
  }
  DALLOC(c, c->ctrl, 4);
  DALLOC(c, c->fpi, (size_t)c->npad * 4); DALLOC(c, c->fpj, (size_t)c->npad * 4);
  DALLOC(c, c->fpi_ew, (size_t)c->npad * 2); DALLOC(c, c->fpj_ew, (size_t)c->npad * 2);
  DALLOC(c, c->red_partials, 2 * HD_REDUCE_BLOCKS);
  DALLOC(c, c->fin_partials, 2 * (size_t)hd_finalize_blocks((unsigned)c->n, true));
  DALLOC(c, c->scal, 8);
  DALLOC(c, c->factors, 2);
  DALLOC(c, c->recip, 2);
  DALLOC(c, c->w_share, 1);
  int r;
  if ((r = upload(c, c->posq, (const float4*)c->h_posq.data(), n))) return r;
  if ((r = upload(c, c->vx, c->h_vx.data(), n))) return r;
  if ((r = upload(c, c->vy, c->h_vy.data(), n))) return r;
  if ((r = upload(c, c->vz, c->h_vz.data(), n))) return r;
  if ((r = upload(c, c->mass, c->h_mass.data(), n))) return r;
  if ((r = upload(c, c->ix, c->h_ix.data(), n))) return r;
  if ((r = upload(c, c->iy, c->h_iy.data(), n))) return r;
  if ((r = upload(c, c->iz, c->h_iz.data(), n))) return r;
  if ((r = upload(c, c->type, c->h_type.data(), n))) return r;
  if ((r = upload(c, c->key, c->h_key.data(), n))) return r;
  if ((r = upload(c, c->voxel_rank, rank.data(), c->nvox))) return r;
  {
    float lf[6];
    for (int i = 0; i < 6; ++i) lf[i] = (float)c->lattice[i];
    if ((r = upload(c, c->d_lattice, lf, 6))) return r;
    float ones[2] = {1.f, 1.f};
    if ((r = upload(c, c->factors, ones, 2))) return r;
  }
  if (!c->h_groups.empty()) {
    DALLOC(c, c->d_groups, c->h_groups.size());
    if ((r = upload(c, c->d_groups, c->h_groups.data(), c->h_groups.size()))) return r;
    if (c->settle_kinds.size() > 1) {
      DALLOC(c, c->d_gkind, c->h_gkind.size());
      if ((r = upload(c, c->d_gkind, c->h_gkind.data(), c->h_gkind.size()))) return r;
    }
  }
  if (!c->h_ang_k.empty()) {
    DALLOC(c, c->d_angles, c->h_angles.size());
    DALLOC(c, c->d_ang_k, c->h_ang_k.size());
    DALLOC(c, c->d_ang_t0, c->h_ang_t0.size());
    if ((r = upload(c, c->d_angles, c->h_angles.data(), c->h_angles.size()))) return r;
    if ((r = upload(c, c->d_ang_k, c->h_ang_k.data(), c->h_ang_k.size()))) return r;
    if ((r = upload(c, c->d_ang_t0, c->h_ang_t0.data(), c->h_ang_t0.size()))) return r;
  }
  if (!c->h_bond_kind.empty()) {
    DALLOC(c, c->d_bonds, c->h_bonds.size());
    DALLOC(c, c->d_bond_kind, c->h_bond_kind.size());
    DALLOC(c, c->d_bond_prm, c->h_bond_prm.size());
    if ((r = upload(c, c->d_bonds, c->h_bonds.data(), c->h_bonds.size()))) return r;
    if ((r = upload(c, c->d_bond_kind, c->h_bond_kind.data(), c->h_bond_kind.size()))) return r;
    if ((r = upload(c, c->d_bond_prm, c->h_bond_prm.data(), c->h_bond_prm.size()))) return r;
  }
  if (c->have_andersen) {
    if (c->have_lan) return fail(c, HD_ERR_STATE, "Langevin BAOAB owns the thermostat");
    DALLOC(c, c->and_counter, 1);
  }
  if (c->rescale) {
    if (c->have_lan) return fail(c, HD_ERR_STATE, "Langevin BAOAB owns the thermostat");
    DALLOC(c, c->rs_counter, 1);
    DALLOC(c, c->rs_injection, 1);
    DALLOC(c, c->rs_partials, HD_CSVR_BLOCKS);
  }
  if (!c->h_dih_prm.empty()) {
    DALLOC(c, c->d_dih, c->h_dih.size());
    DALLOC(c, c->d_dih_prm, c->h_dih_prm.size());
    if ((r = upload(c, c->d_dih, c->h_dih.data(), c->h_dih.size()))) return r;
    if ((r = upload(c, c->d_dih_prm, c->h_dih_prm.data(), c->h_dih_prm.size()))) return r;
  }
  if (!c->h_excl.empty()) {
    int nt = c->ntypes;
    if (c->lj_sigma.empty()) return fail(c, HD_ERR_STATE, "exclusions need hd_set_lj");
    for (auto& e : c->h_excl) {
      size_t k = (size_t)c->h_type[e.i] * nt + c->h_type[e.j];
      e.sigma = c->lj_sigma[k];
      e.eps = c->lj_eps[k];
      e.cut2 = c->lj_cut[k] * c->lj_cut[k];
      e.rsw2 = c->lj_rsw[k] * c->lj_rsw[k];
    }
    DALLOC(c, c->d_excl, c->h_excl.size());
    if ((r = upload(c, c->d_excl, c->h_excl.data(), c->h_excl.size()))) return r;
  }
  if (has_shake(c)) {
    DALLOC(c, c->d_sh_goff, c->h_sh_goff.size());
    DALLOC(c, c->d_sh_atoms, c->h_sh_atoms.size());
    DALLOC(c, c->d_sh_coff, c->h_sh_coff.size());
    DALLOC(c, c->d_sh_cij, c->h_sh_cij.size());
    DALLOC(c, c->d_sh_r2, c->h_sh_r2.size());
    DALLOC(c, c->d_sh_stats, 6);
    {
      std::vector<uint32_t> cg(c->h_sh_cij.size());
      for (size_t g = 0; g + 1 < c->h_sh_coff.size(); ++g)
        for (uint32_t k = c->h_sh_coff[g]; k < c->h_sh_coff[g + 1]; ++k) cg[k] = (uint32_t)g;
      DALLOC(c, c->d_sh_cgrp, cg.size());
      DALLOC(c, c->d_sh_nu, cg.size());
      if (!cg.empty() && (r = upload(c, c->d_sh_cgrp, cg.data(), cg.size()))) return r;
    }
    DALLOC(c, c->d_sh_color, c->h_sh_color.size());
    if (!c->h_sh_color.empty() &&
        (r = upload(c, c->d_sh_color, c->h_sh_color.data(), c->h_sh_color.size())))
      return r;
    {
      std::vector<uint32_t> small, big;
      for (size_t g = 0; g + 1 < c->h_sh_goff.size(); ++g) {
        uint32_t na = c->h_sh_goff[g + 1] - c->h_sh_goff[g];
        uint32_t nc = c->h_sh_coff[g + 1] - c->h_sh_coff[g];
        (na <= 4 && nc <= 6 ? small : big).push_back((uint32_t)g);
      }
      if (!small.empty()) {
        DALLOC(c, c->d_sh_small, small.size());
        DALLOC(c, c->d_sh_big, big.size());
        if ((r = upload(c, c->d_sh_small, small.data(), small.size()))) return r;
        if (!big.empty() && (r = upload(c, c->d_sh_big, big.data(), big.size()))) return r;
        c->sh_n_small = (unsigned)small.size();
        c->sh_n_big = (unsigned)big.size();
      }
    }
    if ((r = upload(c, c->d_sh_goff, c->h_sh_goff.data(), c->h_sh_goff.size()))) return r;
    if ((r = upload(c, c->d_sh_atoms, c->h_sh_atoms.data(), c->h_sh_atoms.size()))) return r;
    if ((r = upload(c, c->d_sh_coff, c->h_sh_coff.data(), c->h_sh_coff.size()))) return r;
    if (!c->h_sh_cij.empty()) {
      if ((r = upload(c, c->d_sh_cij, c->h_sh_cij.data(), c->h_sh_cij.size()))) return r;
      if ((r = upload(c, c->d_sh_r2, c->h_sh_r2.data(), c->h_sh_r2.size()))) return r;
    }
  }
  if (c->have_spme) {
    size_t M = (size_t)c->grid[0] * c->grid[1] * c->grid[2];
    size_t Mc = (size_t)c->grid[0] * c->grid[1] * (c->grid[2] / 2 + 1);
    DALLOC(c, c->rho_fixed, M);
    DALLOC(c, c->cgrid, Mc);
    DALLOC(c, c->vgrid, M);
    DALLOC(c, c->spl, c->npad);
    DALLOC(c, c->sg, c->npad);
    auto ta = twiddles(c->grid[0]), tb = twiddles(c->grid[1]), tc = twiddles(c->grid[2]);
    DALLOC(c, c->tw_a, ta.size()); DALLOC(c, c->tw_b, tb.size()); DALLOC(c, c->tw_c, tc.size());
    if ((r = upload(c, c->tw_a, ta.data(), ta.size()))) return r;
    if ((r = upload(c, c->tw_b, tb.data(), tb.size()))) return r;
    if ((r = upload(c, c->tw_c, tc.data(), tc.size()))) return r;
    int p = c->spline_order;
    auto ba = b_factors(c->grid[0], p), bb = b_factors(c->grid[1], p), bc = b_factors(c->grid[2], p);
    DALLOC(c, c->bf_a, ba.size()); DALLOC(c, c->bf_b, bb.size()); DALLOC(c, c->bf_c, bc.size());
    if ((r = upload(c, c->bf_a, ba.data(), ba.size()))) return r;
    if ((r = upload(c, c->bf_b, bb.data(), bb.size()))) return r;
    if ((r = upload(c, c->bf_c, bc.data(), bc.size()))) return r;
    DALLOC(c, c->recip_partials, std::max(hd_spme_partials_count(c->grid[1], c->grid[2]),
                                          hd_dft_partials_count(c->grid[1], c->grid[2])));
    if (c->dft_mode == 1 && (r = build_dft_mats(c))) return r;
  }
  if (c->have_nhc) {
    int m = c->nhc_chain;
    std::vector<double> st(3 * m, 0.0);
    long long dof = std::max(3LL * (long long)c->n - (long long)c->nhc_nc - 3LL, 0LL);
    double t2 = c->nhc_tau * c->nhc_tau;
    st[2 * m] = (double)dof * c->nhc_kt * t2;
    for (int j = 1; j < m; ++j) st[2 * m + j] = c->nhc_kt * t2;
    DALLOC(c, c->nhc_state, 3 * m);
    if ((r = upload(c, c->nhc_state, st.data(), st.size()))) return r;
  }
  if ((r = size_list(c))) return r;
  c->ready = true;
  return HD_OK;
}

static int check_status(hd_ctx* c) {
  unsigned ctrl[4];
  HIPCHK(c, hipMemcpyAsync(ctrl, c->ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ctrl[0] & 8u) {
    return fail(c, HD_ERR_OVERFLOW,
                "neighbour list overflow (a block needed %u entries, capacity %u)", ctrl[2],
                c->cap);
  }
  if (ctrl[0] & 2u) {
    // high water: grow the entry regions and rebuild before the next step
    unsigned cap = (unsigned)ceil(ctrl[2] * 1.5) + 2;
    drop_graphs(c);
    int r = alloc_entries(c, std::max(cap, c->cap));
    if (r) return r;
    unsigned zero = 0;
    HIPCHK(c, hipMemcpy(c->ctrl, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
    if ((r = build_list(c))) return r;
    HIPCHK(c, hipMemcpy(c->ctrl, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
  }
  return HD_OK;
}

int hd_force_eval(hd_ctx* c, int want_scalars) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "hd_force_eval before hd_setup");
  HIPCHK(c, hipSetDevice(c->device));
  enqueue_forces(c, want_scalars != 0, false, false);
  c->forces_valid = true;
  if (want_scalars)
    hd_launch_ke(c->stream, (unsigned)c->n, c->vx, c->vy, c->vz, c->mass, c->red_partials,
                 c->scal + 0, nullptr, 0.f, nullptr);
  HIPCHK(c, hipGetLastError());
  return check_status(c);
}

// Position-only constraint projection of the minimizer
// (settle.cu:511-583, shake.cu:562-679): the solvers with the moved
// positions as their direction frame and no time step, so velocities and
// the constraint virial are left alone.
static void enqueue_project_positions(hd_ctx* c) {
  if (!c->h_groups.empty()) {
    HdSettleArgs sa = settle_args(c);
    sa.dt = 0.f;
    hd_launch_settle_positions(c->stream, sa);
  }
  if (has_shake(c)) {
    HdShakeArgs sa = shake_args(c, 0.f);
    hd_launch_shake_positions(c->stream, sa);
  }
}

// Philox4x32-10 on the host (Salmon et al., SC11): the MC barostat's two
// uniforms come from one block, as the reference's host mirror draws them
// (src/integrator/philox.rs:21-73)
static void philox_host(uint32_t k0, uint32_t k1, uint32_t c[4]) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)c[0] * 0xD2511F53u, p2 = (uint64_t)c[2] * 0xCD9E8D57u;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi2 = (uint32_t)(p2 >> 32), lo2 = (uint32_t)p2;
    uint32_t n0 = hi2 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo2;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

int hd_set_molecules(hd_ctx* c, uint64_t n_mol, const uint32_t* offsets, const uint32_t* atoms) {
  if (!c || !n_mol || !offsets || !atoms) return HD_ERR_ARG;
  if (c->ready) return fail(c, HD_ERR_STATE, "hd_set_molecules after hd_setup");
  if (offsets[0] != 0 || offsets[n_mol] != c->n)
    return fail(c, HD_ERR_ARG, "molecule offsets must run from 0 to n");
  std::vector<char> seen(c->n, 0);
  for (uint64_t m = 0; m < n_mol; ++m) {
    if (offsets[m + 1] <= offsets[m]) return fail(c, HD_ERR_ARG, "molecule %llu is empty",
                                                  (unsigned long long)m);
    for (uint32_t k = offsets[m]; k < offsets[m + 1]; ++k) {
      if (atoms[k] >= c->n || seen[atoms[k]])
        return fail(c, HD_ERR_ARG, "molecule atoms must list every atom once");
      seen[atoms[k]] = 1;
    }
  }
  c->h_mol_off.assign(offsets, offsets + n_mol + 1);
  c->h_mol_atoms.assign(atoms, atoms + c->n);
  return HD_OK;
}

// One Monte-Carlo volume move (src/integrator/mc_barostat.rs:218-330):
// U_old from a force evaluation at the current state; dV = max_dv (2 u1 - 1)
// and u2 from Philox (seed, counter, 0, 0); early rejection when V_new <= 0
// or the scaled box would be narrower than twice the search radius;
// otherwise the box and the molecule centres scale by (V_new/V_old)^(1/3),
// U_new is evaluated and the move kept when w = dU + P dV
// - N_mol kT ln(V_new/V_old) <= 0 or u2 < exp(-w/kT); a rejected move
// restores positions, forces, scalars and the box.  out[0] accepted,
// out[1] V_old, out[2] V after the move, out[3] U_new - U_old (0 when
// rejected early).
int hd_mc_barostat_move(hd_ctx* c, double pressure, double kt, double max_dv, uint64_t seed,
                        uint64_t counter, double out[4]) {
  if (!c || !out || !(kt > 0) || !(max_dv > 0)) return HD_ERR_ARG;
  if (!c->ready) return fail(c, HD_ERR_STATE, "hd_mc_barostat_move before hd_setup");
  if (c->have_baro) return fail(c, HD_ERR_STATE, "a C-rescale or Berendsen barostat is set");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned n = (unsigned)c->n;
  bool mol = !c->h_mol_off.empty();
  unsigned n_mol = mol ? (unsigned)(c->h_mol_off.size() - 1) : n;
  if (!c->mc_pos) {
    DALLOC(c, c->mc_pos, c->n);
    DALLOC(c, c->mc_f, 3 * c->n);
    DALLOC(c, c->mc_scal, 8);
    if (mol) {
      DALLOC(c, c->d_mol_off, c->h_mol_off.size());
      DALLOC(c, c->d_mol_atoms, c->h_mol_atoms.size());
      int r;
      if ((r = upload(c, c->d_mol_off, c->h_mol_off.data(), c->h_mol_off.size()))) return r;
      if ((r = upload(c, c->d_mol_atoms, c->h_mol_atoms.data(), c->h_mol_atoms.size()))) return r;
    }
  }
  // U_old at the current positions
  enqueue_forces(c, 1, false, false);
  c->forces_valid = true;
  int r = check_status(c);
  if (r) return r;
  float lat[6];
  double u_old = 0.0;
  HIPCHK(c, hipMemcpyAsync(lat, c->d_lattice, sizeof(lat), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&u_old, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double v_old = (double)lat[0] * (double)lat[1] * (double)lat[2];
  double latd[6];
  for (int k = 0; k < 6; ++k) latd[k] = lat[k];
  double min_w = 1.0 / std::max(std::max(inv_row_norm(latd, 0), inv_row_norm(latd, 1)),
                                inv_row_norm(latd, 2));
  uint32_t ctr[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), 0u, 0u};
  philox_host((uint32_t)seed, (uint32_t)(seed >> 32), ctr);
  double u1 = ((double)ctr[0] + 0.5) / 4294967296.0, u2 = ((double)ctr[1] + 0.5) / 4294967296.0;
  double dv = max_dv * (2.0 * u1 - 1.0);
  double v_new = v_old + dv;
  out[0] = 0.0;
  out[1] = v_old;
  out[2] = v_old;
  out[3] = 0.0;
  double r_search = c->rmax + c->r_skin;
  if (v_new <= 0.0 || cbrt(v_new / v_old) * min_w < 2.0 * r_search) return HD_OK;
  double scale = cbrt(v_new / v_old);
  // snapshot, then the trial box and centres
  HIPCHK(c, hipMemcpyAsync(c->mc_pos, c->posq, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->mc_f, c->fx, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->mc_f + n, c->fy, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->mc_f + 2 * n, c->fz, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->mc_scal, c->scal, 8 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hd_launch_mc_scale_com(c->stream, c->posq, c->d_mol_off, mol ? c->d_mol_atoms : nullptr,
                         c->mass, c->d_lattice, (float)scale, n_mol);
  float lat_new[6];
  for (int k = 0; k < 6; ++k) lat_new[k] = lat[k] * (float)scale;
  unsigned force = NB_STATUS_FORCE;
  HIPCHK(c, hipMemcpyAsync(c->d_lattice, lat_new, sizeof(lat_new), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ctrl, &force, sizeof(unsigned), hipMemcpyHostToDevice, c->stream));
  enqueue_forces(c, 1, false, false);
  if ((r = check_status(c))) return r;
  double u_new = 0.0;
  HIPCHK(c, hipMemcpyAsync(&u_new, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double v_trial = (double)lat_new[0] * (double)lat_new[1] * (double)lat_new[2];
  double w = (u_new - u_old) + pressure * dv - (double)n_mol * kt * log(v_new / v_old);
  bool accept = w <= 0.0 || u2 < exp(-w / kt);
  out[3] = u_new - u_old;
  if (accept) {
    for (int k = 0; k < 6; ++k) c->lattice[k] = (double)lat_new[k];
    out[0] = 1.0;
    out[2] = v_trial;
    return HD_OK;
  }
  HIPCHK(c, hipMemcpyAsync(c->posq, c->mc_pos, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->fx, c->mc_f, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->fy, c->mc_f + n, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->fz, c->mc_f + 2 * n, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->scal, c->mc_scal, 8 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_lattice, lat, sizeof(lat), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ctrl, &force, sizeof(unsigned), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

static int sd_eval(hd_ctx* c, double* energy, double* fmax) {
  enqueue_forces(c, 1, false, false);
  c->forces_valid = true;
  hd_launch_fmax2(c->stream, c->fx, c->fy, c->fz, (unsigned)c->n, c->sd_fmax2);
  HIPCHK(c, hipGetLastError());
  int r = check_status(c);
  if (r) return r;
  double pe = 0.0, f2 = 0.0;
  HIPCHK(c, hipMemcpyAsync(&pe, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&f2, c->sd_fmax2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *energy = pe;
  *fmax = sqrt(f2);
  return HD_OK;
}

int hd_minimize_begin(hd_ctx* c, double initial_step, double out[4]) {
  if (!c || !out || !(initial_step > 0)) return HD_ERR_ARG;
  if (!c->ready) return fail(c, HD_ERR_STATE, "hd_minimize_begin before hd_setup");
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->sd_save) {
    DALLOC(c, c->sd_save, c->n);
    DALLOC(c, c->sd_fmax2, 1);
  }
  if (has_constraints(c) && !c->snap) return fail(c, HD_ERR_STATE, "no constraint frame");
  double e, f;
  int r = sd_eval(c, &e, &f);
  if (r) return r;
  c->sd_step = initial_step;
  c->sd_energy = e;
  c->sd_begun = true;
  out[0] = e;
  out[1] = f;
  out[2] = initial_step;
  out[3] = 1.0;
  return HD_OK;
}

int hd_minimize_step(hd_ctx* c, double max_step, double step_increase, double step_decrease,
                     double out[4]) {
  if (!c || !out) return HD_ERR_ARG;
  if (!c->ready || !c->sd_begun) return fail(c, HD_ERR_STATE, "hd_minimize_step before hd_minimize_begin");
  HIPCHK(c, hipSetDevice(c->device));
  double prev = c->sd_energy, used = c->sd_step;
  double f2 = 0.0;
  HIPCHK(c, hipMemcpyAsync(&f2, c->sd_fmax2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double fmax = sqrt(f2);
  out[2] = used;
  if (!(fmax > 0.0)) {
    out[0] = prev;
    out[1] = 0.0;
    out[3] = 0.0;
    return HD_OK;
  }
  unsigned n = (unsigned)c->n;
  float4* frame = has_constraints(c) ? c->snap : nullptr;
  hd_launch_sd_move(c->stream, c->posq, c->fx, c->fy, c->fz, c->sd_save, frame,
                    (float)(used / fmax), n);
  if (frame) enqueue_project_positions(c);
  double e, f;
  int r = sd_eval(c, &e, &f);
  if (r) return r;
  if (e < prev) {
    c->sd_energy = e;
    c->sd_step = std::min(c->sd_step * step_increase, max_step);
    out[0] = e;
    out[1] = f;
    out[3] = 1.0;
  } else {
    hd_launch_sd_restore(c->stream, c->posq, c->sd_save, n);
    r = sd_eval(c, &e, &f);
    if (r) return r;
    c->sd_step = c->sd_step * step_decrease;
    out[0] = prev;
    out[1] = f;
    out[3] = 0.0;
  }
  return HD_OK;
}

int hd_set_positions(hd_ctx* c, const float* posq) {
  if (!c || !posq) return HD_ERR_ARG;
  if (!c->ready) return fail(c, HD_ERR_STATE, "hd_set_positions before hd_setup");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpyAsync(c->posq, posq, c->n * sizeof(float4), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // charges travel with the positions and feed the SPME pipeline as given;
  // forces belong to the old positions
  c->forces_valid = false;
  return HD_OK;
}

int hd_neighbor_pre_step(hd_ctx* c, uint32_t* status_out) {
  if (!c || !status_out) return HD_ERR_ARG;
  if (!c->ready) return fail(c, HD_ERR_STATE, "hd_neighbor_pre_step before hd_setup");
  HIPCHK(c, hipSetDevice(c->device));
  HdNeighborArgs na = nb_args(c);
  hd_launch_disp_check(c->stream, na);
  HIPCHK(c, hipGetLastError());
  unsigned ctrl[4];
  HIPCHK(c, hipMemcpyAsync(ctrl, c->ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *status_out = ctrl[0];
  return HD_OK;
}

int hd_constrain_velocities(hd_ctx* c) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "hd_constrain_velocities before hd_setup");
  if (has_shake(c)) {
    // Constraint::apply_initial_velocity_projection (shake.rs): RATTLE, dt = 0
    hd_launch_rattle_velocities(c->stream, shake_args(c, 0.f), false);
    HIPCHK(c, hipGetLastError());
  }
  if (c->h_groups.empty()) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HD_OK;
  }
  HdSettleArgs sa;
  memset(&sa, 0, sizeof(sa));
  sa.n_groups = (unsigned)(c->h_groups.size() / 3);
  sa.groups = c->d_groups;
  sa.d_oh = c->d_oh;
  sa.d_hh = c->d_hh;
  sa.m_o = (float)c->m_o;
  sa.m_h = (float)c->m_h;
  settle_kinds_args(c, sa);
  sa.dt = 0.f;
  sa.posq = c->posq;
  sa.vx = c->vx;
  sa.vy = c->vy;
  sa.vz = c->vz;
  sa.w_cons = c->w_cons;
  sa.lattice = c->d_lattice;
  hd_launch_settle_velocities(c->stream, sa, false);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

#define HD_GRAPH_STEPS 8

// nsteps steps, the last with scalars when ev.  Every path out of here
// leaves the stream out of capture mode and *out either null or a complete
// executable graph: a failure inside the capture ends the capture and
// frees the partial graph instead of returning with the stream still
// capturing (which would pull the caller's next launches into a graph
// that is never instantiated).
static int capture(hd_ctx* c, bool ev, hipGraphExec_t* out, int nsteps = 1) {
  *out = nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(c, hipStreamIsCapturing(c->stream, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return fail(c, HD_ERR_STATE, "context stream is already capturing");
  // captured from an idle context: nothing queued earlier on either stream
  // can be mistaken for part of the graph
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipStreamSynchronize(c->side));
  (void)hipGetLastError();
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < nsteps; ++k) enqueue_step(c, ev && k == nsteps - 1, false);
  hipError_t le = hipGetLastError();
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(c->stream, &g);
  if (le != hipSuccess || ee != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return fail(c, HD_ERR_HIP, "step capture failed: %s",
                hipGetErrorString(le != hipSuccess ? le : ee));
  }
  hipGraphExec_t x = nullptr;
  hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess || !x)
    return fail(c, HD_ERR_HIP, "step graph instantiation failed: %s", hipGetErrorString(ie));
  *out = x;
  return HD_OK;
}

// Status polling inside a long hd_step call (the reference reads the
// status word once per batch, src/forces/neighbor_list.rs:1300-1340).
// After each graph launch the status word is copied into pinned slot
// k & 1 and an event marks the copy; the slot of launch k - 1 is read once
// launch k is queued, so the device always has a launch in hand while the
// host waits.  Returns 1 when a high-water or overflow bit is set: the
// caller then synchronises and lets check_status grow the list (or fail)
// before any further step runs.
static int poll_status(hd_ctx* c, unsigned k, int* flagged) {
  unsigned slot = k & 1u;
  *flagged = 0;
  HIPCHK(c, hipMemcpyAsync(c->h_stat + 4 * slot, c->ctrl, 4 * sizeof(unsigned),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev_stat[slot], c->stream));
  if (k == 0) return HD_OK;
  unsigned prev = slot ^ 1u;
  HIPCHK(c, hipEventSynchronize(c->ev_stat[prev]));
  if (c->h_stat[4 * prev] & (NB_STATUS_HIGHWATER | NB_STATUS_OVERFLOW)) *flagged = 1;
  return HD_OK;
}

// MTK-NPT steps, the symmetric Trotter splitting of mtk_npt.rs:311-470 with
// the host holding the chains and the barostat momentum p_eps: K and W are
// read back after their device reductions (FP64 trees), so each step
// synchronises a few times and runs eagerly, as the reference's does.
static double sinh_over_x(double a) { return fabs(a) < 1e-6 ? 1.0 + a * a / 6.0 : sinh(a) / a; }

static int mtk_ke(hd_ctx* c, double* ke) {
  hd_launch_ke(c->stream, (unsigned)c->n, c->vx, c->vy, c->vz, c->mass, c->red_partials,
               c->scal + 0, nullptr, 0.f, nullptr);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(ke, c->scal + 0, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

static int mtk_virial(hd_ctx* c, double* w) {
  HIPCHK(c, hipMemcpyAsync(w, c->scal + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

static void mtk_cell_chain(hd_ctx* c, double dt) {
  const double* yw = yoshida_w(c->mtk_yoshida);
  for (int a = 0; a < c->mtk_yoshida; ++a)
    for (int r = 0; r < c->mtk_nresp; ++r) {
      double h = yw[a] * dt / (2.0 * c->mtk_nresp);
      double f = chain_sub_step(c->mtk_xc, c->mtk_pc, c->mtk_qc, h,
                                c->mtk_peps * c->mtk_peps / c->mtk_w, 1.0, c->mtk_kt);
      c->mtk_peps *= f;
    }
}

// particle chain: each sub-step's factor, rounded to f32, scales the
// velocities; k follows in FP64 with the rounded factor
static void mtk_particle_chain(hd_ctx* c, double dt, double* k) {
  const double* yw = yoshida_w(c->mtk_yoshida);
  for (int a = 0; a < c->mtk_yoshida; ++a)
    for (int r = 0; r < c->mtk_nresp; ++r) {
      double h = yw[a] * dt / (2.0 * c->mtk_nresp);
      float f = (float)chain_sub_step(c->mtk_xp, c->mtk_pp, c->mtk_qp, h, 2.0 * *k, c->mtk_g,
                                      c->mtk_kt);
      hd_launch_scale_velocities(c->stream, c->vx, c->vy, c->vz, f, (unsigned)c->n);
      *k *= (double)f * (double)f;
    }
}

static void mtk_kick(hd_ctx* c, double dt) {
  double alpha = (1.0 + 3.0 / c->mtk_g) * (c->mtk_peps / c->mtk_w);
  double e = exp(-alpha * dt / 2.0);
  double phi = 0.5 * dt * sinh_over_x(alpha * dt / 4.0) * exp(-alpha * dt / 4.0);
  hd_launch_mtk_kick(c->stream, c->vx, c->vy, c->vz, c->fx, c->fy, c->fz, c->mass, (float)e,
                     (float)phi, (unsigned)c->n);
}

static int mtk_steps(hd_ctx* c, int n_steps) {
  const double dt = (double)(float)c->dt;  // the f32 time step, widened
  float lat[6];
  HIPCHK(c, hipMemcpyAsync(lat, c->d_lattice, sizeof(lat), hipMemcpyDeviceToHost, c->stream));
  // W of the current positions (the forces are recomputed identically)
  enqueue_forces(c, 1, false, false);
  c->forces_valid = true;
  int r = check_status(c);
  if (r) return r;
  for (int s = 0; s < n_steps; ++s) {
    double k, w;
    if ((r = mtk_ke(c, &k)) || (r = mtk_virial(c, &w))) return r;
    double vol = (double)(lat[0] * lat[1] * lat[2]);
    double pres = (2.0 * k + w) / (3.0 * vol);
    mtk_cell_chain(c, dt);
    mtk_particle_chain(c, dt, &k);
    c->mtk_peps += 0.5 * dt * (3.0 * vol * (pres - c->mtk_p) + 6.0 / c->mtk_g * k);
    mtk_kick(c, dt);
    double beta = c->mtk_peps / c->mtk_w;
    double eb = exp(beta * dt);
    double phx = dt * sinh_over_x(beta * dt / 2.0) * exp(beta * dt / 2.0);
    float mu = (float)eb;
    for (int q = 0; q < 6; ++q) lat[q] = lat[q] * mu;
    HIPCHK(c, hipMemcpyAsync(c->d_lattice, lat, sizeof(lat), hipMemcpyHostToDevice, c->stream));
    hd_launch_mtk_drift(c->stream, c->posq, c->ix, c->iy, c->iz, c->vx, c->vy, c->vz,
                        c->d_lattice, (float)eb, (float)phx, (unsigned)c->n);
    c->mtk_eps += beta * dt;
    enqueue_forces(c, 1, false, false);
    if ((r = check_status(c))) return r;
    if ((r = mtk_ke(c, &k)) || (r = mtk_virial(c, &w))) return r;
    vol = (double)(lat[0] * lat[1] * lat[2]);
    pres = (2.0 * k + w) / (3.0 * vol);
    mtk_kick(c, dt);
    double kk;
    if ((r = mtk_ke(c, &kk))) return r;
    c->mtk_peps += 0.5 * dt * (3.0 * vol * (pres - c->mtk_p) + 6.0 / c->mtk_g * kk);
    mtk_particle_chain(c, dt, &kk);
    c->mtk_last_p = pres;
    c->mtk_last_v = vol;
    mtk_cell_chain(c, dt);
  }
  for (int q = 0; q < 6; ++q) c->lattice[q] = (double)lat[q];
  // the kinetic energy of the final velocities for the caller's scalars
  double k;
  return mtk_ke(c, &k);
}

int hd_get_mtk(hd_ctx* c, double* out) {
  if (!c || !out) return HD_ERR_ARG;
  if (!c->have_mtk || !c->ready) return fail(c, HD_ERR_STATE, "no MTK-NPT integrator set up");
  // the conserved quantity less K + U (mtk_npt.rs:246-268)
  double h = c->mtk_p * c->mtk_last_v + 0.5 * c->mtk_peps * c->mtk_peps / c->mtk_w;
  for (int j = 0; j < c->mtk_chain; ++j) {
    h += c->mtk_pp[j] * c->mtk_pp[j] / (2.0 * c->mtk_qp[j]);
    h += c->mtk_pc[j] * c->mtk_pc[j] / (2.0 * c->mtk_qc[j]);
    h += (j == 0 ? c->mtk_g : 1.0) * c->mtk_kt * c->mtk_xp[j];
    h += c->mtk_kt * c->mtk_xc[j];
  }
  out[0] = c->mtk_last_p;
  out[1] = c->mtk_last_v;
  out[2] = h;
  out[3] = c->mtk_peps;
  out[4] = c->mtk_eps;
  return HD_OK;
}

int hd_step(hd_ctx* c, int n_steps, int scalars_last) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "hd_step before hd_setup");
  if (n_steps < 0) return fail(c, HD_ERR_ARG, "n_steps < 0");
  if (!(c->dt > 0)) return fail(c, HD_ERR_STATE, "hd_step before hd_set_integrator");
  HIPCHK(c, hipSetDevice(c->device));
  bool timed = c->timing && !c->use_graph;
  if (!c->forces_valid) {
    // first forces of new positions: their rebuild's status is read before
    // any step runs on the list
    enqueue_forces(c, false, false, false);
    c->forces_valid = true;
    int r = check_status(c);
    if (r) {
      c->forces_valid = false;
      return r;
    }
  }
  if (c->have_mtk) return mtk_steps(c, n_steps);
  // constraint iteration counters cover this call's steps
  if (c->d_sh_stats)
    HIPCHK(c, hipMemsetAsync(c->d_sh_stats, 0, 6 * sizeof(unsigned), c->stream));
  unsigned polls = 0;
  for (int s = 0; s < n_steps; ++s) {
    bool ev = scalars_last && s == n_steps - 1;
    // whole groups of steps without scalars as one graph launch
    int plain_left = n_steps - s - (scalars_last ? 1 : 0);
    bool poll = false;
    if (c->use_graph && plain_left >= HD_GRAPH_STEPS) {
      if (!c->graph_fk) {
        int r = capture(c, false, &c->graph_fk, HD_GRAPH_STEPS);
        if (r) return r;
      }
      HIPCHK(c, hipGraphLaunch(c->graph_fk, c->stream));
      s += HD_GRAPH_STEPS - 1;
      poll = true;
    } else if (c->use_graph) {
      hipGraphExec_t* gx = ev ? &c->graph_fev : &c->graph_f;
      if (!*gx) {
        int r = capture(c, ev, gx);
        if (r) return r;
      }
      HIPCHK(c, hipGraphLaunch(*gx, c->stream));
    } else {
      if (timed && s == 0) c->n_stage_used = 0;
      enqueue_step(c, ev, timed);
      HIPCHK(c, hipGetLastError());
      poll = (s + 1) % HD_GRAPH_STEPS == 0;
    }
    if (poll && s + 1 < n_steps) {
      int flagged = 0;
      int r = poll_status(c, polls++, &flagged);
      if (r) return r;
      if (flagged) {
        // grow (or report the overflow) before the next step
        if ((r = check_status(c))) return r;
        polls = 0;
      }
    }
  }
  int r = check_status(c);
  if (r) return r;
  if (timed) {
    // mean device time per stage occurrence over this call's steps
    c->stage_names.clear();
    c->stage_ms.clear();
    c->stage_count.clear();
    for (size_t k = 0; k < c->n_stage_used; ++k) {
      const Stage& st = c->stages[k];
      float ms = 0;
      (void)hipEventElapsedTime(&ms, st.a, st.b);
      size_t j = 0;
      while (j < c->stage_names.size() && c->stage_names[j] != st.name) ++j;
      if (j == c->stage_names.size()) {
        c->stage_names.push_back(st.name);
        c->stage_ms.push_back(0.0);
        c->stage_count.push_back(0);
      }
      c->stage_ms[j] += ms;
      c->stage_count[j] += 1;
    }
    for (size_t j = 0; j < c->stage_ms.size(); ++j) c->stage_ms[j] /= (double)c->stage_count[j];
  }
  return HD_OK;
}

int hd_negate_velocities(hd_ctx* c) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "hd_negate_velocities before hd_setup");
  hd_launch_negate_velocities(c->stream, c->vx, c->vy, c->vz, c->lossless ? c->vxlo : nullptr,
                              c->lossless ? c->vylo : nullptr, c->lossless ? c->vzlo : nullptr,
                              (unsigned)c->n);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

int hd_set_fused_settle(hd_ctx* c, int enable) {
  if (!c) return HD_ERR_ARG;
  if (c->ready) {
    HIPCHK(c, hipSetDevice(c->device));
    drop_graphs(c);
  }
  c->fused_settle = enable != 0;
  return HD_OK;
}

int hd_set_small_neighbor(hd_ctx* c, int enable) {
  if (!c) return HD_ERR_ARG;
  if (c->ready) {
    HIPCHK(c, hipSetDevice(c->device));
    drop_graphs(c);
  }
  c->small_neighbor = enable != 0;
  return HD_OK;
}

int hd_use_graph(hd_ctx* c, int enable) {
  if (!c) return HD_ERR_ARG;
  c->use_graph = enable != 0;
  return HD_OK;
}

int hd_set_timing(hd_ctx* c, int enable) {
  if (!c) return HD_ERR_ARG;
  c->timing = enable != 0;
  return HD_OK;
}

int hd_stage_times(hd_ctx* c, char* names, int name_len, double* ms, int cap) {
  if (!c || !names || !ms || name_len <= 0) return HD_ERR_ARG;
  int k = 0;
  for (size_t i = 0; i < c->stage_ms.size() && k < cap; ++i, ++k) {
    snprintf(names + (size_t)k * name_len, name_len, "%s", c->stage_names[i].c_str());
    ms[k] = c->stage_ms[i];
  }
  return k;
}

int hd_sync(hd_ctx* c) {
  if (!c) return HD_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

int hd_get_positions(hd_ctx* c, float* posq, int32_t* ix, int32_t* iy, int32_t* iz) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "not set up");
  int r;
  if ((r = download(c, (float4*)posq, c->posq, c->n))) return r;
  if ((r = download(c, ix, c->ix, c->n))) return r;
  if ((r = download(c, iy, c->iy, c->n))) return r;
  if ((r = download(c, iz, c->iz, c->n))) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

int hd_get_velocities(hd_ctx* c, float* vx, float* vy, float* vz) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "not set up");
  int r;
  if ((r = download(c, vx, c->vx, c->n))) return r;
  if ((r = download(c, vy, c->vy, c->n))) return r;
  if ((r = download(c, vz, c->vz, c->n))) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

int hd_get_forces(hd_ctx* c, float* fx, float* fy, float* fz) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "not set up");
  int r;
  if ((r = download(c, fx, c->fx, c->n))) return r;
  if ((r = download(c, fy, c->fy, c->n))) return r;
  if ((r = download(c, fz, c->fz, c->n))) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

int hd_get_per_atom(hd_ctx* c, float* pe, float* vir) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "not set up");
  int r;
  if ((r = download(c, pe, c->pe, c->n))) return r;
  if ((r = download(c, vir, c->vir, c->n))) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HD_OK;
}

int hd_get_lattice(hd_ctx* c, double* L) {
  if (!c || !L) return HD_ERR_ARG;
  if (!c->ready) {
    memcpy(L, c->lattice, sizeof(double) * 6);
    return HD_OK;
  }
  float lf[6];
  HIPCHK(c, hipMemcpyAsync(lf, c->d_lattice, sizeof(lf), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < 6; ++i) L[i] = lf[i];
  return HD_OK;
}

int hd_get_scalars(hd_ctx* c, double* out) {
  if (!c || !c->ready || !out) return fail(c, HD_ERR_STATE, "not set up");
  double s[8];
  HIPCHK(c, hipMemcpyAsync(s, c->scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  double rc[2] = {0, 0};
  if (c->have_spme) HIPCHK(c, hipMemcpyAsync(rc, c->recip, sizeof(rc), hipMemcpyDeviceToHost, c->stream));
  unsigned ctrl[4];
  HIPCHK(c, hipMemcpyAsync(ctrl, c->ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> st;
  if (c->have_nhc) {
    st.resize(3 * c->nhc_chain);
    HIPCHK(c, hipMemcpyAsync(st.data(), c->nhc_state, st.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[0] = s[0];
  out[1] = s[1];
  out[2] = s[2];
  out[3] = rc[0];
  out[4] = rc[1];
  double chain = 0.0;
  if (c->have_nhc) {
    // conserved-quantity chain term (nose_hoover_chain.rs:365-381)
    int m = c->nhc_chain;
    long long dof = std::max(3LL * (long long)c->n - (long long)c->nhc_nc - 3LL, 0LL);
    for (int j = 0; j < m; ++j) chain += st[m + j] * st[m + j] / (2.0 * st[2 * m + j]);
    chain += (double)dof * c->nhc_kt * st[0];
    for (int j = 1; j < m; ++j) chain += c->nhc_kt * st[j];
  }
  if (c->rescale) {
    // CSVR / Berendsen: the conserved quantity is K + U minus the energy
    // the rescales injected (csvr.rs:266-280, berendsen.rs:191-200)
    double inj = 0.0;
    HIPCHK(c, hipMemcpy(&inj, c->rs_injection, sizeof(inj), hipMemcpyDeviceToHost));
    chain = -inj;
  }
  out[5] = chain;
  out[6] = ctrl[1];
  out[7] = ctrl[0];
  return HD_OK;
}

int hd_get_pairs(hd_ctx* c, uint32_t* out, uint64_t cap, uint64_t* n_pairs) {
  if (!c || !c->ready || !n_pairs) return fail(c, HD_ERR_STATE, "not set up");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<unsigned> ents((size_t)c->nblocks * c->cap * 64), cnt(c->nblocks), sid(c->npad);
  HIPCHK(c, hipMemcpy(ents.data(), c->entry_atoms, ents.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(cnt.data(), c->entry_count, cnt.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(sid.data(), c->sorted_id, sid.size() * 4, hipMemcpyDeviceToHost));
  std::vector<uint64_t> pairs;
  unsigned n = (unsigned)c->n;
  if (c->atom_lists) {
    // row k of block b: the k-th neighbour of each of the block's atoms
    for (unsigned b = 0; b < c->nblocks; ++b)
      for (unsigned k = 0; k < cnt[b]; ++k)
        for (unsigned l = 0; l < 64; ++l) {
          unsigned js = ents[((size_t)b * c->cap + k) * 64 + l], is = b * 64 + l;
          if (js >= n || is >= n) continue;
          uint32_t iid = sid[is], jid = sid[js];
          uint32_t lo = std::min(iid, jid), hi = std::max(iid, jid);
          pairs.push_back(((uint64_t)lo << 32) | hi);
        }
  }
  for (unsigned b = 0; b < c->nblocks && !c->atom_lists; ++b) {
    for (unsigned e = 0; e < cnt[b]; ++e)
      for (unsigned l = 0; l < 64; ++l) {
        unsigned js = ents[((size_t)b * c->cap + e) * 64 + l];
        if (js >= n) continue;
        unsigned jid = sid[js];
        for (unsigned m = 0; m < 64; ++m) {
          unsigned is = b * 64 + m;
          if (is >= n || is == js) continue;
          unsigned iid = sid[is];
          uint32_t lo = std::min(iid, jid), hi = std::max(iid, jid);
          pairs.push_back(((uint64_t)lo << 32) | hi);
        }
      }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  *n_pairs = pairs.size();
  if (out)
    for (uint64_t k = 0; k < std::min<uint64_t>(cap, pairs.size()); ++k) {
      out[2 * k] = (uint32_t)(pairs[k] >> 32);
      out[2 * k + 1] = (uint32_t)(pairs[k] & 0xFFFFFFFFu);
    }
  return HD_OK;
}

int hd_count_pairs(hd_ctx* c, uint64_t* out) {
  if (!c || !c->ready || !out) return fail(c, HD_ERR_STATE, "not set up");
  unsigned long long* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, sizeof(unsigned long long)));
  hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long), c->stream);
  HdPairArgs pa;
  memset(&pa, 0, sizeof(pa));
  pa.n = (unsigned)c->n;
  pa.npad = c->npad;
  pa.cap = c->cap;
  pa.rmax = c->rmax;
  pa.lattice = c->d_lattice;
  pa.tile_posq = c->tile_posq;
  pa.tile_key = c->tile_key;
  pa.excl_span = c->excl_span;
  pa.entry_atoms = c->entry_atoms;
  pa.entry_count = c->entry_count;
  pa.block_order = c->block_order;
  pa.atom_lists = c->atom_lists;
  pa.ortho = box_ortho(c) ? 1 : 0;
  if (e == hipSuccess) {
    hd_launch_tile_gather(c->stream, c->posq, c->sorted_id, c->d_lattice, c->tile_posq, c->fpj,
                          c->fpj_ew, (unsigned)c->n, c->npad);
    hd_launch_count_pairs(c->stream, pa, d);
    e = hipGetLastError();
  }
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, HD_ERR_HIP, "count pairs: %s", hipGetErrorString(e));
  *out = h;
  return HD_OK;
}

int hd_get_barostat(hd_ctx* c, double* out) {
  if (!c || !c->ready || !out) return fail(c, HD_ERR_STATE, "not set up");
  float mu = 1.f;
  double d[4] = {0, 0, 0, 0};
  unsigned long long ctr = 0;
  HIPCHK(c, hipMemcpyAsync(&mu, c->baro_mu, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(d, c->baro_diag, sizeof(d), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&ctr, c->baro_counter, sizeof(ctr), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[0] = mu;
  out[1] = d[0];
  out[2] = d[1];
  out[3] = d[2];
  out[4] = (double)ctr;
  return HD_OK;
}

int hd_get_neighbor_stats(hd_ctx* c, uint64_t* out) {
  if (!c || !c->ready || !out) return fail(c, HD_ERR_STATE, "not set up");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<unsigned> cnt(c->nblocks);
  HIPCHK(c, hipMemcpy(cnt.data(), c->entry_count, cnt.size() * 4, hipMemcpyDeviceToHost));
  unsigned ctrl[4];
  HIPCHK(c, hipMemcpy(ctrl, c->ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
  uint64_t tot = 0, mx = 0;
  for (unsigned v : cnt) {
    tot += v;
    mx = std::max<uint64_t>(mx, v);
  }
  out[0] = c->nblocks;
  out[1] = c->cap;
  out[2] = mx;
  out[3] = tot;
  out[4] = ctrl[1];
  out[5] = ctrl[0];
  return HD_OK;
}

int hd_get_spme_grids(hd_ctx* c, float* rho, float* v) {
  if (!c || !c->ready) return fail(c, HD_ERR_STATE, "not set up");
  if (!c->have_spme) return fail(c, HD_ERR_STATE, "no SPME configured");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  size_t M = (size_t)c->grid[0] * c->grid[1] * c->grid[2];
  if (rho) {
    std::vector<long long> fx(M);
    HIPCHK(c, hipMemcpy(fx.data(), c->rho_fixed, M * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < M; ++i) rho[i] = (float)fx[i] * (1.0f / 4294967296.0f);
  }
  if (v) HIPCHK(c, hipMemcpy(v, c->vgrid, M * 4, hipMemcpyDeviceToHost));
  return HD_OK;
}

int hd_selftest_dpp(hd_ctx* c, int32_t* out) {
  if (!c || !out) return HD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  int* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, 64 * sizeof(int)));
  hd_launch_selftest_dpp(c->stream, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, 64 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, HD_ERR_HIP, "selftest: %s", hipGetErrorString(e));
  return HD_OK;
}

// ---------------------------------------------------------------------
// Batched replicas: one hipGraph per step for many independent contexts.
// The graph forks from the batch's stream into every replica's stream
// (event record / wait), captures that replica's step there, and joins
// back, so a replay is one launch and the replicas' branches are
// independent nodes the runtime may run side by side.  Every replica runs
// exactly the kernels of its own hd_step, on its own buffers, so its
// trajectory is byte-identical to a solo run.
// ---------------------------------------------------------------------
struct hd_batch {
  std::vector<hd_ctx*> ctx;
  std::vector<unsigned> gen;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int device = 0;
  int lanes = 0;  // parallel chains in the step graph (0 = one per replica)
  char err[512] = {0};
};

static int bfail(hd_batch* b, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b->err, sizeof(b->err), fmt, ap);
  va_end(ap);
  return code;
}

#define BCHK(b, expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) return bfail((b), HD_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

static void batch_drop(hd_batch* b) {
  (void)hipStreamSynchronize(b->stream);
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

int hd_batch_create(hd_ctx* const* ctxs, int n, hd_batch** out) {
  if (!out || n < 1 || !ctxs) return HD_ERR_ARG;
  *out = nullptr;
  hd_batch* b = new hd_batch();
  for (int i = 0; i < n; ++i) {
    hd_ctx* c = ctxs[i];
    if (!c || !c->ready) {
      delete b;
      return HD_ERR_STATE;
    }
    if (i > 0 && c->device != b->device) {
      delete b;
      return HD_ERR_ARG;
    }
    b->device = c->device;
    b->ctx.push_back(c);
  }
  if (hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
    delete b;
    return HD_ERR_HIP;
  }
  b->join.resize(n, nullptr);
  for (int i = 0; i < n; ++i)
    if (hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess) {
      hd_batch_destroy(b);
      return HD_ERR_HIP;
    }
  b->gen.assign(n, ~0u);
  *out = b;
  return HD_OK;
}

void hd_batch_destroy(hd_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop(b);
  for (auto e : b->join)
    if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* hd_batch_last_error(const hd_batch* b) { return b ? b->err : "null batch"; }

int hd_batch_set_lanes(hd_batch* b, int lanes) {
  if (!b || lanes < 0) return HD_ERR_ARG;
  BCHK(b, hipStreamSynchronize(b->stream));
  batch_drop(b);
  b->lanes = lanes;
  return HD_OK;
}

// Relative cost of one replica step in a batch graph: launches dominate
// small replicas (about 4 us each), atoms and SPME grid points the large
// ones (per-atom and per-point rates of the large single-system step).
static double batch_step_cost(const hd_ctx* c) {
  double launches = c->have_spme ? 11.0 : 5.0;
  double us = 4.0 * launches + 0.002 * (double)c->n;
  if (c->have_spme) us += 0.0005 * (double)c->grid[0] * c->grid[1] * c->grid[2];
  return us;
}

static int batch_capture(hd_batch* b, bool ev, hipGraphExec_t* out) {
  *out = nullptr;
  for (auto c : b->ctx) {
    BCHK(b, hipStreamSynchronize(c->stream));
    BCHK(b, hipStreamSynchronize(c->side));
  }
  (void)hipGetLastError();
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  // errors inside the capture are collected, and the capture is always
  // ended, so no stream is left capturing after a failure
  hipError_t e = hipEventRecord(b->fork, b->stream);
  // replicas form `lanes` chains, packed by estimated step cost: longest
  // first, each onto the chain with the least cost so far, so the chains
  // end together (the graph takes as long as its longest chain)
  size_t n = b->ctx.size();
  size_t lanes = (b->lanes <= 0 || (size_t)b->lanes >= n) ? n : (size_t)b->lanes;
  std::vector<size_t> by_cost(n);
  std::vector<double> cost(n);
  for (size_t i = 0; i < n; ++i) {
    by_cost[i] = i;
    cost[i] = batch_step_cost(b->ctx[i]);
  }
  std::stable_sort(by_cost.begin(), by_cost.end(),
                   [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  std::vector<double> load(lanes, 0.0);
  std::vector<long> tail(lanes, -1);  // last replica of each chain
  for (size_t k = 0; k < n && e == hipSuccess; ++k) {
    size_t i = by_cost[k];
    size_t l = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
    load[l] += cost[i];
    hd_ctx* c = b->ctx[i];
    e = hipStreamWaitEvent(c->stream, tail[l] < 0 ? b->fork : b->join[(size_t)tail[l]], 0);
    if (e != hipSuccess) break;
    tail[l] = (long)i;
    c->spme_fork = false;
    enqueue_step(c, ev, false);
    c->spme_fork = true;
    e = hipEventRecord(b->join[i], c->stream);
  }
  for (size_t l = 0; l < lanes && e == hipSuccess; ++l)
    if (tail[l] >= 0) e = hipStreamWaitEvent(b->stream, b->join[(size_t)tail[l]], 0);
  if (e == hipSuccess) e = hipGetLastError();
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || ee != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(b, HD_ERR_HIP, "batch capture failed: %s",
                 hipGetErrorString(e != hipSuccess ? e : ee));
  }
  hipGraphExec_t x = nullptr;
  hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess || !x)
    return bfail(b, HD_ERR_HIP, "batch graph instantiation failed: %s", hipGetErrorString(ie));
  *out = x;
  return HD_OK;
}

int hd_batch_step(hd_batch* b, int n_steps, int scalars_last) {
  if (!b || n_steps < 0) return HD_ERR_ARG;
  BCHK(b, hipSetDevice(b->device));
  // forces of the current positions must exist before the first kick
  bool stale = false;
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    hd_ctx* c = b->ctx[i];
    if (!c->forces_valid) {
      enqueue_forces(c, false, false, false);
      c->forces_valid = true;
      BCHK(b, hipStreamSynchronize(c->stream));
    }
    if (c->d_sh_stats) BCHK(b, hipMemsetAsync(c->d_sh_stats, 0, 6 * sizeof(unsigned), c->stream));
    if (b->gen[i] != c->graph_gen) stale = true;
  }
  if (stale) {
    batch_drop(b);
    for (size_t i = 0; i < b->ctx.size(); ++i) b->gen[i] = b->ctx[i]->graph_gen;
  }
  for (auto c : b->ctx) BCHK(b, hipStreamSynchronize(c->stream));
  for (int s = 0; s < n_steps; ++s) {
    bool ev = scalars_last && s == n_steps - 1;
    hipGraphExec_t* gx = &b->graph[ev ? 1 : 0];
    if (!*gx) {
      int r = batch_capture(b, ev, gx);
      if (r) return r;
    }
    BCHK(b, hipGraphLaunch(*gx, b->stream));
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    int r = check_status(b->ctx[i]);
    if (r) return bfail(b, r, "replica %zu: %s", i, b->ctx[i]->err.c_str());
  }
  return HD_OK;
}

int hd_selftest_mfma_k16(hd_ctx* c, int bf16, float* out) {
  if (!c || !out) return HD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  float* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, 1024 * sizeof(float)));
  hd_launch_selftest_mfma_k16(c->stream, d, bf16 != 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, 1024 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, HD_ERR_HIP, "selftest: %s", hipGetErrorString(e));
  return HD_OK;
}

int hd_selftest_mfma(hd_ctx* c, float* out) {
  if (!c || !out) return HD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  float* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, 1024 * sizeof(float)));
  hd_launch_selftest_mfma(c->stream, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, 1024 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, HD_ERR_HIP, "selftest: %s", hipGetErrorString(e));
  return HD_OK;
}

}  // extern "C"
